"""tigerpm_amd -- MI355X-native TreePM gravity kick.

Python view of the C ABI in include/tigerpm_hip.h (libtigerpm_hip.so, built
in-tree by tigerpm_amd.build).  There is no CPU fallback: importing works
without the library, but creating a Context raises if the HIP library is
missing or no GPU is visible.
"""
from ._lib import (  # noqa: F401
    TPM_CIC, TPM_TSC, TPM_CLOUD4, TPM_SKIP_PP, TPM_SKIP_PM, TPM_SKIP_KICK, TPM_KICK_PME, TPM_KICK_TREE,
    TREE_NODE, TREE_BUCKET, TPM_ORDER_DEVICE, TPM_ORDER_REFERENCE,
    Context, TpmError, TpmParams, TpmStats, TpmMoments, TpmTuning, lib, lib_path, exported_symbols, DECLARED_SYMBOLS,
    TPM_BATCH_DRIFT, Batch, TpmBatchLimits, TpmBoxSpec, TpmBatchTuning, TpmBatchStats, box_spec, batch_arena_bytes,
)
from .workloads import baseline_params, reference_params  # noqa: F401
