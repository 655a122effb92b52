"""C ABI and host-side checks that need no GPU."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import tigerpm_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    declared = set(T.DECLARED_SYMBOLS)
    assert len(declared) >= 20
    assert declared <= set(T.exported_symbols())
    T.lib()  # loads, every signature resolves


def test_context_without_gpu_fails_cleanly():
    import torch  # noqa: F401  (only to know whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(T.TpmError):
        T.Context(0)


def test_workload_parameters():
    p = T.baseline_params(512, 512, 3)
    assert p.chain_dim == 91 and abs(p.rs * 512 - 1.25) < 1e-12 and abs(p.rcut / p.rs - 4.5) < 1e-12
    assert abs(p.GM * 512 ** 3 - 1) < 1e-12 and abs(p.hsoft - 2 / (25 * 512)) < 1e-15
    assert T.baseline_params(64, 64, 2).chain_dim == 11
    assert T.baseline_params(256, 256, 3).chain_dim == 45
    r = T.reference_params(130 - 130 % 15)
    assert r.mesh_dim * 3 == 120 and r.chain_dim == 8 and r.order == 4
    with pytest.raises(ValueError):
        T.reference_params(130)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_header_and_binding_agree():
    txt = open(os.path.join(ROOT, "include", "tigerpm_hip.h")).read()
    for name in T.DECLARED_SYMBOLS:
        assert name in open(os.path.join(ROOT, "tigerpm_amd", "_lib.py")).read(), name
    assert "extern \"C\"" in txt


def test_reference_named_cxx_layer_exports():
    lib = os.path.join(ROOT, "tigerpm_amd", "libtigerpm_ref.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    for sig in ["kick_treepm_begin(int, double, double, bool)", "kick_treepm_end()", "gravity_long_compute(int)",
                "chainmesh_create()", "chainmesh_get(std::array<int, 3ul> const&)", "kick_pm()",
                "gravity_short_ewald_compare(int)", "gravity_long_force_at(std::array<double, 3ul> const&)",
                "particles_random_init()", "initialize()", "kick_pme_begin(int, double, double, bool)"]:
        assert sig in out, sig
    assert "particles_X" in out


def test_driver_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    drv = os.path.join(ROOT, "tigerpm_amd", "bin", "tigerpm")
    r = subprocess.run([drv, "--test=treepm", "--parts_dim=60"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "tpm_ctx_create" in r.stderr


def test_struct_layouts_match_header(tmp_path):
    """sizeof / offsetof of every ABI struct, from the C header compiled by
    gcc, equal the ctypes mirrors."""
    structs = {"tpm_params": T.TpmParams, "tpm_stats": T.TpmStats, "tpm_tuning": T.TpmTuning,
               "tpm_moments": T.TpmMoments,
               "tpm_batch_limits": T.TpmBatchLimits, "tpm_box_spec": T.TpmBoxSpec,
               "tpm_batch_tuning": T.TpmBatchTuning, "tpm_batch_stats": T.TpmBatchStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tigerpm_hip.h"', "int main(void) {"]
    for c, py in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (c, c))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (c, f, c, f))
    lines.append("return 0; }")
    src = tmp_path / "l.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "l"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(ln.split() for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                    check=True).stdout.splitlines())
    import ctypes
    for c, py in structs.items():
        assert int(got[c]) == ctypes.sizeof(py), c
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (c, f)]) == getattr(py, f).offset, (c, f)


def test_tree_record_layouts_match_header(tmp_path):
    """The numpy records of the tree download equal tpm_tree_node /
    tpm_tree_bucket as gcc lays them out, field by field."""
    recs = {"tpm_tree_node": T.TREE_NODE, "tpm_tree_bucket": T.TREE_BUCKET}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tigerpm_hip.h"', "int main(void) {"]
    for c, dt in recs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (c, c))
        for f in dt.names:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (c, f, c, f))
    lines.append("return 0; }")
    src = tmp_path / "t.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "t"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(ln.split() for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                    check=True).stdout.splitlines())
    for c, dt in recs.items():
        assert int(got[c]) == dt.itemsize, c
        for f in dt.names:
            assert int(got["%s.%s" % (c, f)]) == dt.fields[f][1], (c, f)
    assert T.TREE_NODE.itemsize == O.TREE_NODE.itemsize == 48


def test_batch_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    assert T.batch_arena_bytes(1 << 20, 1 << 20, 8, 3, 1) > 100 * (1 << 20)
    with pytest.raises(T.TpmError):
        T.Batch(1 << 16, 1 << 16, 4, 3, 0)


def test_bench_useful_pair_count_matches_uniform_density():
    """bench.py's pairs-inside-rcut count, on uniform particles sorted into
    the chain mesh, equals the closed form (4/3) pi rcut^3 N per particle
    within its sampling error."""
    sys.path.insert(0, ROOT)
    import bench
    p = T.baseline_params(32, 32, 2)
    n = 32 ** 3
    rng = np.random.default_rng(3)
    x, y, z = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    key = O.chain_keys(x, y, z, p.chain_dim)
    o = np.argsort(key, kind="stable")
    x, y, z, key = x[o], y[o], z[o], key[o]
    cb = np.searchsorted(key // 64, np.arange(p.chain_dim ** 3 + 1)).astype(np.int64)

    class Resident:
        params = p

        def download(self):
            return {"x": x, "y": y, "z": z}

        def size(self):
            return n

        def chainmesh_cells(self):
            return cb

    r = bench.useful_pairs(Resident(), p, 1024, 1)
    expect = 4.0 / 3.0 * np.pi * p.rcut ** 3 * n
    # Poisson error of the mean over 1024 sinks is ~1 %
    assert abs(r["per_particle"] / expect - 1) < 0.03
