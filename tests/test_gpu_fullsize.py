"""The benchmark box at full size (512^3 Zel'dovich particles, 512^3 mesh,
TSC), checked through properties that do not need an O(N) oracle run:
the force against the FP64 Ewald sum on sampled particles, Newton's third
law for the pair forces (every pair is evaluated from both ends with exactly
negated offsets), every particle kicked, the candidate-pair count equal to
the closed form from the cell counts, a repeated kick's forces
bit-identical, and the tree walk finishing with its multipole error."""
import numpy as np
import pytest

import oracle as O
import tigerpm_amd as T

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def test_zeldovich512_tsc_full_size():
    with T.Context(0) as ctx:
        run_full_size(ctx)


def run_full_size(gpu_ctx):
    n_side = 512
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
    gpu_ctx.ic_zeldovich(n_side, 3, 2.0, -2.0)
    n = gpu_ctx.size()
    print("512^3 pair forces only", flush=True)
    st = gpu_ctx.kick_treepm(0, 1.0, 1.0, True)  # chain mesh and PM for the state below
    cb = gpu_ctx.chainmesh_cells()
    assert st.candidate_pairs == O.candidate_pairs(p.chain_dim, cb)
    gpu_ctx.kick(flags=T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
    d = gpu_ctx.download()
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1).astype(np.float64)
    assert np.all(np.abs(G).sum(1) > 0)
    # Newton's third law: the sum cancels to FP32 accumulation rounding
    total = np.abs(G.sum(0)).max()
    scale = np.abs(G).sum(0).max()
    assert total < 1e-6 * scale
    del G, d
    print("512^3 full kick vs Ewald", flush=True)
    gpu_ctx.kick_treepm(0, 1.0, 1.0, True)
    d = gpu_ctx.download()
    s = np.sort(np.random.default_rng(11).choice(n, 1024, replace=False))
    ph, gx, gy, gz = gpu_ctx.ewald(s)
    G = np.stack([d["gx"][s], d["gy"][s], d["gz"][s]], 1)
    assert rel_rms(G, np.stack([gx, gy, gz], 1) * p.GM) < 1e-3
    # the same step again (positions unchanged, there is no drift): the same
    # order, forces and potential, bit for bit
    gpu_ctx.kick_treepm(0, 1.0, 1.0, True)
    e = gpu_ctx.download()
    for k in ("x", "y", "z", "gx", "gy", "gz", "pot"):
        assert np.array_equal(d[k], e[k]), k
    # the reference's tree walk on the same particles: it finishes at this
    # size (no stack or list exhaustion) and carries its multipole error
    print("512^3 tree kick", flush=True)
    gpu_ctx.chainmesh_create()
    gpu_ctx.gravity_long_compute()
    st = gpu_ctx.kick(flags=T.TPM_KICK_TREE | T.TPM_SKIP_KICK)
    assert st.pc_interactions > 0 and st.executed_pairs > 0
    t = gpu_ctx.download()
    Gt = np.stack([t["gx"][s], t["gy"][s], t["gz"][s]], 1)
    err = rel_rms(Gt, np.stack([gx, gy, gz], 1) * p.GM)
    assert 1e-3 < err < 4e-3
