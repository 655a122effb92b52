"""GPU parity: every stage of the HIP kick against the CPU oracle.

Integer stages (chain mesh, deposit) must be bit-exact; FP32 stages are
compared as relative rms over particles with the tolerance written in each
test.  All calls go through the C ABI (tigerpm_amd.Context).
"""
import numpy as np
import pytest

import oracle as O
import tigerpm_amd as T

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def uniform(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]


def clustered(n, seed, nclump=20):
    """Uniform background plus tight Gaussian clumps (uneven cell occupancy)."""
    rng = np.random.default_rng(seed)
    m = n // 2
    pos = [rng.random(n - m) for _ in range(3)]
    centres = rng.random((nclump, 3))
    which = rng.integers(0, nclump, m)
    for d in range(3):
        c = (centres[which, d] + 0.01 * rng.standard_normal(m)) % 1.0
        pos[d] = np.concatenate([pos[d], c])
    return [np.minimum(np.floor(p * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint32) for p in pos]


def oparams(p):
    return O.make_params(p.mesh_dim, p.chain_dim, p.order, p.rs, p.rcut, p.hsoft, p.GM, p.eta, p.n_total, p.do_phi,
                         p.interlace)


@pytest.mark.parametrize("dist,n_side,C", [("uniform", 24, 5), ("clustered", 32, 7), ("uniform", 16, 3)])
def test_chainmesh_bit_exact(gpu_ctx, dist, n_side, C):
    n = n_side ** 3
    x, y, z = uniform(n, 11) if dist == "uniform" else clustered(n, 12)
    p = T.baseline_params(n_side, n_side, 3)
    p.chain_dim = C
    p.rcut = min(p.rcut, 0.99 / C)
    p = gpu_ctx.set_params(p)
    gpu_ctx.upload(x, y, z)
    gpu_ctx.chainmesh_create()
    perm = gpu_ctx.chainmesh_perm()
    cb = gpu_ctx.chainmesh_cells()
    operm, ocb = O.chainmesh_sort(x, y, z, C)
    assert np.array_equal(perm, operm)
    assert np.array_equal(cb, ocb)
    d = gpu_ctx.download()
    assert np.array_equal(d["x"], x[operm]) and np.array_equal(d["z"], z[operm])
    # the reference's recursive partition gives the same cell ranges
    _, rb, re_ = O.chainmesh_partition_ref(x, y, z, C)
    assert np.array_equal(rb, ocb[:-1]) and np.array_equal(re_, ocb[1:])
    b, e = gpu_ctx.chainmesh_get(1, 2, 0)
    cid = (0 * C + 2) * C + 1
    assert (b, e) == (ocb[cid], ocb[cid + 1])


@pytest.mark.parametrize("order", [2, 3, 4])
def test_deposit_bit_exact(gpu_ctx, order):
    n_side, M = 24, 20
    x, y, z = clustered(n_side ** 3, 21 + order)
    p = T.baseline_params(n_side, M, order)
    p = gpu_ctx.set_params(p)
    gpu_ctx.upload(x, y, z)
    gpu_ctx.gravity_long_compute()
    g = gpu_ctx.get_density()
    og = O.deposit(x, y, z, M, order)
    assert np.array_equal(g, og)
    assert int(g.sum(dtype=np.uint64)) == len(x) << 32


@pytest.mark.parametrize("M,order", [(32, 2), (24, 3), (30, 4)])
def test_green_and_phi(gpu_ctx, M, order):
    n_side = 24
    x, y, z = uniform(n_side ** 3, 31)
    p = gpu_ctx.set_params(T.baseline_params(n_side, M, order))
    gpu_ctx.upload(x, y, z)
    gpu_ctx.gravity_long_compute()
    tab = gpu_ctx.get_green()
    otab = O.green_table(M, p.rs, order)
    # FP64 build on both sides, one float rounding: allow 1 ulp
    ulp = np.abs(tab.view(np.int32).astype(np.int64) - otab.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    phi = gpu_ctx.get_phi()
    rho = O.density_float(O.deposit(x, y, z, M, order))
    ophi = O.pm_phi(M, rho, otab)
    assert rel_rms(phi, ophi) < 2e-6


@pytest.mark.parametrize("order,sorted_first", [(2, False), (3, True), (4, True)])
def test_interlaced_pm(gpu_ctx, order, sorted_first):
    """Second deposit on the half-cell grid is bit-exact; the combined
    potential matches the oracle's FP64 interlaced solve."""
    n_side, M = 24, 24
    x, y, z = clustered(n_side ** 3, 81 + order)
    p = gpu_ctx.set_params(T.baseline_params(n_side, M, order, interlace=1))
    gpu_ctx.upload(x, y, z)
    if sorted_first:
        gpu_ctx.chainmesh_create()  # blocked (LDS) deposit path
    gpu_ctx.gravity_long_compute()
    g1 = gpu_ctx.get_density()  # the last deposit: the shifted grid
    assert np.array_equal(g1, O.deposit_shifted(x, y, z, M, order, 1 << 31))
    rho0 = O.density_float(O.deposit(x, y, z, M, order))
    rho1 = O.density_float(g1)
    ophi = O.pm_phi_interlaced(M, rho0, rho1, O.green_table(M, p.rs, order))
    assert rel_rms(gpu_ctx.get_phi(), ophi) < 2e-6


def test_full_kick_interlaced_vs_oracle(gpu_ctx):
    n_side = 32
    n = n_side ** 3
    x, y, z = uniform(n, 91)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 2, interlace=1))
    gpu_ctx.upload(x, y, z)
    gpu_ctx.kick_treepm()
    d = gpu_ctx.download()
    zf = np.zeros(n, np.float32)
    r = O.kick_full(oparams(p), x, y, z, zf, zf, zf, np.zeros(n, np.int8))
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)
    assert rel_rms(G, np.stack([r["gx"], r["gy"], r["gz"]], 1)) < 1e-5
    # interlaced CIC is within the 1e-3 accuracy target (plain CIC is not)
    s = np.sort(np.random.default_rng(4).choice(n, 512, replace=False))
    ph, gx, gy, gz = gpu_ctx.ewald(s)
    assert rel_rms(G[s], np.stack([gx, gy, gz], 1) * p.GM) < 1e-3


@pytest.mark.parametrize("dist,tile,lists", [("uniform", 32, 0), ("clustered", 32, 0), ("clustered", 16, 0),
                                             ("uniform", 64, 0), ("clustered", 8, 0), ("uniform", 16, 1),
                                             ("clustered", 16, 1)])
def test_pp_only_vs_oracle(gpu_ctx, dist, tile, lists):
    n_side = 32
    n = n_side ** 3
    x, y, z = uniform(n, 41) if dist == "uniform" else clustered(n, 42)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
    gpu_ctx.set_tuning(tile_sinks=tile, pp_lists=lists)
    gpu_ctx.upload(x, y, z)
    gpu_ctx.chainmesh_create()
    st = gpu_ctx.kick(flags=T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
    gpu_ctx.set_tuning(tile_sinks=32, pp_lists=0)
    d = gpu_ctx.download()
    operm, ocb = O.chainmesh_sort(x, y, z, p.chain_dim)
    xs, ys, zs = x[operm], y[operm], z[operm]
    op = oparams(p)
    (gx, gy, gz, pot), cand = O.pp_sum(op, xs, ys, zs, ocb, np.arange(n), fp64=False)
    assert st.candidate_pairs == cand
    assert st.executed_pairs <= cand
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)
    # GPU stores GM * (sum) and GM * (phi0 + sum)
    OG = np.stack([gx, gy, gz], 1) * p.GM
    assert rel_rms(G, OG) < 1e-5
    phi0 = p.n_total * 4 * np.pi * p.rs ** 2 + (35.0 / 16.0) / p.hsoft
    assert rel_rms(d["pot"], (pot + phi0) * p.GM) < 1e-5
    # and against the FP64 truth of the same pair formula
    (hx, hy, hz, hp), _ = O.pp_sum(op, xs, ys, zs, ocb, np.arange(n), fp64=True)
    assert rel_rms(G, np.stack([hx, hy, hz], 1) * p.GM) < 2e-5


def test_pp_on_a_larger_chain_mesh(gpu_ctx):
    """chain_dim 17: tiles laid out in 8^3-cell blocks with partial edge
    blocks.  Every particle is kicked by exactly one tile (all forces
    nonzero; velocities of sampled sinks equal one half kick, not two), and
    the sampled sinks' forces match the oracle."""
    n_side = 96
    n = n_side ** 3
    x, y, z = clustered(n, 43)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
    assert p.chain_dim == 17
    zf = np.zeros(n, np.float32)
    gpu_ctx.upload(x, y, z, zf, zf, zf, np.zeros(n, np.int8))
    gpu_ctx.chainmesh_create()
    print("pp chain_dim 17", flush=True)
    st = gpu_ctx.kick(0, 1.0, 1.0, True, flags=T.TPM_SKIP_PM)
    d = gpu_ctx.download()
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)
    assert np.all(np.abs(G).sum(1) > 0)
    operm, ocb = O.chainmesh_sort(x, y, z, p.chain_dim)
    assert np.array_equal(gpu_ctx.chainmesh_perm(), operm)
    xs, ys, zs = x[operm], y[operm], z[operm]
    assert st.candidate_pairs == O.candidate_pairs(p.chain_dim, ocb)
    sinks = np.sort(np.random.default_rng(7).choice(n, 3000, replace=False))
    op = oparams(p)
    (gx, gy, gz, pot), _ = O.pp_sum(op, xs, ys, zs, ocb, sinks, fp64=False)
    assert rel_rms(G[sinks], np.stack([gx, gy, gz], 1) * p.GM) < 1e-5
    z3 = np.zeros(len(sinks), np.float32)
    (og, opot, ov, orung) = O.kick_epilogue(op, (gx, gy, gz), pot, (z3, z3, z3), np.zeros(len(sinks), np.int8), 0,
                                            1.0, 1.0, True)
    assert rel_rms(d["pot"][sinks], opot) < 1e-5
    same = d["rung"][sinks] == orung
    assert same.mean() > 0.99
    V = np.stack([d["vx"], d["vy"], d["vz"]], 1)[sinks][same]
    assert rel_rms(V, np.stack(ov, 1)[same]) < 1e-5


def test_mesh_term_with_and_without_the_lds_cube(gpu_ctx):
    """The mesh force pass reads a chain cell's potential neighbourhood into
    LDS when it fits 32 KiB (chain_dim 11 at a 64^3 mesh) and the potential
    directly otherwise (chain_dim 3): the same arithmetic, so the mesh term of
    every particle is the same bits either way, and matches the oracle's
    6-point interpolation."""
    n_side, M = 24, 64
    n = n_side ** 3
    x, y, z = clustered(n, 61)
    out = {}
    for C in (11, 3):
        p = T.baseline_params(n_side, M, 3)
        p.chain_dim = C
        p.rcut = min(p.rcut, 0.99 / C)
        p = gpu_ctx.set_params(p)
        gpu_ctx.upload(x, y, z)
        gpu_ctx.chainmesh_create()
        gpu_ctx.gravity_long_compute()
        gpu_ctx.kick(flags=T.TPM_SKIP_PP | T.TPM_SKIP_KICK)
        d = gpu_ctx.download()
        perm = gpu_ctx.chainmesh_perm()  # perm[new] = old
        out[C] = {k: np.empty_like(d[k]) for k in ("gx", "gy", "gz", "pot")}
        for k in out[C]:
            out[C][k][perm] = d[k]
        phi = gpu_ctx.get_phi().reshape(M, M, M)
    for k in ("gx", "gy", "gz", "pot"):
        assert np.array_equal(out[11][k], out[3][k]), k
    ogx, ogy, ogz, _ = O.interp6(M, phi.ravel(), x, y, z)
    G = np.stack([out[3]["gx"], out[3]["gy"], out[3]["gz"]], 1)
    assert rel_rms(G, np.stack([ogx, ogy, ogz], 1) * p.GM) < 1e-5


@pytest.mark.parametrize("order", [2, 3])
def test_full_kick_vs_oracle(gpu_ctx, order):
    n_side = 32
    n = n_side ** 3
    x, y, z = clustered(n, 51)
    rng = np.random.default_rng(5)
    vx, vy, vz = [rng.standard_normal(n).astype(np.float32) * 1e-3 for _ in range(3)]
    rung = rng.integers(0, 3, n).astype(np.int8)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, order))
    gpu_ctx.upload(x, y, z, vx, vy, vz, rung)
    gpu_ctx.kick_treepm(min_rung=1, scale=1.0, t0=0.5, first_call=False)
    d = gpu_ctx.download()
    r = O.kick_full(oparams(p), x, y, z, vx, vy, vz, rung, min_rung=1, scale=1.0, t0=0.5, first_call=False)
    for k in ("x", "y", "z"):
        assert np.array_equal(d[k], r[k])
    operm, _ = O.chainmesh_sort(x, y, z, p.chain_dim)
    act = rung[operm] >= 1
    assert 0 < act.sum() < n
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)[act]
    OG = np.stack([r["gx"], r["gy"], r["gz"]], 1)[act]
    assert rel_rms(G, OG) < 1e-5
    assert rel_rms(d["pot"][act], r["pot"][act]) < 1e-5
    V = np.stack([d["vx"], d["vy"], d["vz"]], 1)
    OV = np.stack([r["vx"], r["vy"], r["vz"]], 1)
    assert np.array_equal(V[~act], OV[~act])  # inactive sinks untouched
    assert rel_rms(V[act], OV[act]) < 1e-5
    # rungs identical except where log2 of the time step sits at an integer
    mism = d["rung"] != r["rung"]
    assert mism.mean() < 1e-3, mism.mean()
    assert np.array_equal(d["rung"][~act], r["rung"][~act])


def test_ewald_gpu_vs_oracle(gpu_ctx):
    n = 4096
    x, y, z = clustered(n, 61)
    p = gpu_ctx.set_params(T.baseline_params(16, 32, 3))
    gpu_ctx.upload(x, y, z)
    s = np.arange(0, n, 97)
    ph, gx, gy, gz = gpu_ctx.ewald(s)
    oph, ogx, ogy, ogz = O.ewald_sf(x, y, z, s, p.hsoft, 9.6)
    assert rel_rms(np.stack([gx, gy, gz], 1), np.stack([ogx, ogy, ogz], 1)) < 1e-10
    assert rel_rms(ph, oph) < 1e-10


@pytest.mark.parametrize("order,tol", [(3, 1e-3), (4, 1e-3)])
def test_kick_vs_ewald(gpu_ctx, order, tol):
    n_side = 32
    n = n_side ** 3
    x, y, z = uniform(n, 71)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, order))
    gpu_ctx.upload(x, y, z)
    gpu_ctx.kick_treepm()
    d = gpu_ctx.download()
    rng = np.random.default_rng(3)
    s = np.sort(rng.choice(n, 512, replace=False))
    ph, gx, gy, gz = gpu_ctx.ewald(s)
    G = np.stack([d["gx"][s], d["gy"][s], d["gz"][s]], 1)
    E = np.stack([gx, gy, gz], 1) * p.GM
    assert rel_rms(G, E) < tol


def test_edge_positions_vs_oracle(gpu_ctx):
    """Coincident particles (zero force and the self-like d0 potential,
    as the reference's r = 0 handling gives), particles at raw 0 and
    2^32 - 1, particles on and beside every chain-cell split plane, and a
    half-empty box: full kick and chain mesh against the oracle."""
    n_side, C = 24, 4
    rng = np.random.default_rng(123)
    p = T.baseline_params(n_side, n_side, 3)
    p.chain_dim = C
    p.rcut = min(p.rcut, 0.99 / C)
    planes = [int(4294967296.0 * (k * (1.0 / C))) for k in range(C)]
    pts = []
    for b in planes:
        for d in (-1, 0, 1):
            pts.append((b + d) % 2 ** 32)
    pts += [0, 1, 2 ** 32 - 1, 2 ** 32 - 2]
    edge = np.array(pts, np.uint64)
    m = len(edge)
    n_rest = n_side ** 3 - 3 * m - 64
    # the rest in one half of the box along x
    xs = rng.integers(0, 2 ** 31, n_rest, dtype=np.uint64)
    ys = rng.integers(0, 2 ** 32, n_rest, dtype=np.uint64)
    zs = rng.integers(0, 2 ** 32, n_rest, dtype=np.uint64)
    ex = np.concatenate([edge, rng.integers(0, 2 ** 32, m, dtype=np.uint64), rng.integers(0, 2 ** 32, m, dtype=np.uint64)])
    ey = np.concatenate([rng.integers(0, 2 ** 32, m, dtype=np.uint64), edge, rng.integers(0, 2 ** 32, m, dtype=np.uint64)])
    ez = np.concatenate([rng.integers(0, 2 ** 32, m, dtype=np.uint64), rng.integers(0, 2 ** 32, m, dtype=np.uint64), edge])
    x = np.concatenate([xs, ex]).astype(np.uint32)
    y = np.concatenate([ys, ey]).astype(np.uint32)
    z = np.concatenate([zs, ez]).astype(np.uint32)
    # 32 coincident pairs
    k = rng.choice(len(x), 32, replace=False)
    x, y, z = np.concatenate([x, x[k]]), np.concatenate([y, y[k]]), np.concatenate([z, z[k]])
    n = len(x)
    p.n_total = float(n)
    p.GM = 1.0 / n
    p = gpu_ctx.set_params(p)
    gpu_ctx.upload(x, y, z)
    gpu_ctx.kick_treepm()
    d = gpu_ctx.download()
    operm, ocb = O.chainmesh_sort(x, y, z, C)
    assert np.array_equal(gpu_ctx.chainmesh_perm(), operm)
    assert np.array_equal(gpu_ctx.chainmesh_cells(), ocb)
    zf = np.zeros(n, np.float32)
    r = O.kick_full(oparams(p), x, y, z, zf, zf, zf, np.zeros(n, np.int8))
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)
    OG = np.stack([r["gx"], r["gy"], r["gz"]], 1)
    assert np.all(np.isfinite(G)) and np.all(np.isfinite(d["pot"]))
    assert rel_rms(G, OG) < 1e-5
    assert rel_rms(d["pot"], r["pot"]) < 1e-5


@pytest.mark.parametrize("flags_extra", [T.TPM_SKIP_PM | T.TPM_SKIP_KICK, 0])
def test_kick_pme_law_vs_oracle(gpu_ctx, flags_extra):
    """TPM_KICK_PME: kick_pme's pair law (every pair of the 27 cells, plain
    softening inside h, phi without phi0) against the oracle's restatement,
    PP alone and the whole kick with mixed rungs."""
    n_side = 28
    n = n_side ** 3
    x, y, z = clustered(n, 101)
    rng = np.random.default_rng(7)
    vx, vy, vz = [rng.standard_normal(n).astype(np.float32) * 1e-3 for _ in range(3)]
    rung = rng.integers(0, 3, n).astype(np.int8)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 4))
    gpu_ctx.upload(x, y, z, vx, vy, vz, rung)
    gpu_ctx.chainmesh_create()
    gpu_ctx.gravity_long_compute()
    st = gpu_ctx.kick(min_rung=1, scale=1.0, t0=0.5, first_call=False, flags=T.TPM_KICK_PME | flags_extra)
    d = gpu_ctx.download()
    op = oparams(p)
    op.pme = 1
    operm, ocb = O.chainmesh_sort(x, y, z, p.chain_dim)
    act = rung[operm] >= 1
    # no culling without a cutoff: every active sink meets all 27 cells
    C = p.chain_dim
    cnt = np.diff(ocb).reshape(C, C, C)
    nsum = sum(np.roll(cnt, (a, b, c), (0, 1, 2)) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1))
    cell_of = np.repeat(np.arange(C ** 3), np.diff(ocb))
    assert st.executed_pairs == float(nsum.ravel()[cell_of][act].sum())
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)[act]
    if flags_extra:
        xs, ys, zs = x[operm], y[operm], z[operm]
        (gx, gy, gz, pot), _ = O.pp_sum(op, xs, ys, zs, ocb, np.arange(n), fp64=False)
        OG = np.stack([gx, gy, gz], 1)[act] * p.GM
        assert rel_rms(G, OG) < 1e-5
        assert rel_rms(d["pot"][act], pot[act] * p.GM) < 1e-5
    else:
        r = O.kick_full(op, x, y, z, vx, vy, vz, rung, min_rung=1, scale=1.0, t0=0.5, first_call=False)
        OG = np.stack([r["gx"], r["gy"], r["gz"]], 1)[act]
        assert rel_rms(G, OG) < 1e-5
        assert rel_rms(d["pot"][act], r["pot"][act]) < 1e-5
        V = np.stack([d["vx"], d["vy"], d["vz"]], 1)
        assert rel_rms(V[act], np.stack([r["vx"], r["vy"], r["vz"]], 1)[act]) < 1e-5


def test_kick_is_deterministic(gpu_ctx):
    """Two whole steps from the same particles give the same bits: the
    deposit is integer, the PP sums have a fixed order, nothing is summed by
    float atomics."""
    n_side = 32
    x, y, z = clustered(n_side ** 3, 41)
    out = []
    for _ in range(2):
        gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
        gpu_ctx.upload(x, y, z)
        gpu_ctx.kick_treepm(0, 1.0, 1.0, True)
        out.append(gpu_ctx.download())
    for k in ("x", "gx", "gy", "gz", "pot", "vx", "rung"):
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("ts,lists", [(16, 0), (32, 0), (16, 1)])
def test_tiles_per_workgroup_same_bits(gpu_ctx, ts, lists):
    """One tile per workgroup or four (pp_waves): every wave computes its own
    tile alone either way, so every array comes out the same bits, for a
    tile count that is not a multiple of 4 x 8."""
    n_side = 30
    x, y, z = clustered(n_side ** 3, 43)
    prev = gpu_ctx.get_tuning()
    out = []
    for waves in (1, 4):
        gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
        gpu_ctx.set_tuning(tile_sinks=ts, pp_waves=waves, pp_lists=lists)
        assert gpu_ctx.get_tuning()["pp_waves"] == waves and gpu_ctx.get_tuning()["pp_lists"] == lists
        gpu_ctx.upload(x, y, z)
        print("pp_waves", waves, "tile", ts, flush=True)
        st = gpu_ctx.kick_treepm(0, 1.0, 1.0, True)
        out.append((st, gpu_ctx.download()))
    gpu_ctx.set_tuning(**prev)
    assert out[0][0].executed_pairs == out[1][0].executed_pairs
    assert out[0][0].max_rung == out[1][0].max_rung
    for k in ("x", "gx", "gy", "gz", "pot", "vx", "vy", "vz", "rung"):
        assert np.array_equal(out[0][1][k], out[1][1][k]), k


@pytest.mark.parametrize("dist", ["uniform", "clustered"])
def test_per_sink_lists_equal_tile_lists(gpu_ctx, dist):
    """Per-sink source lists evaluate only pairs near each sink: fewer
    executed pairs, and the same forces and potentials as one list per tile
    up to FP32 summation order (every pair with r < rcut is still
    evaluated); rungs and the maximum rung agree."""
    n_side = 48
    n = n_side ** 3
    x, y, z = uniform(n, 51) if dist == "uniform" else clustered(n, 52)
    prev = gpu_ctx.get_tuning()
    out = []
    for lists in (0, 1):
        gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
        gpu_ctx.set_tuning(tile_sinks=16, pp_lists=lists)
        gpu_ctx.upload(x, y, z)
        print("pp_lists", lists, dist, flush=True)
        out.append((gpu_ctx.kick_treepm(0, 1.0, 1.0, True), gpu_ctx.download()))
    gpu_ctx.set_tuning(**prev)
    (s0, d0), (s1, d1) = out
    assert s1.candidate_pairs == s0.candidate_pairs
    assert s1.executed_pairs < 0.8 * s0.executed_pairs
    print("executed pairs: tile lists %.4g, per-sink lists %.4g" % (s0.executed_pairs, s1.executed_pairs))
    assert np.array_equal(d0["x"], d1["x"])
    G0 = np.stack([d0["gx"], d0["gy"], d0["gz"]], 1)
    G1 = np.stack([d1["gx"], d1["gy"], d1["gz"]], 1)
    # FP32 sums of ~10^3 terms in another order (measured: 1.05e-6 for the
    # potential of the uniform box)
    assert rel_rms(G1, G0) < 5e-6
    assert rel_rms(d1["pot"], d0["pot"]) < 5e-6
    assert (d0["rung"] == d1["rung"]).mean() > 0.999
    assert s0.max_rung == s1.max_rung


@pytest.mark.parametrize("lists", [0, 1])
def test_sparse_cells_on_a_small_chain_mesh(gpu_ctx, lists):
    """chain_dim 5 with a few particles per cell: one tile spans a whole
    cell, so sinks sit up to a cell from the tile's first particle.  No
    stand-in source may then act on a sink (a source placed at a fixed
    offset can come within rcut of such a sink when C < 8): the pair sums
    equal the oracle's."""
    n_side = 32
    n = 900
    x, y, z = uniform(n, 61)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3, n_total=n))
    assert p.chain_dim == 5
    prev = gpu_ctx.get_tuning()
    gpu_ctx.set_tuning(tile_sinks=16, pp_lists=lists)
    gpu_ctx.upload(x, y, z)
    gpu_ctx.chainmesh_create()
    st = gpu_ctx.kick(flags=T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
    gpu_ctx.set_tuning(**prev)
    d = gpu_ctx.download()
    operm, ocb = O.chainmesh_sort(x, y, z, p.chain_dim)
    (gx, gy, gz, pot), cand = O.pp_sum(oparams(p), x[operm], y[operm], z[operm], ocb, np.arange(n), fp64=False)
    assert st.candidate_pairs == cand
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)
    assert rel_rms(G, np.stack([gx, gy, gz], 1) * p.GM) < 1e-5
    # no sink off by a stray term either
    err = np.linalg.norm(G - np.stack([gx, gy, gz], 1) * p.GM, axis=1) / np.linalg.norm(G, axis=1).mean()
    assert err.max() < 1e-4


@pytest.mark.parametrize("dist,n_side", [("uniform", 48), ("clustered", 48), ("clustered", 64)])
def test_pairs_once_equal_pairs_from_both_ends(gpu_ctx, dist, n_side):
    """pp_newton: each pair evaluated once and added to both ends.  Fewer
    pair evaluations (32-source tiles cull more loosely than 16-sink ones,
    so not half); forces, potentials, velocities and rungs equal the
    two-sided kernel's up to FP32 summation order; the whole kick within
    1e-5 of the oracle; a repeated kick gives the same bits.  The clustered
    boxes put thousands of particles in one cell, so cells are split into
    several LDS chunks."""
    n = n_side ** 3
    x, y, z = uniform(n, 71) if dist == "uniform" else clustered(n, 72)
    prev = gpu_ctx.get_tuning()
    out = []
    for newton in (0, 1, 1):
        p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
        assert p.chain_dim >= 8
        gpu_ctx.set_tuning(tile_sinks=16, pp_newton=newton)
        gpu_ctx.upload(x, y, z)
        print("pp_newton", newton, dist, n_side, flush=True)
        out.append((gpu_ctx.kick_treepm(0, 1.0, 1.0, True), gpu_ctx.download()))
    gpu_ctx.set_tuning(**prev)
    (s0, d0), (s1, d1), (s2, d2) = out
    print("pair evaluations: both ends %.4g, once %.4g" % (s0.executed_pairs, s1.executed_pairs))
    assert s1.candidate_pairs == s0.candidate_pairs
    assert s1.executed_pairs < 0.75 * s0.executed_pairs
    for k in ("x", "gx", "gy", "gz", "pot", "vx", "vy", "vz", "rung"):
        assert np.array_equal(d1[k], d2[k]), k
    assert np.array_equal(d0["x"], d1["x"])
    G0 = np.stack([d0["gx"], d0["gy"], d0["gz"]], 1)
    G1 = np.stack([d1["gx"], d1["gy"], d1["gz"]], 1)
    assert rel_rms(G1, G0) < 5e-6
    assert rel_rms(d1["pot"], d0["pot"]) < 5e-6
    assert (d0["rung"] == d1["rung"]).mean() > 0.999
    assert s0.max_rung == s1.max_rung
    zf = np.zeros(n, np.float32)
    r = O.kick_full(oparams(p), x, y, z, zf, zf, zf, np.zeros(n, np.int8))
    assert rel_rms(G1, np.stack([r["gx"], r["gy"], r["gz"]], 1)) < 1e-5
    assert rel_rms(d1["pot"], r["pot"]) < 1e-5


def cells_with(counts, C, seed, background=0):
    """Fixed-point positions with counts[k] particles inside chain cell
    cells k = (x, y, z) (uniform in the cell), plus a uniform background."""
    rng = np.random.default_rng(seed)
    pos = [[], [], []]
    for (cx, cy, cz), m in counts.items():
        for d, c0 in enumerate((cx, cy, cz)):
            pos[d].append((c0 + rng.random(m)) / C)
    for d in range(3):
        pos[d].append(rng.random(background))
    return [np.minimum(np.floor(np.concatenate(p) * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint32) for p in pos]


@pytest.mark.parametrize("case", ["sparse", "chunks"])
def test_pairs_once_edge_cases(gpu_ctx, case):
    """pp_newton on sparse boxes (most cells empty or nearly so) and on
    cells of exactly 255, 256, 257, 512 and 513 particles next to each
    other (LDS chunks of 256 sinks); with the mesh off, with the kick off
    and with both on, the pair-once kick equals the two-sided one to FP32
    summation order.  A kick with min_rung > 0 is not eligible and runs the
    two-sided kernel: the same bits."""
    n_side = 48
    C = T.baseline_params(n_side, n_side, 3).chain_dim
    assert C == 8
    if case == "sparse":
        x, y, z = uniform(3000, 77)
    else:
        counts = {(1, 1, 1): 255, (2, 1, 1): 256, (2, 2, 1): 257, (1, 2, 2): 512, (2, 2, 2): 513, (5, 5, 5): 1}
        x, y, z = cells_with(counts, C, 78, background=20000)
    n = len(x)
    prev = gpu_ctx.get_tuning()
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3, n_total=n))
    rng = np.random.default_rng(79)
    v0 = [rng.standard_normal(n).astype(np.float32) * 1e-3 for _ in range(3)]
    r0 = rng.integers(0, 3, n).astype(np.int8)
    operm, _ = O.chainmesh_sort(x, y, z, C)
    out = {}
    try:
        for flags in (T.TPM_SKIP_PM, T.TPM_SKIP_KICK, 0):
            for newton in (0, 1):
                gpu_ctx.set_tuning(tile_sinks=16, pp_newton=newton)
                gpu_ctx.upload(x, y, z, *v0, r0)
                gpu_ctx.chainmesh_create()
                gpu_ctx.gravity_long_compute()
                print("pp_newton", newton, case, "flags", flags, flush=True)
                st = gpu_ctx.kick(0, 1.0, 1.0, True, flags)
                out[flags, newton] = (st, gpu_ctx.download())
        for newton in (0, 1):
            gpu_ctx.set_tuning(tile_sinks=16, pp_newton=newton)
            gpu_ctx.upload(x, y, z, *v0, r0)
            gpu_ctx.chainmesh_create()
            gpu_ctx.gravity_long_compute()
            print("pp_newton", newton, case, "min_rung 1", flush=True)
            out["rung1", newton] = (gpu_ctx.kick(1, 1.0, 1.0, True, 0), gpu_ctx.download())
    finally:
        gpu_ctx.set_tuning(**prev)
    for flags in (T.TPM_SKIP_PM, T.TPM_SKIP_KICK, 0):
        (s0, d0), (s1, d1) = out[flags, 0], out[flags, 1]
        assert s1.candidate_pairs == s0.candidate_pairs
        assert s1.executed_pairs < s0.executed_pairs
        assert np.array_equal(d0["x"], d1["x"])
        G0 = np.stack([d0["gx"], d0["gy"], d0["gz"]], 1)
        G1 = np.stack([d1["gx"], d1["gy"], d1["gz"]], 1)
        assert rel_rms(G1, G0) < 5e-6, flags
        assert rel_rms(d1["pot"], d0["pot"]) < 5e-6, flags
        V0 = np.stack([d0["vx"], d0["vy"], d0["vz"]], 1)
        V1 = np.stack([d1["vx"], d1["vy"], d1["vz"]], 1)
        assert rel_rms(V1, V0) < 5e-6, flags
        if flags & T.TPM_SKIP_KICK:
            assert np.array_equal(d1["vx"], v0[0][operm]) and np.array_equal(d1["rung"], r0[operm])
        assert (d0["rung"] == d1["rung"]).mean() > 0.999
    (s0, d0), (s1, d1) = out["rung1", 0], out["rung1", 1]
    for k in ("gx", "gy", "gz", "pot", "vx", "vy", "vz", "rung"):
        assert np.array_equal(d0[k], d1[k]), k
    assert s0.executed_pairs == s1.executed_pairs
    # the whole kick against the oracle (positions in chain-mesh order)
    d1 = out[0, 1][1]
    r = O.kick_full(oparams(p), x, y, z, v0[0], v0[1], v0[2], r0)
    G1 = np.stack([d1["gx"], d1["gy"], d1["gz"]], 1)
    assert np.array_equal(d1["x"], r["x"])
    assert rel_rms(G1, np.stack([r["gx"], r["gy"], r["gz"]], 1)) < 1e-5
    assert rel_rms(d1["pot"], r["pot"]) < 1e-5


@pytest.mark.parametrize("n", [1, 2, 17])
def test_tiny_boxes_both_kernels(gpu_ctx, n):
    """One, two and seventeen particles in a box with chain_dim 8 (all but
    a few cells empty; n = 2 puts both in one cell, 0.013 h apart): the
    short-range forces of the two-sided and the pair-once kernel equal the
    FP64 pair law's (a single particle feels none, only its own potential
    term).  Deep inside the softening length the law as written loses its
    digits to cancellation in FP32 (the FP32 oracle is 0.7 % off there); the
    kernels evaluate those terms by their series.  The mesh force is left
    out: for a handful of particles it is mostly the assignment's
    self-force, a residual of large cancelling terms that FP32 and FP64
    FFTs resolve differently."""
    rng = np.random.default_rng(100 + n)
    if n == 2:
        x = np.array([0x40000000, 0x40010000], np.uint32)
        y = np.array([0x40000000, 0x40002000], np.uint32)
        z = np.array([0x40000000, 0x3fff0000], np.uint32)
    else:
        x, y, z = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    zf = np.zeros(n, np.float32)
    prev = gpu_ctx.get_tuning()
    try:
        for newton in (0, 1):
            p = gpu_ctx.set_params(T.baseline_params(48, 48, 3, n_total=n))
            gpu_ctx.set_tuning(tile_sinks=16, pp_newton=newton)
            gpu_ctx.upload(x, y, z, zf, zf, zf, np.zeros(n, np.int8))
            gpu_ctx.chainmesh_create()
            print("tiny box", n, "pp_newton", newton, flush=True)
            gpu_ctx.kick(flags=T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
            d = gpu_ctx.download()
            operm, ocb = O.chainmesh_sort(x, y, z, p.chain_dim)
            assert np.array_equal(d["x"], x[operm])
            xs, ys, zs = x[operm], y[operm], z[operm]
            (gx, gy, gz, _), _ = O.pp_sum(oparams(p), xs, ys, zs, ocb, np.arange(n), fp64=True)
            # the potential against the FP32 law, which shares the reference's
            # own-term convention (at r = 0 no erfc correction; the FP64 law
            # takes the r -> 0 limit there)
            (_, _, _, pot), _ = O.pp_sum(oparams(p), xs, ys, zs, ocb, np.arange(n), fp64=False)
            G = np.stack([d["gx"], d["gy"], d["gz"]], 1)
            OG = np.stack([gx, gy, gz], 1) * p.GM
            phi0 = p.n_total * 4 * np.pi * p.rs ** 2 + (35.0 / 16.0) / p.hsoft
            if n == 1:
                assert np.all(G == 0.0)
            else:
                assert rel_rms(G, OG) < 1e-5, newton
            assert rel_rms(d["pot"], (pot + phi0) * p.GM) < 1e-5, newton
    finally:
        gpu_ctx.set_tuning(**prev)


def test_pair_rate_probe(gpu_ctx):
    """The pair-body probe runs with the current parameters and reports a
    rate; out-of-range occupancies and missing parameters are errors."""
    fresh = T.Context()
    try:
        with pytest.raises(T.TpmError, match="parameters"):
            fresh.probe_pair_rate(7)
    finally:
        fresh.close()
    gpu_ctx.set_params(T.baseline_params(64, 64, 3))
    with pytest.raises(T.TpmError, match="waves_per_simd"):
        gpu_ctx.probe_pair_rate(9)
    r4, r8 = gpu_ctx.probe_pair_rate(4), gpu_ctx.probe_pair_rate(8)
    print("pair body rate", r4, r8, flush=True)
    assert 1e11 < r4 < 1e14 and 1e11 < r8 < 1e14
