"""GPU parity of the tree path (TPM_KICK_TREE, tpm_tree_*): the per-cell
trees bit-identical to the oracle's; the walk's decisions identical (the
same number of multipole interactions and pairs); forces and the whole kick
within FP32 tolerance of the oracle's restatement of kick_treepm."""
import numpy as np
import pytest

import oracle as O
import tigerpm_amd as T

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def positions(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    c = rng.random((12, 3))
    k = rng.integers(0, 12, n)
    pos = (c[k] + rng.standard_normal((n, 3)) * 0.03) % 1.0
    return [np.minimum(np.floor(pos[:, d] * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint32) for d in range(3)]


def oparams(p):
    return O.make_params(p.mesh_dim, p.chain_dim, p.order, p.rs, p.rcut, p.hsoft, p.GM, p.eta, p.n_total, p.do_phi,
                         p.interlace)


def setup(ctx, kind, n_side, seed, params=None):
    n = n_side ** 3
    x, y, z = positions(kind, n, seed)
    p = ctx.set_params(params if params is not None else T.baseline_params(n_side, n_side, 4))
    ctx.upload(x, y, z)
    ctx.chainmesh_create()
    operm, ocb = O.chainmesh_sort(x, y, z, p.chain_dim)
    assert np.array_equal(ctx.chainmesh_perm(), operm)
    return p, x[operm], y[operm], z[operm], ocb


@pytest.mark.parametrize("kind,n_side", [("uniform", 24), ("clustered", 24), ("clustered", 40)])
def test_trees_bit_identical(gpu_ctx, kind, n_side):
    p, xs, ys, zs, cb = setup(gpu_ctx, kind, n_side, 3)
    print("trees", kind, n_side, flush=True)
    gpu_ctx.tree_build()
    g = gpu_ctx.tree_download()
    o = O.tree_build(oparams(p), xs, ys, zs, cb)
    assert np.array_equal(g["order"], o["order"])
    assert np.array_equal(g["node_begin"], o["node_begin"])
    assert np.array_equal(g["bucket_begin"], o["bucket_begin"])
    assert g["nodes"].tobytes() == o["nodes"].tobytes()
    for f in ("begin", "end", "x", "radius"):
        assert np.array_equal(g["buckets"][f], o["buckets"][f]), f
    cells = np.repeat(np.arange(p.chain_dim ** 3), np.diff(o["bucket_begin"]))
    assert np.array_equal(g["buckets"]["cell"], cells)


@pytest.mark.parametrize("kind,n_side,min_rung", [("uniform", 32, 0), ("clustered", 32, 1), ("clustered", 24, 0)])
def test_walk_matches_oracle(gpu_ctx, kind, n_side, min_rung):
    """Short range alone (no mesh, no kick): the same decisions as the
    oracle's walk (PC and PP counts equal) and the same sums to FP32
    rounding (the GPU's rsqrt and summation order)."""
    p, xs, ys, zs, cb = setup(gpu_ctx, kind, n_side, 5)
    n = len(xs)
    rung = np.random.default_rng(4).integers(0, 3, n).astype(np.int8)
    gpu_ctx.upload(xs, ys, zs, rung=rung)  # already in chain-mesh order: the sort keeps it
    gpu_ctx.chainmesh_create()
    print("walk", kind, n_side, min_rung, flush=True)
    st = gpu_ctx.kick(min_rung=min_rung, flags=T.TPM_KICK_TREE | T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
    d = gpu_ctx.download()
    (ox, oy, oz, oph), ost = O.tree_sum(oparams(p), xs, ys, zs, cb, rung, min_rung)
    assert st.pc_interactions == ost["pc"] and st.executed_pairs == ost["pp"]
    act = rung >= min_rung
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)[act] / p.GM
    OG = np.stack([ox, oy, oz], 1)[act]
    assert rel_rms(G, OG) < 1e-5
    phi0 = p.n_total * 4.0 * np.pi * p.rs ** 2 + (35.0 / 16.0) / p.hsoft
    assert rel_rms(d["pot"][act] / p.GM - phi0, oph[act]) < 1e-5
    assert ost["pc"] > 0


def test_tree_kick_matches_oracle(gpu_ctx):
    """The whole kick (mesh + tree walk + kick epilogue, mixed rungs)."""
    n_side = 32
    p, xs, ys, zs, cb = setup(gpu_ctx, "clustered", n_side, 9)
    n = len(xs)
    rng = np.random.default_rng(2)
    vx, vy, vz = [rng.standard_normal(n).astype(np.float32) * 1e-3 for _ in range(3)]
    rung = rng.integers(0, 3, n).astype(np.int8)
    gpu_ctx.upload(xs, ys, zs, vx, vy, vz, rung)
    gpu_ctx.chainmesh_create()
    gpu_ctx.gravity_long_compute()
    print("tree kick", flush=True)
    gpu_ctx.kick(min_rung=1, scale=1.0, t0=0.5, first_call=False, flags=T.TPM_KICK_TREE)
    d = gpu_ctx.download()
    r = O.kick_full(oparams(p), xs, ys, zs, vx, vy, vz, rung, min_rung=1, scale=1.0, t0=0.5, first_call=False,
                    tree=True)
    assert np.array_equal(d["x"], r["x"])
    act = rung >= 1
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1)[act]
    OG = np.stack([r["gx"], r["gy"], r["gz"]], 1)[act]
    assert rel_rms(G, OG) < 1e-5
    assert rel_rms(d["pot"][act], r["pot"][act]) < 1e-5
    V = np.stack([d["vx"], d["vy"], d["vz"]], 1)
    OV = np.stack([r["vx"], r["vy"], r["vz"]], 1)
    assert rel_rms(V[act] - np.stack([vx, vy, vz], 1)[act], OV[act] - np.stack([vx, vy, vz], 1)[act]) < 1e-4
    assert np.array_equal(V[~act], np.stack([vx, vy, vz], 1)[~act])
    assert np.mean(d["rung"][act] == r["rung"][act]) > 0.999


def test_tree_and_direct_kicks_differ_by_the_multipole_error(gpu_ctx):
    """The tree kick and the direct kick on the same particles differ by the
    multipole approximation only: by what the oracle's two paths differ."""
    n_side = 32
    p, xs, ys, zs, cb = setup(gpu_ctx, "uniform", n_side, 13)
    gpu_ctx.gravity_long_compute()
    print("tree vs direct", flush=True)
    gpu_ctx.kick(flags=T.TPM_KICK_TREE | T.TPM_SKIP_KICK)
    a = gpu_ctx.download()
    gpu_ctx.kick(flags=T.TPM_SKIP_KICK)
    b = gpu_ctx.download()
    Ga = np.stack([a["gx"], a["gy"], a["gz"]], 1)
    Gb = np.stack([b["gx"], b["gy"], b["gz"]], 1)
    n = len(xs)
    zf = np.zeros(n, np.float32)
    z8 = np.zeros(n, np.int8)
    ot = O.kick_full(oparams(p), xs, ys, zs, zf, zf, zf, z8, tree=True)
    od = O.kick_full(oparams(p), xs, ys, zs, zf, zf, zf, z8)
    Ot = np.stack([ot["gx"], ot["gy"], ot["gz"]], 1)
    Od = np.stack([od["gx"], od["gy"], od["gz"]], 1)
    gpu_diff, orc_diff = rel_rms(Ga, Gb), rel_rms(Ot, Od)
    assert orc_diff > 1e-4
    assert abs(gpu_diff - orc_diff) < 0.02 * orc_diff


@pytest.mark.parametrize("n,params", [(60, "baseline32"), (45 ** 3, "reference45")])
def test_sparse_and_reference_boxes(gpu_ctx, n, params):
    """A box with mostly empty cells (empty trees and buckets) and the
    reference's own parameters (rcut = 1/chain_dim, C = 3, unit masses)."""
    p = T.baseline_params(32, 32, 4) if params == "baseline32" else T.reference_params(45)
    rng = np.random.default_rng(17)
    x, y, z = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    p = gpu_ctx.set_params(p)
    gpu_ctx.upload(x, y, z)
    gpu_ctx.chainmesh_create()
    operm, ocb = O.chainmesh_sort(x, y, z, p.chain_dim)
    xs, ys, zs = x[operm], y[operm], z[operm]
    print("sparse/reference", n, params, flush=True)
    gpu_ctx.tree_build()
    g = gpu_ctx.tree_download()
    o = O.tree_build(oparams(p), xs, ys, zs, ocb)
    assert g["nodes"].tobytes() == o["nodes"].tobytes()
    st = gpu_ctx.kick(flags=T.TPM_KICK_TREE | T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
    d = gpu_ctx.download()
    (ox, oy, oz, oph), ost = O.tree_sum(oparams(p), xs, ys, zs, ocb, np.zeros(n, np.int8))
    assert st.pc_interactions == ost["pc"] and st.executed_pairs == ost["pp"]
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1) / p.GM
    assert rel_rms(G, np.stack([ox, oy, oz], 1)) < 1e-5


def test_crowded_cell_and_deep_tree(gpu_ctx):
    """A cell of 920 particles (built from device memory instead of LDS) and
    40 particles within 2^16 raw units of each other (a tree 40 levels
    deep): trees bit-identical, walk equal to the oracle's.  The softening
    length is cut to 1e-7: with the default one the close pairs sit where
    the softened (erfc - 1) term of kick_treepm.cu:263-276 cancels to FP32
    rounding noise (r < ~4e-3 h), which no two implementations share."""
    n_side = 24
    rng = np.random.default_rng(5)
    x, y, z = [rng.integers(0, 2 ** 32, n_side ** 3, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    cl = (np.array([0.3, 0.6, 0.1]) + rng.standard_normal((700, 3)) * 0.01) % 1.0
    cl = np.minimum(np.floor(cl * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint32)
    tight = (np.array([0.71, 0.42, 0.83]) * 2.0 ** 32).astype(np.uint64)
    tight = (tight + rng.integers(0, 2 ** 16, (40, 3), dtype=np.uint64)).astype(np.uint32)
    x, y, z = [np.concatenate([a, cl[:, d], tight[:, d]]) for d, a in enumerate((x, y, z))]
    n = len(x)
    p = T.baseline_params(n_side, n_side, 4, n_total=n)
    p.hsoft = 1e-7
    p = gpu_ctx.set_params(p)
    gpu_ctx.upload(x, y, z)
    gpu_ctx.chainmesh_create()
    operm, cb = O.chainmesh_sort(x, y, z, p.chain_dim)
    assert np.diff(cb).max() > 512
    xs, ys, zs = x[operm], y[operm], z[operm]
    print("crowded/deep", flush=True)
    gpu_ctx.tree_build()
    g = gpu_ctx.tree_download()
    o = O.tree_build(oparams(p), xs, ys, zs, cb)
    assert np.array_equal(g["order"], o["order"])
    assert g["nodes"].tobytes() == o["nodes"].tobytes()
    for f in ("begin", "end", "x", "radius"):
        assert np.array_equal(g["buckets"][f], o["buckets"][f]), f
    st = gpu_ctx.kick(flags=T.TPM_KICK_TREE | T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
    d = gpu_ctx.download()
    (ox, oy, oz, oph), ost = O.tree_sum(oparams(p), xs, ys, zs, cb, np.zeros(n, np.int8))
    assert st.pc_interactions == ost["pc"] and st.executed_pairs == ost["pp"]
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1) / p.GM
    assert rel_rms(G, np.stack([ox, oy, oz], 1)) < 1e-5


def test_trees_follow_the_chain_mesh(gpu_ctx):
    """Trees are reused while the chain mesh stands and rebuilt (buffers
    grown) when a larger, different box replaces it in the same context."""
    for kind, n_side, seed in (("uniform", 24, 31), ("clustered", 32, 32)):
        p, xs, ys, zs, cb = setup(gpu_ctx, kind, n_side, seed)
        print("rebuild", kind, n_side, flush=True)
        a = gpu_ctx.kick(flags=T.TPM_KICK_TREE | T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
        b = gpu_ctx.kick(flags=T.TPM_KICK_TREE | T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
        assert (a.executed_pairs, a.pc_interactions) == (b.executed_pairs, b.pc_interactions)
        g = gpu_ctx.tree_download()
        o = O.tree_build(oparams(p), xs, ys, zs, cb)
        assert g["nodes"].tobytes() == o["nodes"].tobytes()
        assert np.array_equal(g["order"], o["order"])


def test_walk_on_a_displaced_lattice(gpu_ctx):
    """Near-lattice particles (Zel'dovich, 0.5 cell rms), where the net
    forces are small and the multipole errors weigh most: the GPU walk still
    equals the oracle's to FP32 rounding."""
    n_side = 40
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
    gpu_ctx.ic_zeldovich(n_side, 7, 0.5, -2.0)
    gpu_ctx.chainmesh_create()
    d0 = gpu_ctx.download()
    xs, ys, zs = d0["x"], d0["y"], d0["z"]
    cb = gpu_ctx.chainmesh_cells()
    print("lattice walk", flush=True)
    st = gpu_ctx.kick(flags=T.TPM_KICK_TREE | T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
    d = gpu_ctx.download()
    n = len(xs)
    (ox, oy, oz, oph), ost = O.tree_sum(oparams(p), xs, ys, zs, cb, np.zeros(n, np.int8))
    assert st.pc_interactions == ost["pc"] and st.executed_pairs == ost["pp"]
    G = np.stack([d["gx"], d["gy"], d["gz"]], 1) / p.GM
    assert rel_rms(G, np.stack([ox, oy, oz], 1)) < 1e-5


def test_tree_api_errors(gpu_ctx):
    """Out-of-order calls and conflicting flags are errors, not results."""
    n_side = 24
    x, y, z = positions("uniform", n_side ** 3, 2)
    gpu_ctx.set_params(T.baseline_params(n_side, n_side, 4))
    gpu_ctx.upload(x, y, z)
    with pytest.raises(T.TpmError, match="chain mesh"):
        gpu_ctx.tree_build()
    with pytest.raises(T.TpmError, match="not built"):
        gpu_ctx.tree_download()
    gpu_ctx.chainmesh_create()
    with pytest.raises(T.TpmError, match="exclude"):
        gpu_ctx.kick(flags=T.TPM_KICK_TREE | T.TPM_KICK_PME | T.TPM_SKIP_PM)
    gpu_ctx.tree_build()
    gpu_ctx.upload(x, y, z)  # new particles: the trees no longer apply
    with pytest.raises(T.TpmError, match="not built"):
        gpu_ctx.tree_download()


def test_depth_limit_is_an_error(gpu_ctx):
    n_side = 24
    n = n_side ** 3
    x, y, z = positions("uniform", n, 1)
    x[:40], y[:40], z[:40] = 123456789, 987654321, 555555555
    gpu_ctx.set_params(T.baseline_params(n_side, n_side, 4))
    gpu_ctx.upload(x, y, z)
    gpu_ctx.chainmesh_create()
    with pytest.raises(T.TpmError, match="depth"):
        gpu_ctx.tree_build()


@pytest.mark.parametrize("kind,n_side", [("clustered", 24), ("uniform", 20)])
def test_trees_from_the_reference_order(kind, n_side):
    """With TPM_ORDER_REFERENCE the trees start from the order the reference
    partition leaves the caller's arrays in, as tree_create does
    (kick_treepm.cpp:26-44 after chainmesh_create): node for node and in
    particle order bit-identical to the oracle's two-pointer trees of that
    order; tpm_tree_sort_particles then leaves the caller's arrays in tree
    order, and a kick's outputs follow it."""
    n = n_side ** 3
    x, y, z = positions(kind, n, 11)
    with T.Context(0) as ctx:
        p = ctx.set_params(T.baseline_params(n_side, n_side, 4))
        ctx.set_host_order(T.TPM_ORDER_REFERENCE)
        ctx.upload(x, y, z)
        ctx.chainmesh_create()
        idx, rb, _ = O.chainmesh_partition_ref(x, y, z, p.chain_dim)
        cb = np.concatenate([rb, [n]]).astype(np.int64)
        xr, yr, zr = x[idx], y[idx], z[idx]
        ctx.tree_build()
        g = ctx.tree_download()
        o = O.tree_build(oparams(p), xr, yr, zr, cb)
        assert np.array_equal(g["order"], o["order"])
        assert np.array_equal(g["node_begin"], o["node_begin"])
        assert g["nodes"].tobytes() == o["nodes"].tobytes()
        ctx.tree_sort_particles()
        d = ctx.download()
        for k, t in (("x", "tx"), ("y", "ty"), ("z", "tz")):
            assert np.array_equal(d[k], o[t]), k
        # the tree order again is its own partition: rebuilt trees keep it
        ctx.chainmesh_create()
        assert np.array_equal(ctx.chainmesh_perm(), np.arange(n))
        ctx.tree_build()
        assert np.array_equal(ctx.tree_download()["order"], np.arange(n))
        ctx.gravity_long_compute()
        ctx.kick(0, 1.0, 1.0, True, flags=T.TPM_KICK_TREE)
        k1 = ctx.download()
    assert np.array_equal(k1["x"], o["tx"])
