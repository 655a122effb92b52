"""GPU tests of the drop-in boundary: the reference partition's particle
order (TPM_ORDER_REFERENCE), the FP64 momentum / energy reductions and the
Ewald sum at arbitrary positions.  All calls go through the C ABI."""
import numpy as np
import pytest

import oracle as O
import tigerpm_amd as T

pytestmark = pytest.mark.gpu


def uniform(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]


def clustered(n, seed, nclump=20):
    rng = np.random.default_rng(seed)
    m = n // 2
    pos = [rng.random(n - m) for _ in range(3)]
    centres = rng.random((nclump, 3))
    which = rng.integers(0, nclump, m)
    for d in range(3):
        c = (centres[which, d] + 0.01 * rng.standard_normal(m)) % 1.0
        pos[d] = np.concatenate([pos[d], c])
    return [np.minimum(np.floor(p * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint32) for p in pos]


def plane_edges(n, C, seed):
    """Uniform particles, a third of them exactly on, one below or one above
    a split plane fixed32(k / C) on some axis."""
    x, y, z = uniform(n, seed)
    b = np.empty(C + 1, np.uint64)
    O.lib().orc_chain_bounds(C, b)
    rng = np.random.default_rng(seed + 1)
    sel = rng.choice(n, n // 3, replace=False)
    k = rng.integers(0, C, len(sel))
    off = rng.integers(-1, 2, len(sel))
    v = ((b[k].astype(np.int64) + off) % 2 ** 32).astype(np.uint32)
    for a, part in zip((x, y, z), np.array_split(sel, 3)):
        a[part] = v[:len(part)]
        v = v[len(part):]
    return x, y, z


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def small_params(ctx, n_side, C):
    p = T.baseline_params(n_side, n_side, 3)
    p.chain_dim = C
    p.rcut = min(p.rcut, 0.99 / C)
    return ctx.set_params(p)


@pytest.mark.parametrize("dist,n_side,C", [("uniform", 24, 5), ("clustered", 32, 7), ("edges", 20, 6),
                                           ("uniform", 40, 11)])
def test_reference_partition_order_bit_exact(dist, n_side, C):
    """After tpm_chainmesh_create with TPM_ORDER_REFERENCE the caller sees
    the particles exactly where chainmesh.cpp's recursive two-pointer
    partition leaves them, and the permutation is the reference's."""
    n = n_side ** 3
    x, y, z = {"uniform": lambda: uniform(n, 31), "clustered": lambda: clustered(n, 32),
               "edges": lambda: plane_edges(n, C, 33)}[dist]()
    rng = np.random.default_rng(34)
    vx = rng.standard_normal(n).astype(np.float32)
    rung = rng.integers(0, 5, n).astype(np.int8)
    with T.Context(0) as ctx:
        small_params(ctx, n_side, C)
        ctx.set_host_order(T.TPM_ORDER_REFERENCE)
        ctx.upload(x, y, z, vx=vx, rung=rung)
        ctx.chainmesh_create()
        idx, rb, re_ = O.chainmesh_partition_ref(x, y, z, C)
        assert np.array_equal(ctx.chainmesh_perm(), idx)
        d = ctx.download()
        for k, a in (("x", x), ("y", y), ("z", z), ("vx", vx), ("rung", rung)):
            assert np.array_equal(d[k], a[idx]), k
        cb = ctx.chainmesh_cells()
        assert np.array_equal(cb[:-1], rb) and np.array_equal(cb[1:], re_)
        # a second build partitions the caller's (reference) order again:
        # already partitioned, so nothing moves
        ctx.chainmesh_create()
        assert np.array_equal(ctx.chainmesh_perm(), np.arange(n))
        d2 = ctx.download()
        assert np.array_equal(d2["x"], d["x"]) and np.array_equal(d2["z"], d["z"])


def test_reference_order_after_a_drift_and_kick(gpu_ctx):
    """The reference partition of a moved particle set starts from the
    caller's previous (reference) order, and the kick's outputs come back in
    the caller's order: equal to the same kick in device order mapped by the
    permutations."""
    n_side, C = 24, 5
    n = n_side ** 3
    x, y, z = clustered(n, 41)
    with T.Context(0) as ctx:
        p = small_params(ctx, n_side, C)
        ctx.set_host_order(T.TPM_ORDER_REFERENCE)
        ctx.upload(x, y, z)
        ctx.chainmesh_create()
        a = ctx.download()
        # move every particle a little (as a drift would) without an upload:
        # re-upload in the caller's order, then build twice
        rng = np.random.default_rng(42)
        step = rng.integers(-2 ** 26, 2 ** 26, (3, n))
        xs = [((a[k].astype(np.int64) + step[i]) % 2 ** 32).astype(np.uint32) for i, k in enumerate("xyz")]
        ctx.upload(*xs)
        ctx.chainmesh_create()
        idx, _, _ = O.chainmesh_partition_ref(*xs, C)
        b = ctx.download()
        assert np.array_equal(b["x"], xs[0][idx]) and np.array_equal(b["y"], xs[1][idx])
        ctx.gravity_long_compute()
        ctx.kick(0, 1.0, 1.0, True)
        g_ref = ctx.download()
    # the same particles in device order
    gpu_ctx.set_params(p)
    gpu_ctx.set_host_order(T.TPM_ORDER_DEVICE)
    gpu_ctx.upload(*xs)
    gpu_ctx.chainmesh_create()
    dperm = gpu_ctx.chainmesh_perm()
    gpu_ctx.gravity_long_compute()
    gpu_ctx.kick(0, 1.0, 1.0, True)
    g_dev = gpu_ctx.download()
    inv = np.empty(n, np.int64)
    inv[dperm] = np.arange(n)
    where = inv[idx]  # device position of reference position r
    for k in ("x", "gx", "gy", "gz", "pot", "vx", "rung"):
        assert np.array_equal(g_ref[k], g_dev[k][where]), k


def test_ewald_samples_follow_the_callers_order():
    n_side, C = 16, 3
    n = n_side ** 3
    x, y, z = uniform(n, 51)
    with T.Context(0) as ctx:
        p = small_params(ctx, n_side, C)
        ctx.set_host_order(T.TPM_ORDER_REFERENCE)
        ctx.upload(x, y, z)
        ctx.chainmesh_create()
        d = ctx.download()
        s = np.array([0, 5, 77, n - 1], np.int64)
        got = ctx.ewald(s)
    ref = O.ewald_sf(d["x"], d["y"], d["z"], s, p.hsoft, 9.6)
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()


@pytest.mark.parametrize("min_rung", [0, 2])
def test_moments_fp64(gpu_ctx, min_rung):
    """tpm_stats.moments after a kick equal numpy's FP64 sums of the
    downloaded arrays (<= 1e-12 of the scale), are bit-identical on a second
    evaluation, and Newton's third law makes the pair forces' sum vanish."""
    n_side = 32
    n = n_side ** 3
    x, y, z = clustered(n, 61)
    rng = np.random.default_rng(62)
    v = [rng.standard_normal(n).astype(np.float32) * 1e-3 for _ in range(3)]
    rung = rng.integers(0, 4, n).astype(np.int8)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
    gpu_ctx.set_host_order(T.TPM_ORDER_DEVICE)
    gpu_ctx.upload(x, y, z, *v, rung=rung)
    gpu_ctx.chainmesh_create()
    gpu_ctx.gravity_long_compute()
    st = gpu_ctx.kick(min_rung, 1.0, 1e-2, False)
    d = gpu_ctx.download()
    m = p.GM
    vv = np.stack([d["vx"], d["vy"], d["vz"]]).astype(np.float64)
    gg = np.stack([d["gx"], d["gy"], d["gz"]]).astype(np.float64)
    mom = st.moments.as_dict()
    want = dict(momentum=m * vv.sum(1), force=m * gg.sum(1), kinetic=0.5 * m * (vv ** 2).sum(),
                potential=0.5 * m * d["pot"].astype(np.float64).sum())
    scale_v = m * np.abs(vv).sum()
    scale_g = m * np.abs(gg).sum()
    assert np.all(np.abs(np.array(mom["momentum"]) - want["momentum"]) <= 1e-12 * scale_v)
    assert np.all(np.abs(np.array(mom["force"]) - want["force"]) <= 1e-12 * scale_g)
    assert abs(mom["kinetic"] - want["kinetic"]) <= 1e-12 * want["kinetic"]
    assert abs(mom["potential"] - want["potential"]) <= 1e-12 * abs(want["potential"])
    again = gpu_ctx.moments()
    assert again == mom  # same particles, same order: same bits
    # pair forces alone obey Newton's third law
    gpu_ctx.kick(0, 1.0, 1e-2, True, flags=T.TPM_SKIP_PM | T.TPM_SKIP_KICK)
    pp = gpu_ctx.moments()
    gp = np.stack([gpu_ctx.download()[k] for k in ("gx", "gy", "gz")]).astype(np.float64)
    assert np.abs(pp["force"]).max() <= 1e-6 * m * np.abs(gp).sum()


def test_batch_moments_equal_the_box_alone(gpu_ctx):
    """Per-box moments of a batch equal the same box's moments kicked alone."""
    n_side = 24
    p = T.baseline_params(n_side, n_side, 4)
    with T.Batch(3 * n_side ** 3, 3 * n_side ** 3, 4, 4) as bt:
        ids = [bt.add_zeldovich(T.box_spec(p), n_side, 70 + k, 0.8) for k in range(3)]
        bt.step()
        for i in ids:
            mb = bt.moments(i)
            bt.to_ctx(i, gpu_ctx)
            gpu_ctx.set_host_order(T.TPM_ORDER_DEVICE)
            assert gpu_ctx.moments() == mb
            d = bt.download(i)
            k = 0.5 * p.GM * sum((d[c].astype(np.float64) ** 2).sum() for c in ("vx", "vy", "vz"))
            assert abs(mb["kinetic"] - k) <= 1e-12 * k


def test_ewald_at_positions(gpu_ctx):
    """tpm_ewald_at against the oracle's structure-factor sum at arbitrary
    positions, including positions on particles and inside the softening."""
    n_side = 24
    n = n_side ** 3
    x, y, z = uniform(n, 81)
    p = gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
    gpu_ctx.set_host_order(T.TPM_ORDER_DEVICE)
    gpu_ctx.upload(x, y, z)
    sx, sy, sz = uniform(200, 82)
    sx[:4], sy[:4], sz[:4] = x[:4], y[:4], z[:4]
    sx[4] = x[10] + 1000
    sy[4], sz[4] = y[10], z[10]
    got = gpu_ctx.ewald_at(sx, sy, sz)
    ref = O.ewald_sf_at(x, y, z, sx, sy, sz, p.hsoft, 9.6)
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    # on a particle: the per-particle sum plus its softened self potential
    own = gpu_ctx.ewald(np.arange(4))
    # (FP64 sums accumulated by atomics: their order varies, so 1e-10)
    assert np.allclose(got[0][:4], own[0] - 35.0 / 16.0 / p.hsoft, rtol=1e-10, atol=0)


def _kick_state(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = clustered(n, seed)
    v = [rng.standard_normal(n).astype(np.float32) * 1e-3 for _ in range(3)]
    rung = rng.integers(0, 4, n).astype(np.int8)
    return (x, y, z), v, rung


def _fresh(ctx, p, pos, v, rung):
    ctx.set_params(p)
    ctx.set_host_order(T.TPM_ORDER_DEVICE)
    ctx.set_budget(0)
    ctx.upload(*pos, *v, rung=rung)
    ctx.chainmesh_create()
    ctx.gravity_long_compute()


KEYS = ("x", "y", "z", "vx", "vy", "vz", "rung", "gx", "gy", "gz", "pot")


@pytest.mark.parametrize("flags", [0, T.TPM_KICK_PME, T.TPM_KICK_TREE])
def test_kick_box_equals_the_whole_kick_on_its_sinks(gpu_ctx, flags):
    """kick_pme(range<int> box) / kick_treepm's box: the sinks of the box get
    the bits of the whole kick, the others are untouched."""
    n_side = 32
    n = n_side ** 3
    pos, v, rung = _kick_state(n, 91)
    p = T.baseline_params(n_side, n_side, 3)
    _fresh(gpu_ctx, p, pos, v, rung)
    full_st = gpu_ctx.kick(1, 1.0, 1e-2, False, flags=flags)
    full = gpu_ctx.download()
    cb = gpu_ctx.chainmesh_cells()
    C = gpu_ctx.params.chain_dim
    _fresh(gpu_ctx, p, pos, v, rung)
    before = gpu_ctx.download()
    lo, hi = (1, 0, 2), (4, C, 3)
    st = gpu_ctx.kick_box(lo, hi, 1, 1.0, 1e-2, False, flags=flags)
    got = gpu_ctx.download()
    inbox = np.zeros(n, bool)
    for zc in range(lo[2], hi[2]):
        for yc in range(lo[1], hi[1]):
            for xc in range(lo[0], hi[0]):
                c = (zc * C + yc) * C + xc
                inbox[cb[c]:cb[c + 1]] = True
    assert inbox.any() and not inbox.all()
    for k in KEYS:
        assert np.array_equal(got[k][inbox], full[k][inbox]), k
        assert np.array_equal(got[k][~inbox], before[k][~inbox]), k
    assert 0 < st.candidate_pairs < full_st.candidate_pairs
    assert 0 < st.executed_pairs < full_st.executed_pairs


@pytest.mark.parametrize("flags", [0, T.TPM_KICK_PME, T.TPM_KICK_TREE])
def test_budget_split_is_bit_identical(gpu_ctx, flags):
    """A device budget far below the whole box's working set splits the kick
    into sub-boxes (kick_treepm.cu:815-823): same g, phi, v and rung bits,
    same pair counts and maximum rung."""
    n_side = 32
    n = n_side ** 3
    pos, v, rung = _kick_state(n, 92)
    p = T.baseline_params(n_side, n_side, 3)
    _fresh(gpu_ctx, p, pos, v, rung)
    st0 = gpu_ctx.kick(0, 1.0, 1e-2, False, flags=flags)
    ref = gpu_ctx.download()
    for budget in (2_000_000, 400_000):
        _fresh(gpu_ctx, p, pos, v, rung)
        gpu_ctx.set_budget(budget)
        st = gpu_ctx.kick(0, 1.0, 1e-2, False, flags=flags)
        got = gpu_ctx.download()
        for k in KEYS:
            assert np.array_equal(got[k], ref[k]), (budget, k)
        assert st.executed_pairs == st0.executed_pairs and st.candidate_pairs == st0.candidate_pairs
        assert st.max_rung == st0.max_rung == int(got["rung"].max())
        assert st.pc_interactions == st0.pc_interactions
    gpu_ctx.set_budget(0)


def test_budget_below_one_cell_fails(gpu_ctx):
    n_side = 24
    pos, v, rung = _kick_state(n_side ** 3, 93)
    _fresh(gpu_ctx, T.baseline_params(n_side, n_side, 3), pos, v, rung)
    gpu_ctx.set_budget(1000)
    with pytest.raises(T.TpmError) as e:
        gpu_ctx.kick(0, 1.0, 1e-2, True)
    assert e.value.status == -5
    gpu_ctx.set_budget(0)
    gpu_ctx.kick(0, 1.0, 1e-2, True)
