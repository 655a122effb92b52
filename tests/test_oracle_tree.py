"""CPU oracle: the per-cell trees and kick_treepm's tree walk with multipole
(PC) interactions (tree.cpp:7-245, kernels.hpp:13-152,
kick_treepm.cu:226-240, 586-713).

Pinned by: structural invariants of the trees (the ranges partition, the
split planes of the longest-side halving, every particle inside its nodes'
radii, moments against FP64 sums); the PC formula against the pair law
(monopole), against FP64 direct sums (the dipole potential converges at
second order) and against the derivative-tensor form of the reference's
generated kernel; the walk's coverage (with the multipoles summed pair by
pair it reproduces the direct 27-cell sum); an independent numpy
classification of every node the walk meets; the depth limit."""
import math

import numpy as np
import pytest

import oracle as O
from tigerpm_amd.workloads import baseline_params

LEAF, BUCKET, THETA = 32, 128, np.float32(0.4)


def oparams(p, n):
    return O.make_params(p.mesh_dim, p.chain_dim, p.order, p.rs, p.rcut, p.hsoft, p.GM, p.eta, n)


def rel_rms(a, b):
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, -1)) / np.mean(np.sum(b ** 2, -1))))


def clustered(n, seed, nclump=12, width=0.03):
    rng = np.random.default_rng(seed)
    c = rng.random((nclump, 3))
    k = rng.integers(0, nclump, n)
    pos = (c[k] + rng.standard_normal((n, 3)) * width) % 1.0
    return [(pos[:, d] * 2.0 ** 32).astype(np.uint64).astype(np.uint32) for d in range(3)]


def sorted_box(n_side, seed, kind="clustered"):
    p = baseline_params(n_side, n_side, 4)
    n = n_side ** 3
    if kind == "clustered":
        x, y, z = clustered(n, seed)
    else:
        rng = np.random.default_rng(seed)
        x, y, z = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    P = oparams(p, n)
    perm, cb = O.chainmesh_sort(x, y, z, p.chain_dim)
    return p, P, x[perm], y[perm], z[perm], cb


@pytest.fixture(scope="module")
def forest():
    p, P, xs, ys, zs, cb = sorted_box(24, 3)
    return p, P, xs, ys, zs, cb, O.tree_build(P, xs, ys, zs, cb)


def fixed(v):
    return int(math.floor(v * 2.0 ** 32)) & 0xFFFFFFFF


def test_tree_order_and_ranges(forest):
    p, P, xs, ys, zs, cb, t = forest
    n = len(xs)
    order = t["order"].astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(n))
    for c in range(len(cb) - 1):  # particles never leave their cell
        seg = order[cb[c]:cb[c + 1]]
        assert seg.size == 0 or (seg.min() >= cb[c] and seg.max() < cb[c + 1])
    assert np.array_equal(t["tx"], xs[order]) and np.array_equal(t["ty"], ys[order])
    assert np.array_equal(t["tz"], zs[order])


def walk_nodes(nodes, i, box, out):
    """Preorder traversal with each node's box from halving the longest side."""
    out.append((i, box))
    nd = nodes[i]
    if nd["child"][0] >= 0:
        b, e = box
        span = [e[d] - b[d] for d in range(3)]
        dim = int(np.argmax(span))  # first of the longest sides
        mid = (e[dim] + b[dim]) / 2.0
        le, rb = list(e), list(b)
        le[dim] = rb[dim] = mid
        walk_nodes(nodes, nd["child"][0], (b, le), out)
        walk_nodes(nodes, nd["child"][1], (rb, e), out)


def test_tree_structure_and_geometry(forest):
    p, P, xs, ys, zs, cb, t = forest
    C = p.chain_dim
    pos = np.stack([t["tx"], t["ty"], t["tz"]], 1).astype(np.float64) * 2.0 ** -32
    checked_split = 0
    for c in range(C ** 3):
        nodes = t["nodes"][t["node_begin"][c]:t["node_begin"][c + 1]]
        ic = (c % C, (c // C) % C, c // (C * C))
        box = ([ic[d] / C for d in range(3)], [(ic[d] + 1) / C for d in range(3)])
        assert (nodes[0]["begin"], nodes[0]["end"]) == (cb[c], cb[c + 1])
        visited = []
        walk_nodes(nodes, 0, box, visited)
        assert sorted(i for i, _ in visited) == list(range(len(nodes)))  # every node reached once
        for i, (b, e) in visited:
            nd = nodes[i]
            cnt = nd["end"] - nd["begin"]
            assert nd["m"][0] == cnt
            if nd["child"][0] >= 0:
                c0, c1 = nodes[nd["child"][0]], nodes[nd["child"][1]]
                assert cnt > LEAF and nd["child"][0] == i + 1
                assert (c0["begin"], c0["end"], c1["begin"], c1["end"]) == (nd["begin"], c1["begin"], c0["end"],
                                                                            nd["end"])
                span = [e[d] - b[d] for d in range(3)]
                dim = int(np.argmax(span))
                thr = fixed((e[dim] + b[dim]) / 2.0)
                raw = np.stack([t["tx"], t["ty"], t["tz"]], 1)[:, dim]
                assert np.all(raw[c0["begin"]:c0["end"]] < thr) and np.all(raw[c1["begin"]:c1["end"]] >= thr)
                checked_split += 1
            else:
                assert cnt <= LEAF
            ctr = nd["x"].astype(np.float64) * 2.0 ** -32
            if cnt:
                d = pos[nd["begin"]:nd["end"]] - ctr
                d -= np.rint(d)
                dist = np.sqrt((d ** 2).sum(1))
                assert dist.max() <= nd["radius"] - p.hsoft + 1e-6
                if nd["child"][0] < 0:  # a leaf's radius is its farthest particle + hsoft
                    assert dist.max() >= nd["radius"] - p.hsoft - 1e-6
                # moments about the centre: {N, -sum of offsets} (FP64)
                dip = -d.sum(0)
                assert np.abs(nd["m"][1:] - dip).max() <= 1e-5 * cnt * nd["radius"] + 1e-9
            else:
                assert nd["radius"] == np.float32(p.hsoft)
    assert checked_split > 50


def test_sink_buckets(forest):
    p, P, xs, ys, zs, cb, t = forest
    C = p.chain_dim
    pos = np.stack([t["tx"], t["ty"], t["tz"]], 1).astype(np.float64) * 2.0 ** -32
    for c in range(C ** 3):
        bks = t["buckets"][t["bucket_begin"][c]:t["bucket_begin"][c + 1]]
        nodes = t["nodes"][t["node_begin"][c]:t["node_begin"][c + 1]]
        assert len(bks) >= 1
        assert bks[0]["begin"] == cb[c] and bks[-1]["end"] == cb[c + 1]
        assert np.all(bks["begin"][1:] == bks["end"][:-1])
        parent_count = {0: None}
        for i, nd in enumerate(nodes):
            if nd["child"][0] >= 0:
                parent_count[nd["child"][0]] = parent_count[nd["child"][1]] = nd["end"] - nd["begin"]
        ranges = {}
        for i, nd in enumerate(nodes):
            ranges.setdefault((nd["begin"], nd["end"]), []).append(i)
        for bk in bks:
            # a bucket is a node of at most 128 sinks, the first such on its path
            assert bk["end"] - bk["begin"] <= BUCKET
            assert any(parent_count[i] is None or parent_count[i] > BUCKET for i in ranges[(bk["begin"], bk["end"])])
            if bk["end"] > bk["begin"]:
                d = pos[bk["begin"]:bk["end"]] - bk["x"].astype(np.float64) * 2.0 ** -32
                d -= np.rint(d)
                assert np.sqrt((d ** 2).sum(1)).max() <= bk["radius"] - p.hsoft + 1e-6


def test_pc_monopole_is_the_pair_law():
    p = baseline_params(32, 32, 4)
    P = oparams(p, 32 ** 3)
    rng = np.random.default_rng(4)
    for _ in range(40):
        r = rng.uniform(3 * p.hsoft, 0.98 * p.rcut)
        v = rng.standard_normal(3)
        v *= r / np.linalg.norm(v)
        m = float(rng.integers(1, 40))
        pc = O.pc_pair(P, *v, [m, 0, 0, 0]).astype(np.float64)
        exact = O.pp_pair(P, *v, True) * m
        assert np.abs(pc - exact).max() <= 2e-5 * np.abs(exact).max()


def pair_fp64(P, d):
    return O.pp_pair(P, float(d[0]), float(d[1]), float(d[2]), True)


def test_pc_dipole_potential_is_second_order():
    """The potential of a cluster from its monopole and dipole about a centre
    off the cluster's mean: the error vanishes as the square of the size."""
    p = baseline_params(32, 32, 4)
    P = oparams(p, 32 ** 3)
    rng = np.random.default_rng(6)
    R = np.array([0.09, -0.04, 0.05])
    delta = rng.standard_normal((5, 3)) * 0.02
    cen = rng.standard_normal(3) * 0.01
    errs = []
    for s in (1 / 8, 1 / 16, 1 / 32):
        off = (delta - cen) * s
        M = np.concatenate([[len(off)], -off.sum(0)])
        pc = O.pc_pair(P, *R, M)
        exact = sum(pair_fp64(P, R - o) for o in off)
        errs.append(abs(pc[3] - exact[3]) / abs(exact[3]))
    ratios = [errs[k] / errs[k + 1] for k in range(2)]
    assert all(3.3 < q < 4.7 for q in ratios), (errs, ratios)


def reference_tensor_pc(P, R, M, inv2rs):
    """compute_pc_interaction in derivative-tensor form, FP64: f(r) =
    -erfc(a r)/r, D_i = n_i d1, D_ij = n_i n_j d2 + delta_ij d1/r with d0,
    d1, d2 as kernels.hpp:13-32 defines them (d2 = -3 d1/r + e0 c0 with
    e0 = (2/sqrt(pi)) exp(-a^2 r^2) a / r and c0 = -2 r a)."""
    a = inv2rs
    r = float(np.linalg.norm(R))
    n = R / r
    tosp = 2.0 / math.sqrt(math.pi)
    ex = math.exp(-(a * r) ** 2)
    d0 = -math.erfc(a * r) / r
    e0 = tosp * ex * a / r
    d1 = -d0 / r + e0
    d2 = -3.0 * d1 / r + e0 * (-2.0 * r * a)
    D1 = n * d1
    D2 = np.outer(n, n) * d2 + np.eye(3) * d1 / r
    phi = M[0] * d0 + np.dot(M[1:], D1)
    g = -(M[0] * D1 + D2 @ M[1:])
    return np.concatenate([g, [phi]])


def test_pc_matches_the_reference_derivative_tensors():
    p = baseline_params(32, 32, 4)
    P = oparams(p, 32 ** 3)
    inv2rs = float(np.float32(1.0) / np.float32(p.rs) / np.float32(2.0))
    rng = np.random.default_rng(8)
    for _ in range(30):
        R = rng.standard_normal(3)
        R *= rng.uniform(0.03, 0.17) / np.linalg.norm(R)
        R = R.astype(np.float32).astype(np.float64)
        M = np.concatenate([[rng.integers(1, 30)], rng.standard_normal(3) * 0.01]).astype(np.float32)
        got = O.pc_pair(P, *R, M).astype(np.float64)
        want = reference_tensor_pc(P, R, M.astype(np.float64), inv2rs)
        assert np.abs(got - want).max() <= 3e-5 * np.abs(want).max()


def test_reference_second_derivative_differs_from_the_kernel():
    """Records a property of the restated formula: the generated d2 of
    kernels.hpp:24-30 is not f'' - f'/r of f = -erfc(a r)/r (its Gaussian
    term carries a^2 where the derivative has a^3), so the dipole's force
    terms do not converge like the potential's."""
    a, r = 12.8, 0.11
    f = lambda q: -math.erfc(a * q) / q
    h = 1e-5
    f1 = (f(r + h) - f(r - h)) / (2 * h)
    f2 = (f(r + h) - 2 * f(r) + f(r - h)) / h ** 2
    tosp = 2 / math.sqrt(math.pi)
    ex = math.exp(-(a * r) ** 2)
    d1 = math.erfc(a * r) / r ** 2 + tosp * ex * a / r
    assert abs(d1 - f1) < 1e-6 * abs(f1)
    truth = f2 - f1 / r
    generated = -3 * d1 / r - 2 * tosp * ex * a ** 2
    assert abs(-3 * d1 / r - 2 * tosp * ex * a ** 3 - truth) < 1e-4 * abs(truth)
    assert abs(generated - truth) > 0.3 * abs(truth)


@pytest.mark.parametrize("kind", ["uniform", "clustered"])
def test_walk_covers_every_pair_once(kind):
    """With the multipole-list nodes summed pair by pair, the walk gives the
    direct 27-cell sum: no node is dropped inside rcut or met twice."""
    p, P, xs, ys, zs, cb = sorted_box(24, 11, kind)
    n = len(xs)
    rung = np.random.default_rng(2).integers(0, 3, n).astype(np.int8)
    (gx, gy, gz, ph), st = O.tree_sum(P, xs, ys, zs, cb, rung, 1, pc_exact=True)
    act = rung >= 1
    (dx, dy, dz, dp), _ = O.pp_sum(P, xs, ys, zs, cb, np.flatnonzero(act), False)
    G = np.stack([gx, gy, gz], 1)
    assert np.all(G[~act] == 0) and np.all(ph[~act] == 0)
    assert rel_rms(G[act], np.stack([dx, dy, dz], 1)) < 2e-6
    assert rel_rms(ph[act], dp) < 2e-6
    assert st["multi"] > 0 and st["part"] > 0


def test_walk_multipole_error_is_small():
    p, P, xs, ys, zs, cb = sorted_box(32, 5, "uniform")
    rung = np.zeros(len(xs), np.int8)
    (gx, gy, gz, ph), st = O.tree_sum(P, xs, ys, zs, cb, rung)
    (ex, ey, ez, ep), _ = O.tree_sum(P, xs, ys, zs, cb, rung, pc_exact=True)
    assert st["pc"] > 0.01 * st["pp"]
    assert rel_rms(np.stack([gx, gy, gz], 1), np.stack([ex, ey, ez], 1)) < 5e-3
    assert rel_rms(ph, ep) < 2e-3


def classify(P, p, t, rung, min_rung):
    """numpy restatement of the walk's decisions (kick_treepm.cu:586-682) for
    every bucket and neighbour tree: returns (multipole nodes, PP leaves,
    PC interactions, PP pairs) summed over the box."""
    C = p.chain_dim
    f32 = np.float32
    rcut, t2 = f32(p.rcut), f32(1.0) / (THETA * THETA)
    tp = np.stack([t["tx"], t["ty"], t["tz"]], 1)
    act = rung[t["order"]] >= min_rung

    def dist(a, b):
        return ((a.astype(np.int64) - b.astype(np.int64) + 2 ** 31) % 2 ** 32 - 2 ** 31).astype(f32) * f32(2.0 ** -32)

    def fma(a, b, c):  # FP32 fma: the product is exact in FP64, then one rounding to FP32
        return (a.astype(np.float64) * b + c).astype(f32)

    def r2(d):
        return fma(d[..., 0], d[..., 0], fma(d[..., 1], d[..., 1], d[..., 2] * d[..., 2]))

    nm = npart = npc = npp = 0
    for c in range(C ** 3):
        cx, cy, cz = c % C, (c // C) % C, c // (C * C)
        for bk in t["buckets"][t["bucket_begin"][c]:t["bucket_begin"][c + 1]]:
            sinks = np.arange(bk["begin"], bk["end"])
            sinks = sinks[act[sinks]]
            if len(sinks) == 0:
                continue
            for ox in (-1, 0, 1):
                for oy in (-1, 0, 1):
                    for oz in (-1, 0, 1):
                        tc = (((cz + oz) % C) * C + (cy + oy) % C) * C + (cx + ox) % C
                        nodes = t["nodes"][t["node_begin"][tc]:t["node_begin"][tc + 1]]
                        stack = [0]
                        while stack:
                            nd = nodes[stack.pop()]
                            d = dist(bk["x"], nd["x"])
                            R2 = r2(d)
                            R = np.sqrt(R2)
                            if R - nd["radius"] - bk["radius"] > rcut:
                                continue
                            rr = nd["radius"] + bk["radius"]
                            cnt = int(nd["end"] - nd["begin"])
                            if R2 > rr * rr * t2:
                                nm += 1
                                npc += len(sinks)
                            elif cnt <= LEAF:
                                near = ~(r2(dist(tp[sinks], nd["x"])) > nd["radius"] * nd["radius"] * t2)
                                if near.any():
                                    npart += 1
                                    npp += len(sinks) * cnt
                                else:
                                    nm += 1
                                    npc += len(sinks)
                            else:
                                stack += [int(nd["child"][0]), int(nd["child"][1])]
    return nm, npart, npc, npp


def test_walk_decisions_against_numpy():
    p, P, xs, ys, zs, cb = sorted_box(16, 21)
    rung = np.random.default_rng(9).integers(0, 2, len(xs)).astype(np.int8)
    t = O.tree_build(P, xs, ys, zs, cb)
    _, st = O.tree_sum(P, xs, ys, zs, cb, rung, 1)
    nm, npart, npc, npp = classify(P, p, t, rung, 1)
    assert (st["multi"], st["part"], st["pc"], st["pp"]) == (nm, npart, npc, npp)
    assert nm > 0 and npart > 0


def test_depth_limit_is_an_error():
    """40 particles at one point cannot be split into leaves of 32: the
    reference aborts past depth 50 (tree.cpp:16-19), the oracle raises."""
    p = baseline_params(16, 16, 4)
    n = 16 ** 3
    rng = np.random.default_rng(1)
    x, y, z = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    x[:40], y[:40], z[:40] = 123456789, 987654321, 555555555
    P = oparams(p, n)
    perm, cb = O.chainmesh_sort(x, y, z, p.chain_dim)
    with pytest.raises(RuntimeError):
        O.tree_build(P, x[perm], y[perm], z[perm], cb)


def test_tree_leaf_order_hand_worked():
    """The particle order a split leaves, worked out by hand from the
    two-pointer particles_sort (particles.cpp:414-445) that tree.cpp:101-107
    calls.  One chain cell [0, 1/3)^3 (C = 3), 34 particles, so the root
    splits once, on x (the first of three equal sides) at x = 1/6:
    particles 0 and 33 lie right of the plane, 1..32 left.  lo meets 0 at
    position 0 and hi, walking down from 34, finds 32 (33 is right): they
    trade places; lo then walks to 32 = hi and stops.  Order:
    [32, 1, 2, ..., 31 | 0, 33] (a stable partition would give
    [1, ..., 32 | 0, 33])."""
    C = 3
    n = 34
    left_x, right_x = fixed(0.05), fixed(0.25)
    rng = np.random.default_rng(5)
    x = np.full(n, left_x, np.uint32)
    x[0] = x[33] = right_x
    x[1:33] += np.arange(32, dtype=np.uint32) * 1000  # distinct positions
    y = (rng.random(n) * 0.3 * 2 ** 32).astype(np.uint32)
    z = (rng.random(n) * 0.3 * 2 ** 32).astype(np.uint32)
    cb = np.zeros(C ** 3 + 1, np.int64)
    cb[1:] = n
    p = baseline_params(C * 8, C * 8, 4)
    P = O.make_params(p.mesh_dim, C, p.order, p.rs, 0.3, p.hsoft, p.GM, p.eta, n)
    t = O.tree_build(P, x, y, z, cb)
    assert list(t["order"]) == [32] + list(range(1, 32)) + [0, 33]
    root = t["nodes"][0]
    assert (root["begin"], root["end"]) == (0, 34)
    kids = [t["nodes"][c] for c in root["child"]]
    assert [(k["begin"], k["end"]) for k in kids] == [(0, 32), (32, 34)]
