"""The reference's --test= scenarios (test.cpp:120-346) run through the C++
layer with the reference's entry-point names (libtigerpm_ref.so), on the GPU."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tigerpm_amd", "bin", "tigerpm")


def run(*args):
    r = subprocess.run([DRIVER] + list(args), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = dict(re.findall(r"^(Force RMS Error|Force RMS Vector|Force Max Error)\s*= (\S+)", r.stdout, re.M))
    return r.stdout, {k: float(v) for k, v in vals.items()}


def test_treepm_scenario_baseline_parameters():
    # 64^3 uniform particles, 64^3 mesh, rs = 1.25 cells, rcut = 4.5 rs, TSC, m = 1/N
    out, e = run("--test=treepm", "--parts_dim=64", "--four_dim=64", "--rs_cells=1.25", "--rcut_rs=4.5",
                 "--order=3", "--mass=1/N")
    assert "KICK" in out and "COMPARE" in out
    assert e["Force RMS Vector"] < 1e-3 and e["Force RMS Error"] < 1e-3


def test_treepm_scenario_reference_defaults():
    # reference defaults (options.cpp:81-96): rs = 1 mesh cell, rcut = 1/chain_dim, cloud4, unit masses
    out, e = run("--test=treepm", "--parts_dim=90")
    assert e["Force RMS Error"] < 1e-2


def test_treepm_scenario_tree_walk():
    # --short=tree: the reference's own short-range algorithm (per-cell trees,
    # theta = 0.4 monopole + dipole, PP for near leaves) at its defaults
    out, e = run("--test=treepm", "--parts_dim=90", "--short=tree")
    m = re.search(r"multipole interactions (\S+)", out)
    assert m and float(m.group(1)) > 0
    _, d = run("--test=treepm", "--parts_dim=90")
    # the multipoles add their approximation error to the direct sum's
    assert d["Force RMS Error"] < e["Force RMS Error"] < 5e-2


def test_kick_pm_scenario():
    # kick_pm_test (test.cpp:120-143): the PM force alone against Ewald
    out, e = run("--test=kick_pm", "--parts_dim=60")
    # at the reference defaults (rs = 1 cell of a 20^3 mesh for 60^3 random
    # particles) the mesh force is a small part of the total; the
    # interpolator itself is checked in test_ref_layer_force_at
    assert "Force RMS Error" in out and 0.0 < e["Force RMS Error"] < 1.0


def test_kick_pme_scenario():
    # kick_pme_test (test.cpp:182-235): random particles, kick_pme_begin;
    # the law itself is checked against the oracle in test_ref_layer_kick_pme
    out, e = run("--test=kick_pme", "--parts_dim=60")
    assert "KICK" in out and "COMPARE" in out
    assert 0.0 < e["Force RMS Error"] < 0.5


# ---------------------------------------------------------------------------
# The C++ layer called as reference code calls it (tests/ref_layer_check.cpp,
# built here with g++ against libtigerpm_ref.so), its global arrays checked
# against the oracle.
import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
import tigerpm_amd as T  # noqa: E402


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("refcheck")
    exe = str(d / "ref_layer_check")
    lib = os.path.join(ROOT, "tigerpm_amd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "ref_layer_check.cpp"), "-o", exe, "-L" + lib, "-ltigerpm_ref",
                    "-ltigerpm_hip", "-Wl,-rpath," + lib], check=True)

    def go(mode, parts_dim):
        out = d / mode
        out.mkdir(exist_ok=True)
        r = subprocess.run([exe, mode, str(out), str(parts_dim)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr

        def load(name, dt):
            return np.fromfile(str(out / name), dt)
        return load
    return go


def particles(load, tag):
    d = {}
    for k in range(3):
        d["x%d" % k] = load("%s_x%d" % (tag, k), np.uint32)
        d["u%d" % k] = load("%s_u%d" % (tag, k), np.float32)
        d["g%d" % k] = load("%s_g%d" % (tag, k), np.float32)
    d["p"] = load(tag + "_p", np.float32)
    d["r"] = load(tag + "_r", np.int8)
    return d


def key(d, prefix="x"):
    return (d[prefix + "0"].astype(np.uint64) << np.uint64(32)) | d[prefix + "1"].astype(np.uint64)


def match(a_keys, b_keys):
    """Index into b of every entry of a (positions are distinct)."""
    ob = np.argsort(b_keys, kind="stable")
    idx = np.searchsorted(b_keys[ob], a_keys)
    out = ob[idx]
    assert np.array_equal(b_keys[out], a_keys)
    return out


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def oracle_params(n_side, pme=0):
    p = T.baseline_params(n_side, n_side, 3)
    return O.make_params(p.mesh_dim, p.chain_dim, p.order, p.rs, p.rcut, p.hsoft, p.GM, p.eta, p.n_total, 1, 0,
                         pme), p


def test_ref_layer_force_at(checker):
    """gravity_long_force_at (gravity_long.cpp:52-125): at a mesh node the
    4-point cubic weights are (0, 1, 0, 0) and their derivatives
    (-1/2, 0, 1/2, 0), so phi is the node value and g the central
    difference -M (phi[i+1] - phi[i-1]) / 2 (closed forms); between nodes it
    agrees with the 6-point quintic of the kick kernels on the same mesh to
    the interpolation error of a mesh smoothed on the scale rs; kick_pm
    stores it at every particle."""
    M = 32
    load = checker("force_at", M)
    phi = load("phi", np.float32).reshape(M, M, M)
    q = load("query", np.float64).reshape(-1, 3)
    f = load("force_at", np.float64).reshape(-1, 4)
    nodes = np.rint(q[:200] * M).astype(np.int64) % M
    i, j, k = nodes.T
    assert np.array_equal(f[:200, 0], phi[i, j, k].astype(np.float64))
    ph = phi.astype(np.float64)
    for d in range(3):
        up, dn = [i, j, k], [i, j, k]
        up[d] = (up[d] + 1) % M
        dn[d] = (dn[d] - 1) % M
        want = -(ph[tuple(up)] - ph[tuple(dn)]) * M / 2
        assert np.allclose(f[:200, 1 + d], want, rtol=1e-6, atol=1e-6 * np.abs(want).max())
    raw = [np.minimum(np.floor(q[200:, d] * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint32) for d in range(3)]
    g6 = O.interp6(M, phi.ravel(), *raw)
    g4 = f[200:, 1:].T
    err = rel(g4, np.stack(g6[:3]))
    print("force_at 4-point vs 6-point interpolation: rel rms %.3e" % err)
    assert err < 2e-2
    assert rel(f[200:, 0], g6[3]) < 1e-3
    pm = particles(load, "pm")
    gp = O.interp6(M, phi.ravel(), pm["x0"], pm["x1"], pm["x2"])
    assert rel(np.stack([pm["g0"], pm["g1"], pm["g2"]]), np.stack(gp[:3])) < 2e-2


def test_ref_layer_kick_pme(checker):
    """chainmesh_create leaves the reference partition's order (a second
    partition changes nothing); kick_pme(box) kicks exactly the box's sinks
    with kick_pme's law (the oracle's, <= 1e-5); the whole box split by a
    small device budget gives the same bits."""
    n_side = 32
    load = checker("kick_pme", n_side)
    s = particles(load, "sorted")
    C = int(np.floor(1.0 / (4.5 * 1.25 / n_side)))
    idx, cb, ce = O.chainmesh_partition_ref(s["x0"], s["x1"], s["x2"], C)
    assert np.array_equal(idx, np.arange(len(idx)))
    sub, spl = particles(load, "sub"), particles(load, "split")
    inbox = np.zeros(len(idx), bool)
    for zc in range(2, 3):
        for yc in range(0, C):
            for xc in range(1, C - 1):
                c = (zc * C + yc) * C + xc
                inbox[cb[c]:ce[c]] = True
    for k in ("g0", "g1", "g2", "p"):
        assert np.array_equal(sub[k][inbox], spl[k][inbox]), k
        assert np.all(sub[k][~inbox] == 0), k
    for k in ("u0", "r"):
        assert np.array_equal(sub[k][~inbox], s[k][~inbox])
    P, p = oracle_params(n_side, pme=1)
    o = O.kick_full(P, s["x0"], s["x1"], s["x2"], s["u0"], s["u1"], s["u2"], s["r"], min_rung=1, scale=1.0, t0=1e-2,
                    first_call=False)
    w = match(key(s), (o["x"].astype(np.uint64) << np.uint64(32)) | o["y"].astype(np.uint64))
    act = s["r"] >= 1  # kicked sinks (min_rung 1); the oracle leaves no output for the others
    assert act.any() and not act.all()
    G = np.stack([spl["g0"], spl["g1"], spl["g2"]])[:, act]
    OG = np.stack([o["gx"][w], o["gy"][w], o["gz"][w]])[:, act]
    assert rel(G, OG) < 1e-5
    assert rel(spl["p"][act], o["pot"][w][act]) < 1e-5
    assert np.all(spl["g0"][~act] == 0) and np.array_equal(spl["u0"][~act], s["u0"][~act])
    # the first kick's velocities: box sinks from the sub-box kick, the rest from the split one
    V = np.where(inbox, sub["u0"], spl["u0"])
    assert rel(V[act], o["vx"][w][act]) < 1e-5


def test_ref_layer_tree_create_and_kick_treepm(checker):
    """tree_create leaves each cell's particles in the order the reference's
    two-pointer partitions make of the reference chain-mesh order (the
    oracle's trees of that order, bit for bit, with the same node and bucket
    counts); kick_treepm(trees, buckets, box) keeps that order and its forces
    are the oracle's tree walk's (<= 1e-5)."""
    n_side = 24
    load = checker("trees", n_side)
    s = particles(load, "sorted")
    P, p = oracle_params(n_side)
    C = p.chain_dim
    idx, cb, _ = O.chainmesh_partition_ref(s["x0"], s["x1"], s["x2"], C)
    assert np.array_equal(idx, np.arange(len(idx)))
    cbeg = np.concatenate([cb, [len(idx)]]).astype(np.int64)
    t = O.tree_build(P, s["x0"], s["x1"], s["x2"], cbeg)
    tr = particles(load, "treeorder")
    for d, k in enumerate(("tx", "ty", "tz")):
        assert np.array_equal(tr["x%d" % d], t[k])
    assert np.array_equal(tr["u0"], s["u0"][t["order"]]) and np.array_equal(tr["r"], s["r"][t["order"]])
    counts = load("tree_counts", np.int32).reshape(C, C, C, 2)  # [x][y][z], range::index order
    nodes = np.diff(t["node_begin"]).reshape(C, C, C)  # [z][y][x]
    buckets = np.diff(t["bucket_begin"]).reshape(C, C, C)
    assert np.array_equal(counts[..., 0].transpose(2, 1, 0), nodes)
    assert np.array_equal(counts[..., 1].transpose(2, 1, 0), buckets)
    kk = particles(load, "kicked")
    for d in range(3):
        assert np.array_equal(kk["x%d" % d], tr["x%d" % d])
    o = O.kick_full(P, s["x0"], s["x1"], s["x2"], s["u0"], s["u1"], s["u2"], s["r"], tree=True)
    w = match(key(kk), (o["x"].astype(np.uint64) << np.uint64(32)) | o["y"].astype(np.uint64))
    assert rel(np.stack([kk["g0"], kk["g1"], kk["g2"]]), np.stack([o["gx"][w], o["gy"][w], o["gz"][w]])) < 1e-5
