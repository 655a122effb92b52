"""Pins the CPU oracle (oracle/tpm_oracle.c) without the GPU.

The reference has no tests, golden vectors or Python (its only CPU force
path is PM-only and needs FFTW/HPX), so each oracle function is pinned by
closed forms, invariants, special cases that reduce to textbook routines,
independent numpy/scipy reimplementations and brute force on tiny inputs.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.special import erfc as sp_erfc

import oracle as O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "analytic_values.json")))
TWO32 = 2.0 ** 32


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / ((b ** 2).sum())))


def rand_u32(n, seed):
    return np.random.default_rng(seed).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)


# --------------------------------------------------------------------- a1
def test_fixed_encoding_closed_forms():
    L = O.lib()
    assert L.orc_fixed_from_double(0.5) == 2 ** 31
    assert L.orc_fixed_from_double(0.25) == 2 ** 30
    assert L.orc_fixed_from_double(1.0 - 2.0 ** -32) == 2 ** 32 - 1
    assert L.orc_fixed_from_double(1.0 / 3.0) == 1431655765  # floor(2^32 / 3): truncation
    assert L.orc_fixed_to_float(2 ** 31) == 0.5
    assert L.orc_fixed_to_double(1) == 2.0 ** -32


def test_distance_is_minimum_image():
    L = O.lib()
    a, b = rand_u32(2000, 1), rand_u32(2000, 2)
    for ai, bi in zip(a[:500], b[:500]):
        d = L.orc_distance(int(ai), int(bi))
        mi = ((float(ai) - float(bi)) / TWO32 + 0.5) % 1.0 - 0.5
        assert abs(d - mi) <= 2 ** -24 * max(abs(mi), 2 ** -32) + 1e-12
        if abs(mi) < 0.49:
            assert d == -L.orc_distance(int(bi), int(ai))
    assert abs(L.orc_distance(L.orc_fixed_from_double(0.95), L.orc_fixed_from_double(0.05)) + 0.1) < 1e-7


# --------------------------------------------------------------------- a3
@pytest.mark.parametrize("C", [3, 7, 11, 45])
def test_chain_coord_matches_partition_planes(C):
    L = O.lib()
    b = np.empty(C + 1, np.uint64)
    L.orc_chain_bounds(C, b)
    # planes are fixed32(k / C) truncated: within one raw unit of k 2^32 / C
    k = np.arange(C)
    assert np.all(np.abs(b[:C].astype(np.float64) - k * TWO32 / C) <= 1.0)
    probes = []
    for kk in range(1, C):
        probes += [int(b[kk]) - 1, int(b[kk]), int(b[kk]) + 1]
    probes += [0, 2 ** 32 - 1] + list(rand_u32(300, C))
    for r in probes:
        c = L.orc_chain_coord(r, C)
        assert b[c] <= r and (c == C - 1 or r < b[c + 1])


@pytest.mark.parametrize("C,n", [(3, 500), (7, 4000), (11, 20000)])
def test_chainmesh_sort_equals_reference_partition(C, n):
    rng = np.random.default_rng(C)
    x, y, z = rand_u32(n, 10 + C), rand_u32(n, 20 + C), rand_u32(n, 30 + C)
    # particles exactly on and next to split planes
    b = np.empty(C + 1, np.uint64)
    O.lib().orc_chain_bounds(C, b)
    x[:C - 1] = b[1:C].astype(np.uint32)
    y[C:2 * C - 1] = (b[1:C] - 1).astype(np.uint32)
    perm, cb = O.chainmesh_sort(x, y, z, C)
    assert np.array_equal(np.sort(perm), np.arange(n))
    idx, rb, re_ = O.chainmesh_partition_ref(x, y, z, C)
    assert np.array_equal(rb, cb[:-1]) and np.array_equal(re_, cb[1:])
    for cid in rng.choice(C ** 3, 20):
        assert set(perm[cb[cid]:cb[cid + 1]]) == set(idx[rb[cid]:re_[cid]])
    # keys sorted, ties in previous order (stable)
    k = O.chain_keys(x, y, z, C)[perm]
    assert np.all(np.diff(k.astype(np.int64)) >= 0)
    tie = np.diff(k.astype(np.int64)) == 0
    assert np.all(np.diff(perm)[tie] > 0)


def test_reference_partition_hand_worked():
    """The reference partition's order, worked out by hand from the
    two-pointer particles_sort (particles.cpp:414-445) and the recursion of
    chainmesh.cpp:168-213 (z split first, then y, then x; C = 2, split plane
    fixed32(0.5) = 2^31; a coordinate >= the plane goes right)."""
    L, H = 0x10000000, 0x90000000
    # (1) one particle per cell.  z split of [0,6): right = {0, 2}; left
    # count 4; misplaced right at 0, 2 trade with left at 5, 4 ->
    # [5, 1, 4, 3, 2, 0].  y split of [5, 1, 4, 3]: lo meets 1 at position 1,
    # hi finds 4 at position 2 -> [5, 4, 1, 3]; y split of [2, 0]: no move.
    # x splits: [5, 4] -> [4, 5]; [1, 3] and [2, 0] stay.
    x = np.array([H, L, L, H, L, H], np.uint32)
    y = np.array([L, H, L, H, L, L], np.uint32)
    z = np.array([H, L, H, L, L, L], np.uint32)
    idx, cb, ce = O.chainmesh_partition_ref(x, y, z, 2)
    assert list(idx) == [4, 5, 1, 3, 2, 0]
    assert list(cb[:6]) == [0, 1, 2, 3, 4, 5] and list(ce[:6]) == [1, 2, 3, 4, 5, 6]
    # (2) several particles per cell, all at x, y < 1/2: z classes
    # [R, R, L, L, R, L] -> positions 0, 1 trade with 5, 3 -> [5, 3, 2, 1, 4, 0]
    # (not the stable order [2, 3, 5 | 0, 1, 4]); the y and x splits keep it.
    # A particle exactly on the plane (raw 2^31) goes right.
    x = np.full(6, L, np.uint32)
    y = np.full(6, L, np.uint32)
    z = np.array([H, 2 ** 31, L, L, H, L], np.uint32)
    idx, cb, ce = O.chainmesh_partition_ref(x, y, z, 2)
    assert list(idx) == [5, 3, 2, 1, 4, 0]
    assert (cb[0], ce[0]) == (0, 3) and (cb[4], ce[4]) == (3, 6)


def test_chainmesh_edge_cases():
    e = np.zeros(0, np.uint32)
    perm, cb = O.chainmesh_sort(e, e, e, 5)
    assert len(perm) == 0 and np.all(cb == 0)
    one = np.full(50, 123456789, np.uint32)
    perm, cb = O.chainmesh_sort(one, one, one, 5)
    assert np.array_equal(perm, np.arange(50))
    assert np.count_nonzero(np.diff(cb)) == 1


# --------------------------------------------------------------------- a4
@pytest.mark.parametrize("order", [2, 3, 4])
def test_deposit_conserves_mass_exactly(order):
    x, y, z = rand_u32(5000, 1), rand_u32(5000, 2), rand_u32(5000, 3)
    x[:3] = [0, 2 ** 32 - 1, 2 ** 31]
    g = O.deposit(x, y, z, 12, order)
    assert int(g.sum(dtype=np.uint64)) == 5000 << 32


@pytest.mark.parametrize("order", [2, 3, 4])
def test_deposit_is_linear_and_shift_covariant(order):
    M = 16
    x, y, z = rand_u32(3000, 4), rand_u32(3000, 5), rand_u32(3000, 6)
    g = O.deposit(x, y, z, M, order).reshape(M, M, M)
    ga = O.deposit(x[:1000], y[:1000], z[:1000], M, order).reshape(M, M, M)
    gb = O.deposit(x[1000:], y[1000:], z[1000:], M, order).reshape(M, M, M)
    assert np.array_equal(g, ga + gb)
    step = np.uint32(2 ** 32 // M)  # exactly one mesh cell for M a power of two
    gs = O.deposit(x + step, y, z, M, order).reshape(M, M, M)
    assert np.array_equal(gs, np.roll(g, 1, axis=0))


@pytest.mark.parametrize("order", [2, 3, 4])
def test_weights_match_assignment_functions(order):
    """1-D weights equal the textbook CIC/TSC/cubic B-spline values."""
    M = 10
    for raw in list(rand_u32(400, order)) + [0, 2 ** 32 - 1]:
        base, w, q = O.weights_1d(order, int(raw), M)
        X = int(raw) * M / TWO32
        nodes = base + np.arange(len(w))
        d = X - nodes  # particle minus node, in cells
        if order == 2:
            ref = np.maximum(1 - np.abs(d), 0)
        elif order == 3:
            ad = np.abs(d)
            ref = np.where(ad < 0.5, 0.75 - d * d, np.where(ad < 1.5, 0.5 * (1.5 - ad) ** 2, 0.0))
        else:
            ad = np.abs(d)
            ref = np.where(ad < 1, (4 - 6 * ad ** 2 + 3 * ad ** 3) / 6,
                           np.where(ad < 2, (2 - ad) ** 3 / 6, 0.0))
        assert int(w.sum()) == 1 << q
        assert np.all(np.abs(w / 2.0 ** q - ref) < 4 * 2.0 ** -q)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_shifted_deposit_is_half_cell_translation(order):
    """The interlacing grid (shift 2^31 in raw * M) equals depositing the
    particles moved by +1/2 cell, exact when 2^32 / (2M) is an integer."""
    M = 16
    x, y, z = rand_u32(3000, 61), rand_u32(3000, 62), rand_u32(3000, 63)
    half = np.uint32(2 ** 31 // M)
    a = O.deposit_shifted(x, y, z, M, order, 1 << 31)
    b = O.deposit(x + half, y + half, z + half, M, order)
    assert np.array_equal(a, b)


def test_interlaced_phi_smooth_mode_unchanged():
    """For a band-limited density both grids sample the same field, so the
    interlaced combination reproduces the single-grid potential."""
    M = 16
    j = np.arange(M)
    f = lambda s: np.cos(2 * np.pi * (j + s) / M)[:, None, None] * np.cos(2 * np.pi * 2 * (j + s) / M)[None, :, None]
    rho0 = (f(0.0) * np.ones((1, 1, M))).astype(np.float32).reshape(-1)
    rho1 = (f(-0.5) * np.ones((1, 1, M))).astype(np.float32).reshape(-1)  # node j at (j - 1/2) / M
    t = O.green_table(M, 1.25 / M, 2)
    a = O.pm_phi(M, rho0, t)
    b = O.pm_phi_interlaced(M, rho0, rho1, t)
    assert rel_rms(b, a) < 1e-5


def test_single_particle_tsc_at_node():
    M = 8
    raw = 3 * 2 ** 32 // M  # exactly on node 3
    g = O.deposit([raw], [raw], [raw], M, 3).reshape(M, M, M) / TWO32
    w1 = {2: 0.125, 3: 0.75, 4: 0.125}
    for i in (2, 3, 4):
        for j in (2, 3, 4):
            for k in (2, 3, 4):
                assert abs(g[i, j, k] - w1[i] * w1[j] * w1[k]) < 1e-8
    assert abs(g.sum() - 1.0) == 0.0


# --------------------------------------------------------------------- a5
@pytest.mark.parametrize("M", [8, 12, 16, 7])
def test_fft_matches_numpy(M):
    a = np.random.default_rng(M).standard_normal((M, M, M))
    c = O.fft3d_r2c(a)
    assert np.abs(c - np.fft.rfftn(a)).max() < 1e-10 * np.abs(c).max()
    assert np.abs(O.fft3d_c2r(c, M) / M ** 3 - a).max() < 1e-12


# --------------------------------------------------------------------- a6
@pytest.mark.parametrize("M,order", [(16, 2), (20, 3), (12, 4)])
def test_green_table_values(M, order):
    rs = 1.25 / M
    t = O.green_table(M, rs, order).reshape(M, M, M // 2 + 1)
    assert t[0, 0, 0] == 0.0
    assert np.all(t.reshape(-1)[1:] <= 0.0)
    for (ix, iy, iz) in [(1, 0, 0), (0, 3, 2), (M - 1, 2, M // 2), (M // 2, M // 2, 1)]:
        k = [2 * math.pi * (i if 2 * i < M else i - M) / M for i in (ix, iy, iz)]
        k2 = sum(kk * kk for kk in k)
        s = 1.0
        for kk in k:
            s *= math.sin(kk / 2) / (kk / 2) if kk else 1.0
        ref = -4 * math.pi * M / M ** 3 / k2 * math.exp(-k2 * (rs * M) ** 2) * s ** (-order)
        assert abs(t[ix, iy, iz] - ref) <= 1e-7 * abs(ref)


def test_pm_plane_wave_closed_form():
    """One cosine mode through the whole PM pipeline: phi = G(k) M^3 rho."""
    M, order = 16, 2
    rs = 1.25 / M
    j = np.arange(M)
    rho = np.cos(2 * np.pi * 2 * j / M)[:, None, None] * np.ones((1, M, M))
    phi = O.pm_phi(M, rho.astype(np.float32).reshape(-1), O.green_table(M, rs, order)).reshape(M, M, M)
    k = 2 * np.pi * 2 / M
    expect = -4 * np.pi * M / k ** 2 * np.exp(-k ** 2 * (rs * M) ** 2) / (np.sin(k / 2) / (k / 2)) ** order
    assert np.abs(phi - expect * rho).max() < 1e-5 * abs(expect)


# --------------------------------------------------------------------- a8
def test_interp6_weights_properties():
    w0, dw0 = O.interp6_weights(0.0)
    assert np.allclose(w0, [0, 0, 1, 0, 0, 0], atol=0)
    # at a node the derivative weights are the 4th-order central difference
    assert np.allclose(dw0, GOLDEN["central_difference_4th_order"] + [0.0], atol=1e-7)
    for x in np.random.default_rng(0).random(50):
        w, dw = O.interp6_weights(x)
        nodes = np.arange(-2, 4)
        # FP32 evaluation of the reference's expanded polynomials: ~1e-6
        assert abs(w.sum() - 1) < 5e-6 and abs(dw.sum()) < 5e-6
        # exact for polynomials up to degree 3 (value) / derivative of degree 2
        for p in range(4):
            assert abs((w * nodes ** p).sum() - x ** p) < 2e-5
        for p in range(1, 3):
            assert abs((dw * nodes ** p).sum() - p * x ** (p - 1)) < 2e-5


def test_interp6_smooth_field():
    M = 64
    j = np.arange(M)
    phi = np.cos(2 * np.pi * j / M)[:, None, None] * np.ones((1, M, M))
    x, y, z = rand_u32(200, 9), rand_u32(200, 10), rand_u32(200, 11)
    gx, gy, gz, p = O.interp6(M, phi.astype(np.float32).reshape(-1), x, y, z)
    X = x / TWO32
    assert np.abs(p - np.cos(2 * np.pi * X)).max() < 2e-6
    # derivative of a 6-point interpolant: O((2 pi / M)^4) relative error
    assert np.abs(gx - 2 * np.pi * np.sin(2 * np.pi * X)).max() < 5e-5 * 2 * np.pi
    # transverse components: FP32 rounding of sum(dw) ~ 1e-6 times M
    assert np.abs(gy).max() < 3e-4 and np.abs(gz).max() < 3e-4


# --------------------------------------------------------------------- a9
def pp_params(M=64, n_side=64):
    rs = 1.25 / M
    return O.make_params(M, int(1 / (4.5 * rs)), 3, rs, 4.5 * rs, 2 / (25 * n_side), 1.0, 0.1, n_side ** 3)


def test_pp_pair_fp64_matches_scipy():
    p = pp_params()
    for r in np.geomspace(p.hsoft * 1.01, p.rcut * 0.999, 40):
        x = r / (2 * p.rs)
        f = (sp_erfc(x) + 2 / np.sqrt(np.pi) * x * np.exp(-x * x)) / r ** 3
        out = O.pp_pair(p, r, 0.0, 0.0, True)
        assert abs(out[0] + r * f) <= 1e-12 * r * f
        assert abs(out[3] + sp_erfc(x) / r) <= 1e-12 * sp_erfc(x) / r
    for r in np.linspace(p.hsoft * 0.3, p.hsoft * 0.99, 10):
        x = r / (2 * p.rs)
        q2 = (r / p.hsoft) ** 2
        d1 = (15 / 8 * q2 ** 2 - 21 / 4 * q2 + 35 / 8) / p.hsoft ** 3
        f = d1 + (sp_erfc(x) - 1 + 2 / np.sqrt(np.pi) * x * np.exp(-x * x)) / r ** 3
        out = O.pp_pair(p, 0.0, r, 0.0, True)
        assert abs(out[1] + r * f) <= 1e-7 * r * f


def test_pp_pair_fp32_formula_accuracy_and_edges():
    p = pp_params()
    rs = np.geomspace(p.hsoft * 0.1, p.rcut * 0.999, 300)
    for r in rs:
        a = O.pp_pair(p, r * 0.6, r * 0.8, 0.0, False)
        b = O.pp_pair(p, r * 0.6, r * 0.8, 0.0, True)
        # Abramowitz & Stegun 7.1.26 carries |erfc error| <= 1.5e-7, plus
        # ~2.5e-7 of FP32 rounding in its 5-term polynomial; inside the
        # softening the reference multiplies (erfc - 1) by r^-3, so the FP32
        # formula's own error there is 4e-7 / r^3 (r times that in g).
        # The same absolute erfc error sets the floor outside h (erfc/r in
        # phi, erfc/r^2 in g).
        assert np.abs(a[:3] - b[:3]).max() <= 2e-6 * np.abs(b[:3]).max() + 4e-7 / r ** 2
        assert abs(a[3] - b[3]) <= 2e-6 * abs(b[3]) + 4e-7 / r
    assert np.all(O.pp_pair(p, p.rcut * 1.0001, 0, 0, False) == 0)
    self_ = O.pp_pair(p, 0.0, 0.0, 0.0, False)
    assert np.all(self_[:3] == 0) and abs(self_[3] + 35 / 16 / p.hsoft) < 1e-6 / p.hsoft
    # force continuous across r = h
    lo = O.pp_pair(p, p.hsoft * (1 - 1e-6), 0, 0, True)[0]
    hi = O.pp_pair(p, p.hsoft * (1 + 1e-6), 0, 0, True)[0]
    assert abs(lo - hi) < 1e-4 * abs(hi)


def brute_pp(x, y, z, p):
    """O(N^2) minimum-image sum with scipy erfc, softened inside h."""
    P = np.stack([x, y, z], 1).astype(np.float64) / TWO32
    g = np.zeros_like(P)
    h = p.hsoft
    for i in range(len(P)):
        d = (P[i] - P + 0.5) % 1.0 - 0.5
        r = np.sqrt((d * d).sum(1))
        m = (r > 0) & (r < p.rcut)
        rm = r[m]
        xx = rm / (2 * p.rs)
        F = sp_erfc(xx) + 2 / np.sqrt(np.pi) * xx * np.exp(-xx * xx)
        q2 = (rm / h) ** 2
        soft = (15 / 8 * q2 ** 2 - 21 / 4 * q2 + 35 / 8) / h ** 3 + (F - 1) / rm ** 3
        f = np.where(rm > h, F / rm ** 3, soft)
        g[i] = -(d[m] * f[:, None]).sum(0)
    return g


def test_pp_sum_equals_brute_force():
    n_side = 12
    n = 2000
    x, y, z = rand_u32(n, 31), rand_u32(n, 32), rand_u32(n, 33)
    x[1], y[1], z[1] = x[0] + 3000, y[0], z[0]  # a pair inside the softening
    p = pp_params(32, n_side)
    assert p.chain_dim == 5
    perm, cb = O.chainmesh_sort(x, y, z, p.chain_dim)
    xs, ys, zs = x[perm], y[perm], z[perm]
    (gx, gy, gz, _), cand = O.pp_sum(p, xs, ys, zs, cb, np.arange(n), fp64=True)
    assert cand == O.candidate_pairs(p.chain_dim, cb)
    G = brute_pp(xs, ys, zs, p)
    assert rel_rms(np.stack([gx, gy, gz], 1), G) < 1e-6
    # momentum: the short-range forces sum to zero
    assert abs(gx.astype(np.float64).sum()) < 1e-6 * np.abs(gx).sum()


# --------------------------------------------------------------------- a12
def test_kick_epilogue_semantics():
    p = pp_params()
    n = 2000
    rng = np.random.default_rng(4)
    g = [rng.standard_normal(n).astype(np.float32) * 10 for _ in range(3)]
    pot = rng.standard_normal(n).astype(np.float32)
    v0 = [np.zeros(n, np.float32)] * 3
    rung = rng.integers(0, 5, n).astype(np.int8)
    (gx, gy, gz), pot2, (vx, vy, vz), r2 = O.kick_epilogue(p, g, pot, v0, rung, 0, 1.0, 1.0, True)
    gm = np.sqrt(gx.astype(np.float64) ** 2 + gy ** 2 + gz ** 2)
    dt_req = np.minimum(p.eta * np.sqrt(p.hsoft) / np.sqrt(gm), 1.0)
    dt_rung = 2.0 ** -r2.astype(np.float64)
    free = r2 > rung - 1  # not held by the "drop at most one rung" rule
    assert np.all(dt_rung[free] <= dt_req[free] * (1 + 1e-6))
    assert np.all(2 * dt_rung[free] > dt_req[free] * (1 - 1e-6))
    assert np.all(r2 >= rung - 1)
    assert np.allclose(vx, gx * 0.5 * dt_rung, rtol=1e-6)
    # inactive particles untouched
    (hx, _, _), _, (ux, _, _), r3 = O.kick_epilogue(p, g, pot, v0, rung, 10, 1.0, 1.0, True)
    assert np.array_equal(hx, g[0]) and np.all(ux == 0) and np.array_equal(r3, rung)


# --------------------------------------------------------------------- a14
def test_ewald_madelung_and_alpha_independence():
    xi = O.ewald_psi(0, 0, 0, 2.0, 4, 4, True)[0]
    assert abs(xi - GOLDEN["madelung_sc_unit_cube"]) < 1e-10
    ref = O.ewald_psi(0.1, 0.2, -0.3, 2.0, 4, 4, False)
    for a in (1.5, 3.0):
        assert np.abs(O.ewald_psi(0.1, 0.2, -0.3, a, 5, 8, False) - ref).max() < 1e-11
    # small separations: psi = -1/r + xi - (2 pi / 3) r^2 + O(r^4) (the
    # neutralizing background of density -1: laplacian = -4 pi; cubic
    # symmetry kills the other r^2 harmonics), so grad psi = 1/r^2 - (4 pi/3) r
    for r in (1e-3, 1e-2):
        g = O.ewald_psi(r, 0, 0, 2.0, 4, 4, False)
        assert abs(g[1] - 1 / r ** 2 + 4 * math.pi / 3 * r) < 1e-6 + 5e-11 / r ** 2 + 20 * r ** 3


def test_ewald_symmetric_configurations_have_no_force():
    h = 1e-3
    L = [int(v) for v in (0, 2 ** 31)]
    x = np.array([a for a in L for b in L for c in L], np.uint32)
    y = np.array([b for a in L for b in L for c in L], np.uint32)
    z = np.array([c for a in L for b in L for c in L], np.uint32)
    s = np.arange(8)
    for fn in (lambda: O.ewald_direct(x, y, z, s, h), lambda: O.ewald_sf(x, y, z, s, h, 10.0)):
        phi, gx, gy, gz = fn()
        assert np.abs(np.stack([gx, gy, gz])).max() < 1e-9
        assert np.ptp(phi) < 1e-9


def test_ewald_structure_factor_equals_direct():
    n = 300
    x, y, z = rand_u32(n, 41), rand_u32(n, 42), rand_u32(n, 43)
    x[1], y[1], z[1] = x[0] + 50000, y[0], z[0]  # one pair inside the softening
    s = np.arange(0, n, 7)
    a = O.ewald_direct(x, y, z, s, 1e-4)
    b = O.ewald_sf(x, y, z, s, 1e-4, 9.6)
    assert rel_rms(np.stack(b[1:]), np.stack(a[1:])) < 1e-9
    assert rel_rms(b[0], a[0]) < 1e-9


def test_ewald_at_positions():
    """The field of all particles at arbitrary positions: structure-factor
    form equals the pair-by-pair sum; at a particle's own position it is the
    per-particle sum plus that particle's softened self potential
    (-35/16 h^-1, no force); and g = -grad phi (central differences)."""
    n, h = 200, 2e-4
    x, y, z = rand_u32(n, 61), rand_u32(n, 62), rand_u32(n, 63)
    x[1], y[1], z[1] = x[0] + 20000, y[0], z[0]  # a pair inside the softening
    sx, sy, sz = rand_u32(40, 64), rand_u32(40, 65), rand_u32(40, 66)
    sx[:3], sy[:3], sz[:3] = x[:3], y[:3], z[:3]  # on particles
    sx[3], sy[3], sz[3] = x[0] + 3000, y[0] - 7000, z[0]  # inside the softening of two
    a = O.ewald_direct_at(x, y, z, sx, sy, sz, h)
    b = O.ewald_sf_at(x, y, z, sx, sy, sz, h, 9.6)
    assert rel_rms(np.stack(b[1:]), np.stack(a[1:])) < 1e-9
    assert rel_rms(b[0], a[0]) < 1e-9
    own = O.ewald_sf(x, y, z, np.arange(3), h, 9.6)
    assert np.allclose(b[0][:3], own[0] - 35.0 / 16.0 / h, rtol=0, atol=1e-8 * abs(own[0]).max())
    for d in range(3):
        assert np.allclose(b[1 + d][:3], own[1 + d], rtol=1e-10, atol=1e-9 * np.abs(own[1 + d]).max())
    # g = -grad phi away from particles: step of 2^12 raw units
    far = slice(4, 40)
    step = 4096
    for d, arr in enumerate((sx, sy, sz)):
        up = [sx.copy(), sy.copy(), sz.copy()]
        dn = [sx.copy(), sy.copy(), sz.copy()]
        up[d] = (arr.astype(np.int64) + step).astype(np.uint32)
        dn[d] = (arr.astype(np.int64) - step).astype(np.uint32)
        pu = O.ewald_sf_at(x, y, z, *up, h, 9.6)[0]
        pd = O.ewald_sf_at(x, y, z, *dn, h, 9.6)[0]
        fd = -(pu - pd) / (2 * step * 2.0 ** -32)
        assert rel_rms(fd[far], b[1 + d][far]) < 1e-4


# ------------------------------------------------------------ whole kick
@pytest.mark.parametrize("order,interlace", [(3, 0), (2, 1)])
def test_cpu_kick_meets_ewald_tolerance(order, interlace):
    """The oracle's full TreePM kick (TSC, or CIC with the interlaced
    deposit) is within 1e-3 rms of FP64 Ewald."""
    n_side = 24
    n = n_side ** 3
    x, y, z = rand_u32(n, 51), rand_u32(n, 52), rand_u32(n, 53)
    rs = 1.25 / n_side
    p = O.make_params(n_side, int(1 / (4.5 * rs)), order, rs, 4.5 * rs, 2 / (25 * n_side), 1.0 / n, 0.1, n, 1,
                      interlace)
    zf = np.zeros(n, np.float32)
    r = O.kick_full(p, x, y, z, zf, zf, zf, np.zeros(n, np.int8))
    s = np.sort(np.random.default_rng(5).choice(n, 200, replace=False))
    phi, gx, gy, gz = O.ewald_sf(r["x"], r["y"], r["z"], s, p.hsoft, 9.6)
    G = np.stack([r["gx"][s], r["gy"][s], r["gz"][s]], 1)
    assert rel_rms(G, np.stack([gx, gy, gz], 1) / n) < 1e-3


def test_drift_closed_forms():
    """The fixed-point drift: exact integer displacements, periodic wrap,
    round-to-nearest-even of half units, zero velocity, and two half-step
    drifts equal one full drift when every displacement is an integer."""
    rng = np.random.default_rng(17)
    n = 4096
    x, y, z = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    zero = np.zeros(n, np.float32)
    for a, b in zip(O.drift(x, y, z, zero, zero, zero, 0.37), (x, y, z)):
        assert np.array_equal(a, b)
    t0 = 2.0 ** -3
    k = rng.integers(-2 ** 20, 2 ** 20, n)
    v = (k.astype(np.float64) * 2.0 ** -29).astype(np.float32)  # v t0 = k raw units exactly
    nx, _, _ = O.drift(x, y, z, v, zero, zero, t0)
    assert np.array_equal(nx, ((x.astype(np.int64) + k) % 2 ** 32).astype(np.uint32))
    # wrap at both ends
    xe = np.array([2 ** 32 - 5, 3], np.uint32)
    ve = np.array([10, -10], np.float32) * np.float32(2.0 ** -29)
    we, _, _ = O.drift(xe, xe, xe, ve, np.zeros(2, np.float32), np.zeros(2, np.float32), t0)
    assert we.tolist() == [5, 2 ** 32 - 7]
    # half units round to even
    h = np.array([0.5, 1.5, 2.5, -0.5, -1.5], np.float64)
    vh = (h * 2.0 ** -29).astype(np.float32)
    xh = np.full(5, 1000, np.uint32)
    rh, _, _ = O.drift(xh, xh, xh, vh, np.zeros(5, np.float32), np.zeros(5, np.float32), t0)
    assert rh.tolist() == [1000, 1002, 1002, 1000, 998]
    # composition with even integer displacements
    k2 = 2 * rng.integers(-2 ** 19, 2 ** 19, n)
    v2 = (k2.astype(np.float64) * 2.0 ** -29).astype(np.float32)
    one = O.drift(x, y, z, v2, v2, v2, t0)
    half = O.drift(*O.drift(x, y, z, v2, v2, v2, t0 / 2), v2, v2, v2, t0 / 2)
    for a, b in zip(one, half):
        assert np.array_equal(a, b)


def test_pme_pair_law_closed_forms():
    """kick_pme's pair law (kick_pme.cu:280-310): inside h the plain softened
    polynomials (no erfc terms), at r = 0 exactly -35/16 h^-1 and no force,
    beyond h the erfc-split force with no cutoff (checked against scipy's
    erfc past rcut, where the kick_treepm law gives zero)."""
    from scipy.special import erfc
    rs, rcut, h = 0.01, 0.045, 0.002
    p = O.make_params(32, 6, 3, rs, rcut, h, 1.0, 0.14, 1000.0, pme=1)
    q = O.make_params(32, 6, 3, rs, rcut, h, 1.0, 0.14, 1000.0, pme=0)
    g = O.pp_pair(p, 0.0, 0.0, 0.0, False)
    assert g[0] == 0.0 and g[1] == 0.0 and g[2] == 0.0
    assert abs(g[3] + 35.0 / 16.0 / h) <= 1e-6 * 35.0 / 16.0 / h
    for r in (0.3 * h, 0.7 * h, 0.99 * h):
        u = r / h
        f = (15.0 / 8.0 * u ** 4 - 21.0 / 4.0 * u ** 2 + 35.0 / 8.0) / h ** 3 * r
        ph = -(-5.0 / 16.0 * u ** 6 + 21.0 / 16.0 * u ** 4 - 35.0 / 16.0 * u ** 2 + 35.0 / 16.0) / h
        out = O.pp_pair(p, r, 0.0, 0.0, False)
        assert abs(-out[0] - f) < 2e-6 * f and abs(out[3] - ph) < 2e-6 * abs(ph)
    for r in (0.5 * rcut, 1.2 * rcut, 1.8 * rcut):
        x = r / (2 * rs)
        f = (erfc(x) + 2 * x / np.sqrt(np.pi) * np.exp(-x * x)) / r ** 2
        out = O.pp_pair(p, 0.0, r, 0.0, False)
        assert abs(-out[1] - f) < 5e-6 / r ** 2 + 1e-5 * f
        assert abs(out[3] + erfc(x) / r) < 5e-6 / r
        if r > rcut:
            assert O.pp_pair(q, 0.0, r, 0.0, False)[1] == 0.0
