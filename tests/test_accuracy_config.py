"""The force error of small, weakly clustered boxes is the configuration's.

The workloads fix rcut = 4.5 rs.  Beyond rcut only the mesh force remains,
and in a weakly clustered box the net force is a small residual, so the
truncated short-range tail shows in the rms error against Ewald.  These CPU
tests pin that with the oracle alone, the direct sum in FP64 (no FP32, no
GPU): on Zel'dovich-displaced lattices of 24^3 particles, rcut = 4.5 rs gives
well over 1e-3 at small displacements, rcut = 6 rs about half of it, and the
error falls as the displacement (the clustering) grows.
"""
import numpy as np

import oracle as O


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def zeldovich_lattice(n_side, rms_cells, seed):
    """A cell-centred lattice displaced by the gradient of a Gaussian field
    with P(k) ~ k^-2 exp(-(k / k_Nyq)^2), scaled to the given rms
    displacement in lattice cells (numpy, for these tests only)."""
    rng = np.random.default_rng(seed)
    N = n_side
    k = np.fft.fftfreq(N) * N
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    delta = np.fft.fftn(rng.standard_normal((N, N, N))) * np.sqrt(np.exp(-k2 / (N / 2) ** 2) / k2)
    delta[0, 0, 0] = 0.0
    psi = [np.real(np.fft.ifftn(1j * kk * delta / k2)) for kk in (kx, ky, kz)]
    scale = rms_cells / np.sqrt(np.mean(psi[0] ** 2 + psi[1] ** 2 + psi[2] ** 2))
    g = np.arange(N) + 0.5
    q = np.meshgrid(g, g, g, indexing="ij")
    return [np.minimum(np.floor(((qq + p * scale) / N % 1.0).ravel() * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint32)
            for qq, p in zip(q, psi)]


def kick_error(x, y, z, n_side, rcut_over_rs):
    """rms vector force error of the oracle's kick (FP64 direct sum,
    interlaced TSC) against FP64 Ewald on 256 sampled particles."""
    n = len(x)
    rs = 1.25 / n_side
    rcut = rcut_over_rs * rs
    p = O.make_params(n_side, int(1 / rcut), 3, rs, rcut, 2 / (25 * n_side), 1.0 / n, 0.1, n, 1, 1)
    zf = np.zeros(n, np.float32)
    r = O.kick_full(p, x, y, z, zf, zf, zf, np.zeros(n, np.int8), fp64_pp=True)
    s = np.sort(np.random.default_rng(5).choice(n, 256, replace=False))
    _, gx, gy, gz = O.ewald_sf(r["x"], r["y"], r["z"], s, p.hsoft, 9.6)
    G = np.stack([r["gx"][s], r["gy"][s], r["gz"][s]], 1).astype(np.float64)
    return rel_rms(G, np.stack([gx, gy, gz], 1) / n)


def test_small_weakly_clustered_box_error_is_the_rcut_truncation():
    n_side = 24
    x, y, z = zeldovich_lattice(n_side, 0.3, 11)
    e45 = kick_error(x, y, z, n_side, 4.5)
    e60 = kick_error(x, y, z, n_side, 6.0)
    print("rms displacement 0.3 cells: rcut = 4.5 rs %.3e, 6 rs %.3e" % (e45, e60))
    assert e45 > 1.3e-3
    assert e60 < 0.65 * e45


def test_error_falls_with_clustering():
    n_side = 24
    weak = kick_error(*zeldovich_lattice(n_side, 0.3, 11), n_side, 4.5)
    more = kick_error(*zeldovich_lattice(n_side, 1.0, 11), n_side, 4.5)
    print("rcut = 4.5 rs: rms displacement 0.3 cells %.3e, 1 cell %.3e" % (weak, more))
    assert more < 0.75 * weak
