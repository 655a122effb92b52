"""GPU: the mesh FFTs (fft3d_execute / fft3d_inv_execute, fft.cpp:51-84).

The LDS-tiled transform (power-of-two M in [16, 1024]) and rocFFT (every
other M, or tuning fft=1) are checked against numpy's FP64 transforms, each
other, closed-form plane waves and round trips.  Tolerances are FP32
transform errors: relative rms 2e-6 of the spectrum or field.
"""
import numpy as np
import pytest

import tigerpm_amd as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fctx():
    ctx = T.Context(0)
    ctx.set_params(T.baseline_params(32, 32, 2))
    yield ctx
    ctx.close()


def rel_rms(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.sqrt((np.abs(a - b) ** 2).sum() / max((np.abs(b) ** 2).sum(), 1e-300)))


def field(M, seed):
    return np.random.default_rng(seed).standard_normal((M, M, M)).astype(np.float32)


@pytest.mark.parametrize("M", [16, 32, 64, 128, 256, 24, 40])
@pytest.mark.parametrize("backend", [0, 1])
def test_r2c_c2r_match_numpy(fctx, M, backend):
    fctx.set_tuning(fft=backend)
    a = field(M, M + 7)
    ref = np.fft.rfftn(a.astype(np.float64))
    c = fctx.fft3d_r2c(a)
    assert c.shape == ref.shape
    assert rel_rms(c, ref) < 2e-6
    # the inverse of an exact Hermitian spectrum, unnormalised
    b = fctx.fft3d_c2r(ref.astype(np.complex64))
    assert rel_rms(b, a.astype(np.float64) * M ** 3) < 2e-6
    fctx.set_tuning()


@pytest.mark.parametrize("M", [32, 128])
def test_backends_agree(fctx, M):
    a = field(M, 3 * M)
    fctx.set_tuning(fft=1)
    c_roc = fctx.fft3d_r2c(a)
    b_roc = fctx.fft3d_c2r(c_roc)
    fctx.set_tuning(fft=0)
    c_lds = fctx.fft3d_r2c(a)
    b_lds = fctx.fft3d_c2r(c_lds)
    assert rel_rms(c_lds, c_roc) < 2e-6
    assert rel_rms(b_lds, b_roc) < 2e-6
    assert rel_rms(b_lds / M ** 3, a) < 2e-6
    # the same call twice gives the same bits
    assert np.array_equal(fctx.fft3d_r2c(a), c_lds)


@pytest.mark.parametrize("M", [16, 512, 1024])
def test_plane_waves_closed_form(fctx, M):
    """cos(2 pi (a x + b y + c z) / M) with 0 < c < M/2 has the single half-
    spectrum entry M^3 / 2 at (a, b, c); two such waves, two entries."""
    fctx.set_tuning()
    waves = [(3, M - 5, 2, 1.0), (M // 2, 1, M // 2 - 1, -0.5)]
    ix = np.arange(M, dtype=np.int64)
    f = np.zeros((M, M, M), np.float32)
    for (a, b, c, amp) in waves:
        # separable phases, a slab of x planes at a time (M = 1024: 1e9 nodes)
        ey = np.exp(2j * np.pi * ((b * ix) % M) / M)
        ez = np.exp(2j * np.pi * ((c * ix) % M) / M)
        eyz = ey[:, None] * ez[None, :]
        for x0 in range(0, M, 64):
            ex = np.exp(2j * np.pi * ((a * ix[x0:x0 + 64]) % M) / M)
            f[x0:x0 + 64] += (amp * (ex[:, None, None] * eyz[None, :, :]).real).astype(np.float32)
    s = fctx.fft3d_r2c(f)
    exp = {(a % M, b % M, c): amp * M ** 3 / 2 for (a, b, c, amp) in waves}
    scale = M ** 3 / 2
    for key, v in exp.items():
        assert abs(complex(s[key]) - v) < 1e-5 * scale, (key, s[key], v)
        s[key] = 0
    assert float(np.abs(s).max()) < 1e-5 * scale
    # back to the field
    spec = np.zeros((M, M, M // 2 + 1), np.complex64)
    for key, v in exp.items():
        spec[key] = v
    g = fctx.fft3d_c2r(spec)
    err = float(np.abs(g.astype(np.float64) / M ** 3 - f).max())
    assert err < 2e-5, err


def test_pm_potential_same_for_both_backends(fctx):
    """The PM potential through either transform agrees to FP32 rounding."""
    n_side, M = 32, 64
    rng = np.random.default_rng(5)
    x, y, z = [rng.integers(0, 2 ** 32, n_side ** 3, dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    fctx.set_params(T.baseline_params(n_side, M, 3))
    fctx.upload(x, y, z)
    phis = []
    for backend in (1, 0):
        fctx.set_tuning(fft=backend)
        fctx.gravity_long_compute()
        phis.append(fctx.get_phi())
    fctx.set_tuning()
    assert rel_rms(phis[1], phis[0]) < 2e-6
