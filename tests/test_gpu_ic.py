"""GPU tests of the Zel'dovich initial conditions (tpm_ic_zeldovich):
statistical, as the field is random -- the rms displacement is the requested
one, the displacement is a gradient field (no transverse part), and its
shell-averaged spectrum follows k^n exp(-(k/k_Nyq)^2) / k^2 within the
sampling error of each shell."""
import numpy as np
import pytest

import tigerpm_amd as T

pytestmark = pytest.mark.gpu


def displacements(n, seed, rms, index=-2.0):
    with T.Context(0) as c:
        c.set_params(T.baseline_params(n, n, 3))
        c.ic_zeldovich(n, seed, rms, index)
        d = c.download()  # lattice order: index l = (i n + j) n + k
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    out = []
    for pos, q in zip((d["x"], d["y"], d["z"]), (i, j, k)):
        u = pos.astype(np.float64).reshape(n, n, n) * 2.0 ** -32 * n - (q + 0.5)
        out.append(u - np.rint(u / n) * n)  # periodic, in lattice cells
    return out


@pytest.mark.parametrize("n,rms", [(32, 0.5), (48, 2.0)])
def test_rms_displacement_and_gradient_field(n, rms):
    psi = displacements(n, 5, rms)
    got = np.sqrt(np.mean(psi[0] ** 2 + psi[1] ** 2 + psi[2] ** 2))
    assert abs(got / rms - 1) < 1e-5  # positions are rounded to 2^-32 only
    ph = [np.fft.rfftn(p) for p in psi]
    k1 = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    K = np.meshgrid(k1, k1, kz, indexing="ij")
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    k2[0, 0, 0] = 1.0
    dot = sum(Kd * pd for Kd, pd in zip(K, ph))
    trans = sum(np.abs(pd - Kd * dot / k2) ** 2 for Kd, pd in zip(K, ph)).sum()
    total = sum(np.abs(pd) ** 2 for pd in ph).sum()
    assert trans / total < 1e-8


def test_spectrum_follows_the_power_law():
    n, index = 48, -2.0
    psi = displacements(n, 9, 1.0, index)
    ph = [np.fft.fftn(p) for p in psi]
    k1 = np.fft.fftfreq(n) * 2 * np.pi
    K = np.meshgrid(k1, k1, k1, indexing="ij")
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    kk = np.sqrt(k2)
    power = sum(np.abs(p) ** 2 for p in ph)
    nyq = np.zeros_like(kk, bool)
    for Kd in K:
        nyq |= np.isclose(np.abs(Kd), np.pi)
    ok = (k2 > 0) & ~nyq
    # |psi_hat|^2 = P(k) / k^2 |noise_hat|^2, P = k^n exp(-(k / k_Nyq)^2)
    model = np.where(ok, kk ** index * np.exp(-k2 / np.pi ** 2) / np.where(ok, k2, 1.0), 0.0)
    assert power[~ok].max() < 1e-10 * power.max()  # k = 0 and Nyquist planes carry nothing
    edges = np.linspace(0.2, 2.6, 9)
    ratios, sig = [], []
    for a, b in zip(edges[:-1], edges[1:]):
        m = ok & (kk >= a) & (kk < b)
        r = power[m] / model[m]
        ratios.append(r.mean())
        sig.append(r.std() / np.sqrt(m.sum() / 2))  # conjugate pairs are one mode
    ratios, sig = np.array(ratios), np.array(sig)
    amp = np.average(ratios, weights=1 / sig ** 2)
    # one amplitude fits every shell within 4 sigma of its sampling error
    assert np.all(np.abs(ratios - amp) < 4 * sig), (ratios / amp, sig / amp)
