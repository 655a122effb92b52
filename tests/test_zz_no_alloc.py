"""tpm_ewald_at_noalloc, the Ewald sum at arbitrary positions that allocates
nothing: through the C ABI, on device buffers, against tpm_ewald_at on the
same positions, on two calls in a row (the second with fewer samples, so the
shared work space keeps the first call's contents), plus its limits."""
import ctypes

import numpy as np
import pytest

import tigerpm_amd as T

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class DeviceBuffers:
    """Plain device allocations for the test's inputs and outputs (the entry
    point itself allocates nothing)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(int(nbytes), 8)) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(d, a.ctypes.data, a.nbytes, _H2D) == 0
        return d

    def get(self, d, n, dtype):
        a = np.empty(n, dtype)
        assert self.hip.hipMemcpy(a.ctypes.data, d, a.nbytes, _D2H) == 0
        return a

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def uniform(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]


def run_noalloc(ctx, buf, sx, sy, sz, alpha):
    n = len(sx)
    d_in = [buf.put(a) for a in (sx, sy, sz)]
    d_out = [buf.alloc(8 * n) for _ in range(4)]
    print("tpm_ewald_at_noalloc: n=%d alpha=%g" % (n, alpha), flush=True)
    ctx.ewald_at_noalloc(n, *d_in, *d_out, alpha=alpha)
    ctx.sync()
    return [buf.get(d, n, np.float64) for d in d_out]


def test_noalloc_equals_ewald_at_on_two_calls(gpu_ctx):
    n_side = 24
    x, y, z = uniform(n_side ** 3, 91)
    gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
    gpu_ctx.set_host_order(T.TPM_ORDER_DEVICE)
    gpu_ctx.upload(x, y, z)
    buf = DeviceBuffers()
    try:
        for ns, alpha, seed in ((700, 0.0, 92), (130, 11.0, 93)):
            sx, sy, sz = uniform(ns, seed)
            # positions on particles and next to one (inside the softening)
            sx[:3], sy[:3], sz[:3] = x[:3], y[:3], z[:3]
            sx[3], sy[3], sz[3] = x[7] + 999, y[7], z[7]
            got = run_noalloc(gpu_ctx, buf, sx, sy, sz, alpha)
            ref = gpu_ctx.ewald_at(sx, sy, sz, alpha=alpha)
            # FP64 sums accumulated by atomics: their order varies between
            # calls, so equal to 1e-10 of the largest value, not bit for bit
            for a, b in zip(got, ref):
                assert np.all(np.isfinite(a))
                assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    finally:
        buf.close()


def test_noalloc_limits(gpu_ctx):
    n_side = 24
    x, y, z = uniform(n_side ** 3, 94)
    gpu_ctx.set_params(T.baseline_params(n_side, n_side, 3))
    gpu_ctx.upload(x, y, z)
    # rejected before any work is enqueued: no device buffer is touched
    with pytest.raises(T.TpmError) as e:
        gpu_ctx.ewald_at_noalloc(1, 8, 8, 8, 8, 8, 8, 8, alpha=12.5)
    assert "wave vectors" in str(e.value)
    with pytest.raises(T.TpmError) as e:
        gpu_ctx.ewald_at_noalloc((1 << 20) + 1, 8, 8, 8, 8, 8, 8, 8)
    assert "samples" in str(e.value)
    with pytest.raises(T.TpmError):
        gpu_ctx.ewald_at_noalloc(1, 0, 8, 8, 8, 8, 8, 8)
    # zero samples: nothing to do
    gpu_ctx.ewald_at_noalloc(0, 0, 0, 0, 0, 0, 0, 0)
