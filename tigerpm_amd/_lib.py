"""ctypes binding of libtigerpm_hip.so (C ABI: include/tigerpm_hip.h)."""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
_LIB = os.path.join(_HERE, "libtigerpm_hip.so")
_HEADER = os.path.join(_ROOT, "include", "tigerpm_hip.h")

TPM_CIC, TPM_TSC, TPM_CLOUD4 = 2, 3, 4
TPM_ORDER_DEVICE, TPM_ORDER_REFERENCE = 0, 1
TPM_SKIP_PP, TPM_SKIP_PM, TPM_SKIP_KICK = 1, 2, 4
TPM_KICK_PME = 16  # kick_pme's pair law (no cutoff, plain softening, no phi0)
TPM_KICK_TREE = 32  # kick_treepm's tree walk: per-cell trees, theta = 0.4 multipoles + PP

# tpm_tree_node / tpm_tree_bucket (include/tigerpm_hip.h)
TREE_NODE = np.dtype([("x", "<u4", 3), ("radius", "<f4"), ("child", "<i4", 2), ("begin", "<i4"), ("end", "<i4"),
                      ("m", "<f4", 4)])
TREE_BUCKET = np.dtype([("begin", "<i4"), ("end", "<i4"), ("x", "<u4", 3), ("radius", "<f4"), ("cell", "<i4"),
                        ("pad_", "<i4")])

_STATUS = {0: "ok", -1: "HIP error", -2: "invalid argument", -3: "call out of order", -4: "rocFFT error",
           -5: "device memory budget exceeded"}


class TpmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("tigerpm_hip: %s (%d): %s" % (_STATUS.get(status, "error"), status, msg))
        self.status = status


class TpmParams(ctypes.Structure):
    _fields_ = [("mesh_dim", ctypes.c_int), ("chain_dim", ctypes.c_int), ("order", ctypes.c_int),
                ("do_phi", ctypes.c_int), ("long_type", ctypes.c_int), ("interlace", ctypes.c_int),
                ("rs", ctypes.c_double), ("rcut", ctypes.c_double),
                ("hsoft", ctypes.c_double), ("GM", ctypes.c_double), ("eta", ctypes.c_double),
                ("n_total", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TpmMoments(ctypes.Structure):
    _fields_ = [("momentum", ctypes.c_double * 3), ("force", ctypes.c_double * 3), ("kinetic", ctypes.c_double),
                ("potential", ctypes.c_double)]

    def as_dict(self):
        return dict(momentum=list(self.momentum), force=list(self.force), kinetic=self.kinetic,
                    potential=self.potential)


class TpmStats(ctypes.Structure):
    _fields_ = [("candidate_pairs", ctypes.c_double), ("executed_pairs", ctypes.c_double),
                ("ms_chainmesh", ctypes.c_double), ("ms_pm", ctypes.c_double), ("ms_pp", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("n_active", ctypes.c_int64), ("n_tiles", ctypes.c_int64),
                ("max_rung", ctypes.c_int), ("pad_", ctypes.c_int), ("pc_interactions", ctypes.c_double),
                ("moments", TpmMoments)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pad_", "moments")}
        d["moments"] = self.moments.as_dict()
        return d


class TpmTuning(ctypes.Structure):
    _fields_ = [("tile_sinks", ctypes.c_int), ("deposit_block", ctypes.c_int), ("pp_waves", ctypes.c_int),
                ("pp_lists", ctypes.c_int), ("pp_newton", ctypes.c_int), ("fft", ctypes.c_int)]


TPM_BATCH_DRIFT = 8


class TpmBatchLimits(ctypes.Structure):
    _fields_ = [("max_particles", ctypes.c_int64), ("max_mesh_nodes", ctypes.c_int64),
                ("max_cells", ctypes.c_int64), ("max_boxes", ctypes.c_int32), ("order", ctypes.c_int32),
                ("interlace", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class TpmBoxSpec(ctypes.Structure):
    _fields_ = [("params", TpmParams), ("t0", ctypes.c_double), ("scale", ctypes.c_double),
                ("min_rung", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class TpmBatchTuning(ctypes.Structure):
    _fields_ = [("tile_sinks", ctypes.c_int), ("deposit_block", ctypes.c_int), ("sched", ctypes.c_int),
                ("graph", ctypes.c_int), ("pp_lists", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TpmBatchStats(ctypes.Structure):
    _fields_ = [("candidate_pairs", ctypes.c_double), ("executed_pairs", ctypes.c_double),
                ("ms_step", ctypes.c_double), ("ms_chainmesh", ctypes.c_double), ("ms_pm", ctypes.c_double),
                ("ms_pp", ctypes.c_double), ("n_particles", ctypes.c_int64), ("n_boxes", ctypes.c_int32),
                ("max_rung", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def box_spec(params, t0=1.0, scale=1.0, min_rung=0):
    """A tpm_box_spec: the box's parameters and its time stepping."""
    return TpmBoxSpec(params=params, t0=float(t0), scale=float(scale), min_rung=int(min_rung))


_lib = None
_P = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib_path():
    return _LIB


def _declared_symbols():
    txt = open(_HEADER).read()
    return sorted(set(re.findall(r"^\s*TPM_API\s+(?:tpm_status|const char\*|int64_t)\s+(tpm_\w+)\s*\(", txt, re.M)))


DECLARED_SYMBOLS = _declared_symbols()


def exported_symbols():
    """Names exported by the built library (via nm -D)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", _LIB], capture_output=True, text=True, check=True).stdout
    return sorted({ln.split()[-1] for ln in out.splitlines() if " T " in ln and ln.split()[-1].startswith("tpm_")})


def lib():
    """Load libtigerpm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise TpmError(-3, "libtigerpm_hip.so not built (run python -m tigerpm_amd.build)")
    L = ctypes.CDLL(_LIB)
    vp = ctypes.c_void_p
    sig = {
        "tpm_ctx_create": (_int, [_int, ctypes.POINTER(vp)]),
        "tpm_ctx_destroy": (_int, [vp]),
        "tpm_last_error": (ctypes.c_char_p, [vp]),
        "tpm_sync": (_int, [vp]),
        "tpm_set_params": (_int, [vp, ctypes.POINTER(TpmParams)]),
        "tpm_get_params": (_int, [vp, ctypes.POINTER(TpmParams)]),
        "tpm_particles_upload": (_int, [vp, _i64, vp, vp, vp, vp, vp, vp, vp]),
        "tpm_particles_download": (_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "tpm_particles_size": (_i64, [vp]),
        "tpm_chainmesh_create": (_int, [vp]),
        "tpm_chainmesh_get": (_int, [vp, _int, _int, _int, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
        "tpm_chainmesh_cells": (_int, [vp, _i64p]),
        "tpm_chainmesh_perm": (_int, [vp, _i64p]),
        "tpm_gravity_long_compute": (_int, [vp]),
        "tpm_gravity_long_get_phi": (_int, [vp, _f32p]),
        "tpm_gravity_long_get_density": (_int, [vp, _u64p]),
        "tpm_gravity_long_get_green": (_int, [vp, _f32p]),
        "tpm_fft3d_r2c": (_int, [vp, _int, _f32p, _f32p]),
        "tpm_fft3d_c2r": (_int, [vp, _int, _f32p, _f32p]),
        "tpm_kick": (_int, [vp, _int, _dbl, _dbl, _int, ctypes.c_uint, ctypes.POINTER(TpmStats)]),
        "tpm_kick_treepm": (_int, [vp, _int, _dbl, _dbl, _int, ctypes.POINTER(TpmStats)]),
        "tpm_tree_build": (_int, [vp]),
        "tpm_tree_info": (_int, [vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
        "tpm_tree_download": (_int, [vp, vp, vp, vp, vp, vp]),
        "tpm_ewald": (_int, [vp, _i64, _i64p, _dbl, _f64p, _f64p, _f64p, _f64p]),
        "tpm_ewald_at": (_int, [vp, _i64, _u32p, _u32p, _u32p, _dbl, _f64p, _f64p, _f64p, _f64p]),
        "tpm_ewald_at_noalloc": (_int, [vp, _i64, vp, vp, vp, _dbl, vp, vp, vp, vp]),
        "tpm_set_host_order": (_int, [vp, _int]),
        "tpm_kick_box": (_int, [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _int, _dbl, _dbl,
                                _int, ctypes.c_uint, ctypes.POINTER(TpmStats)]),
        "tpm_set_budget": (_int, [vp, _i64]),
        "tpm_tree_sort_particles": (_int, [vp]),
        "tpm_moments_get": (_int, [vp, ctypes.POINTER(TpmMoments)]),
        "tpm_batch_moments": (_int, [vp, _int, ctypes.POINTER(TpmMoments)]),
        "tpm_ic_zeldovich": (_int, [vp, _int, ctypes.c_uint64, _dbl, _dbl]),
        "tpm_ic_uniform": (_int, [vp, _i64, ctypes.c_uint64]),
        "tpm_set_tuning": (_int, [vp, ctypes.POINTER(TpmTuning)]),
        "tpm_get_tuning": (_int, [vp, ctypes.POINTER(TpmTuning)]),
        "tpm_profile": (_int, [vp, _int]),
        "tpm_profile_read": (_int, [vp, _f64p, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
        "tpm_probe_pair_rate": (_int, [vp, _int, ctypes.POINTER(ctypes.c_double)]),
        "tpm_batch_create": (_int, [_int, ctypes.POINTER(TpmBatchLimits), ctypes.POINTER(vp)]),
        "tpm_batch_destroy": (_int, [vp]),
        "tpm_batch_last_error": (ctypes.c_char_p, [vp]),
        "tpm_batch_arena_bytes": (_int, [ctypes.POINTER(TpmBatchLimits), ctypes.POINTER(_i64)]),
        "tpm_batch_fits": (_int, [vp, ctypes.POINTER(TpmBoxSpec), _i64, ctypes.POINTER(_int)]),
        "tpm_batch_add": (_int, [vp, ctypes.POINTER(TpmBoxSpec), _i64, vp, vp, vp, vp, vp, vp, vp,
                                 ctypes.POINTER(_int)]),
        "tpm_batch_add_zeldovich": (_int, [vp, ctypes.POINTER(TpmBoxSpec), _int, ctypes.c_uint64, _dbl, _dbl,
                                           ctypes.POINTER(_int)]),
        "tpm_batch_remove": (_int, [vp, _int]),
        "tpm_batch_prepare": (_int, [vp, ctypes.POINTER(TpmBoxSpec), _int]),
        "tpm_batch_step": (_int, [vp, ctypes.c_uint, ctypes.POINTER(TpmBatchStats)]),
        "tpm_batch_sync": (_int, [vp]),
        "tpm_batch_box_info": (_int, [vp, _int, ctypes.POINTER(_i64), ctypes.POINTER(_int)]),
        "tpm_batch_download": (_int, [vp, _int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "tpm_batch_to_ctx": (_int, [vp, _int, vp]),
        "tpm_batch_set_tuning": (_int, [vp, ctypes.POINTER(TpmBatchTuning)]),
        "tpm_batch_get_tuning": (_int, [vp, ctypes.POINTER(TpmBatchTuning)]),
        "tpm_batch_autotune": (_int, [vp, _int, ctypes.POINTER(TpmBatchTuning), ctypes.POINTER(_dbl)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One device-resident TreePM state (tpm_ctx)."""

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        s = L.tpm_ctx_create(int(device), ctypes.byref(h))
        if s != 0:
            raise TpmError(s, "tpm_ctx_create(%d) failed (no GPU visible?)" % device)
        self._h = h
        self._L = L
        self.params = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.tpm_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, s):
        if s != 0:
            raise TpmError(s, self._L.tpm_last_error(self._h).decode())

    def set_params(self, p):
        self._chk(self._L.tpm_set_params(self._h, ctypes.byref(p)))
        q = TpmParams()
        self._chk(self._L.tpm_get_params(self._h, ctypes.byref(q)))
        self.params = q
        return q

    def set_tuning(self, tile_sinks=16, deposit_block=0, pp_waves=0, pp_lists=-1, pp_newton=-1, fft=0):
        self._chk(self._L.tpm_set_tuning(self._h, ctypes.byref(TpmTuning(int(tile_sinks), int(deposit_block),
                                                                         int(pp_waves), int(pp_lists),
                                                                         int(pp_newton), int(fft)))))

    def get_tuning(self):
        t = TpmTuning()
        self._chk(self._L.tpm_get_tuning(self._h, ctypes.byref(t)))
        return dict(tile_sinks=t.tile_sinks, deposit_block=t.deposit_block, pp_waves=t.pp_waves,
                    pp_lists=t.pp_lists, pp_newton=t.pp_newton, fft=t.fft)

    def upload(self, x, y, z, vx=None, vy=None, vz=None, rung=None):
        x, y, z = [np.ascontiguousarray(a, np.uint32) for a in (x, y, z)]
        vs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
        r = None if rung is None else np.ascontiguousarray(rung, np.int8)
        self._keep = (x, y, z, vs, r)
        self._chk(self._L.tpm_particles_upload(self._h, len(x), _ptr(x), _ptr(y), _ptr(z), _ptr(vs[0]),
                                               _ptr(vs[1]), _ptr(vs[2]), _ptr(r)))

    def size(self):
        return int(self._L.tpm_particles_size(self._h))

    def download(self):
        n = self.size()
        out = dict(x=np.empty(n, np.uint32), y=np.empty(n, np.uint32), z=np.empty(n, np.uint32),
                   vx=np.empty(n, np.float32), vy=np.empty(n, np.float32), vz=np.empty(n, np.float32),
                   rung=np.empty(n, np.int8), gx=np.empty(n, np.float32), gy=np.empty(n, np.float32),
                   gz=np.empty(n, np.float32), pot=np.empty(n, np.float32))
        k = ["x", "y", "z", "vx", "vy", "vz", "rung", "gx", "gy", "gz", "pot"]
        self._chk(self._L.tpm_particles_download(self._h, *[_ptr(out[a]) for a in k]))
        return out

    def chainmesh_create(self):
        self._chk(self._L.tpm_chainmesh_create(self._h))

    def chainmesh_cells(self):
        C = self.params.chain_dim
        cb = np.empty(C ** 3 + 1, np.int64)
        self._chk(self._L.tpm_chainmesh_cells(self._h, cb))
        return cb

    def chainmesh_get(self, ix, iy, iz):
        b, e = _i64(), _i64()
        self._chk(self._L.tpm_chainmesh_get(self._h, ix, iy, iz, ctypes.byref(b), ctypes.byref(e)))
        return b.value, e.value

    def chainmesh_perm(self):
        p = np.empty(self.size(), np.int64)
        self._chk(self._L.tpm_chainmesh_perm(self._h, p))
        return p

    def gravity_long_compute(self):
        self._chk(self._L.tpm_gravity_long_compute(self._h))

    def get_phi(self):
        M = self.params.mesh_dim
        a = np.empty(M ** 3, np.float32)
        self._chk(self._L.tpm_gravity_long_get_phi(self._h, a))
        return a

    def get_density(self):
        M = self.params.mesh_dim
        a = np.empty(M ** 3, np.uint64)
        self._chk(self._L.tpm_gravity_long_get_density(self._h, a))
        return a

    def get_green(self):
        M = self.params.mesh_dim
        a = np.empty(M * M * (M // 2 + 1), np.float32)
        self._chk(self._L.tpm_gravity_long_get_green(self._h, a))
        return a

    def fft3d_r2c(self, a):
        """R2C of an (M, M, M) float32 mesh: (M, M, M//2+1) complex64, unnormalised."""
        M = a.shape[0]
        a = np.ascontiguousarray(a, np.float32)
        out = np.empty(M * M * (M // 2 + 1) * 2, np.float32)
        self._chk(self._L.tpm_fft3d_r2c(self._h, M, a.ravel(), out))
        return out.view(np.complex64).reshape(M, M, M // 2 + 1)

    def fft3d_c2r(self, c):
        """C2R of an (M, M, M//2+1) spectrum: (M, M, M) float32, unnormalised."""
        M = c.shape[0]
        c = np.ascontiguousarray(c, np.complex64)
        out = np.empty(M ** 3, np.float32)
        self._chk(self._L.tpm_fft3d_c2r(self._h, M, c.view(np.float32).ravel(), out))
        return out.reshape(M, M, M)

    def kick(self, min_rung=0, scale=1.0, t0=1.0, first_call=True, flags=0):
        st = TpmStats()
        self._chk(self._L.tpm_kick(self._h, int(min_rung), float(scale), float(t0), int(bool(first_call)),
                                   int(flags), ctypes.byref(st)))
        return st

    def kick_box(self, lo, hi, min_rung=0, scale=1.0, t0=1.0, first_call=True, flags=0):
        """The kick of the sinks in chain cells [lo, hi) (x, y, z)."""
        st = TpmStats()
        a = (ctypes.c_int32 * 3)(*[int(v) for v in lo])
        b = (ctypes.c_int32 * 3)(*[int(v) for v in hi])
        self._chk(self._L.tpm_kick_box(self._h, a, b, int(min_rung), float(scale), float(t0), int(bool(first_call)),
                                       int(flags), ctypes.byref(st)))
        return st

    def set_budget(self, nbytes):
        """Device budget of one kick's working set (0: none); kicks over it
        are split into sub-boxes."""
        self._chk(self._L.tpm_set_budget(self._h, int(nbytes)))

    def tree_build(self):
        self._chk(self._L.tpm_tree_build(self._h))

    def tree_download(self):
        """The per-cell trees: order (tree position -> chain-mesh index),
        node_begin, bucket_begin, nodes, buckets (as the oracle's tree_build)."""
        nn, nb = _i64(), _i64()
        self._chk(self._L.tpm_tree_info(self._h, ctypes.byref(nn), ctypes.byref(nb)))
        C3 = self.params.chain_dim ** 3
        out = dict(node_begin=np.empty(C3 + 1, np.int64), bucket_begin=np.empty(C3 + 1, np.int64),
                   nodes=np.empty(nn.value, TREE_NODE), buckets=np.empty(nb.value, TREE_BUCKET),
                   order=np.empty(self.size(), np.int32))
        self._chk(self._L.tpm_tree_download(self._h, *[_ptr(out[k]) for k in ("node_begin", "bucket_begin", "nodes",
                                                                             "buckets", "order")]))
        return out

    def tree_sort_particles(self):
        """The caller's order becomes the trees' order (TPM_ORDER_REFERENCE)."""
        self._chk(self._L.tpm_tree_sort_particles(self._h))

    def kick_treepm(self, min_rung=0, scale=1.0, t0=1.0, first_call=True, stats=True):
        st = TpmStats()
        self._chk(self._L.tpm_kick_treepm(self._h, int(min_rung), float(scale), float(t0), int(bool(first_call)),
                                          ctypes.byref(st) if stats else None))
        return st

    def ewald(self, samples, alpha=0.0):
        samples = np.ascontiguousarray(samples, np.int64)
        out = [np.empty(len(samples), np.float64) for _ in range(4)]
        self._chk(self._L.tpm_ewald(self._h, len(samples), samples, float(alpha), *out))
        return out

    def ewald_at(self, sx, sy, sz, alpha=0.0):
        """FP64 Ewald potential and force of all resident particles at fixed32
        positions (nothing left out; softened inside hsoft)."""
        sx, sy, sz = [np.ascontiguousarray(a, np.uint32) for a in (sx, sy, sz)]
        out = [np.empty(len(sx), np.float64) for _ in range(4)]
        self._chk(self._L.tpm_ewald_at(self._h, len(sx), sx, sy, sz, float(alpha), *out))
        return out

    def ewald_at_noalloc(self, n, d_sx, d_sy, d_sz, d_phi, d_gx, d_gy, d_gz, alpha=0.0):
        """tpm_ewald_at_noalloc on device addresses (ints): n positions in,
        four FP64 arrays of n out, enqueued on the context's stream (sync()
        waits)."""
        self._chk(self._L.tpm_ewald_at_noalloc(self._h, int(n), d_sx, d_sy, d_sz, float(alpha), d_phi, d_gx, d_gy,
                                               d_gz))

    def set_host_order(self, order):
        """TPM_ORDER_DEVICE or TPM_ORDER_REFERENCE (the reference partition's
        order after each chain mesh)."""
        self._chk(self._L.tpm_set_host_order(self._h, int(order)))

    def moments(self):
        m = TpmMoments()
        self._chk(self._L.tpm_moments_get(self._h, ctypes.byref(m)))
        return m.as_dict()

    def ic_zeldovich(self, n_side, seed, rms_disp, spectral_index=-2.0):
        self._chk(self._L.tpm_ic_zeldovich(self._h, int(n_side), int(seed), float(rms_disp), float(spectral_index)))

    def ic_uniform(self, n, seed):
        self._chk(self._L.tpm_ic_uniform(self._h, int(n), int(seed)))

    def profile(self, enable):
        self._chk(self._L.tpm_profile(self._h, int(bool(enable))))

    def profile_read(self):
        """Summed ms of recorded launches and their counts."""
        ms = np.zeros(4, np.float64)
        npp, nst = _i64(), _i64()
        self._chk(self._L.tpm_profile_read(self._h, ms, ctypes.byref(npp), ctypes.byref(nst)))
        return dict(chainmesh_ms=ms[0], pm_ms=ms[1], pp_ms=ms[2], step_ms=ms[3], pp_launches=npp.value,
                    steps=nst.value)

    def probe_pair_rate(self, waves_per_simd=8):
        """Pairs/s of the direct kernel's pair body alone (the ceiling of its
        pair loop at that occupancy)."""
        r = ctypes.c_double()
        self._chk(self._L.tpm_probe_pair_rate(self._h, int(waves_per_simd), ctypes.byref(r)))
        return r.value

    def sync(self):
        self._chk(self._L.tpm_sync(self._h))


def batch_arena_bytes(max_particles, max_mesh_nodes, max_boxes, order, interlace=0, max_cells=0):
    """Device bytes of the arenas of a batch with these limits."""
    lim = TpmBatchLimits(max_particles=int(max_particles), max_mesh_nodes=int(max_mesh_nodes),
                         max_cells=int(max_cells), max_boxes=int(max_boxes), order=int(order),
                         interlace=int(interlace))
    b = _i64()
    s = lib().tpm_batch_arena_bytes(ctypes.byref(lim), ctypes.byref(b))
    if s != 0:
        raise TpmError(s, "tpm_batch_arena_bytes")
    return b.value


class Batch:
    """Many independent boxes resident on one device, kicked together
    (tpm_batch)."""

    def __init__(self, max_particles, max_mesh_nodes, max_boxes, order, interlace=0, max_cells=0, device=0):
        L = lib()
        self._L = L
        self._h = None
        self.limits = TpmBatchLimits(max_particles=int(max_particles), max_mesh_nodes=int(max_mesh_nodes),
                                     max_cells=int(max_cells), max_boxes=int(max_boxes), order=int(order),
                                     interlace=int(interlace))
        h = ctypes.c_void_p()
        s = L.tpm_batch_create(int(device), ctypes.byref(self.limits), ctypes.byref(h))
        if s != 0:
            raise TpmError(s, "tpm_batch_create failed (no GPU visible, or the arenas do not fit)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.tpm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, s):
        if s != 0:
            raise TpmError(s, self._L.tpm_batch_last_error(self._h).decode())

    def fits(self, spec, n):
        f = _int()
        self._chk(self._L.tpm_batch_fits(self._h, ctypes.byref(spec), int(n), ctypes.byref(f)))
        return bool(f.value)

    def add(self, spec, x, y, z, vx=None, vy=None, vz=None, rung=None):
        x, y, z = [np.ascontiguousarray(a, np.uint32) for a in (x, y, z)]
        vs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
        r = None if rung is None else np.ascontiguousarray(rung, np.int8)
        i = _int()
        self._chk(self._L.tpm_batch_add(self._h, ctypes.byref(spec), len(x), _ptr(x), _ptr(y), _ptr(z),
                                        _ptr(vs[0]), _ptr(vs[1]), _ptr(vs[2]), _ptr(r), ctypes.byref(i)))
        return i.value

    def add_zeldovich(self, spec, n_side, seed, rms_disp, spectral_index=-2.0):
        i = _int()
        self._chk(self._L.tpm_batch_add_zeldovich(self._h, ctypes.byref(spec), int(n_side), int(seed),
                                                  float(rms_disp), float(spectral_index), ctypes.byref(i)))
        return i.value

    def prepare(self, spec, n_side=0):
        self._chk(self._L.tpm_batch_prepare(self._h, ctypes.byref(spec), int(n_side)))

    def remove(self, box):
        self._chk(self._L.tpm_batch_remove(self._h, int(box)))

    def step(self, flags=0, stats=True):
        st = TpmBatchStats()
        self._chk(self._L.tpm_batch_step(self._h, int(flags), ctypes.byref(st) if stats else None))
        return st if stats else None

    def sync(self):
        self._chk(self._L.tpm_batch_sync(self._h))

    def box_info(self, box):
        n, steps = _i64(), _int()
        self._chk(self._L.tpm_batch_box_info(self._h, int(box), ctypes.byref(n), ctypes.byref(steps)))
        return n.value, steps.value

    def download(self, box):
        n, _ = self.box_info(box)
        out = dict(x=np.empty(n, np.uint32), y=np.empty(n, np.uint32), z=np.empty(n, np.uint32),
                   vx=np.empty(n, np.float32), vy=np.empty(n, np.float32), vz=np.empty(n, np.float32),
                   rung=np.empty(n, np.int8), gx=np.empty(n, np.float32), gy=np.empty(n, np.float32),
                   gz=np.empty(n, np.float32), pot=np.empty(n, np.float32))
        k = ["x", "y", "z", "vx", "vy", "vz", "rung", "gx", "gy", "gz", "pot"]
        self._chk(self._L.tpm_batch_download(self._h, int(box), *[_ptr(out[a]) for a in k]))
        return out

    def moments(self, box):
        m = TpmMoments()
        self._chk(self._L.tpm_batch_moments(self._h, int(box), ctypes.byref(m)))
        return m.as_dict()

    def to_ctx(self, box, ctx):
        self._chk(self._L.tpm_batch_to_ctx(self._h, int(box), ctx._h))
        q = TpmParams()
        ctx._chk(ctx._L.tpm_get_params(ctx._h, ctypes.byref(q)))
        ctx.params = q

    def set_tuning(self, tile_sinks=16, deposit_block=0, sched=0, graph=0, pp_lists=-1):
        t = TpmBatchTuning(int(tile_sinks), int(deposit_block), int(sched), int(graph), int(pp_lists))
        self._chk(self._L.tpm_batch_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self):
        t = TpmBatchTuning()
        self._chk(self._L.tpm_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, reps=2):
        t = TpmBatchTuning()
        ms = _dbl()
        self._chk(self._L.tpm_batch_autotune(self._h, int(reps), ctypes.byref(t), ctypes.byref(ms)))
        return t.as_dict(), ms.value
