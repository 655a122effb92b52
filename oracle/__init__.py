"""CPU oracle for the tigerpm TreePM kick path -- TEST INFRASTRUCTURE ONLY.

Thin ctypes wrapper over ``liboracle_tpm.so`` (built from ``tpm_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline; the product path never calls it.

Every function restates a reference file:line; see ``tpm_oracle.c``.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("TPM_ORACLE_LIB") or os.path.join(_HERE, "liboracle_tpm.so")

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("mesh_dim", c_int),
        ("chain_dim", c_int),
        ("order", c_int),
        ("do_phi", c_int),
        ("interlace", c_int),
        ("pme", c_int),
        ("rs", c_dbl),
        ("rcut", c_dbl),
        ("hsoft", c_dbl),
        ("GM", c_dbl),
        ("eta", c_dbl),
        ("n_total", c_dbl),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_fixed_from_double": (ctypes.c_uint32, [c_dbl]),
            "orc_fixed_to_float": (ctypes.c_float, [ctypes.c_uint32]),
            "orc_fixed_to_double": (c_dbl, [ctypes.c_uint32]),
            "orc_distance": (ctypes.c_float, [ctypes.c_uint32, ctypes.c_uint32]),
            "orc_chain_bounds": (None, [c_int, u64p]),
            "orc_chain_coord": (c_int, [ctypes.c_uint32, c_int]),
            "orc_chain_keys": (None, [c_i64, u32p, u32p, u32p, c_int, u32p]),
            "orc_chainmesh_sort": (None, [c_i64, u32p, u32p, u32p, c_int, i64p, i64p]),
            "orc_chainmesh_partition_ref": (None, [c_i64, u32p, u32p, u32p, c_int, i64p, i64p, i64p]),
            "orc_weights_1d": (c_int, [c_int, ctypes.c_uint32, c_int, ctypes.POINTER(c_i64), u64p,
                                       ctypes.POINTER(c_int)]),
            "orc_deposit": (None, [c_i64, u32p, u32p, u32p, c_int, c_int, u64p]),
            "orc_deposit_shifted": (None, [c_i64, u32p, u32p, u32p, c_int, c_int, ctypes.c_uint64, u64p]),
            "orc_pm_phi_interlaced": (None, [c_int, f32p, f32p, f32p, f32p]),
            "orc_density_float": (None, [c_i64, u64p, f32p]),
            "orc_fft3d_r2c": (None, [c_int, f64p, f64p]),
            "orc_fft3d_c2r": (None, [c_int, f64p, f64p]),
            "orc_green_table": (None, [c_int, c_dbl, c_int, f32p]),
            "orc_pm_phi": (None, [c_int, f32p, f32p, f32p]),
            "orc_interp6_weights": (None, [ctypes.c_float, f32p, f32p]),
            "orc_interp6": (None, [c_int, f32p, c_i64, u32p, u32p, u32p, f32p, f32p, f32p, f32p]),
            "orc_pp_pair": (None, [ctypes.POINTER(OrcParams), c_dbl, c_dbl, c_dbl, c_int, f64p]),
            "orc_pp_sum": (c_dbl, [ctypes.POINTER(OrcParams), c_i64, u32p, u32p, u32p, i64p, c_i64, i64p,
                                   c_int, f32p, f32p, f32p, f32p]),
            "orc_candidate_pairs": (c_dbl, [c_int, i64p]),
            "orc_drift": (None, [c_i64, u32p, u32p, u32p, f32p, f32p, f32p, c_dbl]),
            "orc_kick_epilogue": (None, [ctypes.POINTER(OrcParams), c_i64, f32p, f32p, f32p, f32p, f32p, f32p,
                                         f32p, i8p, c_int, c_dbl, c_dbl, c_int]),
            "orc_kick_full": (None, [ctypes.POINTER(OrcParams), c_i64, u32p, u32p, u32p, f32p, f32p, f32p, i8p,
                                     f32p, f32p, f32p, f32p, i64p, c_int, c_dbl, c_dbl, c_int, c_int, f64p]),
            "orc_ewald_psi": (None, [c_dbl, c_dbl, c_dbl, c_dbl, c_int, c_int, c_int, f64p]),
            "orc_ewald_direct": (None, [c_i64, u32p, u32p, u32p, c_i64, i64p, c_dbl, f64p, f64p, f64p, f64p]),
            "orc_ewald_sf": (None, [c_i64, u32p, u32p, u32p, c_i64, i64p, c_dbl, c_dbl, f64p, f64p, f64p,
                                    f64p]),
            "orc_ewald_sf_at": (None, [c_i64, u32p, u32p, u32p, c_i64, u32p, u32p, u32p, c_dbl, c_dbl, f64p, f64p,
                                       f64p, f64p]),
            "orc_ewald_direct_at": (None, [c_i64, u32p, u32p, u32p, c_i64, u32p, u32p, u32p, c_dbl, f64p, f64p,
                                           f64p, f64p]),
            "orc_tree_build": (c_int, [ctypes.POINTER(OrcParams), c_i64, u32p, u32p, u32p, i64p, i32p, u32p, u32p,
                                       u32p, i64p, i64p, ctypes.c_void_p, c_i64, ctypes.c_void_p, c_i64]),
            "orc_tree_sum": (c_int, [ctypes.POINTER(OrcParams), c_i64, u32p, u32p, u32p, i64p, i8p, c_int, c_int,
                                     f32p, f32p, f32p, f32p, f64p]),
            "orc_pc_pair": (None, [ctypes.POINTER(OrcParams), ctypes.c_float, ctypes.c_float, ctypes.c_float, f32p,
                                   f32p]),
            "orc_num_threads": (c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def make_params(mesh_dim, chain_dim, order, rs, rcut, hsoft, GM, eta, n_total, do_phi=1, interlace=0, pme=0):
    """pme = 1: kick_pme's pair law (no cutoff, plain softening, no phi0)."""
    return OrcParams(int(mesh_dim), int(chain_dim), int(order), int(do_phi), int(interlace), int(pme), float(rs),
                     float(rcut), float(hsoft), float(GM), float(eta), float(n_total))


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def chain_keys(x, y, z, C):
    x, y, z = _u32(x), _u32(y), _u32(z)
    key = np.empty(len(x), np.uint32)
    lib().orc_chain_keys(len(x), x, y, z, C, key)
    return key


def chainmesh_sort(x, y, z, C):
    """Returns (perm, cell_begin[C^3+1]); perm[new] = old."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    perm = np.empty(len(x), np.int64)
    cb = np.empty(C ** 3 + 1, np.int64)
    lib().orc_chainmesh_sort(len(x), x, y, z, C, perm, cb)
    return perm, cb


def chainmesh_partition_ref(x, y, z, C):
    """The reference's recursive partition. Returns (idx, cell_begin, cell_end)."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    idx = np.empty(len(x), np.int64)
    cb = np.full(C ** 3, -1, np.int64)
    ce = np.full(C ** 3, -1, np.int64)
    lib().orc_chainmesh_partition_ref(len(x), x, y, z, C, idx, cb, ce)
    return idx, cb, ce


def weights_1d(order, raw, M):
    base = c_i64()
    w = np.zeros(4, np.uint64)
    q = c_int()
    n = lib().orc_weights_1d(order, int(raw), M, ctypes.byref(base), w, ctypes.byref(q))
    return base.value, w[:n].copy(), q.value


def deposit(x, y, z, M, order):
    x, y, z = _u32(x), _u32(y), _u32(z)
    g = np.empty(M ** 3, np.uint64)
    lib().orc_deposit(len(x), x, y, z, M, order, g)
    return g


def deposit_shifted(x, y, z, M, order, shift):
    x, y, z = _u32(x), _u32(y), _u32(z)
    g = np.empty(M ** 3, np.uint64)
    lib().orc_deposit_shifted(len(x), x, y, z, M, order, int(shift), g)
    return g


def pm_phi_interlaced(M, rho0, rho1, table):
    phi = np.empty(M ** 3, np.float32)
    lib().orc_pm_phi_interlaced(M, np.ascontiguousarray(rho0, np.float32), np.ascontiguousarray(rho1, np.float32),
                                np.ascontiguousarray(table, np.float32), phi)
    return phi


def density_float(grid):
    grid = np.ascontiguousarray(grid, np.uint64)
    rho = np.empty(grid.size, np.float32)
    lib().orc_density_float(grid.size, grid, rho)
    return rho


def fft3d_r2c(a):
    M = a.shape[0]
    a = np.ascontiguousarray(a, np.float64)
    out = np.empty(M * M * (M // 2 + 1) * 2, np.float64)
    lib().orc_fft3d_r2c(M, a.reshape(-1), out)
    return out.view(np.complex128).reshape(M, M, M // 2 + 1)


def fft3d_c2r(c, M):
    c = np.ascontiguousarray(c, np.complex128)
    out = np.empty(M ** 3, np.float64)
    lib().orc_fft3d_c2r(M, c.view(np.float64).reshape(-1), out)
    return out.reshape(M, M, M)


def green_table(M, rs, order):
    t = np.empty(M * M * (M // 2 + 1), np.float32)
    lib().orc_green_table(M, rs, order, t)
    return t


def pm_phi(M, rho, table):
    phi = np.empty(M ** 3, np.float32)
    lib().orc_pm_phi(M, np.ascontiguousarray(rho, np.float32), np.ascontiguousarray(table, np.float32), phi)
    return phi


def interp6_weights(x):
    w = np.empty(6, np.float32)
    dw = np.empty(6, np.float32)
    lib().orc_interp6_weights(float(x), w, dw)
    return w, dw


def interp6(M, phi, x, y, z):
    x, y, z = _u32(x), _u32(y), _u32(z)
    n = len(x)
    out = [np.empty(n, np.float32) for _ in range(4)]
    lib().orc_interp6(M, np.ascontiguousarray(phi, np.float32), n, x, y, z, *out)
    return out


def pp_pair(params, dx, dy, dz, fp64):
    out = np.empty(4, np.float64)
    lib().orc_pp_pair(ctypes.byref(params), dx, dy, dz, int(fp64), out)
    return out


def pp_sum(params, x, y, z, cell_begin, sinks, fp64):
    x, y, z = _u32(x), _u32(y), _u32(z)
    sinks = np.ascontiguousarray(sinks, np.int64)
    out = [np.empty(len(sinks), np.float32) for _ in range(4)]
    cand = lib().orc_pp_sum(ctypes.byref(params), len(x), x, y, z, np.ascontiguousarray(cell_begin, np.int64),
                            len(sinks), sinks, int(fp64), *out)
    return out, cand


def candidate_pairs(C, cell_begin):
    return lib().orc_candidate_pairs(C, np.ascontiguousarray(cell_begin, np.int64))


def kick_epilogue(params, g, pot, v, rung, min_rung, scale, t0, first_call):
    """In-place on copies; returns (g[3], pot, v[3], rung)."""
    gx, gy, gz = [np.array(a, np.float32) for a in g]
    pot = np.array(pot, np.float32)
    vx, vy, vz = [np.array(a, np.float32) for a in v]
    rung = np.array(rung, np.int8)
    lib().orc_kick_epilogue(ctypes.byref(params), len(gx), gx, gy, gz, pot, vx, vy, vz, rung, min_rung, scale,
                            t0, int(first_call))
    return (gx, gy, gz), pot, (vx, vy, vz), rung


def drift(x, y, z, vx, vy, vz, t0):
    """Fixed-point drift x + v t0 (periodic); returns new arrays."""
    x, y, z = [np.array(a, np.uint32) for a in (x, y, z)]
    vx, vy, vz = [np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
    lib().orc_drift(len(x), x, y, z, vx, vy, vz, float(t0))
    return x, y, z


def kick_full(params, x, y, z, vx, vy, vz, rung, min_rung=0, scale=1.0, t0=1.0, first_call=True, fp64_pp=False,
              tree=False):
    """Whole CPU kick. Returns dict of sorted arrays and stage times.
    fp64_pp: the direct sum in FP64; tree: kick_treepm's tree walk."""
    x, y, z = [np.array(a, np.uint32) for a in (x, y, z)]
    vx, vy, vz = [np.array(a, np.float32) for a in (vx, vy, vz)]
    rung = np.array(rung, np.int8)
    n = len(x)
    g = [np.zeros(n, np.float32) for _ in range(4)]
    C = params.chain_dim
    cb = np.empty(C ** 3 + 1, np.int64)
    times = np.zeros(5, np.float64)
    lib().orc_kick_full(ctypes.byref(params), n, x, y, z, vx, vy, vz, rung, g[0], g[1], g[2], g[3], cb, min_rung,
                        scale, t0, int(first_call), 2 if tree else int(fp64_pp), times)
    if times[3] < 0:
        raise RuntimeError("tree depth limit exceeded")
    return dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, rung=rung, gx=g[0], gy=g[1], gz=g[2], pot=g[3],
                cell_begin=cb, times=times)


def ewald_psi(dx, dy, dz, alpha=2.0, nimg=4, hmax=4, self_term=False):
    out = np.empty(4, np.float64)
    lib().orc_ewald_psi(dx, dy, dz, alpha, nimg, hmax, int(self_term), out)
    return out


def ewald_direct(x, y, z, samples, hsoft):
    x, y, z = _u32(x), _u32(y), _u32(z)
    samples = np.ascontiguousarray(samples, np.int64)
    out = [np.empty(len(samples), np.float64) for _ in range(4)]
    lib().orc_ewald_direct(len(x), x, y, z, len(samples), samples, hsoft, *out)
    return out


def ewald_sf(x, y, z, samples, hsoft, alpha=10.0):
    x, y, z = _u32(x), _u32(y), _u32(z)
    samples = np.ascontiguousarray(samples, np.int64)
    out = [np.empty(len(samples), np.float64) for _ in range(4)]
    lib().orc_ewald_sf(len(x), x, y, z, len(samples), samples, hsoft, alpha, *out)
    return out


def ewald_sf_at(x, y, z, sx, sy, sz, hsoft, alpha=10.0):
    """Ewald potential and force of all particles at fixed32 positions."""
    x, y, z, sx, sy, sz = _u32(x), _u32(y), _u32(z), _u32(sx), _u32(sy), _u32(sz)
    out = [np.empty(len(sx), np.float64) for _ in range(4)]
    lib().orc_ewald_sf_at(len(x), x, y, z, len(sx), sx, sy, sz, hsoft, alpha, *out)
    return out


def ewald_direct_at(x, y, z, sx, sy, sz, hsoft):
    x, y, z, sx, sy, sz = _u32(x), _u32(y), _u32(z), _u32(sx), _u32(sy), _u32(sz)
    out = [np.empty(len(sx), np.float64) for _ in range(4)]
    lib().orc_ewald_direct_at(len(x), x, y, z, len(sx), sx, sy, sz, hsoft, *out)
    return out


# Tree node and sink bucket of the per-cell trees (tree.hpp:17-36).
TREE_NODE = np.dtype([("x", "<u4", 3), ("radius", "<f4"), ("child", "<i4", 2), ("begin", "<i4"), ("end", "<i4"),
                      ("m", "<f4", 4)])
TREE_BUCKET = np.dtype([("begin", "<i4"), ("end", "<i4"), ("x", "<u4", 3), ("radius", "<f4")])


def tree_build(params, x, y, z, cell_begin):
    """Per-cell trees of the chain-mesh-sorted particles.  Returns a dict:
    order (tree position -> input index), tx/ty/tz, node_begin, bucket_begin,
    nodes, buckets.  Raises on a tree deeper than the reference allows."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    n = len(x)
    C3 = params.chain_dim ** 3
    cb = np.ascontiguousarray(cell_begin, np.int64)
    order = np.empty(n, np.int32)
    t = [np.empty(n, np.uint32) for _ in range(3)]
    nb, bb = np.empty(C3 + 1, np.int64), np.empty(C3 + 1, np.int64)
    rc = lib().orc_tree_build(ctypes.byref(params), n, x, y, z, cb, order, *t, nb, bb, None, 0, None, 0)
    if rc == 1:
        raise RuntimeError("tree depth limit exceeded")
    nodes = np.empty(int(nb[-1]), TREE_NODE)
    buckets = np.empty(int(bb[-1]), TREE_BUCKET)
    rc = lib().orc_tree_build(ctypes.byref(params), n, x, y, z, cb, order, *t, nb, bb, nodes.ctypes.data, len(nodes),
                              buckets.ctypes.data, len(buckets))
    assert rc == 0
    return dict(order=order, tx=t[0], ty=t[1], tz=t[2], node_begin=nb, bucket_begin=bb, nodes=nodes,
                buckets=buckets)


def tree_sum(params, x, y, z, cell_begin, rung, min_rung=0, pc_exact=False):
    """kick_treepm's short-range sum (tree walk, PC + PP) for the active
    sinks of the chain-mesh-sorted particles; unit-mass sums in input order
    and stats {pc, pp, multi, part}.  pc_exact: the multipole-list nodes
    summed pair by pair instead (a check of the walk's coverage)."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    n = len(x)
    out = [np.empty(n, np.float32) for _ in range(4)]
    st = np.zeros(4, np.float64)
    rc = lib().orc_tree_sum(ctypes.byref(params), n, x, y, z, np.ascontiguousarray(cell_begin, np.int64),
                            np.ascontiguousarray(rung, np.int8), int(min_rung), int(pc_exact), *out, st)
    if rc:
        raise RuntimeError("tree depth limit exceeded")
    return out, dict(pc=st[0], pp=st[1], multi=st[2], part=st[3])


def pc_pair(params, dx, dy, dz, M):
    """One multipole (PC) interaction; M = (mass, -dipole)."""
    out = np.empty(4, np.float32)
    lib().orc_pc_pair(ctypes.byref(params), float(dx), float(dy), float(dz), np.ascontiguousarray(M, np.float32), out)
    return out


def num_threads():
    return lib().orc_num_threads()
