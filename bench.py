"""Benchmark: TreePM particle kicks/s on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one full kick of every particle of the box: chain mesh (sort +
cell table), PM (int64 deposit, R2C FFT, Green's function, C2R FFT) and the
fused short-range PP + mesh interpolation + kick kernel, with the particles
already resident in HBM.  value = N * K / (wall time of K steps).

Beside it, in the same run: the PP kernel's roofline (algorithmic FP32
flops = 44 per candidate pair of the 27-cell chain-mesh stencil, divided by
the kernel's HIP-event duration on its stream), the CPU restatement timed on
this host's cores on a bounded sample of the same box, and the rms force
error against the FP64 Ewald sum on sampled particles.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle kicks/s at 512^3 + batched-ensemble vs plain, rms err<1e-3 vs Ewald"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOPS_PER_PAIR = 44


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def cpu_baseline(ctx, p, budget_s, seed):
    """Times the CPU restatement on a bounded sample of the same box."""
    import oracle as O
    n = ctx.size()
    d = ctx.download()
    x, y, z = d["x"], d["y"], d["z"]
    cb = ctx.chainmesh_cells()
    op = O.make_params(p.mesh_dim, p.chain_dim, p.order, p.rs, p.rcut, p.hsoft, p.GM, p.eta, p.n_total, 1,
                       p.interlace)
    rng = np.random.default_rng(seed)
    threads = O.num_threads()
    # PP + interpolation over a random sample of sinks (27-cell stencil, the
    # reference pair formula in FP32), sized to ~budget/2 seconds.
    s0 = min(n, 4096)
    sinks = np.sort(rng.choice(n, s0, replace=False))
    t = time.perf_counter()
    O.pp_sum(op, x, y, z, cb, sinks, fp64=False)
    t_probe = time.perf_counter() - t
    s1 = int(min(n, max(s0, s0 * (0.5 * budget_s) / max(t_probe, 1e-6))))
    sinks = np.sort(rng.choice(n, s1, replace=False))
    t = time.perf_counter()
    O.pp_sum(op, x, y, z, cb, sinks, fp64=False)
    t_pp = (time.perf_counter() - t) / s1 * n
    # deposit over a particle sample, interpolation over the sinks
    M = p.mesh_dim
    nd = int(min(n, 1 << 22))
    sel = np.sort(rng.choice(n, nd, replace=False))
    t = time.perf_counter()
    O.deposit(x[sel], y[sel], z[sel], M, p.order)
    t_dep = (time.perf_counter() - t) / nd * n
    phi = np.zeros(M ** 3, np.float32)
    t = time.perf_counter()
    O.interp6(M, phi, x[sinks], y[sinks], z[sinks])
    t_int = (time.perf_counter() - t) / s1 * n
    # FFT + Green's on a smaller mesh, scaled by M^3 log M
    Ms = min(M, 128)
    rho = rng.random(Ms ** 3).astype(np.float32)
    t = time.perf_counter()
    tab = O.green_table(Ms, p.rs * M / Ms, p.order)
    O.pm_phi(Ms, rho, tab)
    t_fft = (time.perf_counter() - t) * (M / Ms) ** 3 * np.log(M) / np.log(Ms)
    total = t_pp + t_dep + t_int + t_fft
    return {
        "value": n / total, "unit": "particle kicks/s", "cores": threads, "kind": "port",
        "sample": ("oracle (C restatement, %d OpenMP threads) on the same box: PP+interp over %d random sinks "
                   "(27-cell stencil, FP32 reference pair formula), deposit over %d particles, FFT+Green at "
                   "%d^3 scaled by M^3 log M; per-stage seconds extrapolated to all %d particles"
                   % (threads, s1, nd, Ms, n)),
        "est_seconds_per_kick": {"pp": t_pp, "deposit": t_dep, "interp": t_int, "fft_green": t_fft},
    }


def copy_probe(nbytes):
    """One device-to-device hipMemcpy of nbytes (the runtime's copy kernel),
    the known-size reference for FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def accuracy(ctx, p, nsamp, seed):
    d = ctx.download()
    n = ctx.size()
    rng = np.random.default_rng(seed + 1000)
    s = np.sort(rng.choice(n, min(nsamp, n), replace=False))
    t = time.perf_counter()
    ph, gx, gy, gz = ctx.ewald(s)
    t_ew = time.perf_counter() - t
    G = np.stack([d["gx"][s], d["gy"][s], d["gz"][s]], 1).astype(np.float64)
    E = np.stack([gx, gy, gz], 1) * p.GM
    mag = np.linalg.norm(G, axis=1)
    magE = np.linalg.norm(E, axis=1)
    # The kick's potential includes the mesh self term -GM / (sqrt(pi) rs)
    # that the Ewald potential (own image + background) does not.
    P = d["pot"][s].astype(np.float64) + p.GM / (np.sqrt(np.pi) * p.rs)
    return {"samples": int(len(s)), "force_rms_vector": rel_rms(G, E),
            "force_rms_magnitude": float(np.sqrt(((mag - magE) ** 2).sum() / (magE ** 2).sum())),
            "potential_rms": rel_rms(P, ph * p.GM), "ewald_seconds": t_ew}


def main():
    import tigerpm_amd as T
    from tigerpm_amd.workloads import WORKLOADS

    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="zeldovich512_tsc", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=16.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ewald", action="store_true")
    ap.add_argument("--no-pcie", action="store_true", help="skip the host-buffer (PCIe-inclusive) measurement")
    ap.add_argument("--probe-copy", action="store_true",
                    help="also run one plain 1 GiB device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported (multi-GPU sharding is "
              "described in DESIGN.md, not implemented)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    w = WORKLOADS[args.workload]
    ctx = T.Context(0)
    p = ctx.set_params(T.baseline_params(w["n_side"], w["mesh"], w["order"], interlace=w["interlace"]))
    log("workload", args.workload, "params", p.as_dict())
    t = time.perf_counter()
    if w["ic"] == "zeldovich":
        ctx.ic_zeldovich(w["n_side"], w["seed"], w["rms"], -2.0)
    else:
        ctx.ic_uniform(w["n_side"] ** 3, w["seed"])
    ctx.sync()
    n = ctx.size()
    log("initial conditions: %d particles in %.2f s" % (n, time.perf_counter() - t))

    st = None
    for i in range(args.warmup):
        st = ctx.kick_treepm(0, 1.0, 1.0, True, stats=True)
        log("warmup %d: %.1f ms (chain %.1f, pm %.1f, pp %.1f); candidate pairs %.4g, executed %.4g"
            % (i, st.ms_total, st.ms_chainmesh, st.ms_pm, st.ms_pp, st.candidate_pairs, st.executed_pairs))
    if st is None:
        st = ctx.kick_treepm(0, 1.0, 1.0, True, stats=True)

    ctx.profile(True)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.kick_treepm(0, 1.0, 1.0, True, stats=False)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    log("timed %d steps in %.3f s; events: %s" % (args.steps, elapsed, prof))

    ms_step = elapsed / args.steps * 1e3
    pp_ms = prof["pp_ms"] / max(prof["pp_launches"], 1)
    W = st.candidate_pairs
    achieved = FLOPS_PER_PAIR * W / (pp_ms * 1e-3) / 1e12
    executed = FLOPS_PER_PAIR * st.executed_pairs / (pp_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": n * args.steps / elapsed,
        "unit": "particle kicks/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "n_particles": n, "mesh": p.mesh_dim, "chain_dim": p.chain_dim,
                   "assignment": {2: "CIC", 3: "TSC", 4: "cloud4"}[p.order], "rs_cells": p.rs * p.mesh_dim,
                   "rcut_over_rs": p.rcut / p.rs, "ic": w["ic"], "rms_disp_cells": w["rms"], "seed": w["seed"]},
        # Bound vocabulary is hbm|mfma: the PP kernel is FP32 compute bound and
        # its peak is MI355X's FP32 figure (157.3 TFLOP/s, vector and matrix).
        "roofline": {"bound": "mfma", "bound_note": "FP32 vector ALU peak; the kernel issues no MFMA",
                     "kernel": "pp_kick_kernel",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "flops_per_launch": FLOPS_PER_PAIR * W, "kernel_ms": pp_ms,
                     "candidate_pairs": W, "executed_pairs": st.executed_pairs,
                     "executed_tflops": executed},
        "stages_ms": {"chainmesh": prof["chainmesh_ms"] / max(prof["steps"], 1),
                      "pm": prof["pm_ms"] / max(prof["steps"], 1), "pp_kick": pp_ms,
                      "step": prof["step_ms"] / max(prof["steps"], 1)},
    }
    # PCIe-inclusive rate of the host-buffer boundary (not the metric): host
    # SoA arrays in, one full kick, velocities/rungs/forces/potential out.
    if not args.no_pcie:
        h = ctx.download()
        t = time.perf_counter()
        ctx.upload(h["x"], h["y"], h["z"], h["vx"], h["vy"], h["vz"], h["rung"])
        ctx.kick_treepm(0, 1.0, 1.0, True, stats=False)
        ctx.download()
        tp = time.perf_counter() - t
        out["pcie_inclusive"] = {"kicks_per_s": n / tp, "seconds": tp,
                                 "note": "tpm_particles_upload + tpm_kick_treepm + tpm_particles_download, "
                                         "pageable host memory"}
        log("PCIe-inclusive", out["pcie_inclusive"])
        del h
    # HBM bytes per PP launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs (scaled by a plain copy of known size).
    tfile = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        out["roofline"]["traffic"] = t["traffic_bytes_per_launch"]
        out["roofline"]["traffic_source"] = t["source"]
    if not args.no_ewald:
        log("FP64 Ewald on %d samples..." % w["samples"])
        out["accuracy"] = accuracy(ctx, p, w["samples"], w["seed"])
        log("accuracy", out["accuracy"])
    if not args.no_cpu:
        log("CPU baseline (budget %.0f s)..." % args.cpu_budget)
        out["cpu_baseline"] = cpu_baseline(ctx, p, args.cpu_budget, w["seed"])
        log("cpu", out["cpu_baseline"])
    ctx.close()
    if args.probe_copy:
        out["copy_probe_bytes"] = copy_probe(1 << 30)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
