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

Two more legs (keys "ensemble" and "serving" of the same JSON line):
64 independent boxes of N log-uniform in [32^3, 128^3] kicked by the plain
per-box loop and by one batch (autotuned, graph-captured), and a
continuous-batching server under Poisson arrivals of boxes with N
log-uniform in [32^3, 256^3], 10 resident kick + drift steps each, with
p50/p99 latency per arrival rate.
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


def tree_leg(T, ctx, p, steps, nsamp, seed, ewald=True):
    """The same particles through kick_treepm's own algorithm (TPM_KICK_TREE:
    per-cell trees, theta = 0.4 monopole + dipole, PP for near leaves): one
    step = chain mesh + PM + tree build + walk/kick, each stage waited for;
    forces against the FP64 Ewald sum.  Not the benchmark value."""
    rows = []
    st = None
    for k in range(steps + 1):
        ctx.sync()
        t = time.perf_counter()
        ctx.chainmesh_create()
        ctx.gravity_long_compute()
        ctx.sync()
        t1 = time.perf_counter()
        ctx.tree_build()
        ctx.sync()
        t2 = time.perf_counter()
        st = ctx.kick(0, 1.0, 1.0, True, flags=T.TPM_KICK_TREE)
        t3 = time.perf_counter()
        if k:  # the first step is a warm-up
            rows.append((t1 - t, t2 - t1, t3 - t2, st.ms_pp, t3 - t))
    r = np.mean(np.array(rows), 0)
    out = {"steps": steps, "kicks_per_s": ctx.size() / r[4], "ms_per_step": r[4] * 1e3,
           "stages_ms": {"chainmesh_pm": r[0] * 1e3, "tree_build": r[1] * 1e3, "walk_kick": r[2] * 1e3,
                         "walk_kernel_events": r[3]},
           "pp_pairs": st.executed_pairs, "pc_interactions": st.pc_interactions,
           "candidate_pairs": st.candidate_pairs, "sink_buckets": int(st.n_tiles)}
    if ewald:
        out["accuracy"] = accuracy(ctx, p, nsamp, seed)
    return out


def useful_pairs(ctx, p, nsamp, seed):
    """Pairs inside rcut per particle, counted over sampled sinks from the
    downloaded positions (minimum-image int32 offsets, as the kernel forms
    them) and the 27 chain-mesh cells around each sink; beside it the
    kernel's candidate and executed pairs per particle."""
    d = ctx.download()
    n = ctx.size()
    C = p.chain_dim
    cb = ctx.chainmesh_cells()
    x, y, z = d["x"], d["y"], d["z"]
    rng = np.random.default_rng(seed + 2000)
    sinks = np.sort(rng.choice(n, min(nsamp, n), replace=False))
    cell = np.searchsorted(cb, sinks, side="right") - 1
    rc2 = (p.rcut * 2.0 ** 32) ** 2
    counts = np.empty(len(sinks), np.int64)
    for k, (i, c) in enumerate(zip(sinks, cell)):
        cx, cy, cz = c % C, (c // C) % C, c // (C * C)
        idx = np.concatenate([np.arange(cb[j], cb[j + 1]) for j in
                              [(((cz + dz) % C) * C + (cy + dy) % C) * C + (cx + dx) % C
                               for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]])
        r2 = np.zeros(len(idx))
        for a in (x, y, z):
            off = (a[idx] - a[i]).view(np.int32).astype(np.float64)
            r2 += off * off
        counts[k] = int((r2 < rc2).sum()) - 1  # the sink itself
    return {"samples": int(len(sinks)), "per_particle": float(counts.mean()),
            "total_estimate": float(counts.mean()) * n,
            "note": "pairs with r < rcut (sampled sinks); the roofline credits every candidate pair"}


def ewald_error(ctx, d, nsamp, seed):
    """Vector rms force error of downloaded state d against the FP64 Ewald
    sum of the same particles (resident in ctx)."""
    rng = np.random.default_rng(seed)
    s = np.sort(rng.choice(len(d["x"]), min(nsamp, len(d["x"])), replace=False))
    ph, gx, gy, gz = ctx.ewald(s)
    G = np.stack([d["gx"][s], d["gy"][s], d["gz"][s]], 1).astype(np.float64)
    E = np.stack([gx, gy, gz], 1) * ctx.params.GM
    return rel_rms(G, E)


def ensemble_leg(T, steps, seed=1234, nbox=64, graph=1):
    """64 independent boxes, N log-uniform in [32^3, 128^3] (n_side
    log-uniform, mesh = n_side), rms displacement U[0.3, 2] cells, own seeds;
    cloud4 assignment (the reference's default; as accurate as interlaced TSC
    on these boxes with one forward FFT instead of two).  The same kicks through (a) the plain per-box drop-in
    loop -- one resident context per box, tpm_kick_treepm and a wait per box
    -- and (b) one tpm_batch: autotuned launch shape, the step captured as a
    HIP graph and replayed."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(32), np.log(128), nbox))).astype(int)
    rms = rng.uniform(0.3, 2.0, nbox)
    total = int((ns.astype(np.int64) ** 3).sum())
    specs = [T.box_spec(T.baseline_params(int(n), int(n), 4), t0=1.0) for n in ns]
    bt = T.Batch(total, total, nbox, 4, 0)
    t = time.perf_counter()
    ids = [bt.add_zeldovich(sp, int(n), 1000 + k, float(r)) for k, (sp, n, r) in enumerate(zip(specs, ns, rms))]
    bt.sync()
    log("ensemble: %d boxes, %d particles, initial conditions %.2f s" % (nbox, total, time.perf_counter() - t))
    ctxs = []
    for i in ids:
        c = T.Context(0)
        bt.to_ctx(i, c)
        ctxs.append(c)
    # (a) the plain loop
    for c in ctxs:
        c.kick_treepm(0, 1.0, 1.0, True, stats=False)
        c.sync()
    t = time.perf_counter()
    for k in range(steps):
        for c in ctxs:
            c.kick_treepm(0, 1.0, 1.0, False, stats=False)
            c.sync()
    plain = total * steps / (time.perf_counter() - t)
    t = time.perf_counter()
    for k in range(steps):
        for c in ctxs:
            c.kick_treepm(0, 1.0, 1.0, False, stats=False)
    for c in ctxs:
        c.sync()
    plain_async = total * steps / (time.perf_counter() - t)
    # the drop-in boundary with host-resident boxes: upload, kick, download
    host = [c.download() for c in ctxs]
    t = time.perf_counter()
    for c, h in zip(ctxs, host):
        c.upload(h["x"], h["y"], h["z"], h["vx"], h["vy"], h["vz"], h["rung"])
        c.kick_treepm(0, 1.0, 1.0, False, stats=False)
        c.download()
    plain_host = total / (time.perf_counter() - t)
    del host
    for c in ctxs:
        c.close()
    log("ensemble plain loop: %.4g kicks/s (wait per box), %.4g (no wait), host-resident %.4g"
        % (plain, plain_async, plain_host))
    # (b) the batch
    t = time.perf_counter()
    best, ms_best = bt.autotune(reps=2)
    t_tune = time.perf_counter() - t
    stage = bt.step(0, stats=True)  # eager: stage breakdown
    log("ensemble batch step (eager) %.1f ms: chain mesh %.1f, PM %.1f, PP %.1f"
        % (stage.ms_step, stage.ms_chainmesh, stage.ms_pm, stage.ms_pp))
    best["graph"] = graph
    bt.set_tuning(**best)
    bt.step(0, stats=False)
    bt.sync()
    t = time.perf_counter()
    for k in range(steps):
        bt.step(0, stats=False)
    bt.sync()
    batch = total * steps / (time.perf_counter() - t)
    lat = []
    for k in range(steps):
        t = time.perf_counter()
        bt.step(0, stats=False)
        bt.sync()
        lat.append(time.perf_counter() - t)
    st = bt.step(0, stats=True)
    log("ensemble batch: %.4g kicks/s, tuning %s (autotune %.1f s), step %.1f ms"
        % (batch, best, t_tune, st.ms_step))
    # accuracy of four boxes against FP64 Ewald
    order = np.argsort(ns)
    errs = {}
    for k in (order[0], order[nbox // 3], order[2 * nbox // 3], order[-1]):
        with T.Context(0) as c:
            bt.to_ctx(ids[k], c)
            errs["n_side_%d" % ns[k]] = ewald_error(c, bt.download(ids[k]), 512, int(k))
    bt.close()
    log("ensemble accuracy", errs)
    return {
        "boxes": nbox, "n_particles": total, "n_side_min": int(ns.min()), "n_side_max": int(ns.max()),
        "assignment": "cloud4", "steps": steps,
        "batch_kicks_per_s": batch,
        # the drop-in path as the reference is called: particles in host
        # arrays, one kick_treepm per box (upload, kick, download)
        "plain_dropin_loop_kicks_per_s": plain_host, "speedup_vs_plain_dropin_loop": batch / plain_host,
        # the same loop with every box resident in its own context and a
        # wait after each box's kick
        "plain_resident_loop_kicks_per_s": plain, "speedup_vs_plain_resident_loop": batch / plain,
        # ... and without the waits (the contexts' streams overlap)
        "plain_resident_overlapped_kicks_per_s": plain_async,
        "speedup_vs_plain_resident_overlapped": batch / plain_async,
        "tuning": best, "autotune_s": t_tune,
        "eager_step_ms": {"step": stage.ms_step, "chainmesh": stage.ms_chainmesh, "pm": stage.ms_pm,
                          "pp": stage.ms_pp},
        "step_ms_p50": 1e3 * float(np.percentile(lat, 50)),
        "step_ms_p99": 1e3 * float(np.percentile(lat, 99)),
        "candidate_pairs": st.candidate_pairs, "executed_pairs": st.executed_pairs,
        "pp_credited_tflops": FLOPS_PER_PAIR * st.candidate_pairs / (st.ms_step * 1e-3) / 1e12,
        "force_rms_vector_vs_ewald": errs,
    }


def serving_leg(T, rates, count, budget_gib, seed=77):
    """Poisson arrivals at each rate, N log-uniform in [32^3, 256^3], 10
    kick + drift steps per box with the state resident, admission by the
    arena budget; latency = arrival to the end of the box's last step."""
    from tigerpm_amd import serving as S
    out = {"budget_gib": budget_gib, "kicks_per_request": 10, "requests_per_rate": count, "rates": []}
    lim = S.budget_limits(budget_gib * (1 << 30))
    with T.Batch(**lim) as bt:
        srv = S.Server(bt, kicks=10)
        # FFT plans of every size the requests use, made once (not part of
        # any latency)
        out["prepare_s"] = srv.prepare(S.poisson_requests(count, 1.0, seed))
        # capacity: every request waiting at t = 0 (closed loop, the batch
        # always as full as the budget allows)
        cap = srv.run([r._replace(arrival=0.0) for r in S.poisson_requests(count, 1.0, seed)])
        cap.pop("records")
        out["capacity"] = cap
        log("serving capacity (all requests at t = 0): %.4g kicks/s, p50 %.3f s, p99 %.3f s, mean resident %.1f"
            % (cap["kicks_per_s"], cap["latency_p50_s"], cap["latency_p99_s"], cap["mean_resident_boxes"]))
        for rate in rates:
            # the same boxes at every rate; only the arrival times scale
            reqs = S.poisson_requests(count, rate, seed)
            r = srv.run(reqs)
            r.pop("records")
            r["rate_per_s"] = rate
            r["offered_kicks_per_s"] = sum(q.n_side ** 3 for q in reqs) * 10 / max(reqs[-1].arrival, 1e-9)
            r["fraction_of_capacity"] = r["kicks_per_s"] / cap["kicks_per_s"]
            log("serving rate %g/s: p50 %.3f s, p99 %.3f s, %.4g kicks/s (%.2f of capacity), mean resident %.1f"
                % (rate, r["latency_p50_s"], r["latency_p99_s"], r["kicks_per_s"], r["fraction_of_capacity"],
                   r["mean_resident_boxes"]))
            out["rates"].append(r)
    # accuracy after 10 steps for three requests served alone: forces at the
    # final positions (one more step with the kick and drift off) vs Ewald
    reqs = S.poisson_requests(count, rates[0], seed)
    picks = sorted(reqs, key=lambda q: q.n_side)
    errs = {}
    for q in (picks[0], picks[len(picks) // 2], picks[-1]):
        with T.Batch(**lim) as bt:
            sp = S.request_spec(q)
            box = bt.add_zeldovich(sp, q.n_side, q.seed, q.rms)
            for k in range(10):
                bt.step(T.TPM_BATCH_DRIFT, stats=False)
            bt.step(T.TPM_SKIP_KICK, stats=False)
            with T.Context(0) as c:
                bt.to_ctx(box, c)
                errs["n_side_%d" % q.n_side] = ewald_error(c, bt.download(box), 256, q.rid)
    out["force_rms_vector_vs_ewald_after_10_steps"] = errs
    log("serving accuracy", errs)
    return out


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
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-box batched-ensemble leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the continuous-batching serving leg")
    ap.add_argument("--no-tree", action="store_true", help="skip the tree-walk (kick_treepm algorithm) leg")
    ap.add_argument("--ensemble-graph", type=int, default=1, choices=[0, 1],
                    help="replay the ensemble step as a captured graph (0: plain launches, e.g. for tracing)")
    ap.add_argument("--only", choices=["ensemble", "serving", "tree"],
                    help="run only this leg and print its result (for profiling; not the benchmark line)")
    ap.add_argument("--serving-rates", default="4,8,12,16,20,24,32",
                    help="Poisson arrival rates (requests/s); the leg also measures the capacity")
    ap.add_argument("--serving-requests", type=int, default=60)
    ap.add_argument("--serving-budget-gib", type=float, default=16.0)
    ap.add_argument("--deposit-block", type=int, default=0,
                    help="deposit block edge in chain cells (0: automatic, see tpm_tuning)")
    ap.add_argument("--tile-sinks", type=int, default=16, help="sinks per PP tile (8, 16, 32, 64)")
    ap.add_argument("--pp-waves", type=int, default=0, help="PP tiles per workgroup (1, 4; 0: default)")
    ap.add_argument("--pp-lists", type=int, default=-1, help="per-sink source lists (0, 1; -1: default)")
    ap.add_argument("--pp-newton", type=int, default=-1,
                    help="each pair evaluated once, both ends (0, 1; -1: default)")
    ap.add_argument("--fft", type=int, default=0,
                    help="mesh FFTs: 0 = LDS-tiled transform for power-of-two meshes, 1 = rocFFT")
    ap.add_argument("--probe-copy", action="store_true",
                    help="also run one plain 1 GiB device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported (sharding a box over several GPUs "
              "is designed but not implemented)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.only == "ensemble":
        print(json.dumps({"ensemble": ensemble_leg(T, max(args.steps, 3), graph=args.ensemble_graph)}), flush=True)
        return
    if args.only == "serving":
        rates = [float(r) for r in args.serving_rates.split(",") if r]
        print(json.dumps({"serving": serving_leg(T, rates, args.serving_requests, args.serving_budget_gib)}),
              flush=True)
        return
    w = WORKLOADS[args.workload]
    ctx = T.Context(0)
    p = ctx.set_params(T.baseline_params(w["n_side"], w["mesh"], w["order"], interlace=w["interlace"]))
    ctx.set_tuning(tile_sinks=args.tile_sinks, deposit_block=args.deposit_block, pp_waves=args.pp_waves,
                   pp_lists=args.pp_lists, pp_newton=args.pp_newton, fft=args.fft)
    if args.only == "tree":
        if w["ic"] == "zeldovich":
            ctx.ic_zeldovich(w["n_side"], w["seed"], w["rms"], -2.0)
        else:
            ctx.ic_uniform(w["n_side"] ** 3, w["seed"])
        ctx.kick_treepm(0, 1.0, 1.0, True, stats=True)
        print(json.dumps({"tree": tree_leg(T, ctx, p, max(args.steps, 1), w["samples"], w["seed"],
                                           ewald=not args.no_ewald)}), flush=True)
        ctx.close()
        return
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
                   "rcut_over_rs": p.rcut / p.rs, "ic": w["ic"], "rms_disp_cells": w["rms"], "seed": w["seed"],
                   "tuning": ctx.get_tuning()},
        # Bound vocabulary is hbm|mfma: the PP kernel is FP32 compute bound and
        # its peak is MI355X's FP32 figure (157.3 TFLOP/s, vector and matrix).
        # unit_kind says what actually bounds it: FP32 on the vector ALUs.
        "roofline": {"bound": "mfma", "unit_kind": "valu_fp32",
                     "bound_note": "FP32 vector ALU peak; the kernel issues no MFMA instruction",
                     "kernel": "pp_kick_kernel + mesh_force_kernel (the PP stage: pairs, mesh interpolation, kick)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "flops_per_launch": FLOPS_PER_PAIR * W, "kernel_ms": pp_ms,
                     "candidate_pairs": W, "executed_pairs": st.executed_pairs,
                     "executed_tflops": executed, "executed_frac": executed / PEAK_FP32_TFLOPS},
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
    # The pair body timed alone at the kernel's occupancy over sources held
    # in LDS (tpm_probe_pair_rate).  A reference rate, not a ceiling: its
    # uniform sources mix in-cutoff and softened pairs differently from the
    # kernel's, and the kernel has run faster than it.
    body = ctx.probe_pair_rate(8)
    kernel_rate = st.executed_pairs / (pp_ms * 1e-3)
    out["roofline"]["pair_body_rate"] = {
        "pairs_per_s": body, "kernel_executed_pairs_per_s": kernel_rate, "ratio": kernel_rate / body,
        "waves_per_simd": 8,
        "note": "the PP stage's executed pairs/s over the pair body's rate alone (a reference, not a ceiling)"}
    log("pair body alone %.4g pairs/s; kernel %.4g executed pairs/s (%.3f)" % (body, kernel_rate, kernel_rate / body))
    up = useful_pairs(ctx, p, 2048, w["seed"])
    up["candidate_per_particle"] = W / n
    up["executed_per_particle"] = st.executed_pairs / n
    out["roofline"]["useful_pairs"] = up
    log("pairs per particle: useful %.1f, executed %.1f, candidate %.1f"
        % (up["per_particle"], up["executed_per_particle"], up["candidate_per_particle"]))
    if not args.no_ewald:
        log("FP64 Ewald on %d samples..." % w["samples"])
        out["accuracy"] = accuracy(ctx, p, w["samples"], w["seed"])
        log("accuracy", out["accuracy"])
    if not args.no_tree:
        out["tree"] = tree_leg(T, ctx, p, min(args.steps, 3), w["samples"], w["seed"], ewald=not args.no_ewald)
        log("tree", out["tree"])
    if not args.no_cpu:
        log("CPU baseline (budget %.0f s)..." % args.cpu_budget)
        out["cpu_baseline"] = cpu_baseline(ctx, p, args.cpu_budget, w["seed"])
        log("cpu", out["cpu_baseline"])
    ctx.close()
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(T, max(args.steps, 3), graph=args.ensemble_graph)
    if not args.no_serving:
        rates = [float(r) for r in args.serving_rates.split(",") if r]
        out["serving"] = serving_leg(T, rates, args.serving_requests, args.serving_budget_gib)
    if args.probe_copy:
        out["copy_probe_bytes"] = copy_probe(1 << 30)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
