"""Continuous-batching kick server over one device-resident tpm_batch.

A request is one periodic box (Zel'dovich lattice n_side^3 with its own seed
and rms displacement) that wants `kicks` successive kick + drift steps with
its state resident on the device.  The loop

  * moves requests that have arrived (wall clock) into a FIFO queue;
  * admits queued requests, in order, while the batch's fixed arenas (the
    HBM budget) have room for them -- their initial conditions are made on
    the device while the batch's previous step is still running;
  * runs one batched step (every resident box: chain mesh, PM, fused PP +
    interpolation + kick, drift) and waits for it;
  * retires the boxes that have taken all their steps and frees their space
    for the next admissions.

Boxes join and leave between steps, so work that arrives while others are on
the GPU shares their launches (continuous batching, not a queue in front of
a single-box kernel).  Every box's result does not depend on what else is
resident (tpm_batch_* guarantees the same bits as the box alone).

No reference counterpart: the reference kicks one box per call.
"""
import collections
import math
import time

import numpy as np

from ._lib import TPM_BATCH_DRIFT, Batch, box_spec
from .workloads import baseline_params

Request = collections.namedtuple("Request", "rid arrival n_side seed rms")


def poisson_requests(count, rate, seed, n_lo=32, n_hi=256, rms_lo=0.3, rms_hi=2.0):
    """count requests arriving as a Poisson process of `rate` per second
    (the first at t = 0); N = n_side^3 log-uniform in [n_lo^3, n_hi^3]
    (n_side log-uniform, rounded), rms displacement uniform."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    gaps[0] = 0.0
    arrivals = np.cumsum(gaps)
    ns = np.rint(np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), count))).astype(int)
    rms = rng.uniform(rms_lo, rms_hi, count)
    return [Request(i, float(arrivals[i]), int(ns[i]), int(seed) * 100003 + i, float(rms[i])) for i in range(count)]


def request_spec(req, order=4, interlace=0, t0=0.01):
    """Box parameters of a request: mesh = n_side, rs = 1.25 cells,
    rcut = 4.5 rs, masses 1/N."""
    return box_spec(baseline_params(req.n_side, req.n_side, order, interlace=interlace), t0=t0)


def budget_limits(budget_bytes, n_hi=256, order=4, interlace=0, max_boxes=256):
    """Arena capacities for a device-memory budget: particle and mesh slots
    in the ratio of boxes with mesh = n_side, the largest box always fits."""
    from ._lib import batch_arena_bytes
    per = batch_arena_bytes(1 << 20, 1 << 20, max_boxes, order, interlace) / float(1 << 20)
    slots = int(budget_bytes / per)
    slots = max(slots, n_hi ** 3)
    return dict(max_particles=slots, max_mesh_nodes=slots, max_boxes=max_boxes, order=order, interlace=interlace)


class Server:
    """Serves requests with the batch's assignment order and interlacing."""

    def __init__(self, batch, kicks=10, t0=0.01):
        self.batch = batch
        self.kicks = kicks
        self.order = batch.limits.order
        self.interlace = batch.limits.interlace
        self.t0 = t0

    def spec(self, req):
        return request_spec(req, self.order, self.interlace, self.t0)

    def prepare(self, requests):
        """FFT plans and scratch space of every size the requests use."""
        t = time.perf_counter()
        for n in sorted({r.n_side for r in requests}):
            self.batch.prepare(self.spec(Request(-1, 0.0, n, 0, 0.0)), n)
        return time.perf_counter() - t

    def run(self, requests, on_retire=None, clock=time.perf_counter):
        """Serves the requests; returns per-request records and totals.
        on_retire(batch, box_id, request) is called before a box leaves."""
        b = self.batch
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        active = {}
        records = []
        steps = 0
        resident = []
        start = clock()
        while pending or waiting or active:
            now = clock() - start
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            if active:
                # the next step of the resident boxes runs while new arrivals
                # are admitted (their initial conditions are made meanwhile)
                b.step(TPM_BATCH_DRIFT, stats=False)
                steps += 1
                resident.append(len(active))
            self._admit(waiting, active, start, clock)
            if not active and not waiting:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - (clock() - start)))
                continue
            if not active:
                raise RuntimeError("request %d does not fit the arena budget" % waiting[0].rid)
            b.sync()
            t_done = clock() - start
            for box in list(active):
                n, s = b.box_info(box)
                if s >= self.kicks:
                    req, t_adm = active.pop(box)
                    if on_retire is not None:
                        on_retire(b, box, req)
                    b.remove(box)
                    records.append(dict(rid=req.rid, n=n, arrival=req.arrival, admitted=t_adm, done=t_done,
                                        latency=t_done - req.arrival, wait=t_adm - req.arrival))
        makespan = clock() - start
        lat = np.array([r["latency"] for r in records])
        kicks = sum(r["n"] for r in records) * self.kicks
        first = min(r.arrival for r in requests)
        last = max(r["done"] for r in records)
        return {
            "requests": len(records), "steps": steps,
            "mean_resident_boxes": float(np.mean(resident)) if resident else 0.0,
            "max_resident_boxes": int(max(resident)) if resident else 0,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max()),
            "wait_mean_s": float(np.mean([r["wait"] for r in records])),
            "kicks_per_s": kicks / (last - first), "particle_kicks": kicks, "makespan_s": makespan,
            "records": records,
        }

    def _admit(self, waiting, active, start, clock):
        b = self.batch
        while waiting:
            req = waiting[0]
            sp = self.spec(req)
            if not b.fits(sp, req.n_side ** 3):
                break
            box = b.add_zeldovich(sp, req.n_side, req.seed, req.rms)
            active[box] = (req, clock() - start)
            waiting.popleft()
