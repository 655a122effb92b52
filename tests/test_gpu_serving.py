"""GPU test of the continuous-batching server: requests arriving while
others are resident are admitted as space frees up, and every request's
final state is the same as when it is served alone (bit for bit)."""
import numpy as np
import pytest

import tigerpm_amd as T
from tigerpm_amd import serving as S

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "vx", "vy", "vz", "rung", "gx", "gy", "gz", "pot")


def serve(requests, budget_particles, kicks=3, graph=0):
    finals = {}

    def keep(batch, box, req):
        finals[req.rid] = batch.download(box)

    with T.Batch(budget_particles, budget_particles, 32, 3, 1) as bt:
        bt.set_tuning(tile_sinks=16, graph=graph)
        srv = S.Server(bt, kicks=kicks, t0=0.01)
        srv.prepare(requests)
        out = srv.run(requests, on_retire=keep)
    return out, finals


def test_served_alone_equals_served_in_the_mix():
    reqs = S.poisson_requests(10, rate=400.0, seed=3, n_lo=18, n_hi=36)
    budget = 3 * max(r.n_side for r in reqs) ** 3
    mix, finals = serve(reqs, budget, graph=1)
    assert mix["requests"] == len(reqs)
    assert mix["max_resident_boxes"] >= 2  # the work did overlap
    assert mix["latency_p99_s"] >= mix["latency_p50_s"] > 0
    for r in reqs[:4]:
        alone, fa = serve([r._replace(arrival=0.0)], budget)
        assert alone["requests"] == 1
        for k in KEYS:
            assert np.array_equal(finals[r.rid][k], fa[r.rid][k]), (r, k)


def test_budget_too_small_for_a_request_raises():
    reqs = [S.Request(0, 0.0, 30, 1, 1.0)]
    with T.Batch(20 ** 3, 20 ** 3, 4, 3, 1) as bt:
        with pytest.raises(RuntimeError):
            S.Server(bt, kicks=1).run(reqs)
