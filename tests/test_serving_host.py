"""Host logic of the serving loop (no GPU): the request stream and the
arena sizing that admission relies on."""
import numpy as np

import tigerpm_amd as T
from tigerpm_amd.serving import Request, budget_limits, poisson_requests, request_spec


def test_poisson_requests_are_seeded_and_shaped():
    a = poisson_requests(400, 8.0, seed=5)
    b = poisson_requests(400, 8.0, seed=5)
    assert a == b  # same seed, same stream
    assert a[0].arrival == 0.0
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) >= 0)
    # exponential gaps of mean 1 / rate (400 samples: within 15 %)
    assert abs(np.diff(arr).mean() * 8.0 - 1.0) < 0.15
    ns = np.array([r.n_side for r in a])
    assert ns.min() >= 32 and ns.max() <= 256
    # log-uniform n_side: the median sits near the geometric mean of the bounds
    assert abs(np.log(np.median(ns)) - 0.5 * (np.log(32) + np.log(256))) < 0.2
    assert len({r.seed for r in a}) == len(a)
    assert all(0.3 <= r.rms <= 2.0 for r in a)


def test_rates_share_one_draw():
    """Every rate serves the same boxes: only the arrival times scale."""
    a = poisson_requests(50, 4.0, seed=3)
    b = poisson_requests(50, 16.0, seed=3)
    assert [(r.n_side, r.seed, r.rms) for r in a] == [(r.n_side, r.seed, r.rms) for r in b]
    assert np.allclose([r.arrival * 4.0 for r in a], [r.arrival * 16.0 for r in b])


def test_request_spec_follows_the_workload_rules():
    sp = request_spec(Request(0, 0.0, 100, 1, 0.5))
    p = sp.params
    assert p.mesh_dim == 100 and abs(p.rs * 100 - 1.25) < 1e-12 and abs(p.rcut / p.rs - 4.5) < 1e-12
    assert p.chain_dim == int(np.floor(1.0 / p.rcut)) and p.order == 4


def test_budget_limits_fit_the_budget_and_the_largest_box():
    for gib in (2.0, 16.0):
        lim = budget_limits(gib * 2 ** 30)
        assert lim["max_particles"] >= 256 ** 3
        used = T.batch_arena_bytes(lim["max_particles"], lim["max_mesh_nodes"], lim["max_boxes"], lim["order"],
                                   lim["interlace"])
        if lim["max_particles"] > 256 ** 3:  # sized by the budget, not by the largest box
            assert used <= gib * 2 ** 30 * 1.001
