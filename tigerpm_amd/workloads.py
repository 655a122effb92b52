"""Parameter sets of the kick workloads.

reference_params(): the reference's defaults (options.cpp:81-96 and
kick_treepm.cu:833-853): four_dim = parts_dim / parts_o_four, chain_dim =
parts_dim / (parts_o_four * four_o_chain), rs = 1 / four_dim, rcut =
1 / chain_dim, hsoft = 2 / (25 parts_dim), GM = 1 (unit masses), eta =
0.2 / sqrt(2), cloud4 deposit (gravity_long.cpp:127-140).

baseline_params(): the benchmark workloads: rs = 1.25 mesh cells, rcut =
4.5 rs, chain_dim = floor(1 / rcut), hsoft = 2 / (25 n_side) and masses 1/N
(GM = 1/N with unit-mass sums).
"""
import math

from ._lib import TpmParams

WORKLOADS = {
    # n_side^3 particles, mesh^3 PM mesh, assignment order, initial conditions,
    # rms displacement [cells], Ewald samples, seed, interlaced deposit
    "uniform64_cic": dict(n_side=64, mesh=64, order=2, ic="uniform", rms=0.0, samples=1024, seed=1, interlace=0),
    "uniform64_cic_interlaced": dict(n_side=64, mesh=64, order=2, ic="uniform", rms=0.0, samples=1024, seed=1,
                                     interlace=1),
    "zeldovich256_tsc": dict(n_side=256, mesh=256, order=3, ic="zeldovich", rms=0.5, samples=4096, seed=2,
                             interlace=0),
    "zeldovich256_tsc_interlaced": dict(n_side=256, mesh=256, order=3, ic="zeldovich", rms=0.5, samples=4096,
                                        seed=2, interlace=1),
    "zeldovich512_tsc": dict(n_side=512, mesh=512, order=3, ic="zeldovich", rms=2.0, samples=8192, seed=3,
                             interlace=0),
}


def baseline_params(n_side, mesh, order, n_total=None, interlace=0):
    n = n_total if n_total is not None else n_side ** 3
    rs = 1.25 / mesh
    rcut = 4.5 * rs
    return TpmParams(mesh_dim=int(mesh), chain_dim=int(math.floor(1.0 / rcut)), order=int(order), do_phi=1,
                     long_type=1, interlace=int(interlace),
                     rs=rs, rcut=rcut, hsoft=2.0 / (25.0 * n_side), GM=1.0 / n, eta=0.2 / math.sqrt(2.0),
                     n_total=float(n))


def reference_params(parts_dim, parts_o_four=3, four_o_chain=5):
    if parts_dim % (parts_o_four * four_o_chain):
        raise ValueError("parts_dim must be a multiple of %d" % (parts_o_four * four_o_chain))
    four_dim = parts_dim // parts_o_four
    chain_dim = parts_dim // (parts_o_four * four_o_chain)
    return TpmParams(mesh_dim=four_dim, chain_dim=chain_dim, order=4, do_phi=1, long_type=1, rs=1.0 / four_dim,
                     rcut=1.0 / chain_dim, hsoft=2.0 / 25.0 / parts_dim, GM=1.0, eta=0.2 / math.sqrt(2.0),
                     n_total=float(parts_dim) ** 3)


def workload_params(name):
    """The parameters of a named workload (baseline_params with its mesh,
    order and interlacing)."""
    w = WORKLOADS[name]
    return baseline_params(w["n_side"], w["mesh"], w["order"], interlace=w["interlace"])
