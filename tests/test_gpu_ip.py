"""Initial place: the device B2B assembly builds the same CSR and right-hand
side as the sequential host assembly (Eigen setFromTriplets semantics:
duplicates summed in triplet order), so every BiCGSTAB iterate, the IP log
and the final positions agree bit for bit."""
import os

import numpy as np
import pytest

from helpers import design
from replace_amd import netgen
from replace_amd.capi import Placer

pytestmark = pytest.mark.gpu


def _ip(nl, p, host):
    if host:
        os.environ["RP_IP_HOST_ASSEMBLY"] = "1"
    try:
        g = Placer(nl, p)
        st = g.initial_place()
        err, hp = g.ip_records()
        return st.iterations, err.tolist(), hp.tolist(), g.positions()
    finally:
        os.environ.pop("RP_IP_HOST_ASSEMBLY", None)


def _cases():
    out = [design(n) for n in ("simple03", "simple04", "ar01", "incremental01")]
    nl = netgen.generate(3000, 3150, seed=77, num_macros=4, name="ipmix")
    out.append((nl, netgen.params_default(target_density=0.7, initial_place_max_iter=6,
                                          initial_place_max_fanout=30)))
    nl, p = netgen.workload("c0", scale=0.25)
    p.initial_place_max_iter = 5
    out.append((nl, p))
    # 14000 movable rows (grid-wide solver)
    nl = netgen.generate(14000, 14700, seed=78, name="ip14k")
    out.append((nl, netgen.params_default(target_density=0.7, initial_place_max_iter=5)))
    # 30000 movable rows (grid-wide solver, several 1024-row partial blocks)
    nl = netgen.generate(30000, 31500, seed=79, name="ip30k")
    out.append((nl, netgen.params_default(target_density=0.7, initial_place_max_iter=5)))
    return out


@pytest.mark.parametrize("case", range(8))
def test_device_assembly_matches_host_assembly(case):
    nl, p = _cases()[case]
    d = _ip(nl, p, host=False)
    h = _ip(nl, p, host=True)
    assert d[0] == h[0] and d[1] == h[1] and d[2] == h[2]
    assert np.array_equal(d[3][0], h[3][0]) and np.array_equal(d[3][1], h[3][1])


def _ip_solver(nl, p, solver):
    if solver:
        os.environ["RP_IP_SOLVER"] = solver
    try:
        g = Placer(nl, p)
        st = g.initial_place()
        err, hp = g.ip_records()
        return st.iterations, err, hp, g.positions()
    finally:
        os.environ.pop("RP_IP_SOLVER", None)


@pytest.mark.parametrize("case", [0, 1, 2, 3, 4, 5])
def test_one_workgroup_solver_tracks_grid_solver(case):
    """Systems of up to 8192 rows are solved by one workgroup per axis; the
    grid-wide kernels (RP_IP_SOLVER=blocks) run the same BiCGSTAB recurrence
    with other FP64 summation orders.  Both must meet the IP stop rule
    (error <= 1e-5 from the fifth IP iteration on).  On the reference designs
    the solves converge and the IP HPWL agrees to 1e-4 in every iteration.
    On the synthetic cases the solves stop at the solver's iteration cap, and
    the restart rule (|r0.r| below 2^-46 |r0|^2) can fire in one summation
    order and not in the other, which sends the next iterates elsewhere: there
    the IP HPWL must agree to 1e-4 in the first two iterations (before any
    restart can matter) and to 20 % at the end."""
    nl, p = _cases()[case]
    one = _ip_solver(nl, p, None)
    grid = _ip_solver(nl, p, "blocks")
    assert one[0] == grid[0]
    if one[0] == 0:  # a design that runs no IP iteration
        return
    assert (one[1][4:] < 1e-5).all() and (grid[1][4:] < 1e-5).all()
    assert one[2][0] == grid[2][0]
    strict = len(one[2]) if case < 4 else 2
    for a, b in zip(one[2][:strict], grid[2][:strict]):
        assert abs(int(a) - int(b)) <= 1e-4 * int(b)
    assert abs(int(one[2][-1]) - int(grid[2][-1])) <= 0.2 * int(grid[2][-1])
