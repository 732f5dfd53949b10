"""Parity of the HIP path at the sizes the placement workloads run.

* medium01 (the reference's 17,782-instance regression design, 61k fillers,
  256 x 256 bins, one 2,251-pin net): stage-level evaluations against the
  oracle, and full placements against the oracle's.
* nets of more than 512 pins (the wirelength kernel gives such a net a
  workgroup of its own and sums it chunk by chunk): bit-exact sums.
* the 1M-cell workload (1,000,000 cells, 40 fixed macros, 1024 x 1024 bins):
  evaluations with the two-cells-per-thread scatter the library launches
  from 131072 cells on, against the oracle; a whole placement checked
  through properties that hold at any size (stop condition, reproducibility,
  cells inside the core, positions consistent with the final centres).

Bars as in test_gpu_parity: integer work and the sums kept in the
reference's order bit-exact, the Poisson solve and what follows it within
1e-4 relative; end to end, HPWL within 0.5 %.
"""
import os

import numpy as np
import pytest

import oracle
from helpers import GOLDEN, check_eval, clamped_random_centres, hpwl_at
from replace_amd import netgen
from replace_amd.capi import Placer
from replace_amd.netlist import Netlist, params_default

pytestmark = pytest.mark.gpu


def medium01(**flags):
    """medium01 with the flags of test/medium01.tcl (global_placement
    defaults: density 0.7, initial place on)."""
    nl = Netlist.load(os.path.join(GOLDEN, "medium01.npz"), name="medium01")
    f = dict(target_density=0.7)
    f.update(flags)
    return nl, params_default(**f)


def _pair(nl, p):
    g = Placer(nl, p)
    o = oracle.Oracle(nl, p, poisson="auto")
    g.initial_place()
    o.initial_place()
    return g, o, g.setup(), o.setup()


def _eval_both(g, o, cx, cy, lam, wl):
    check_eval(g.eval_gradient(cx, cy, lam, wl, wl), o.eval_gradient(cx, cy, lam, wl, wl), lam,
               g.info, g.nonplace())


def test_medium01_eval_matches_oracle():
    nl, p = medium01(initial_place_max_iter=0)
    assert np.diff(nl.net_pin_start).max() > 512  # the large-net branch runs
    g, o, gi, oi = _pair(nl, p)
    assert (gi.num_gcells, gi.num_fillers, gi.bin_cnt_x, gi.bin_cnt_y) == (
        oi.num_gcells, oi.num_fillers, oi.bin_cnt_x, oi.bin_cnt_y)
    gc, oc = g.gcells(), o.gcells()
    for k in gc:
        assert np.array_equal(gc[k], oc[k]), k
    assert np.array_equal(g.nonplace(), o.nonplace())
    _eval_both(g, o, gc["init_cx"], gc["init_cy"], 0.0, 0.0007)
    cx, cy = clamped_random_centres(gc, oi, 31)
    _eval_both(g, o, cx, cy, 2e-5, 0.0004)


def test_medium01_full_run_matches_oracle():
    """From identical starting positions (no initial place): the first 50
    iterations within 2e-3 overflow / 0.2 % HPWL of the oracle's, the finish
    within +-12 iterations, HPWL within 0.5 % at equal overflow."""
    nl, p = medium01(initial_place_max_iter=0)
    g, o, _, _ = _pair(nl, p)
    code, st = g.nesterov_place()
    ocode, ost = o.nesterov_place()
    assert code == ocode == 0 and st.finished == 1
    rec, orec = g.records(), o.records()
    for i in range(min(50, len(orec))):
        assert abs(rec[i][0] - orec[i][0]) < 2e-3, i
        assert abs(rec[i][1] - orec[i][1]) <= 2e-3 * orec[i][1], i
    assert abs(st.iterations - ost.iterations) <= 12
    assert abs(hpwl_at(rec, ost.final_overflow) - ost.final_hpwl) <= 0.005 * ost.final_hpwl
    assert st.final_overflow <= 0.1


def test_medium01_with_initial_place_matches_oracle():
    """The design's own flow (20 initial-place iterations, then Nesterov):
    the BiCGSTAB iterates are parity-unpinned, so the bar is the end result:
    finished at overflow <= 0.10 with HPWL within 0.5 % of the oracle's."""
    nl, p = medium01()
    g, o, _, _ = _pair(nl, p)
    code, st = g.nesterov_place()
    ocode, ost = o.nesterov_place()
    assert code == ocode == 0 and st.finished == 1 and st.final_overflow <= 0.1
    assert abs(st.final_hpwl - ost.final_hpwl) <= 0.005 * ost.final_hpwl, (st.final_hpwl, ost.final_hpwl)


def _with_large_nets(nl, degrees, seed):
    """nl plus nets of the given degrees over random movable cells (pins at
    random offsets inside their cells), appended after its own nets."""
    rng = np.random.default_rng(seed)
    movable = np.flatnonzero(nl.inst_fixed == 0)
    w = nl.inst_box[:, 2] - nl.inst_box[:, 0]
    h = nl.inst_box[:, 3] - nl.inst_box[:, 1]
    starts = list(nl.net_pin_start)
    pins = [nl.pin_inst]
    offs = [nl.pin_off]
    for d in degrees:
        cells = rng.choice(movable, size=d, replace=False).astype(np.int32)
        ox = (rng.random(d) * w[cells]).astype(np.int32) - w[cells] // 2
        oy = (rng.random(d) * h[cells]).astype(np.int32) - h[cells] // 2
        pins.append(cells)
        offs.append(np.stack([ox, oy], axis=1).astype(np.int32))
        starts.append(starts[-1] + d)
    return Netlist(dbu_per_micron=nl.dbu_per_micron, die=nl.die, site_w=nl.site_w, site_h=nl.site_h,
                   row_box=nl.row_box, inst_box=nl.inst_box, inst_fixed=nl.inst_fixed,
                   net_pin_start=np.array(starts, np.int32), pin_inst=np.concatenate(pins),
                   pin_off=np.concatenate(offs), name="large_nets")


def test_nets_above_workgroup_capacity_bit_exact():
    base = netgen.generate(12000, 12600, seed=41)
    nl = _with_large_nets(base, [513, 1024, 1025, 2600, 7000], seed=42)
    p = netgen.params_default(target_density=0.7, bin_grid_cnt_x=128, bin_grid_cnt_y=128,
                              initial_place_max_iter=0)
    g, o, gi, oi = _pair(nl, p)
    gc = g.gcells()
    _eval_both(g, o, gc["init_cx"], gc["init_cy"], 0.0, 0.0007)
    cx, cy = clamped_random_centres(gc, oi, 43)
    _eval_both(g, o, cx, cy, 5e-5, 0.0002)


@pytest.fixture(scope="module")
def c1():
    nl, p = netgen.workload("c1")
    return nl, p


def test_c1_eval_matches_oracle(c1):
    """The 1M-cell workload's evaluation (two-cells-per-thread scatter)
    against the oracle, at the initial centres and at a random clamped
    iterate."""
    nl, p0 = c1
    p = params_default(**{f: getattr(p0, f) for f, _ in type(p0)._fields_})
    p.initial_place_max_iter = 0
    g, o, gi, oi = _pair(nl, p)
    assert gi.num_gcells >= 1 << 17 and (gi.bin_cnt_x, gi.bin_cnt_y) == (1024, 1024)
    assert np.array_equal(g.nonplace(), o.nonplace())
    gc = g.gcells()
    _eval_both(g, o, gc["init_cx"], gc["init_cy"], 0.0, 0.0007)
    cx, cy = clamped_random_centres(gc, oi, 77)
    _eval_both(g, o, cx, cy, 1e-5, 0.0001)
    g.close()
    o.close()


def test_c1_full_placement_properties(c1):
    """One whole 1M-cell placement (device initial place, then Nesterov) run
    twice: both finish at overflow <= 0.10 after more than 50 iterations,
    the records and positions are bit-identical between the runs, every
    movable instance lies inside the core, fixed ones are untouched, and the
    positions are the final density centres minus half the density size."""
    nl, p = c1
    runs = []
    for _ in range(2):
        g = Placer(nl, p)
        g.initial_place()
        info = g.setup()
        code, st = g.nesterov_place()
        assert code == 0 and st.finished == 1
        assert st.final_overflow <= 0.10 and st.iterations > 50
        rec = g.records()
        assert rec[-1][0] <= 0.10 and all(r[0] > 0.10 for r in rec[51:-1])
        lx, ly = g.positions()
        gc = g.gcells()
        cx, cy = g.gcell_centres()
        runs.append((rec, lx, ly))
        g.close()
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])
    fixed = nl.inst_fixed.astype(bool)
    assert np.array_equal(lx[fixed], nl.inst_box[fixed, 0]) and np.array_equal(ly[fixed], nl.inst_box[fixed, 1])
    mov = np.flatnonzero(~fixed)
    w = nl.inst_box[mov, 2] - nl.inst_box[mov, 0]
    h = nl.inst_box[mov, 3] - nl.inst_box[mov, 1]
    # updateDb: LL = dC - dD/2 of the last evaluated iterate (movable
    # instances are the first gcells, in instance order)
    k = np.arange(len(mov))
    assert np.array_equal(lx[mov], (cx[k].astype(np.int64) - gc["dhalf_x"][k]).astype(np.int32))
    assert np.array_equal(ly[mov], (cy[k].astype(np.int64) - gc["dhalf_y"][k]).astype(np.int32))
    # the density box is at least the cell, so the cell is inside the core too
    assert (lx[mov] >= info.core_lx).all() and (ly[mov] >= info.core_ly).all()
    assert (lx[mov] + w <= info.core_ux).all() and (ly[mov] + h <= info.core_uy).all()
