"""Pins of the oracle (the CPU restatement the HIP path is checked against).

Each check is chosen so that a plausible slip (a dropped term, a swapped
axis, a wrong rounding) fails: the reference's own golden logs, libstdc++'s
mt19937, the reference's own Ooura Poisson output, FP64 closed forms and
single-mode eigenfunctions.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from helpers import GOLDEN, design, golden_lines, nesterov_lines, rel_err, traces
from replace_amd.netlist import Netlist, params_default


def setup_module(module):
    oracle.build()


# --------------------------------------------------------------------------
# engines and closed forms
# --------------------------------------------------------------------------
def test_mt19937_matches_libstdcxx():
    """initFillerGCells draws std::mt19937(0) (src/nesterovBase.cpp:1139)."""
    g = json.load(open(os.path.join(GOLDEN, "mt19937.json")))
    got = oracle.mt19937(0, len(g["draws"]))
    assert [int(v) for v in got] == g["draws"]


def test_fast_exp_closed_form():
    """fastExp = float(1 + a/1024) squared ten times (src/nesterovBase.cpp:1568-1582);
    within the (1+a/1024)^1024 vs exp(a) bound for a in (-300, 0]."""
    for a in [0.0, -0.5, -1.0, -3.7, -10.0, -50.0, -120.0, -299.0]:
        v = oracle.fast_exp(a)
        x = np.float32(1.0 + a / 1024.0)
        for _ in range(10):
            x = np.float32(x * x)
        assert v == x
        # (1+a/n)^n = exp(a - a^2/(2n) + O(a^3/n^2))
        approx = math.exp(a - a * a / 2048.0)
        assert abs(v - approx) <= 0.05 * approx + 1e-30
    assert oracle.fast_exp(0.0) == 1.0


# --------------------------------------------------------------------------
# Poisson solve (src/fft.cpp:112-177)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nx,ny,bsx,bsy", [(8, 8, 100, 100), (16, 32, 3515, 5075), (32, 16, 5368, 3325)])
def test_poisson_fft64_matches_direct_sums(nx, ny, bsx, bsy):
    rng = np.random.default_rng(nx * 100 + ny)
    rho = rng.uniform(0, 2, (ny, nx)).astype(np.float32)
    d = oracle.poisson_direct64(rho, bsx, bsy)
    f = oracle.poisson_fft64(rho, bsx, bsy)
    for a, b in zip(f, d):
        assert rel_err(a, b) < 2e-6


def test_poisson_single_mode_eigenfunction():
    """rho = cos(pi u0 (i+1/2)/Nx) cos(pi v0 (j+1/2)/Ny) gives
    phi = rho / (wx^2 + wy^2), Ex = phi_amp * wx * sin(..)cos(..),
    Ey = phi_amp * wy * cos(..)sin(..) with wy carrying bsy/bsx."""
    nx, ny, bsx, bsy, u0, v0 = 16, 32, 3000, 4500, 3, 5
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    cx = np.cos(np.pi * u0 * (i + 0.5) / nx)
    cy = np.cos(np.pi * v0 * (j + 0.5) / ny)
    sx = np.sin(np.pi * u0 * (i + 0.5) / nx)
    sy = np.sin(np.pi * v0 * (j + 0.5) / ny)
    rho = (cx * cy).astype(np.float32)
    wx = np.pi * u0 / nx
    wy = np.pi * v0 / ny * bsy / bsx
    amp = 1.0 / (wx * wx + wy * wy)
    phi, ex, ey = oracle.poisson_direct64(rho, bsx, bsy)
    assert rel_err(phi, amp * cx * cy) < 1e-5
    assert rel_err(ex, amp * wx * sx * cy) < 1e-5
    assert rel_err(ey, amp * wy * cx * sy) < 1e-5
    f = oracle.poisson_fft64(rho, bsx, bsy)
    assert rel_err(f[0], amp * cx * cy) < 1e-5


def test_poisson_matches_reference_ooura_vectors():
    """Golden vectors written by the reference's own FFT (src/fft.cpp + fftsg)."""
    z = np.load(os.path.join(GOLDEN, "poisson_ref.npz"))
    k = 0
    while f"c{k}_dims" in z.files:
        nx, ny, bsx, bsy = (int(v) for v in z[f"c{k}_dims"])
        got = oracle.poisson_fft64(z[f"c{k}_rho"], bsx, bsy)
        for a, name in zip(got, ("phi", "ex", "ey")):
            assert rel_err(a, z[f"c{k}_{name}"]) < 1e-5, (k, name)
        k += 1
    assert k == 4


# --------------------------------------------------------------------------
# scatter invariants (src/nesterovBase.cpp:745-819)
# --------------------------------------------------------------------------
def _tiny(n_cells=40, seed=3):
    from replace_amd import netgen
    nl = netgen.generate(n_cells, n_cells, seed=seed, util=0.5, td=0.7)
    return nl, params_default(target_density=0.7, bin_grid_cnt_x=8, bin_grid_cnt_y=8,
                              initial_place_max_iter=0)


def test_scatter_box_inside_one_bin():
    """A density box inside one bin adds trunc(float(dDx)*float(dDy)*scale) to it."""
    nl, p = _tiny()
    o = oracle.Oracle(nl, p, poisson="fft64")
    o.initial_place()
    info = o.setup()
    g = o.gcells()
    n = info.num_gcells
    # park every cell in bin (2,3)'s centre: their boxes (2*dhalf <= sqrt2*bin) span <= 2 bins
    bx = info.core_lx + 2 * info.bin_size_x + info.bin_size_x // 2
    by = info.core_ly + 3 * info.bin_size_y + info.bin_size_y // 2
    cx = np.full(n, bx, np.float32)
    cy = np.full(n, by, np.float32)
    r = o.eval_gradient(cx, cy, 0.0, 1e-4, 1e-4)
    # each box straddles at most the neighbouring bins; total equals the per-bin truncated sum
    total = r["bin_inst"].sum() + r["bin_filler"].sum()
    full = (np.float32(2 * g["dhalf_x"].astype(np.float32)) * np.float32(2 * g["dhalf_y"].astype(np.float32))
            * g["scale"]).astype(np.float64)
    assert abs(total - full.sum()) <= 4 * n  # <= one truncation per touched bin
    assert r["bin_inst"].sum() > 0


def test_scatter_total_area_invariant():
    nl, p = _tiny(60, 5)
    o = oracle.Oracle(nl, p, poisson="fft64")
    o.initial_place()
    info = o.setup()
    g = o.gcells()
    rng = np.random.default_rng(0)
    cx = rng.uniform(info.core_lx, info.core_ux, info.num_gcells).astype(np.float32)
    cy = rng.uniform(info.core_ly, info.core_uy, info.num_gcells).astype(np.float32)
    cx = np.clip(cx, info.core_lx + g["dhalf_x"], info.core_ux - g["dhalf_x"]).astype(np.float32)
    cy = np.clip(cy, info.core_ly + g["dhalf_y"], info.core_uy - g["dhalf_y"]).astype(np.float32)
    r = o.eval_gradient(cx, cy, 0.0, 1e-4, 1e-4)
    full = (4.0 * g["dhalf_x"] * g["dhalf_y"] * g["scale"].astype(np.float64)).sum()
    total = r["bin_inst"].sum() + r["bin_filler"].sum()
    assert full - 9 * info.num_gcells <= total <= full + 1


def test_wa_two_pin_net_gradient_limits():
    """2-pin net, large 1/gamma: the WA gradient per pin tends to -/+1 in x
    (min pin pulled right, max pin pulled left)."""
    nl = Netlist(dbu_per_micron=2000, die=(0, 0, 38000, 28000), site_w=380, site_h=2800,
                 row_box=np.array([[0, r * 2800, 38000, (r + 1) * 2800] for r in range(10)]),
                 inst_box=np.array([[0, 0, 380, 2800], [0, 0, 380, 2800]]),
                 inst_fixed=np.zeros(2, np.uint8), net_pin_start=np.array([0, 2]),
                 pin_inst=np.array([0, 1]), pin_off=np.zeros((2, 2), np.int32))
    p = params_default(target_density=0.7, bin_grid_cnt_x=4, bin_grid_cnt_y=4, initial_place_max_iter=0)
    o = oracle.Oracle(nl, p, poisson="fft64")
    o.initial_place()
    info = o.setup()
    n = info.num_gcells
    cx = np.full(n, 19000, np.float32)
    cy = np.full(n, 14000, np.float32)
    cx[0], cx[1] = 10000, 25000
    r = o.eval_gradient(cx, cy, 0.0, 0.05, 0.05)
    g = r["wl_grad"].reshape(-1, 2)
    # reference returns the negative gradient of W (the descent direction)
    assert abs(g[0, 0] - 1.0) < 1e-3 and abs(g[1, 0] + 1.0) < 1e-3
    assert abs(g[0, 1]) < 1e-6 and abs(g[1, 1]) < 1e-6


# --------------------------------------------------------------------------
# end-to-end: the reference's golden logs (test/*.ok)
# --------------------------------------------------------------------------
EXACT = ["simple01", "simple02", "ar01", "ar02", "core01", "incremental01"]


def _run(name, poisson):
    nl, p = design(name)
    o = oracle.Oracle(nl, p, poisson=poisson)
    err, hp = o.initial_place()
    o.setup()
    code, st = o.nesterov_place()
    return o, err, hp, code, st


@pytest.mark.parametrize("name", EXACT)
def test_trace_exact_with_reference_transform(name):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref (reference Ooura build) not present")
    o, err, hp, code, st = _run(name, "reference")
    assert code == 0
    assert nesterov_lines(o.records(), st.final_overflow if st.finished else None) == golden_lines(name)


@pytest.mark.parametrize("name", EXACT)
def test_trace_close_with_own_transform(name):
    """Own FP64 transform: logged overflow within 1e-3 abs and HPWL within
    0.1 % for the first 100 iterations (0.01 / 0.5 % later), stop iteration
    +-12, final HPWL within 0.5 %."""
    o, err, hp, code, st = _run(name, "fft64")
    t = traces()[name]
    rec = o.records()
    for it, ovf, hpwl in t["nesterov"]:
        if it - 1 >= len(rec):
            break
        late = it >= 100  # chaotic dynamics amplify rounding differences late in the run
        assert abs(rec[it - 1][0] - float(ovf)) < (0.01 if late else 1e-3)
        assert abs(rec[it - 1][1] - hpwl) <= (0.005 if late else 1e-3) * hpwl
    last_it, last_o, last_h = t["nesterov"][-1]
    assert abs(st.iterations - last_it) <= 12
    assert abs(hpwl_at_overflow(rec, float(last_o)) - last_h) <= 0.005 * last_h


def hpwl_at_overflow(rec, ovf):
    """HPWL of the first iterate whose overflow is <= ovf (equal-overflow comparison)."""
    for r in rec:
        if r[0] <= ovf:
            return r[1]
    return rec[-1][1]


@pytest.mark.parametrize("name", ["simple03", "simple04"])
def test_trace_with_initial_place(name):
    """IP designs: the first [InitialPlace] HPWL (DB positions) is exact; the
    BiCGSTAB iterates are parity-unpinned (Eigen is absent), so later lines
    are checked within 0.01 % and the Nesterov finish within 0.5 %."""
    o, err, hp, code, st = _run(name, "auto")
    t = traces()[name]
    assert len(hp) == len(t["ip"])
    assert hp[0] == t["ip"][0][2]
    for (it, e, h), got in zip(t["ip"], hp):
        assert abs(got - h) <= 1e-4 * h
    assert all(e < 1e-5 for e in err[4:])
    last_it, last_o, last_h = t["nesterov"][-1]
    assert abs(st.iterations - last_it) <= 12
    assert abs(hpwl_at_overflow(o.records(), float(last_o)) - last_h) <= 0.005 * last_h


def test_diverge01_returns_repl3():
    o, err, hp, code, st = _run("diverge01", "fft64")
    assert traces()["diverge01"]["code"] == 3
    assert code == 3


def test_info_values_match_golden_logs():
    for name in ["simple01", "simple02", "ar01", "simple04"]:
        nl, p = design(name)
        o = oracle.Oracle(nl, p, poisson="fft64")
        o.initial_place()
        info = o.setup()
        g = traces()[name]["info"]
        assert f"({info.bin_cnt_x}, {info.bin_cnt_y})" == g["BinCnt"]
        assert f"({info.bin_size_x}, {info.bin_size_y})" == g["BinSize"]
        assert str(info.num_gcells) == g["FillerInit: NumGCells"]
        assert str(info.std_insts_area) == g["StdInstsArea"]
        assert str(info.num_gnets) == g["FillerInit: NumGNets"]
        assert str(info.num_gpins) == g["FillerInit: NumGPins"]
        assert f"{info.target_density:.6f}" == g["TargetDensity"]
        assert str(info.average_place_inst_area) == g["AveragePlaceInstArea"]
        assert str(info.ideal_bin_area) == g["IdealBinArea"]
        assert str(info.ideal_bin_count) == g["IdealBinCnt"]
        assert str(info.total_bin_area) == g["TotalBinArea"]


# --------------------------------------------------------------------------
# final placements: the reference's golden result DEFs (test/*.defok)
# --------------------------------------------------------------------------
DEFOK_EXACT = EXACT + ["diverge01"]


@pytest.mark.parametrize("name", DEFOK_EXACT)
def test_final_positions_match_reference_defok(name):
    """With the reference's own Ooura transform the oracle's final lower-left
    corners equal, cell for cell, those the reference wrote with write_def
    (test/<design>.defok; updateDb, src/nesterovPlace.cpp:585-595)."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref (reference Ooura build) not present")
    o, err, hp, code, st = _run(name, "reference")
    lx, ly = o.positions()
    ok = np.load(os.path.join(GOLDEN, "defok.npz"))[name]
    assert np.array_equal(lx, ok[:, 0]) and np.array_equal(ly, ok[:, 1])


REF_TEST = "/root/reference/test"


@pytest.mark.parametrize("name", ["simple01", "ar02"])
def test_def_writer_matches_reference_components(name, tmp_path):
    """write_def of the oracle's final positions reproduces the COMPONENTS
    section of the reference's golden DEF line for line (the reference's own
    check diffs the written DEF against test/<design>.defok)."""
    if not oracle.ref_available() or not os.path.exists(os.path.join(REF_TEST, name + ".def")):
        pytest.skip("reference test DEFs not present")
    from replace_amd.lefdef import components_section, write_def
    o, err, hp, code, st = _run(name, "reference")
    lx, ly = o.positions()
    out = str(tmp_path / (name + ".def"))
    write_def(os.path.join(REF_TEST, name + ".def"), out, lx, ly)
    assert components_section(out) == components_section(os.path.join(REF_TEST, name + ".defok"))


def test_state_round_trip_resumes_the_same_run():
    """ro_get_state / ro_set_state: a run stopped after k iterations and
    resumed in a fresh context from its state continues exactly as the
    uninterrupted run (records and final state equal)."""
    nl, p = design("simple01")

    def fresh():
        o = oracle.Oracle(nl, p)
        o.initial_place()
        o.setup()
        assert o.nesterov_init() == 0
        return o

    a = fresh()
    a.nesterov_iterate(40)
    s40 = a.get_state()
    a.nesterov_iterate(15)
    want = a.records()[40:]
    s55 = a.get_state()
    b = fresh()
    b.set_state(s40)
    assert b.get_state()["iter"] == 40
    b.nesterov_iterate(15)
    assert b.records() == want
    got = b.get_state()
    for k, v in s55.items():
        assert np.array_equal(np.asarray(got[k]), np.asarray(v)), k


def test_repl1_paths():
    """REPL-1 (src/placerBase.cpp:952-956 utilisation above 100.1 %,
    src/nesterovBase.cpp:1116-1121 negative filler area)."""
    nl, p = design("simple01")
    # a target density below the utilisation: negative filler area
    p.target_density = 0.3
    o = oracle.Oracle(nl, p)
    with pytest.raises(RuntimeError, match="code 1"):
        o.setup()
    # a quarter of the rows: the cells no longer fit the core
    nl2, p2 = design("simple01")
    rows = nl2.row_box[: len(nl2.row_box) // 4]
    small = Netlist(dbu_per_micron=nl2.dbu_per_micron, die=nl2.die, site_w=nl2.site_w, site_h=nl2.site_h,
                    row_box=rows, inst_box=nl2.inst_box, inst_fixed=nl2.inst_fixed,
                    net_pin_start=nl2.net_pin_start, pin_inst=nl2.pin_inst, pin_off=nl2.pin_off,
                    pin_box=nl2.pin_box)
    assert oracle.Oracle(small, p2).create_code == 1
