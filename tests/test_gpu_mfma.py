"""Matrix-core Poisson (dense transform-matrix GEMMs, split-BF16 operands,
FP32 accumulation) against the FP64 oracle, and end to end against the FP32
FFT path.

Tolerances: potential and field 2e-4 relative (normwise: max abs error /
max abs value of the FP64 result; split BF16 carries ~16 mantissa bits per
operand); final HPWL within 0.5 % of the FFT path at overflow <= 0.10.
"""
import numpy as np
import pytest

import oracle
from helpers import rel_err
from replace_amd import netgen
from replace_amd.capi import Placer

pytestmark = pytest.mark.gpu

TOL = 2e-4


def _placer(nl, nx, ny, mode):
    p = netgen.params_default(target_density=0.7, bin_grid_cnt_x=nx, bin_grid_cnt_y=ny,
                              initial_place_max_iter=0, poisson_transform=mode)
    g = Placer(nl, p)
    g.initial_place()
    return g, g.setup()


@pytest.mark.parametrize("nx,ny", [(128, 128), (256, 128), (128, 512), (1024, 1024), (2048, 2048)])
def test_mfma_poisson_against_fp64(nx, ny):
    # a core large enough for the grid (bins of at least a few DBU)
    if max(nx, ny) <= 256:
        nl = netgen.generate(4000, 4200, seed=11)
    else:
        nl = netgen.workload("c0" if max(nx, ny) <= 512 else "c1")[0]
    rng = np.random.default_rng(nx * 7 + ny)
    rho = rng.uniform(0, 2, (ny, nx)).astype(np.float32)
    errs = {}
    for mode in (0, 1):
        g, info = _placer(nl, nx, ny, mode)
        phi, ex, ey = g.eval_poisson(rho)
        if mode == 0:
            if nx * ny <= 128 * 128:
                ref = oracle.poisson_direct64(rho, info.bin_size_x, info.bin_size_y)
            else:
                ref = oracle.poisson_fft64(rho, info.bin_size_x, info.bin_size_y)
        errs[mode] = [rel_err(a.reshape(ny, nx), b) for a, b in zip((phi, ex, ey), ref)]
        g.close()
    print(f"{nx}x{ny} rel err vs FP64 (phi, ex, ey): fft32 {errs[0]} mfma {errs[1]}")
    assert max(errs[1]) < TOL


def test_mfma_single_mode_amplitudes():
    """rho = cos(pi u0 (2i+1)/2nx) cos(pi v0 (2j+1)/2ny): one spectral mode,
    phi = rho / (wx^2 + wy^2) exactly (closed form)."""
    nx = ny = 256
    nl = netgen.generate(4000, 4200, seed=11)
    g, info = _placer(nl, nx, ny, 1)
    u0, v0 = 5, 17
    i = np.arange(nx)
    j = np.arange(ny)
    rho = (np.cos(np.pi * u0 * (2 * i[None, :] + 1) / (2 * nx)) *
           np.cos(np.pi * v0 * (2 * j[:, None] + 1) / (2 * ny)))
    phi, ex, ey = g.eval_poisson(rho.astype(np.float32))
    asp = info.bin_size_y / info.bin_size_x
    wx, wy = np.pi * u0 / nx, np.pi * v0 / ny * asp
    want = rho / (wx * wx + wy * wy)
    assert rel_err(phi.reshape(ny, nx), want) < TOL


def test_mfma_full_run_tracks_fft_path():
    nl, p = netgen.workload("c0", scale=1.0)
    out = {}
    for mode in (0, 1):
        q = netgen.params_default(**{k: getattr(p, k) for k, _ in p._fields_})
        q.poisson_transform = mode
        q.bin_grid_cnt_x = q.bin_grid_cnt_y = 128
        g = Placer(nl, q)
        g.initial_place()
        g.setup()
        code, st = g.nesterov_place()
        assert code == 0 and st.finished == 1 and st.final_overflow <= 0.1
        out[mode] = st.final_hpwl
        g.close()
    assert abs(out[1] - out[0]) <= 0.005 * out[0], out


def test_c2_gradients_against_fp64():
    """The 1M-cell netlist on 2048 x 2048 bins (the precision workload),
    cells starting at the core centre: the gradients of both GPU transforms
    at the states of the matrix-core run after 0, 150 and 400 iterations,
    evaluated from the same SLP coordinates, penalty and wire-length
    coefficients as the oracle's FP64 FFT solve.  Normwise bars: the field
    (2e-4 for split BF16, the stage bar above; 2e-5 for FP32 FFT), the
    density gradient (cell-area-weighted field: the same bars), and the
    whole preconditioner input wl + lambda density (1e-4 / 1e-5); the
    wire-length gradient is the same FP32 arithmetic on all three (1e-6).
    The measured errors are printed."""
    nl, p = netgen.workload("c1")
    p.bin_grid_cnt_x = p.bin_grid_cnt_y = 2048
    p.initial_place_max_iter = 0
    o = oracle.Oracle(nl, p, poisson="fft64")
    o.initial_place()
    o.setup()
    gs = {}
    for mode in (0, 1):
        q = netgen.params_default(**{k: getattr(p, k) for k, _ in p._fields_})
        q.poisson_transform = mode
        gs[mode] = Placer(nl, q)
        gs[mode].initial_place()
        gs[mode].setup()
    run = gs[1]
    assert run.nesterov_init() == 0
    done = 0
    bars = {0: (2e-5, 1e-5), 1: (2e-4, 1e-4)}
    for k in (0, 150, 400):
        if k > done:
            run.nesterov_iterate(k - done)
            done = k
        s = run.get_state()
        cx = np.ascontiguousarray(s["cur_slp"][0::2])
        cy = np.ascontiguousarray(s["cur_slp"][1::2])
        lam, wlx, wly = s["density_penalty"], s["wl_coef_x"], s["wl_coef_y"]
        ref = o.eval_gradient(cx, cy, lam, wlx, wly)
        for mode in (0, 1):
            got = gs[mode].eval_gradient(cx, cy, lam, wlx, wly)
            e = {key: rel_err(got[key], ref[key]) for key in ("ex", "ey", "dens_grad", "wl_grad", "sum_grad")}
            print(f"c2 iteration {k} mode {mode} rel err vs FP64: {e}")
            f_bar, s_bar = bars[mode]
            assert max(e["ex"], e["ey"], e["dens_grad"]) < f_bar, (k, mode, e)
            assert e["sum_grad"] < s_bar and e["wl_grad"] < 1e-6, (k, mode, e)
    for g in gs.values():
        g.close()
    o.close()
