"""Seeded synthetic netlists for the placement workloads.

Workloads (sizes as named in the benchmark configurations):
  * ``c0``: 20,000 movable std cells (width U{1..8} sites, one row high,
    site 380 x 2800 DBU = 0.19 x 1.4 um at 2000 DBU/um), 21,000 nets,
    square core at 70 % utilisation, 128 x 128 bins.
  * ``c1``: 1,000,000 cells, 1,050,000 nets, 40 fixed macros, 1024 x 1024 bins.
  * ``c3``: batch instances, cells log-uniform in [500, 50k], bin grids
    sized to the cell count (bins_for_cells).

Net degrees: 2 pins 60 %, 3 pins 18 %, U{4..10} 19 %, U{11..50} 3 %.
Pins come from a locality window: cells are shuffled onto a virtual
sqrt(N) x sqrt(N) slot grid, a net picks a seed cell and draws its other
cells uniformly from a window around the seed whose half-width grows with
the net degree (2 * ceil(sqrt(d))), which gives clustered, Rent-like
connectivity.  Pin offsets are uniform inside the cell.

Core sizing: rows of the site height cover the whole core; the core is the
smallest square-ish grid of sites for which the reference's filler-area
formula (src/nesterovBase.cpp:1104-1113, evaluated with its float
rounding) stays non-negative at the target density, so the run never takes
the REPL-1 exit.

Macros (``c1``): 40 FIXED squares.  Fractions f_i ~ U(1 %, 3 %) of the core
are drawn and rescaled so they sum to 40 % of the core (40 literal draws
average 80 %, which does not fit beside the std cells); they are placed on
the site grid by seeded rejection sampling, largest first, without overlap.
The generator uses numpy's PCG64 streams, seeded per workload.
"""
from __future__ import annotations

import math

import numpy as np

from .netlist import Netlist, params_default

SITE_W = 380
SITE_H = 2800
DBU = 2000


def _degrees(rng, n_nets):
    u = rng.random(n_nets)
    d = np.empty(n_nets, np.int64)
    d[u < 0.60] = 2
    m3 = (u >= 0.60) & (u < 0.78)
    d[m3] = 3
    m4 = (u >= 0.78) & (u < 0.97)
    d[m4] = rng.integers(4, 11, size=int(m4.sum()))
    m5 = u >= 0.97
    d[m5] = rng.integers(11, 51, size=int(m5.sum()))
    return d


def _filler_area_ok(core_area, nonplace_area, std_area, td):
    """Reference filler-area test with its float rounding (src/nesterovBase.cpp:1104-1121)."""
    white = core_area - nonplace_area
    movable = int(np.float32(np.float32(white) * np.float32(td)))
    return movable - std_area >= 0


def generate(num_cells: int, num_nets: int, seed: int, util: float = 0.70, td: float = 0.70,
             num_macros: int = 0, macro_total: float = 0.40, name: str = "synthetic",
             slots_out: list | None = None) -> Netlist:
    """The netlist; with `slots_out` (a list) the virtual slot grid side and
    every movable cell's slot are appended to it (for rent_profile)."""
    rng = np.random.default_rng(seed)
    widths = rng.integers(1, 9, size=num_cells).astype(np.int64) * SITE_W
    heights = np.full(num_cells, SITE_H, np.int64)
    std_area = int((widths * heights).sum())

    # core: square grid of sites with std_area <= util * (core - macros)
    need = std_area / util / (1.0 - (macro_total if num_macros else 0.0))
    side = math.sqrt(need)
    n_rows = max(1, int(math.ceil(side / SITE_H)))
    n_sites = max(1, int(math.ceil(need / (n_rows * SITE_H) / SITE_W)))
    core_w, core_h = n_sites * SITE_W, n_rows * SITE_H

    macro_boxes = []
    if num_macros:
        f = rng.uniform(0.01, 0.03, size=num_macros)
        f = f / f.sum() * macro_total
        order = np.argsort(-f)
        occ = np.zeros((n_rows, n_sites), bool)
        core_area = core_w * core_h
        for k in order:
            side_dbu = math.sqrt(f[k] * core_area)
            ms = max(1, int(round(side_dbu / SITE_W)))
            mr = max(1, int(round(side_dbu / SITE_H)))
            ms, mr = min(ms, n_sites), min(mr, n_rows)
            for _ in range(20000):
                sx = int(rng.integers(0, n_sites - ms + 1))
                sy = int(rng.integers(0, n_rows - mr + 1))
                if not occ[sy:sy + mr, sx:sx + ms].any():
                    occ[sy:sy + mr, sx:sx + ms] = True
                    macro_boxes.append((sx * SITE_W, sy * SITE_H, (sx + ms) * SITE_W,
                                        (sy + mr) * SITE_H))
                    break
            else:
                raise RuntimeError("macro placement failed")
    macro_area = sum((b[2] - b[0]) * (b[3] - b[1]) for b in macro_boxes)
    # grow the core (one site column at a time) until the filler area is >= 0
    while not _filler_area_ok(core_w * core_h, macro_area, std_area, td):
        n_sites += 1
        core_w = n_sites * SITE_W

    # instances: movable cells (UNPLACED at the origin) then macros (FIXED)
    n_inst = num_cells + len(macro_boxes)
    inst_box = np.zeros((n_inst, 4), np.int32)
    inst_box[:num_cells, 2] = widths
    inst_box[:num_cells, 3] = heights
    for k, b in enumerate(macro_boxes):
        inst_box[num_cells + k] = b
    inst_fixed = np.zeros(n_inst, np.uint8)
    inst_fixed[num_cells:] = 1

    # nets from locality windows on a shuffled slot grid
    deg = _degrees(rng, num_nets)
    g = int(math.ceil(math.sqrt(num_cells)))
    slot_of = rng.permutation(g * g)[:num_cells]
    cell_at = np.full(g * g, -1, np.int64)
    cell_at[slot_of] = np.arange(num_cells)
    if slots_out is not None:
        slots_out.extend([g, slot_of])
    net_start = np.zeros(num_nets + 1, np.int64)
    np.cumsum(deg, out=net_start[1:])
    npin = int(net_start[-1])
    net_of_pin = np.repeat(np.arange(num_nets), deg)
    first = np.zeros(npin, bool)
    first[net_start[:-1]] = True
    seed_cell = rng.integers(0, num_cells, size=num_nets)
    seed_slot = slot_of[seed_cell]
    sx, sy = seed_slot % g, seed_slot // g
    half = 2 * np.ceil(np.sqrt(deg)).astype(np.int64)
    pin_cell = np.empty(npin, np.int64)
    pin_cell[first] = seed_cell
    rest = ~first
    nr = int(rest.sum())
    hp = half[net_of_pin[rest]]
    for _attempt in range(64):
        ox = rng.integers(-(2 ** 31), 2 ** 31, size=nr) % (2 * hp + 1) - hp
        oy = rng.integers(-(2 ** 31), 2 ** 31, size=nr) % (2 * hp + 1) - hp
        px = np.clip(sx[net_of_pin[rest]] + ox, 0, g - 1)
        py = np.clip(sy[net_of_pin[rest]] + oy, 0, g - 1)
        cand = cell_at[py * g + px]
        if _attempt == 0:
            picked = cand
        else:
            picked = np.where(bad, cand, picked)
        bad = picked < 0
        if not bad.any():
            break
    picked = np.where(picked < 0, rng.integers(0, num_cells, size=nr), picked)
    pin_cell[rest] = picked
    # pin offsets uniform inside the cell, relative to the cell centre
    w = widths[pin_cell]
    h = heights[pin_cell]
    offx = (rng.random(npin) * w).astype(np.int64) - w // 2
    offy = (rng.random(npin) * h).astype(np.int64) - h // 2
    pin_off = np.stack([offx, offy], axis=1).astype(np.int32)

    rows = np.zeros((n_rows, 4), np.int32)
    rows[:, 0] = 0
    rows[:, 1] = np.arange(n_rows) * SITE_H
    rows[:, 2] = core_w
    rows[:, 3] = rows[:, 1] + SITE_H
    return Netlist(
        dbu_per_micron=DBU,
        die=(0, 0, core_w, core_h),
        site_w=SITE_W,
        site_h=SITE_H,
        row_box=rows,
        inst_box=inst_box,
        inst_fixed=inst_fixed,
        net_pin_start=net_start.astype(np.int32),
        pin_inst=pin_cell.astype(np.int32),
        pin_off=pin_off,
        pin_box=None,
        name=name,
    )


def rent_profile(nl: Netlist, g: int, slot_of: np.ndarray, sides=(2, 4, 8, 16, 32, 64)):
    """Rent's rule on the generator's virtual slot grid: for square blocks of
    side L slots (G cells each), T = the nets with pins both inside and
    outside the block, averaged over the blocks; the exponent p is the slope
    of log T against log G (least squares over the sides).  Returns
    (p, [(L, mean G, mean T), ...])."""
    ns = np.asarray(nl.net_pin_start, np.int64)
    pin_cell = np.asarray(nl.pin_inst, np.int64)
    net_of_pin = np.repeat(np.arange(len(ns) - 1), np.diff(ns))
    n_mov = len(slot_of)
    keep = pin_cell < n_mov
    pc, pn = pin_cell[keep], net_of_pin[keep]
    sx, sy = slot_of[pc] % g, slot_of[pc] // g
    pts = []
    for L in sides:
        nb = (g + L - 1) // L
        blk = (sy // L) * nb + (sx // L)
        # nets touching a block (each (net, block) pair once), and the blocks each net touches
        pair = np.unique(pn * (nb * nb) + blk)
        net_blocks = np.bincount(pair // (nb * nb), minlength=len(ns) - 1)
        cut = net_blocks[pair // (nb * nb)] > 1   # the net leaves this block
        t_per_block = np.bincount(pair % (nb * nb), weights=cut, minlength=nb * nb)
        cells = np.bincount((slot_of // g // L) * nb + (slot_of % g) // L, minlength=nb * nb)
        full = cells > 0
        pts.append((L, float(cells[full].mean()), float(t_per_block[full].mean())))
    x = np.log([q[1] for q in pts])
    y = np.log([q[2] for q in pts])
    p = float(np.polyfit(x, y, 1)[0])
    return p, pts


def next_pow2(x: int) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(1, x)))))


def workload(name: str, scale: float = 1.0):
    """(netlist, params) of a named workload.  ``scale`` shrinks cell and
    net counts (tests use it for quick variants of the same shape)."""
    if name == "c0":
        nc = max(200, int(20000 * scale))
        nl = generate(nc, int(nc * 1.05), seed=1, name="c0")
        p = params_default(target_density=0.70, init_density_penalty=8e-5,
                           init_wire_length_coef=0.25, initial_place_max_iter=20,
                           bin_grid_cnt_x=128 if scale == 1.0 else 0,
                           bin_grid_cnt_y=128 if scale == 1.0 else 0, target_overflow=0.10)
        return nl, p
    if name == "c1":
        nc = max(1000, int(1000000 * scale))
        nl = generate(nc, int(nc * 1.05), seed=2, num_macros=40, name="c1")
        nb = 1024 if scale == 1.0 else max(32, next_pow2(int(math.sqrt(nc))))
        p = params_default(target_density=0.70, init_density_penalty=8e-5,
                           init_wire_length_coef=0.25, initial_place_max_iter=20,
                           bin_grid_cnt_x=nb, bin_grid_cnt_y=nb, target_overflow=0.10)
        return nl, p
    raise KeyError(name)


def bins_for_cells(nc: int, lo: int, hi: int) -> int:
    """Bin count per side sized to the cell count: the largest power of two
    not above sqrt(Nc) (what the reference's automatic rule, BinGrid::initBins
    src/nesterovBase.cpp:651-659, gives when utilisation is near the target
    density: at most one bin per cell), clamped to [lo, hi].  A grid with
    more bins than cells (nextPow2 rounding up) leaves many instances unable
    to reach overflow 0.10."""
    return min(hi, max(lo, next_pow2(int(math.ceil(math.sqrt(nc)))) // 2))


def batch_instance(i: int):
    """Instance i of the batched workload: cells log-uniform in [500, 50k],
    target density U[0.6, 0.9], utilisation min(0.70, td - 0.01), bins
    bins_for_cells(Nc, 32, 256) per side, seed 1000 + i."""
    rng = np.random.default_rng(1000 + i)
    nc = int(round(math.exp(rng.uniform(math.log(500), math.log(50000)))))
    td = float(rng.uniform(0.6, 0.9))
    util = min(0.70, td - 0.01)
    nb = bins_for_cells(nc, 32, 256)
    nl = generate(nc, int(round(nc * 1.05)), seed=1000 + i, util=util, td=td, name=f"c3_{i}")
    p = params_default(target_density=td, bin_grid_cnt_x=nb, bin_grid_cnt_y=nb,
                       target_overflow=0.10)
    return nl, p


def large_instance(i: int):
    """A large served job: cells uniform in [200k, 1M], target density
    U[0.6, 0.9] and utilisation min(0.70, td - 0.01) as in the batched
    distribution, bins bins_for_cells(Nc, 512, 1024) per side (512 for every
    size drawn), seed 5000 + i."""
    rng = np.random.default_rng(5000 + i)
    nc = int(rng.integers(200000, 1000001))
    td = float(rng.uniform(0.6, 0.9))
    util = min(0.70, td - 0.01)
    nb = bins_for_cells(nc, 512, 1024)
    nl = generate(nc, int(round(nc * 1.05)), seed=5000 + i, util=util, td=td, name=f"large_{i}")
    p = params_default(target_density=td, bin_grid_cnt_x=nb, bin_grid_cnt_y=nb, target_overflow=0.10)
    return nl, p


def served_stream(num_jobs: int, rate: float, large_frac: float = 0.05, seed: int = 7):
    """Job stream of the served workload: Poisson arrivals at `rate` jobs/s
    (exponential gaps, seed 7); each job is a large instance with probability
    `large_frac`, otherwise an instance of the batched distribution.
    Returns [(arrival_s, kind, index)], kind in {"batch", "large"}."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    nb = nl = 0
    for _ in range(num_jobs):
        t += float(rng.exponential(1.0 / rate))
        if rng.uniform() < large_frac:
            out.append((t, "large", nl))
            nl += 1
        else:
            out.append((t, "batch", nb))
            nb += 1
    return out
