"""Convergence-debugging dumps of a placement on the GPU: bin density,
potential and field as ``.npy`` arrays and PNG pictures, the equivalents of
the reference's CImg plots (PlotEnv, src/plot.cpp:284-426: DrawBinDensity,
DrawGcell / DrawModule, DrawArrowDensity).

  python -m replace_amd.debug --netlist design.npz --iters 1 50 200 --out dumps/

or from Python::

    from replace_amd import debug
    debug.dump(placer, "dumps/", "simple01_it200")

``dump`` re-evaluates the density scatter, Poisson solve and gradients at the
context's last evaluated gcell centres (rp_get_gcell_centres, then the
stage-level rp_eval_gradient) and writes

  <tag>_density.npy / _phi.npy / _ex.npy / _ey.npy   (Ny x Nx float32, bin
      (i, j) at [j, i], the reference's y-major bin order)
  <tag>_density.png   bins shaded as DrawBinDensity: 255 - clamp(50 rho + 20)
  <tag>_cells.png     gcells as DrawGcell: std cells red, macros blue, fillers
                      grey, over the density picture
  <tag>_field.png     one arrow per (Nx/16)-th bin along each axis, scaled by
                      the largest |Ex| and |Ey| as DrawArrowDensity

``trace`` runs fresh placements stopped after each requested iteration count
and dumps each one (the loop runs on the device, so an intermediate state is
reached by capping the iteration count).
"""
from __future__ import annotations

import argparse
import os
import sys
from typing import Iterable

import numpy as np

from .capi import Placer

_PX = 512  # picture side in pixels


def fields(placer: Placer):
    """(density, phi, ex, ey) on the bin grid, Ny x Nx, at the last
    evaluated gcell centres of ``placer``."""
    info = placer.info
    cx, cy = placer.gcell_centres()
    out = placer.eval_gradient(cx, cy, 0.0, 1.0, 1.0)
    shape = (info.bin_cnt_y, info.bin_cnt_x)
    return tuple(out[k].reshape(shape) for k in ("density", "phi", "ex", "ey"))


def _canvas(placer: Placer):
    from PIL import Image
    info = placer.info
    w = max(1, info.core_ux - info.core_lx)
    h = max(1, info.core_uy - info.core_ly)
    sx = (_PX - 1) / w
    sy = (_PX - 1) / h

    def to_px(x, y):  # picture y grows downwards, layout y upwards
        return (x - info.core_lx) * sx, (_PX - 1) - (y - info.core_ly) * sy

    return Image, to_px


def _density_picture(placer: Placer, density: np.ndarray):
    Image, to_px = _canvas(placer)
    from PIL import ImageDraw
    info = placer.info
    img = Image.new("RGB", (_PX, _PX), (255, 255, 255))
    dr = ImageDraw.Draw(img)
    for j in range(info.bin_cnt_y):
        for i in range(info.bin_cnt_x):
            c = int(density[j, i] * 50 + 20)
            c = 255 - min(255, max(20, c))
            lx = info.core_lx + i * info.bin_size_x
            ly = info.core_ly + j * info.bin_size_y
            x0, y1 = to_px(lx, ly)
            x1, y0 = to_px(min(lx + info.bin_size_x, info.core_ux), min(ly + info.bin_size_y, info.core_uy))
            dr.rectangle([x0, y0, x1, y1], fill=(c, c, c))
    return img


def dump(placer: Placer, out_dir: str, tag: str) -> dict:
    """Writes the arrays and pictures listed in the module docstring; returns
    their paths."""
    from PIL import ImageDraw
    os.makedirs(out_dir, exist_ok=True)
    density, phi, ex, ey = fields(placer)
    paths = {}
    for name, a in (("density", density), ("phi", phi), ("ex", ex), ("ey", ey)):
        paths[name] = os.path.join(out_dir, f"{tag}_{name}.npy")
        np.save(paths[name], a)
    info = placer.info
    _, to_px = _canvas(placer)

    img = _density_picture(placer, density)
    paths["density_png"] = os.path.join(out_dir, f"{tag}_density.png")
    img.save(paths["density_png"])

    cells = img.copy()
    dr = ImageDraw.Draw(cells)
    g = placer.gcells()
    cx, cy = placer.gcell_centres()
    colour = {0: (220, 40, 40), 1: (40, 40, 220), 2: (150, 150, 150)}
    for k in range(len(cx)):
        hx, hy = int(g["dx"][k]) // 2, int(g["dy"][k]) // 2
        x0, y1 = to_px(cx[k] - hx, cy[k] - hy)
        x1, y0 = to_px(cx[k] + hx, cy[k] + hy)
        dr.rectangle([x0, y0, max(x0, x1), max(y0, y1)], outline=colour.get(int(g["kind"][k]), (0, 0, 0)))
    paths["cells_png"] = os.path.join(out_dir, f"{tag}_cells.png")
    cells.save(paths["cells_png"])

    arrows = img.copy()
    dr = ImageDraw.Draw(arrows)
    step = max(1, info.bin_cnt_x // 16)
    sub_x = np.abs(ex[::step, ::step])
    sub_y = np.abs(ey[::step, ::step])
    exmax = float(sub_x.max()) or 1.0
    eymax = float(sub_y.max()) or 1.0
    length = float(np.hypot(info.bin_size_x, info.bin_size_y)) * 5
    for j in range(0, info.bin_cnt_y, step):
        for i in range(0, info.bin_cnt_x, step):
            bx = info.core_lx + i * info.bin_size_x + info.bin_size_x / 2
            by = info.core_ly + j * info.bin_size_y + info.bin_size_y / 2
            dx = ex[j, i] / exmax
            dy = ey[j, i] / eymax
            p0 = to_px(bx, by)
            p1 = to_px(bx + dx * length, by + dy * length)
            dr.line([p0, p1], fill=(220, 0, 0), width=2)
            dr.ellipse([p1[0] - 2, p1[1] - 2, p1[0] + 2, p1[1] + 2], fill=(220, 0, 0))
    paths["field_png"] = os.path.join(out_dir, f"{tag}_field.png")
    arrows.save(paths["field_png"])
    return paths


def trace(netlist, params, iterations: Iterable[int], out_dir: str, tag: str) -> list:
    """Places ``netlist`` once per entry of ``iterations`` with the Nesterov
    loop capped there and dumps each state; returns the path dicts."""
    out = []
    for n in iterations:
        p = type(params).from_buffer_copy(params)
        p.nesterov_place_max_iter = int(n)
        g = Placer(netlist, p)
        g.initial_place()
        g.setup()
        g.nesterov_place()
        out.append(dump(g, out_dir, f"{tag}_it{n}"))
        g.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="replace_amd.debug")
    ap.add_argument("--netlist", help="netlist .npz (Netlist.save)")
    ap.add_argument("--workload", help="synthetic workload name (replace_amd.netgen)")
    ap.add_argument("--iters", type=int, nargs="+", default=[1, 50, 200])
    ap.add_argument("--out", default="dumps")
    a = ap.parse_args(sys.argv[1:] if argv is None else argv)
    from .netlist import Netlist, params_default
    if a.netlist:
        nl, p = Netlist.load(a.netlist), params_default()
        tag = os.path.splitext(os.path.basename(a.netlist))[0]
    else:
        from . import netgen
        nl, p = netgen.workload(a.workload or "c0")
        tag = a.workload or "c0"
    for paths in trace(nl, p, a.iters, a.out, tag):
        print(paths["density_png"], paths["cells_png"], paths["field_png"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
