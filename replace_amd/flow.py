"""LEF/DEF in, placed DEF out: the reference's Tcl flow
``read_lef; read_def; global_placement ...; write_def`` (test/*.tcl,
src/replace.tcl:8-97) on this build's placer.

  python -m replace_amd.flow --lef tech.lef --def design.def --out placed.def \
      [-density 0.7] [-init_density_penalty 8e-5] [-skip_initial_place] \
      [-incremental] [-bin_grid_count n] [-overflow 0.1]

The placer runs on the GPU through the C ABI.  The flags keep the Tcl
command's names and defaults (density 0.7, -skip_initial_place = no initial
place iterations, -incremental = incremental mode without them).
"""
from __future__ import annotations

import argparse
import sys

from .lefdef import read_def, read_lef, write_def
from .netlist import params_default


def parse(argv):
    ap = argparse.ArgumentParser(prog="replace_amd.flow")
    ap.add_argument("--lef", required=True, action="append")
    ap.add_argument("--def", dest="defn", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("-density", type=float, default=0.7)
    ap.add_argument("-init_density_penalty", type=float, default=None)
    ap.add_argument("-init_wirelength_coef", type=float, default=None)
    ap.add_argument("-overflow", type=float, default=None)
    ap.add_argument("-bin_grid_count", type=int, default=None)
    ap.add_argument("-initial_place_max_iter", type=int, default=None)
    ap.add_argument("-skip_initial_place", action="store_true")
    ap.add_argument("-incremental", action="store_true")
    return ap.parse_args(argv)


def params_from_flags(a):
    kw = dict(target_density=a.density)
    if a.init_density_penalty is not None:
        kw["init_density_penalty"] = a.init_density_penalty
    if a.init_wirelength_coef is not None:
        kw["init_wire_length_coef"] = a.init_wirelength_coef
    if a.overflow is not None:
        kw["target_overflow"] = a.overflow
    if a.bin_grid_count is not None:
        kw["bin_grid_cnt_x"] = kw["bin_grid_cnt_y"] = a.bin_grid_count
    if a.initial_place_max_iter is not None:
        kw["initial_place_max_iter"] = a.initial_place_max_iter
    if a.skip_initial_place:
        kw["initial_place_max_iter"] = 0
    if a.incremental:
        kw["incremental_place_mode"] = 1
        kw["initial_place_max_iter"] = 0
    return params_default(**kw)


def main(argv=None) -> int:
    a = parse(sys.argv[1:] if argv is None else argv)
    lef = read_lef(a.lef[0])
    for extra in a.lef[1:]:
        more = read_lef(extra, lef.dbu)
        lef.sites.update(more.sites)
        lef.macros.update(more.macros)
    nl = read_def(a.defn, lef)
    p = params_from_flags(a)
    from .capi import Placer
    placer = Placer(nl, p)
    placer.initial_place()
    placer.setup()
    code, st = placer.nesterov_place()
    lx, ly = placer.positions()
    write_def(a.defn, a.out, lx, ly)
    print(f"[NesterovSolve] Finished with Overflow: {st.final_overflow:g}  (code {code}, "
          f"{st.iterations} iterations) -> {a.out}")
    return int(code)


if __name__ == "__main__":
    sys.exit(main())
