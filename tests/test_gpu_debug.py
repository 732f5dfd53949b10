"""Debug dumps (replace_amd.debug): the fields they write are the GPU's own
evaluation at the last evaluated gcell centres, and those centres are the
ones the final positions come from."""
import numpy as np
import pytest

import oracle
from helpers import design
from replace_amd import debug
from replace_amd.capi import Placer

pytestmark = pytest.mark.gpu


def test_dump_fields_match_oracle_at_final_centres(tmp_path):
    nl, p = design("simple01")
    g = Placer(nl, p)
    g.initial_place()
    info = g.setup()
    g.nesterov_place()
    cx, cy = g.gcell_centres()
    gc = g.gcells()
    # updateDb (src/nesterovPlace.cpp:585-595): LL = dC - dD/2 for placed instances
    lx, ly = g.positions()
    movable = np.flatnonzero(nl.inst_fixed == 0)
    npl = info.num_place_insts
    assert np.array_equal(lx[movable], cx[:npl].astype(np.int32) - gc["dhalf_x"][:npl])
    assert np.array_equal(ly[movable], cy[:npl].astype(np.int32) - gc["dhalf_y"][:npl])
    paths = debug.dump(g, str(tmp_path), "simple01")
    dens = np.load(paths["density"])
    o = oracle.Oracle(nl, p, poisson="auto")
    o.initial_place()
    o.setup()
    r = o.eval_gradient(cx, cy, 0.0, 1.0, 1.0)
    assert np.array_equal(dens.reshape(-1), r["density"])  # bin density: bit-exact
    phi = np.load(paths["phi"]).reshape(-1)
    assert np.abs(phi - r["phi"]).max() <= 1e-4 * np.abs(r["phi"]).max()
    for k in ("density_png", "cells_png", "field_png"):
        from PIL import Image
        assert Image.open(paths[k]).size == (512, 512)
