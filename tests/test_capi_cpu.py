"""Host-side checks of the product library that need no GPU: the C ABI
library loads and exports every entry point include/replace/*.h declares,
the ctypes mirrors match the C struct layouts, and the host helpers work."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import ROOT, design
from replace_amd import capi, lefdef, netgen
from replace_amd.netlist import (Netlist, RpEvalOut, RpIterRecord, RpNbInfo, RpNesterovStats,
                                 RpNetlist, RpParams)


def _declared(header):
    txt = open(os.path.join(ROOT, "include", "replace", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rp_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = _declared("rp_capi.h")
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_facade_symbols_exported():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    out = os.popen(f"nm -DC {capi.LIB_PATH}").read()
    for m in ["replace::Replace::doInitialPlace()", "replace::Replace::doNesterovPlace()",
              "replace::Replace::setTargetDensity(float)", "replace::Replace::setBinGridCntX(int)",
              "replace::Replace::setDb(odb::dbDatabase*)", "replace::Replace::reset()"]:
        assert m in out, m


def test_params_default_matches_library():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    p = RpParams()
    capi.lib().rp_params_default(ctypes.byref(p))
    q = capi.params_default()
    for name, _ in RpParams._fields_:
        assert getattr(p, name) == getattr(q, name), name


def test_server_opts_default():
    """The service defaults (rp_server_opts_default fills the struct only,
    no device): the initial place of jobs above 8192 movable instances in
    shared rounds (batched_ip 2), 16 set-up workers, graph replay on."""
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("library not built")
    from replace_amd.netlist import RpServerOpts
    o = RpServerOpts()
    capi.lib().rp_server_opts_default(ctypes.byref(o))
    assert o.batched_ip == 2
    assert o.setup_threads == 16 and o.steps_per_poll == 8 and o.use_graph == 1


def test_struct_sizes_match_header_layout():
    # sizes of the C structs on x86-64 (pointers 8 bytes, natural alignment)
    assert ctypes.sizeof(RpParams) == 22 * 4
    assert ctypes.sizeof(RpIterRecord) == 24
    assert ctypes.sizeof(RpNesterovStats) == 48
    assert ctypes.sizeof(RpEvalOut) == 14 * 8
    assert ctypes.sizeof(RpNbInfo) == 15 * 4 + 4 + 4 * 8 + 8  # int32 x15, pad, int64 x4, float + pad
    assert ctypes.sizeof(RpNetlist) % 8 == 0


def test_netlist_roundtrip(tmp_path):
    nl, _ = design("simple02")
    path = os.path.join(tmp_path, "n.npz")
    nl.save(path)
    nl2 = Netlist.load(path)
    for a in ("row_box", "inst_box", "inst_fixed", "net_pin_start", "pin_inst", "pin_off", "pin_box",
              "inst_pin_start", "inst_pins"):
        assert np.array_equal(getattr(nl, a), getattr(nl2, a)), a
    nl.write_bin(os.path.join(tmp_path, "n.bin"))
    head = np.fromfile(os.path.join(tmp_path, "n.bin"), dtype=np.int32, count=17)
    assert head[0] == 0x52504E4C and head[10] == nl.num_insts and head[12] == nl.num_pins


def test_generator_workload_shapes():
    nl, p = netgen.workload("c0")
    assert nl.num_insts == 20000 and nl.num_nets == 21000
    deg = np.diff(nl.net_pin_start)
    assert abs((deg == 2).mean() - 0.60) < 0.02 and abs((deg == 3).mean() - 0.18) < 0.02
    assert deg.max() <= 50 and deg.min() >= 2
    assert 3.8 < deg.mean() < 4.2
    w = nl.inst_box[:, 2] - nl.inst_box[:, 0]
    assert set(np.unique(w)) <= set(range(380, 3041, 380))
    rb = nl.row_box.astype(np.int64)
    core = (rb[:, 2].max() - rb[:, 0].min()) * (rb[:, 3].max() - rb[:, 1].min())
    util = (w.astype(np.int64) * 2800).sum() / core
    assert 0.69 < util <= 0.70
    assert p.bin_grid_cnt_x == 128 and p.target_density == np.float32(0.7)


def test_generator_macros_do_not_overlap():
    nl = netgen.generate(3000, 3150, seed=2, num_macros=40)
    m = nl.inst_box[nl.inst_fixed == 1]
    assert len(m) == 40
    for a in range(len(m)):
        for b in range(a + 1, len(m)):
            ox = min(m[a, 2], m[b, 2]) - max(m[a, 0], m[b, 0])
            oy = min(m[a, 3], m[b, 3]) - max(m[a, 1], m[b, 1])
            assert ox <= 0 or oy <= 0


def test_lefdef_reader_counts():
    from helpers import traces
    nl, _ = design("simple01")
    info = traces()["simple01"]["info"]
    assert str(nl.num_insts) == info["NumInstances"]
    assert str(nl.num_pins) == info["NumPins"]
    assert str(nl.num_nets) == info["NumNets"]
