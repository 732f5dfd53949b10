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


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror against the C header: a probe compiled with gcc
    prints sizeof and the offset of every field of each rp_capi.h struct."""
    from replace_amd import netlist as N
    mirrors = {"rp_netlist": N.RpNetlist, "rp_params": N.RpParams, "rp_iter_record": N.RpIterRecord,
               "rp_nesterov_stats": N.RpNesterovStats, "rp_ip_stats": N.RpIpStats, "rp_nb_info": N.RpNbInfo,
               "rp_pb_info": N.RpPbInfo, "rp_eval_out": N.RpEvalOut, "rp_eval_record": N.RpEvalRecord,
               "rp_nesterov_state": N.RpNesterovState, "rp_server_opts": N.RpServerOpts,
               "rp_tile_choice": N.RpTileChoice, "rp_job_result": N.RpJobResult,
               "rp_server_stats": N.RpServerStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "replace/rp_capi.h"', "int main(void) {"]
    for cname, m in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in m._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0;\n}")
    src = os.path.join(tmp_path, "probe.c")
    exe = os.path.join(tmp_path, "probe")
    open(src, "w").write("\n".join(lines))
    import subprocess
    subprocess.check_call(["gcc", "-I" + os.path.join(ROOT, "include"), src, "-o", exe])
    got = {}
    for line in subprocess.check_output([exe]).decode().splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, m in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(m), cname
        for f, _ in m._fields_:
            assert got[(cname, f)] == getattr(m, f).offset, (cname, f)


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
    assert head[0] == 0x52504E4C and head[1] == 2 and head[10] == nl.num_insts and head[12] == nl.num_pins
    assert nl2.pin_names == nl.pin_names and sum(1 for v in nl.pin_names if v) == 54


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


def test_generator_rent_exponent():
    """The 20k-cell workload's connectivity follows Rent's rule with p near
    0.6 (the configuration's clustered, Rent-like netlist): on the
    generator's virtual slot grid, blocks of 4 to ~2200 cells have on average
    T = t G^p external nets with p in [0.5, 0.7], and T grows monotonically
    with G."""
    out = []
    nl = netgen.generate(20000, 21000, seed=1, name="c0", slots_out=out)
    ref, _ = netgen.workload("c0")
    assert np.array_equal(nl.pin_inst, ref.pin_inst)  # the workload itself
    p, pts = netgen.rent_profile(nl, out[0], out[1])
    assert 0.5 <= p <= 0.7, (p, pts)
    ts = [t for _, _, t in pts]
    assert all(a < b for a, b in zip(ts, ts[1:]))
    # a block of one slot side 2 holds 4 cells with ~4 pins each, nearly all
    # on nets that leave it
    assert 10.0 < pts[0][2] < 17.0


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
