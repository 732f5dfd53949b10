"""Shared helpers of the test suite (fixture loading, trace formatting)."""
import json
import os

import numpy as np

from replace_amd.netlist import Netlist, params_default

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def traces():
    with open(os.path.join(GOLDEN, "traces.json")) as f:
        return json.load(f)


def design(name):
    """(netlist, params) of a reference regression design, flags of its .tcl."""
    t = traces()[name]
    nl = Netlist.load(os.path.join(GOLDEN, name + ".npz"), name=name)
    return nl, params_default(**t["flags"])


def nesterov_lines(records, finished=None):
    """[NesterovSolve] lines as the reference prints them (src/nesterovPlace.cpp:421-423, 460)."""
    out = []
    for i, r in enumerate(records):
        if i == 0 or (i + 1) % 10 == 0:
            out.append(f"[NesterovSolve] Iter: {i + 1} overflow: {r[0]:g} HPWL: {r[1]}")
    if finished is not None:
        out.append(f"[NesterovSolve] Finished with Overflow: {finished:g}")
    return out


def golden_lines(name):
    t = traces()[name]
    out = [f"[NesterovSolve] Iter: {i} overflow: {o} HPWL: {h}" for i, o, h in t["nesterov"]]
    if t["finished"] is not None:
        out.append(f"[NesterovSolve] Finished with Overflow: {t['finished']}")
    return out


def rel_err(a, b):
    """max |a-b| / max |b| (normwise relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(b).max()
    return float(np.abs(a - b).max() / (d if d > 0 else 1.0))
