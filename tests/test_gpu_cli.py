"""The command-line front end (replace_cli, the `global_placement` flags of
src/replace.tcl) run as a child process on the reference's regression
designs, its output diffed against the reference's golden logs
(test/<design>.ok, recorded line by line in tests/golden/traces.json):
  * every [INFO] / [WARN] line of PlacerBase and NesterovBase, in order,
    exactly (src/placerBase.cpp:609-626, 394-399, 902-950,
    src/nesterovBase.cpp:645-713, 1003-1005);
  * the [NesterovSolve] lines: the same iteration numbers up to the finish
    (within +-12 iterations), overflow within 2e-3 and HPWL within 0.2 % over
    the first 50 iterations (5e-3 / 0.5 % after the parity-unpinned initial
    place);
  * the [ERROR] lines and the exit code (REPL-3 for diverge01).
The verbose-3 detail has no golden log; its lines are checked against the
run's own records (structure, HPWL and overflow per iteration).
"""
import math
import os
import re
import subprocess

import pytest

from helpers import design, traces
from replace_amd import build

pytestmark = pytest.mark.gpu

CLI = build.CLI


def _flags(flags):
    out = []
    for k, v in flags.items():
        if k == "target_density":
            out += ["-density", repr(float(v))]
        elif k == "init_density_penalty":
            out += ["-init_density_penalty", repr(float(v))]
        elif k == "initial_place_max_iter" and v == 0 and not flags.get("incremental_place_mode"):
            out += ["-skip_initial_place"]
        elif k == "incremental_place_mode":
            out += ["-incremental"]
    return out


def _run(name, tmp_path, extra=()):
    nl, _ = design(name)
    path = os.path.join(tmp_path, name + ".bin")
    nl.write_bin(path)
    t = traces()[name]
    r = subprocess.run([CLI, path, *_flags(t["flags"]), *extra], capture_output=True, text=True, timeout=120)
    return r, t


def _kind(lines, prefixes):
    """The lines of the given kinds, with the continuation line of a port
    warning."""
    return [ln for ln in lines if ln.startswith(prefixes) or ln.startswith("       Replace will regard")]


def _nesterov(lines):
    rows = []
    for ln in lines:
        m = re.match(r"\[NesterovSolve\] Iter: (\d+) overflow: (\S+) HPWL: (\d+)", ln)
        if m:
            rows.append((int(m.group(1)), float(m.group(2)), int(m.group(3))))
    return rows


@pytest.mark.parametrize("name", ["simple01", "simple02", "simple04", "ar01", "core01", "incremental01"])
def test_cli_log_matches_reference_log(name, tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("replace_cli not built")
    r, t = _run(name, tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = r.stdout.splitlines()
    want = t["log"]
    info = ("[INFO]", "[WARN]")
    assert _kind(got, info) == _kind(want, info)
    g, w = _nesterov(got), _nesterov(want)
    ip = t["flags"].get("initial_place_max_iter", 20) > 0
    tol_o, tol_h = (5e-3, 5e-3) if ip else (2e-3, 2e-3)
    for (gi, go, gh), (wi, wo, wh) in zip(g, w):
        assert gi == wi
        if gi <= 50:
            assert abs(go - wo) < tol_o and abs(gh - wh) <= tol_h * wh, (gi, go, gh, wo, wh)
    assert abs(g[-1][0] - w[-1][0]) <= 12
    fin = [ln for ln in got if ln.startswith("[NesterovSolve] Finished with Overflow:")]
    assert len(fin) == 1 and float(fin[0].split(":")[1]) <= 0.1


def test_cli_diverge01_exits_with_repl3(tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("replace_cli not built")
    r, t = _run("diverge01", tmp_path)
    assert r.returncode == 3
    got = r.stdout.splitlines()
    assert _kind(got, ("[INFO]", "[WARN]")) == _kind(t["log"], ("[INFO]", "[WARN]"))
    assert got[-2:] == t["log"][-2:]  # the [ERROR] ... (REPL-3) lines


def test_cli_checks_positive_values(tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("replace_cli not built")
    nl, _ = design("simple01")
    path = os.path.join(tmp_path, "s.bin")
    nl.write_bin(path)
    for flag, val in (("-density", "-0.5"), ("-bin_grid_count", "abc"), ("-initial_place_max_iter", "1.5")):
        r = subprocess.run([CLI, path, flag, val], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and r.stdout.startswith(f"Error: {flag} '{val}' is not a positive"), r.stdout
    # accepted and unused, as in the reference's Tcl procedure
    r = subprocess.run([CLI, path, "-timing_driven", "-wire_res", "1.0", "-wire_cap", "2.0", "-skip_initial_place",
                        "-nesterov_max_iter", "3"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout[-1000:]


def test_cli_verbose3_trace_is_consistent(tmp_path):
    """-verbose_level 3: the NesterovInit block and one block per iteration
    (src/nesterovPlace.cpp:64-175, 315-471, 514-583), whose HPWL / overflow
    values equal the [NesterovSolve] lines of the same run."""
    if not os.path.exists(CLI):
        pytest.skip("replace_cli not built")
    r, t = _run("simple01", tmp_path, ["-verbose_level", "3"])
    assert r.returncode == 0, r.stdout[-2000:]
    lines = r.stdout.splitlines()
    assert "[PROC] Begin NesterovInit ..." in lines and "[PROC] End NesterovInit" in lines
    init = lines[lines.index("[PROC] Begin NesterovInit ..."):lines.index("[PROC] End NesterovInit")]
    keys = [ln.split(" = ")[0] for ln in init[1:]]
    assert keys[:4] == ["[INFO] InitialHPWL", "[INFO] BaseWireLengthCoef", "[INFO] InitSumOverflow",
                        "[INFO] WireLengthCoef"]
    assert keys[-1] == "[INFO] InitialStepLength"
    # the verbose-3 FillerInit lines precede the unconditional ones
    fi = [ln for ln in lines if ln.startswith("[INFO] FillerInit:")]
    assert fi[0].startswith("[INFO] FillerInit: CoreArea") and fi[-3].startswith("[INFO] FillerInit: NumGCells")
    iters = [int(ln.split(" = ")[1]) for ln in lines if ln.startswith("[INFO] Iter = ")]
    solve = _nesterov(lines)
    assert iters == list(range(1, len(iters) + 1)) and len(iters) >= solve[-1][0]
    new_hpwl = [int(ln.split(" = ")[1]) for ln in lines if ln.startswith("[INFO]   NewHPWL = ")]
    ovf = [float(ln.split(" = ")[1]) for ln in lines if ln.startswith("[INFO]   Overflow = ")]
    phi = [float(ln.split(" = ")[1]) for ln in lines if ln.startswith("[INFO]   Phi = ")]
    assert len(new_hpwl) == len(ovf) == len(phi) == len(iters)
    for it, o, h in solve:
        assert new_hpwl[it - 1] == h and abs(ovf[it - 1] - o) <= 1e-6 * max(1.0, o)
    assert all(math.isfinite(v) for v in phi)
    bt = [int(ln.split(" = ")[1]) for ln in lines if ln.startswith("[INFO]   NumBackTrak = ")]
    steps = [ln for ln in lines if ln.startswith("[INFO]   NewStepLength = ")]
    assert len(bt) == len(iters) and sum(bt) == len(steps)
