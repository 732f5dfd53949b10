"""Profiling helpers for the committed rocprofv3 summaries under profiles/.

  python -m replace_amd.profiling copy [--mb 512] [--reps 20]
      a plain device-to-device copy of known size (torch), the calibration
      kernel for the FETCH_SIZE / WRITE_SIZE counters
  python -m replace_amd.profiling stats DIR [--out FILE]
      markdown summary of a rocprofv3 --kernel-trace --stats run
  python -m replace_amd.profiling pmc DIR COUNTER [--regex R]
      mean of one counter per dispatch and kernel from a --pmc run
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys
from collections import defaultdict


def _find(d, pattern):
    hits = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not hits:
        raise FileNotFoundError(f"{pattern} under {d}")
    return hits


def copy_probe(mb: int, reps: int) -> None:
    import torch
    n = mb * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy: {reps} x {n * 4} bytes read + {n * 4} bytes written per dispatch")


def _short(name: str) -> str:
    n = name.replace("(anonymous namespace)::", "").split("(")[0]
    return n.replace("void ", "")[:70]


def _kernel_rows(d: str):
    """(name, durations_ns list) from rocpd databases or kernel_trace CSVs."""
    acc = defaultdict(list)
    dbs = glob.glob(os.path.join(d, "**", "*results.db"), recursive=True)
    if dbs:
        import sqlite3
        for db in dbs:
            con = sqlite3.connect(db)
            for name, dur in con.execute("select name, duration from kernels"):
                acc[_short(name)].append(float(dur))
        return acc
    for path in _find(d, "*kernel_trace.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                acc[_short(r["Kernel_Name"])].append(float(r["End_Timestamp"]) - float(r["Start_Timestamp"]))
    return acc


def stats_summary(d: str) -> str:
    acc = _kernel_rows(d)
    total = sum(sum(v) for v in acc.values()) or 1.0
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for name, v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| {name} | {len(v)} | {sum(v) / 1e6:.2f} | {sum(v) / len(v) / 1e3:.2f} | "
                   f"{min(v) / 1e3:.2f} | {max(v) / 1e3:.2f} | {100 * sum(v) / total:.1f} |")
    return "\n".join(out)


def pmc_means(d: str, counter: str, regex: str = ""):
    import re
    acc = defaultdict(list)
    dbs = glob.glob(os.path.join(d, "**", "*results.db"), recursive=True)
    if dbs:
        import sqlite3
        for db in dbs:
            con = sqlite3.connect(db)
            q = "select kernel_name, value from counters_collection where counter_name = ?"
            for name, val in con.execute(q, (counter,)):
                n = _short(name)
                if regex and not re.search(regex, n):
                    continue
                acc[n].append(float(val))
    else:
        for path in _find(d, "*counter_collection.csv"):
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name") != counter:
                        continue
                    n = _short(r["Kernel_Name"])
                    if regex and not re.search(regex, n):
                        continue
                    acc[n].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mb", type=int, default=512)
    c.add_argument("--reps", type=int, default=20)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--out")
    p = sub.add_parser("pmc")
    p.add_argument("dir")
    p.add_argument("counter")
    p.add_argument("--regex", default="")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_probe(a.mb, a.reps)
    elif a.cmd == "stats":
        txt = stats_summary(a.dir)
        if a.out:
            open(a.out, "w").write(txt + "\n")
        print(txt)
    else:
        for k, (m, n) in sorted(pmc_means(a.dir, a.counter, a.regex).items()):
            print(f"{k}\t{a.counter}\tmean={m:.1f}\tdispatches={n}")


if __name__ == "__main__":
    main(sys.argv[1:])
