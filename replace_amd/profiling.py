"""Profiling helpers for the committed rocprofv3 summaries under profiles/.

  python -m replace_amd.profiling copy [--mb 512] [--reps 20]
      a plain device-to-device copy of known size (torch), the calibration
      kernel for the FETCH_SIZE / WRITE_SIZE counters
  python -m replace_amd.profiling stats DIR [--out FILE]
      markdown summary of a rocprofv3 --kernel-trace --stats run
  python -m replace_amd.profiling pmc DIR COUNTER [--regex R]
      mean of one counter per dispatch and kernel from a --pmc run
  python -m replace_amd.profiling gemms DIR --n 2048
      the five products of the matrix-core Poisson solve (nx = ny = n): mean
      time and TFLOP/s of each product position, for the library's own GEMM
      kernel (k_gemm_nt) and for hipBLASLt's (Cijk_*), from the kernel trace
  python -m replace_amd.profiling traffic FETCH_DIR WRITE_DIR COPYF_DIR COPYW_DIR --copy-mb 512 --out F
      HBM bytes per launch and kernel: FETCH_SIZE and WRITE_SIZE (KiB per
      dispatch) each scaled by what the same counter reads for a device copy
      of known size, written as JSON {kernel: bytes}
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


KERNEL_SOURCES = ("rp_kernels.hip", "rp_dev_struct.h", "rp_device.h")


def kernel_source_hash() -> str:
    """First 16 hex digits of the SHA-256 over the placement kernels' sources:
    a counter profile is valid for the kernels it was measured on."""
    import hashlib
    h = hashlib.sha256()
    base = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
    for name in KERNEL_SOURCES:
        with open(os.path.join(base, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def calibrated_traffic(fetch_dir, write_dir, copyf_dir, copyw_dir, copy_mb):
    """{kernel: (bytes per launch, fetch scale, write scale)} from two --pmc
    runs of the workload and two of the calibration copy."""
    nbytes = copy_mb * (1 << 20)

    def copy_kib(d, counter):
        m = pmc_means(d, counter, "copyBuffer|elementwise|copy")
        best = max(m.values(), key=lambda v: v[0])
        return best[0]

    fs = nbytes / (copy_kib(copyf_dir, "FETCH_SIZE") * 1024.0)
    ws = nbytes / (copy_kib(copyw_dir, "WRITE_SIZE") * 1024.0)
    f = pmc_means(fetch_dir, "FETCH_SIZE")
    w = pmc_means(write_dir, "WRITE_SIZE")
    out = {}
    for k in sorted(set(f) & set(w)):
        out[k] = {"hbm_bytes_per_launch": int(f[k][0] * 1024 * fs + w[k][0] * 1024 * ws),
                  "fetch_kib": round(f[k][0], 1), "write_kib": round(w[k][0], 1)}
    return out, fs, ws


def gemm_positions(d: str, n: int) -> str:
    """Per product position (G1..G5 of one solve, phi not computed): mean time
    and TFLOP/s, for k_gemm_nt and for hipBLASLt's kernels.  Launches are taken
    in start order and assigned to positions cyclically, after dropping the
    launches of hipBLASLt's set-up timing (a multiple of 5 is kept from the end)."""
    rows = []
    for path in _find(d, "*kernel_trace.csv"):
        with open(path) as f:
            rows += list(csv.DictReader(f))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    k3 = 3 * n
    flops = [2 * n * n * k3, 2 * n * n * k3, 4 * n * n * k3, 2 * n * n * k3, 2 * n * n * k3]
    out = ["| kernel | product | launches | mean us | TFLOP/s |", "|---|---|---|---|---|"]
    for label, key in (("k_gemm_nt", "k_gemm_nt"), ("hipBLASLt", "Cijk")):
        sel = [r for r in rows if key in r["Kernel_Name"]]
        if label == "hipBLASLt":
            # the set-up's timing launches come first; the solves' five per solve follow
            lt_solves = len(sel) // 5
            sel = sel[len(sel) - 5 * lt_solves:]
        if not sel:
            continue
        pos = defaultdict(list)
        for k, r in enumerate(sel):
            pos[k % 5].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        total = 0.0
        for k in range(5):
            m = sum(pos[k]) / len(pos[k])
            total += m
            if label == "k_gemm_nt":  # hipBLASLt's set-up launches leave its positions unaligned
                out.append(f"| {label} | G{k + 1} | {len(pos[k])} | {m:.2f} | {flops[k] / (m * 1e-6) / 1e12:.0f} |")
        out.append(f"| {label} | all five | | {total:.2f} | {sum(flops) / (total * 1e-6) / 1e12:.0f} |")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("copyf_dir")
    t.add_argument("copyw_dir")
    t.add_argument("--copy-mb", type=int, default=512)
    t.add_argument("--workload", default="")
    t.add_argument("--out")
    c = sub.add_parser("copy")
    c.add_argument("--mb", type=int, default=512)
    c.add_argument("--reps", type=int, default=20)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--out")
    gp = sub.add_parser("gemms")
    gp.add_argument("dir")
    gp.add_argument("--n", type=int, default=2048)
    p = sub.add_parser("pmc")
    p.add_argument("dir")
    p.add_argument("counter")
    p.add_argument("--regex", default="")
    a = ap.parse_args(argv)
    if a.cmd == "traffic":
        import json
        k, fs, ws = calibrated_traffic(a.fetch_dir, a.write_dir, a.copyf_dir, a.copyw_dir, a.copy_mb)
        doc = {"workload": a.workload, "source_sha16": kernel_source_hash(),
               "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
               "source": "rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE, KiB per dispatch, scaled by "
                         f"the same counters on a {a.copy_mb} MiB device copy", "kernels": k}
        txt = json.dumps(doc, indent=1)
        if a.out:
            open(a.out, "w").write(txt + "\n")
        print(txt)
    elif a.cmd == "gemms":
        print(gemm_positions(a.dir, a.n))
    elif a.cmd == "copy":
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
