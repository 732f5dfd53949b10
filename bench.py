#!/usr/bin/env python3
"""Benchmark of the placement hot path on one MI355X.

Metric: seconds from the start of doNesterovPlace (netlist resident in HBM,
after the initial place) to the first iteration with overflow <= 0.10, on
the 1M-cell workload (1,000,000 movable cells, 1,050,000 nets, ~4.2M pins,
40 fixed macros, 1024 x 1024 bins, target density 0.70, FP32 with int64
density atomics).  One step = one full Nesterov placement to overflow 0.10.

Besides the contract fields the JSON line carries
  roofline      the dominant kernel's algorithmic bytes per launch / its mean
                launch time (HIP events on the library stream, timed steps)
  cpu_baseline  the oracle (single-threaded CPU restatement, the reference's
                own Ooura FFT when oracle/_ref is built) pinned to one core:
                seconds per gradient evaluation on a bounded sample x the
                GPU run's evaluation count
  batched       (default workload) 256 uneven instances through the
                placement service against the same jobs back to back on the
                plain path (--workload c3 prints the full record)
  served        (default workload) a Poisson stream with 5% large jobs
                through the service against a FIFO replay on the plain path
                (--workload c4 prints the full record)
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s to overflow≤0.10 at 1M cells; served placements/s + p99 at ≤0.5% HPWL"
HBM_PEAK_GBS = 8000.0
# timing slot -> kernel name in rocprofv3 output (profiles/traffic_*.json)
ROCPROF_NAME = {"scatter": "k_scatter", "scatter_wa": "k_scatter_wa", "control": "k_control_main"}


def algorithmic_bytes(nc, npin, nn, nb, sort_bits=16):
    """Bytes one launch of each kernel must move: its inputs read once and
    its outputs written once.
      step           56 B/cell (read cur SLP, gradient, coordinate, half
                     size; write next coordinate, next SLP, integer centre)
      scatter        25 B/cell (order index, integer centre, half size,
                     scale, kind) + 16 B/bin (two int64 accumulators)
      pin_pos        reserved slot (0: the wirelength kernel computes the
                     pin centres)
      binfin_rowdct  40 B/bin (read the inst and filler int64 accumulators and
                     the float nonPlace area; write the row transform and the
                     zeroed accumulators; rho itself is stored only for the
                     matrix-core transform)
      col_spectral   12 B/bin (read the row transform, write Y1, Y2)
      row_inverse    16 B/bin (read Y1, Y2; write Ex, Ey)
      mfma_poisson   the matrix-core solve (flops, see the precision leg)
      wa             20 B/pin (cell id, offset, per-pin gradient out) +
                     8 B/cell (the integer centres the pin centres are made
                     of) + 4 B/net
      scatter_wa     scatter + wa: the single-instance path runs both in one
                     launch
      grad           64 B/cell + 12 B/pin (pin index, per-pin gradient)
                     + 8 B/bin (field)
      control        5 doubles per gradient block
      sort_refresh   12 B/cell (centre in, key out) + 16 B/cell per 8-bit
                     radix pass (key and value in and out); it runs once per
                     batch of queued slots, so its per-slot figure is amortised
    """
    return {
        "step": 56 * nc,
        "scatter": 25 * nc + 16 * nb,
        "pin_pos": 0,
        "binfin_rowdct": 40 * nb,
        "mfma_poisson": 0,
        "col_spectral": 12 * nb,
        "row_inverse": 16 * nb,
        "wa": 20 * npin + 8 * nc + 4 * nn,
        "grad": 64 * nc + 12 * npin + 8 * nb,
        "control": 8 * 5 * ((nc + 255) // 256),
        "sort_refresh": (12 + 16 * ((sort_bits + 7) // 8)) * nc,
    }


def copy_bandwidth_gbs(mib=1024, reps=10):
    """Achievable HBM bandwidth on this box: a device-to-device copy of `mib`
    MiB (read + write counted), best of `reps`, timed with events."""
    import torch
    a = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del a, b
    torch.cuda.empty_cache()
    return round(2 * (mib << 20) / (best * 1e-3) / 1e9, 1)


def cpu_baseline(workload, gpu_evals, sample_evals, core):
    code = f"""
import json, os, sys, time
sys.path.insert(0, {ROOT!r})
os.sched_setaffinity(0, {{{core}}})
import oracle
from replace_amd import netgen
nl, p = netgen.workload({workload!r})
p.initial_place_max_iter = 0
o = oracle.Oracle(nl, p, poisson='auto')
o.initial_place()
o.setup()
t0 = time.perf_counter()
o.nesterov_init()
t1 = time.perf_counter()
o.nesterov_iterate({sample_evals})
t2 = time.perf_counter()
c, st = o.nesterov_iterate(0)
print(json.dumps(dict(init_s=t1 - t0, iter_s=t2 - t1, evals=st.evals, poisson=o.poisson, stages=o.stage_times())))
"""
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        raise RuntimeError(out.stderr[-2000:])
    r = json.loads(out.stdout.strip().splitlines()[-1])
    loop_evals = r["evals"] - 2
    per_eval = r["iter_s"] / max(1, loop_evals)
    est = r["init_s"] + per_eval * max(0, gpu_evals - 2)
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    # where the restatement's time goes, per evaluation of the sample (its own
    # per-stage timers: the two init evaluations and the loop's)
    split = {k: round(v / max(1, r["evals"]), 5) for k, v in r["stages"].items()}
    return {
        "value": round(est, 3),
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "split_s_per_eval": split,
        "sample": (f"oracle {workload}: init (2 evals) {r['init_s']:.2f}s + {loop_evals} loop evals "
                   f"{r['iter_s']:.2f}s = {per_eval:.3f}s/eval on 1 core ({cpu}); value = init + "
                   f"s/eval x {gpu_evals - 2} GPU loop evals; Poisson: {r['poisson']}"),
    }


_PHASES = {"create": 0.0, "initial_place": 0.0, "setup": 0.0, "nesterov": 0.0}
SETUP_THREADS = [16]  # host workers of the placement service (--setup-threads)
BATCHED_IP = [2]  # the service's batched initial place (--batched-ip)
TILES = [None]  # tile shapes the service's tuner chose


def _plain_job(nl, p):
    """One job on the plain path (create, initial place, set-up, Nesterov)."""
    from replace_amd.capi import Placer
    t = time.perf_counter()
    g = Placer(nl, p)
    t1 = time.perf_counter()
    g.initial_place()
    t2 = time.perf_counter()
    g.setup()
    t3 = time.perf_counter()
    code, st = g.nesterov_place()
    g.close()
    t4 = time.perf_counter()
    _PHASES["create"] += t1 - t
    _PHASES["initial_place"] += t2 - t1
    _PHASES["setup"] += t3 - t2
    _PHASES["nesterov"] += t4 - t3
    return (code, int(st.final_hpwl), float(st.final_overflow), int(st.iterations)), t4 - t


def _served_common(jobs, arrivals, steps_per_poll):
    """Runs `jobs` through the placement service and back to back on the
    plain path.  Returns (served results, served wall s, plain results, plain
    service times)."""
    from replace_amd.capi import Server
    _plain_job(*jobs[0])  # warm-up: module load, first allocations
    srv = Server(steps_per_poll=steps_per_poll, setup_threads=SETUP_THREADS[0], batched_ip=BATCHED_IP[0])
    ids = [srv.submit(nl, p, a) for (nl, p), a in zip(jobs, arrivals)]
    t = time.perf_counter()
    srv.run()
    wall = time.perf_counter() - t
    res = [srv.result(j) for j in ids]
    stats = srv.stats()
    TILES[0] = [t for t in srv.tuning() if t["tuned"]]
    srv.close()
    plain, svc = [], []
    for k in _PHASES:
        _PHASES[k] = 0.0
    for nl, p in jobs:
        r, dt = _plain_job(nl, p)
        plain.append(r)
        svc.append(dt)
    return res, wall, stats, plain, svc


def _parity(res, plain, jobs):
    diffs = [abs(r.stats.final_hpwl - q[1]) / max(1, q[1]) for r, q in zip(res, plain)]
    unfinished = [{"job": k, "code": int(r.code), "iterations": int(r.stats.iterations),
                   "overflow": round(float(r.stats.final_overflow), 4), "cells": int(jobs[k][0].num_insts),
                   "td": round(float(jobs[k][1].target_density), 3)}
                  for k, r in enumerate(res) if not (r.done == 1 and r.stats.finished == 1)]
    fin = [float(r.stats.final_overflow) for r in res if r.done == 1 and r.stats.finished == 1]
    return {"hpwl_max_rel_diff_vs_plain": max(diffs),
            "max_final_overflow_finished": max(fin) if fin else None,
            "jobs_finished": len(fin), "all_finished": not unfinished, "unfinished": unfinished[:20]}


def _finished(res):
    """Jobs that reached the overflow target (the placements counted in the
    throughput; jobs that stop on the reference's divergence rules, the same
    way on the plain path, are listed separately)."""
    return sum(1 for r in res if r.done == 1 and r.stats.finished == 1)


def _timeline(res):
    import numpy as np
    setup = np.array([r.start_s - r.admit_s for r in res])
    run = np.array([r.done_s - r.start_s for r in res])
    return {"setup_s": {"p50": round(float(np.percentile(setup, 50)), 4), "max": round(float(setup.max()), 4)},
            "on_gpu_s": {"p50": round(float(np.percentile(run, 50)), 4), "max": round(float(run.max()), 4)},
            "last_start_s": round(max(r.start_s for r in res), 4),
            "last_done_s": round(max(r.done_s for r in res), 4)}


def bench_batch(args):
    """256 uneven instances submitted together: batched service vs back to back."""
    from replace_amd import netgen
    n = args.jobs or 256
    t = time.perf_counter()
    jobs = [netgen.batch_instance(i) for i in range(n)]
    gen_s = time.perf_counter() - t
    res, wall, stats, plain, svc = _served_common(jobs, [0.0] * n, args.steps_per_poll)
    plain_s = sum(svc)
    cells = sum(nl.num_insts for nl, _ in jobs)
    out = {
        "metric": "placements/s, 256 uneven instances (batched service; jobs reaching overflow 0.10)",
        "value": round(_finished(res) / wall, 3), "unit": "placements/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} instances, cells log-uniform 500-50k (sum {cells}), td U[0.6,0.9], "
                               "bins clamp(nextPow2(ceil(sqrt(Nc)))/2,32,256)",
                   "steps_per_poll": args.steps_per_poll},
        "wall_s": round(wall, 3),
        "plain_back_to_back": {"value": round(_finished(res) / plain_s, 3), "unit": "placements/s",
                               "wall_s": round(plain_s, 3),
                               "phases_s": {k: round(v, 3) for k, v in _PHASES.items()}},
        "speedup_vs_plain": round(plain_s / wall, 2),
        "batched_steps": int(stats.steps), "table_rebuilds": int(stats.rebuilds),
        "netgen_s": round(gen_s, 2), "timeline": _timeline(res), "setup_threads": args.setup_threads,
        "batched_ip": args.batched_ip,
        "tiles": TILES[0],
    }
    out.update(_parity(res, plain, jobs))
    return out


def bench_serve(args):
    """Poisson arrivals (batched-distribution jobs + 5% large): served latency
    and throughput vs the same stream back to back on the plain path (FIFO
    single server, timeline from each job's measured plain service time)."""
    from replace_amd import netgen
    n = args.jobs or 200
    stream = netgen.served_stream(n, args.rate)
    t = time.perf_counter()
    jobs = [netgen.batch_instance(i) if kind == "batch" else netgen.large_instance(i) for _, kind, i in stream]
    gen_s = time.perf_counter() - t
    arrivals = [a for a, _, _ in stream]
    res, wall, stats, plain, svc = _served_common(jobs, arrivals, args.steps_per_poll)
    import numpy as np
    lat = np.array([r.done_s - r.arrival_s for r in res])
    done = 0.0
    plat = []
    for a, dt in zip(arrivals, svc):
        done = max(done, a) + dt
        plat.append(done - a)
    plat = np.array(plat)
    makespan = max(r.done_s for r in res)
    out = {
        "metric": "served placements/s + p99 latency (Poisson arrivals, 5% large jobs; jobs reaching overflow 0.10)",
        "value": round(_finished(res) / makespan, 3), "unit": "placements/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} jobs, Poisson rate {args.rate}/s (seed 7), 95% batched distribution, "
                               "5% large (200k-1M cells, 512x512 bins)",
                   "large_jobs": sum(1 for _, k, _ in stream if k == "large"),
                   "steps_per_poll": args.steps_per_poll},
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                      "max": round(float(lat.max()), 4)},
        "plain_back_to_back": {"value": round(_finished(res) / done, 3), "unit": "placements/s",
                               "latency_s": {"p50": round(float(np.percentile(plat, 50)), 4),
                                             "p99": round(float(np.percentile(plat, 99)), 4)},
                               "service_s_total": round(sum(svc), 3)},
        "speedup_vs_plain": round(done / makespan, 2),
        "batched_steps": int(stats.steps), "table_rebuilds": int(stats.rebuilds),
        "netgen_s": round(gen_s, 2), "timeline": _timeline(res), "batched_ip": args.batched_ip,
    }
    out.update(_parity(res, plain, jobs))
    return out


def bench_precision(args):
    """1M-cell netlist on a 2048 x 2048 grid: time to overflow 0.10 with the
    Poisson solve as split-BF16 GEMMs on matrix cores vs the FP32 FFT path."""
    import torch
    from replace_amd import netgen
    from replace_amd.capi import Placer
    nl, p = netgen.workload("c1")
    p.bin_grid_cnt_x = p.bin_grid_cnt_y = 2048
    out = {}
    # 0: FP32 FFT; 1: matrix cores, products by the library's own GEMM kernel;
    # "lt": the same products through hipBLASLt (a reference point)
    for mode in (0, 1, "lt"):
        p.poisson_transform = 0 if mode == 0 else 1
        if mode == "lt":
            os.environ["RP_MFMA_GEMM"] = "hipblaslt"
        g = Placer(nl, p)
        g.initial_place()
        info = g.setup()  # the matrix-core set-up reads RP_MFMA_GEMM
        os.environ.pop("RP_MFMA_GEMM", None)
        for _ in range(args.warmup):
            g.nesterov_place()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            code, st = g.nesterov_place()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        g.set_kernel_timing(True)  # per-kernel times from one more, separately timed placement
        g.nesterov_place()
        g.set_kernel_timing(False)
        kt = g.kernel_timing()
        out[mode] = {"s": dt, "code": code, "iterations": st.iterations, "evals": st.evals,
                     "finished": int(st.finished), "final_overflow": float(st.final_overflow),
                     "final_hpwl": int(st.final_hpwl),
                     "kernels_us": {k: round(v[0] / max(1, v[1]) * 1e3, 2) for k, v in kt.items()}}
        g.close()
    N = 2048
    # 5 products without phi: N^3 + N^3 + 2 N^3 + N^3 + N^3 = 6 N^3 multiply-adds, each for
    # 3 split components (hi.hi, hi.lo, lo.hi), 2 flops per multiply-add
    flops = 3 * 2 * 6 * N ** 3
    poisson_us = out[1]["kernels_us"]["mfma_poisson"]
    achieved = flops / (poisson_us * 1e-6) / 1e12
    res = {
        "metric": "s to overflow<=0.10 at 1M cells, 2048x2048 bins (matrix-core split-BF16 Poisson)",
        "value": round(out[1]["s"], 4), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "bf16x2 (split) GEMM, f32 accumulate", "data": "synthetic",
        "config": {"workload": "1M-cell netlist, 2048x2048 bins, td 0.70", "steps": args.steps},
        "fft_path": {"s": round(out[0]["s"], 4), "iterations": out[0]["iterations"],
                     "final_hpwl": out[0]["final_hpwl"], "kernels_us": out[0]["kernels_us"]},
        "mfma_path": {"iterations": out[1]["iterations"], "final_hpwl": out[1]["final_hpwl"],
                      "final_overflow": round(out[1]["final_overflow"], 5), "kernels_us": out[1]["kernels_us"]},
        "mfma_hipblaslt": {"s": round(out["lt"]["s"], 4), "iterations": out["lt"]["iterations"],
                           "final_hpwl": out["lt"]["final_hpwl"],
                           "mfma_poisson_us": out["lt"]["kernels_us"]["mfma_poisson"]},
        "capped": {"finished": [out[0]["finished"], out[1]["finished"]],
                   "note": "neither transform reaches overflow 0.10 on 2048x2048 bins (4.2M bins for 1M cells): "
                           "the overflow plateaus near 0.12 and the run stops at the 2000-iteration cap; "
                           "value is the time of the capped run"},
        "speedup_vs_fft": round(out[0]["s"] / out[1]["s"], 3),
        "hpwl_rel_diff_vs_fft": abs(out[1]["final_hpwl"] - out[0]["final_hpwl"]) / out[0]["final_hpwl"],
        "roofline": {"bound": "mfma", "kernel": "poisson_mfma", "achieved": round(achieved, 1),
                     "peak": 2500.0, "unit": "TFLOP/s", "frac": round(achieved / 2500.0, 4),
                     "flops_per_launch": flops, "avg_launch_us": poisson_us},
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c2", "c3", "c4"])
    ap.add_argument("--cpu-sample-evals", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--jobs", type=int, default=None, help="c3/c4: number of jobs")
    ap.add_argument("--no-service", action="store_true",
                    help="c1: leave out the batched (c3) and served (c4) summaries of the default line")
    ap.add_argument("--rate", type=float, default=27.0,
                    help="c4: arrival rate, jobs/s (default: ~80%% of the measured 34/s batched capacity)")
    ap.add_argument("--steps-per-poll", type=int, default=8, help="c3/c4: batched steps per graph replay")
    ap.add_argument("--setup-threads", type=int, default=16, help="c3/c4: host set-up workers of the service")
    ap.add_argument("--batched-ip", type=int, default=2, choices=[0, 1, 2],
                    help="c3/c4: initial place in shared launches for the jobs above 8192 movable "
                         "instances (2, default), for every job (1) or none (0)")
    args = ap.parse_args()
    SETUP_THREADS[0] = args.setup_threads
    BATCHED_IP[0] = args.batched_ip
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build places one instance per GPU and "
              "has no multi-GPU path (see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c2":
        import torch
        torch.cuda.set_device(0)
        bench_precision(args)
        return
    if args.workload in ("c3", "c4"):
        import torch
        torch.cuda.set_device(0)
        print(json.dumps((bench_batch if args.workload == "c3" else bench_serve)(args)), flush=True)
        return

    import torch
    torch.cuda.set_device(0)
    from replace_amd import netgen
    from replace_amd.capi import Placer

    t = time.perf_counter()
    nl, p = netgen.workload(args.workload)
    gen_s = time.perf_counter() - t
    placer = Placer(nl, p)
    t = time.perf_counter()
    ip = placer.initial_place()
    ip_s = time.perf_counter() - t
    t = time.perf_counter()
    info = placer.setup()
    setup_s = time.perf_counter() - t
    for _ in range(args.warmup):
        placer.nesterov_place()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    runs = []
    for _ in range(args.steps):
        code, st = placer.nesterov_place()
        runs.append((code, st.iterations, st.evals, st.final_overflow, st.final_hpwl, st.seconds,
                     st.finished))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) / args.steps * 1e3
    # Per-kernel times come from one more placement, untimed above: the HIP
    # events bracketing every kernel add their own cost to the run.
    placer.set_kernel_timing(True)
    placer.nesterov_place()
    placer.set_kernel_timing(False)
    kt = placer.kernel_timing()
    nb = info.bin_cnt_x * info.bin_cnt_y
    abytes = algorithmic_bytes(info.num_gcells, info.num_gpins, info.num_gnets, nb)
    if kt.get("wa", (0.0, 0))[0] == 0.0 and kt.get("scatter", (0.0, 0))[0] > 0.0:
        # one launch does the scatter and the wirelength (the scatter slot
        # timed it): its bytes are the two kernels' bytes
        kt = dict(kt)
        kt["scatter_wa"] = kt.pop("scatter")
        kt.pop("wa", None)
        abytes["scatter_wa"] = abytes["scatter"] + abytes["wa"]
    kernels = {}
    for k, (tot, n) in kt.items():
        avg = tot / max(1, n)
        if k == "sort_refresh":  # once per batch of queued slots: amortised per slot
            kernels[k] = {"avg_us_per_slot": round(avg * 1e3, 2), "total_ms": round(tot, 2)}
            continue
        kernels[k] = {"avg_us": round(avg * 1e3, 2), "launches": n,
                      "GBs": round(abytes[k] / (avg * 1e-3) / 1e9, 1) if avg > 0 and abytes[k] else None}
    dom = max((k for k in kt if k != "sort_refresh"), key=lambda k: kt[k][0])
    copy_gbs = copy_bandwidth_gbs()
    avg_ms = kt[dom][0] / max(1, kt[dom][1])
    achieved = abytes[dom] / (avg_ms * 1e-3) / 1e9
    traffic, traffic_note = None, "no counter profile for this workload"
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath):
        from replace_amd.profiling import kernel_source_hash
        doc = json.load(open(tpath))
        if doc.get("source_sha16") != kernel_source_hash():
            traffic_note = (f"{os.path.basename(tpath)} was measured on other kernel sources "
                            f"({doc.get('source_sha16')} != {kernel_source_hash()}): not used")
        else:
            name = ROCPROF_NAME.get(dom, "k_" + dom)
            kern = doc["kernels"]
            # the kernel, or the template instance of it this workload launches
            ent = kern.get(name) or next((v for k, v in kern.items() if k.startswith(name + "<")), None)
            traffic = ent["hbm_bytes_per_launch"] if ent else None
            traffic_note = f"{os.path.basename(tpath)}, kernel sources {doc['source_sha16']}"
    code, iters, evals, ovf, hpwl, dev_s, fin = runs[-1]
    result = {
        "metric": METRIC,
        "value": round(ms_per_step / 1e3, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("1M-cell plain placement: 1,000,000 movable cells, 1,050,000 nets, 40 fixed "
                         "macros, 1024x1024 bins, td 0.70, stop at overflow 0.10")
            if args.workload == "c1" else
            "20k-cell plain placement: 20,000 cells, 21,000 nets, 128x128 bins, td 0.70",
            "cells": int(info.num_place_insts), "fillers": int(info.num_fillers),
            "nets": int(info.num_gnets), "pins": int(info.num_gpins),
            "bins": f"{info.bin_cnt_x}x{info.bin_cnt_y}", "parallelism": "none (1 GPU)",
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_note,
                     "algorithmic_bytes_per_launch": abytes[dom], "avg_launch_us": round(avg_ms * 1e3, 2),
                     "copy_gbs": copy_gbs},
        "run": {"code": code, "finished": fin, "iterations": iters, "evals": evals,
                "final_overflow": round(ovf, 5), "final_hpwl": hpwl, "device_s": round(dev_s, 4),
                "initial_place_s": round(ip_s, 3), "initial_place_iters": ip.iterations,
                "with_initial_place_s": round(ms_per_step / 1e3 + ip_s, 4),
                "nesterov_setup_s": round(setup_s, 3), "netgen_s": round(gen_s, 2)},
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        n = args.cpu_sample_evals or (8 if args.workload == "c1" else 150)
        try:
            result["cpu_baseline"] = cpu_baseline(args.workload, evals, n, core=0)
        except Exception as e:  # reported, never fatal for the GPU number
            result["cpu_baseline"] = {"value": None, "unit": "s", "cores": 1, "kind": "port",
                                      "sample": f"failed: {e}"[:300]}
    placer.close()
    if args.workload == "c1" and not args.no_service:
        # the batching and serving forms on their workloads, summarised; each
        # runs in a child process, so a failure there is reported in its key
        # and never costs the headline line
        b = _service_leg("c3", args)
        if "error" in b:
            result["batched"] = b
        else:
            _summarise_batch(result, b)
        v = _service_leg("c4", args)
        if "error" in v:
            result["served"] = v
        else:
            _summarise_serve(result, v)
    print(json.dumps(result), flush=True)


def _service_leg(workload, args):
    """`bench.py --workload c3|c4` as a child process; its JSON line, or
    {"error": ...} when it fails."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--workload", workload,
           "--steps-per-poll", str(args.steps_per_poll), "--setup-threads", str(args.setup_threads),
           "--batched-ip", str(args.batched_ip), "--rate", str(args.rate)]
    if args.jobs:
        cmd += ["--jobs", str(args.jobs)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    except subprocess.TimeoutExpired:
        return {"error": f"{workload}: no result within 900 s"}
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"{workload}: exit status {r.returncode}", "stderr_tail": r.stderr[-600:]}
    return json.loads(lines[-1])


def _summarise_batch(result, b):
    result["batched"] = {k: b[k] for k in ("metric", "value", "unit", "speedup_vs_plain", "wall_s", "jobs_finished",
                                            "unfinished", "max_final_overflow_finished",
                                            "hpwl_max_rel_diff_vs_plain")}
    result["batched"]["plain_back_to_back"] = {k: b["plain_back_to_back"][k] for k in ("value", "wall_s")}
    result["batched"]["config"] = b["config"]["workload"]


def _summarise_serve(result, v):
    result["served"] = {k: v[k] for k in ("metric", "value", "unit", "speedup_vs_plain", "latency_s", "jobs_finished",
                                          "unfinished", "max_final_overflow_finished",
                                          "hpwl_max_rel_diff_vs_plain")}
    result["served"]["plain_back_to_back"] = v["plain_back_to_back"]
    result["served"]["config"] = v["config"]["workload"]


if __name__ == "__main__":
    main()
