#!/usr/bin/env python3
"""Benchmark of the FP64 RAM split step on one MI355X.

Workload (the multi-species run): e-, H+, He+, O+ (composition
1:0.7:0.05:0.25) on the 20x25x35x72 grid, dt = 5 s split steps, Kp ramping
linearly 2 -> 6 over the timed steps, nightside boundary flux rescaled every
step by a seeded LogNormal(0, 0.3) factor.  One step = one ram_run call over
all four species (R,P,E,Mu drifts, charge exchange, atmospheric loss,
Mu,E,P,R drifts, SUMRC bookkeeping, seam copy and magnetopause mask).  F2 and
the whole Kp/boundary schedule are resident on the device before the timed
region; the device recomputes VT and the Courant-number tables every step.

Prints one JSON line.  value = cell-updates/s (NS*NR*NT*NE*NPA cells
advanced by one split step, per second of device time over the timed
steps).  The roofline object is for the dominant kernel, timed with HIP
events on the library's stream; the cpu_baseline object is the oracle's
C restatement (OpenMP over species, as the reference) on a bounded sample of
the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path[:0] = [str(REPO), str(REPO / "ram-scb_amd")]

import numpy as np  # noqa: E402

HBM_PEAK_GBPS = 8000.0       # MI355X HBM3E, AMD specification
BYTES_PER_CELL_PASS = 16     # one FP64 read + one FP64 write of F2 per cell
ALGO_BYTES_PER_CELL_UPDATE = 128  # 8 directional sweeps x (read + write) of F2
METRIC = "FP64 RAM split-step cell-updates/s + HBM roofline %; batched & adjoint vs plain"
TRAFFIC_FILE = REPO / "profiles" / "traffic_multi_species.txt"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def schedule(nsteps: int, seed: int, ns: int):
    rng = np.random.default_rng(seed)
    kp = np.array([2.0 + 4.0 * n / max(nsteps - 1, 1) for n in range(nsteps)])
    bscale = np.repeat(np.exp(0.3 * rng.standard_normal(nsteps)), ns)
    return kp, bscale


def measured_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch="):
            return float(line.split("=", 1)[1])
    return None


def cpu_baseline(problem, target_s: float = 15.0):
    """The oracle (C restatement, OpenMP over species as the reference's
    src/ModRamRun.f90:64) on a bounded sample of the same workload."""
    from oracle.oracle import Oracle
    from ramgpu import inputs

    ns = problem.grid.ns
    cells = ns * problem.grid.nr * problem.grid.nt * problem.grid.ne * problem.grid.npa
    threads = min(ns, os.cpu_count() or 1)
    o = Oracle(problem)
    kp, bs = schedule(720, seed=2, ns=ns)
    t0 = time.perf_counter()
    o.set_vt(inputs.volland_stern(problem.grid, kp[0]))
    o.set_fgeos(problem.fgeos * bs[0])
    o.ram_run(threads)
    one = time.perf_counter() - t0
    n = int(max(1, min(60, target_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    for s in range(1, n + 1):
        o.set_vt(inputs.volland_stern(problem.grid, kp[s]))
        o.set_fgeos(problem.fgeos * bs[s * ns])
        o.ram_run(threads)
    dt = time.perf_counter() - t0
    return {"value": cells * n / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} split steps of the same 4-species 20x25x35x72 Kp-ramp workload "
                      f"({dt:.1f} s, oracle C restatement, OpenMP over species)"}


def calibrate_copy(nbytes: int = 1 << 29, reps: int = 5):
    """A plain device copy (k_copy) of known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run, and the measured
    copy bandwidth."""
    from ramgpu.capi import copy_probe

    gbps = copy_probe(nbytes, reps)
    print(json.dumps({"calibration_copy_bytes": nbytes, "reps": reps + 1, "copy_GBps": gbps}))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=720)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "species and ensemble members shard across GPUs without communication, "
              "but no multi-GPU path is built.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.steps < 1 or args.warmup < 0:
        ap.error("--steps must be >= 1 and --warmup >= 0")

    from ramgpu import inputs
    from ramgpu.capi import RamGpu

    problem = inputs.multi_species(seed=args.seed)
    g = problem.grid
    ns = g.ns
    cells = ns * g.nr * g.nt * g.ne * g.npa
    cells_kl = ns * (g.nr - 1) * g.nt * g.ne * g.npa

    log(f"bench: 4 species {g.nr}x{g.nt}x{g.ne}x{g.npa} ({cells} cells), "
        f"warmup {args.warmup}, timed {args.steps} split steps")
    dev = RamGpu.from_problem(problem)
    dev.set_kp(2.0)
    if args.warmup:
        kpw, bsw = schedule(args.warmup, args.seed + 100, ns)
        dev.ram_run_schedule(kpw, bsw)
    kp, bs = schedule(args.steps, args.seed + 1, ns)

    from ramgpu.capi import copy_probe
    copy_peak = copy_probe(1 << 30, 10)
    log(f"bench: plain-copy bandwidth {copy_peak:.0f} GB/s (read+write, 1 GiB)")
    # timed region: every launch bracketed by HIP events on the library stream
    dev.kernel_timing(True)
    t0 = time.perf_counter()
    dev.ram_run_schedule(kp, bs)     # returns after a stream synchronize
    wall = time.perf_counter() - t0
    dev_ms = dev.last_run_ms()
    ktimes = dev.kernel_times()
    dev.kernel_timing(False)
    log(f"bench: device {dev_ms:.3f} ms, wall {wall * 1e3:.3f} ms for {args.steps} steps")
    for k, (ms, n) in ktimes.items():
        if n:
            log(f"bench:   {k:9s} {n:6d} launches  {ms / n * 1e3:9.2f} us avg  {ms:9.3f} ms total")

    ms_per_step = dev_ms / args.steps
    value = cells * args.steps / (dev_ms * 1e-3)
    # dominant kernel by total device time
    dom = max(("ij", "kl"), key=lambda k: ktimes[k][0])
    ms_tot, nlaunch = ktimes[dom]
    avg_s = ms_tot / nlaunch * 1e-3
    per_launch = BYTES_PER_CELL_PASS * (cells_kl if dom == "kl" else cells)
    achieved = per_launch / avg_s / 1e9
    kname = {"kl": "k_kl (energy + pitch-angle sweeps, loss operators)",
             "ij": "k_ij (radial + azimuthal sweeps)"}[dom]
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS, "traffic": measured_traffic(),
                "kernel": kname, "bytes_per_launch": per_launch,
                "avg_launch_us": avg_s * 1e6}
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain multi-species: e-,H+,He+,O+ (1:0.7:0.05:0.25) on "
                               "NR=20,NT=25,NE=35,NPA=72, dt=5 s, Kp 2->6 ramp, "
                               "boundary scale LogNormal(0,0.3) per step",
                   "species": ns, "grid": [g.nr, g.nt, g.ne, g.npa], "cells": cells,
                   "parallelism": "single GPU"},
        "step_algorithmic_GBps": ALGO_BYTES_PER_CELL_UPDATE * value / 1e9,
        "copy_peak_GBps": copy_peak,
        "roofline": roofline,
    }
    dev.close()
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(problem)
        log(f"bench: CPU baseline {out['cpu_baseline']['value']:.3e} cell-updates/s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
