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
steps).  The roofline object is for the dominant kernel, its launches timed
with HIP event pairs on the library's stream in a second pass of the same
steps from the same state (the event pairs add idle time between launches,
so the throughput is timed without them); the cpu_baseline object is the oracle's
C restatement (OpenMP over species, as the reference) on a bounded sample of
the same workload.

More objects in the same line, each on a bounded sample:
  "batched": an ensemble of uneven members (half 20x25x35x72, half
    40x48x48x90, 4 species each, per-member Kp ~ U(1,7) and boundary scale
    ~ LogNormal(0,0.5)) stepped as one launch per kernel under a captured
    HIP graph with autotuned launch grids, against the same members stepped
    one after another (one ram_run call each);
  "adjoint": H+ on 20x25x35x72, 60 steps, virtual-satellite misfit at 8
    points, gradient w.r.t. the initial F2 and every step's FGEOS with
    checkpoints every 10 steps, against the 60-step forward alone;
  "batched_adjoint": the same misfit gradient for an ensemble of 4-species
    40x96x64x90 members (Kp ~ U(1,8) per member), 120 steps with checkpoints
    every 20, in one launch per kernel, against the ensemble's forward;
  "moments": the block after ram_run's species loop on the multi-species
    F2 -- ANISCH of every species (PPERT, PPART and the densities), alone and
    with FLUX -- per launch.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--members M]
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
# algorithmic F2 passes per launch of each sweep kernel: k_r one (DRIFTR),
# k_p one (DRIFTP, rows I>=2), k_plane one plane pass covering DRIFTE,
# DRIFTMU, the loss chain, DRIFTMU, DRIFTE (rows I>=2)
KERNEL_NAMES = {"r": "k_r (radial sweep, DRIFTR)", "p": "k_p (azimuthal sweep, DRIFTP)",
                "plane": "k_plane (energy + pitch-angle sweeps and loss chain, one (K,L) plane per block)"}
ALGO_BYTES_PER_CELL_UPDATE = 128  # 8 directional sweeps x (read + write) of F2
METRIC = "FP64 RAM split-step cell-updates/s + HBM roofline %; batched & adjoint vs plain"
TRAFFIC_FILE = REPO / "profiles" / "traffic_multi_species.txt"
SQ_FILE = REPO / "profiles" / "sq_counters_multi_species.txt"
# MI355X vector issue: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock, one
# wave64 VALU instruction (FP64 included) per 4 cycles per SIMD
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4


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


def measured_valu(kernel: str = "ramgpu::k_plane<5, 8>"):
    """SQ_INSTS_VALU (wave instructions per launch) of a kernel, from the
    committed rocprofv3 counter summary."""
    if not SQ_FILE.exists():
        return None
    cur = None
    for line in SQ_FILE.read_text().splitlines():
        if line and not line.startswith((" ", "#")):
            cur = line.strip()
        elif cur == kernel and line.strip().startswith("SQ_INSTS_VALU "):
            return float(line.split()[1])
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


def bench_batched(members: int, steps: int, seed: int):
    """Uneven ensemble in one launch per kernel (graph-captured step, tuned
    grids) vs the same members stepped one after another."""
    from ramgpu import inputs
    from ramgpu.capi import RamBatch, RamGpu

    nsmall = members - members // 2
    shapes = [(20, 25, 35, 72)] * nsmall + [(40, 48, 48, 90)] * (members // 2)
    protos = {sh: inputs.make_problem(*sh, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=seed + 5)
              for sh in set(shapes)}
    rng = np.random.default_rng(seed + 9)
    kpm = rng.uniform(1.0, 7.0, members)
    scm = np.exp(0.5 * rng.standard_normal(members))
    log(f"bench: ensemble of {members} members ({nsmall} x 20x25x35x72, {members // 2} x 40x48x48x90, "
        f"4 species), {steps} steps")
    mem = []
    for m, sh in enumerate(shapes):
        g = RamGpu.from_problem(protos[sh])
        g.set_kp(kpm[m])
        g.set_boundary_scale(scm[m])
        mem.append(g)
    cells = sum(g.ns * g.nr * g.nt * g.ne * g.npa for g in mem)
    log(f"bench: {members} members on the device ({cells * 8 / 1e9:.1f} GB of F2)")
    b = RamBatch(mem)
    best = b.autotune(reps=2)
    log(f"bench: autotuned grids {b.shapes()}")
    kp = np.tile(kpm, (steps, 1))
    bs = np.repeat(np.tile(scm, (steps, 1)), 4, axis=1)
    b.ram_run_schedule(kp[:1], bs[:1])          # warm: capture and replay once
    b.ram_run_schedule(kp, bs)
    ms_batch = b.last_run_ms()
    log(f"bench: ensemble run {ms_batch:.1f} ms; now the members one after another")
    # the same work, member after member: one ram_run_schedule call per member
    t0 = time.perf_counter()
    dev_ms = 0.0
    for m, g in enumerate(mem):
        g.ram_run_schedule(kp[:, m], bs[:, 4 * m:4 * m + 4].ravel())
        dev_ms += g.last_run_ms()
        if (m + 1) % 64 == 0 and m + 1 < members:
            log(f"bench:   {m + 1} members done")
    wall_seq = (time.perf_counter() - t0) * 1e3
    # one plain step of every member, one after another (what a per-member
    # host loop does each step)
    t0 = time.perf_counter()
    for g in mem:
        g.ram_run(1)
    wall_one = (time.perf_counter() - t0) * 1e3
    b.close()
    for g in mem:
        g.close()
    res = {"members": members, "steps": steps, "cells": cells,
           "value": cells * steps / (ms_batch * 1e-3), "unit": "cell-updates/s",
           "ms_per_step": ms_batch / steps,
           "sequential_device_ms_per_step": dev_ms / steps,
           "sequential_wall_ms_per_step": wall_seq / steps,
           "one_step_member_loop_wall_ms": wall_one,
           "speedup_vs_sequential_device": dev_ms / ms_batch,
           "speedup_vs_member_loop_wall": wall_one / (ms_batch / steps),
           "autotune_best_ms": best, "graph": True}
    log(f"bench: ensemble {ms_batch / steps:.3f} ms/step, sequential device {dev_ms / steps:.3f} ms/step, "
        f"member loop wall {wall_one:.3f} ms/step")
    return res


def bench_adjoint(seed: int, nsteps: int = 60, ckpt: int = 10):
    """Checkpointed adjoint of the 60-step H+ run vs the forward alone."""
    from ramgpu import inputs, species as sp
    from ramgpu.capi import RamGpu

    p = inputs.make_problem(20, 25, 35, 72, (sp.HYDROGEN,), (1.0,), kp=3.0, seed=seed + 11)
    pts, w = inputs.satellite_points(p, 8, seed=seed + 11)
    rng = np.random.default_rng(seed + 12)
    g = RamGpu.from_problem(p)
    g.put_f2(p.f2 * (1.0 + 0.2 * rng.uniform(-1, 1, p.f2.shape)))  # truth run
    target = np.zeros((nsteps, len(pts)))
    for n in range(nsteps):
        g.ram_run(1)
        target[n] = np.log(inputs.satellite_flux(g.get_f2(), pts, w))
    g.put_f2(p.f2)
    g.ram_run(nsteps)                      # warm
    g.put_f2(p.f2)
    g.ram_run(nsteps)
    fwd_ms = g.last_run_ms()
    g.put_f2(p.f2)
    g.adjoint_run(nsteps, pts, w, target, ckpt_every=ckpt)   # warm
    g.put_f2(p.f2)
    loss, gf, gb = g.adjoint_run(nsteps, pts, w, target, ckpt_every=ckpt)
    adj_ms = g.last_run_ms()
    g.close()
    log(f"bench: adjoint {adj_ms:.3f} ms (forward {fwd_ms:.3f} ms, ratio {adj_ms / fwd_ms:.2f}), loss {loss:.6e}")
    return {"config": "H+ 20x25x35x72, 60 steps, 8 satellite points, checkpoint every 10",
            "forward_ms": fwd_ms, "adjoint_ms": adj_ms, "ratio": adj_ms / fwd_ms, "loss": loss,
            "grad_f2_norm": float(np.linalg.norm(gf)), "grad_fgeos_norm": float(np.linalg.norm(gb))}


def bench_batched_adjoint(members: int, nsteps: int, ckpt: int, seed: int):
    """Ensemble adjoint at high resolution: members of 4 species on
    40x96x64x90 with per-member Kp ~ U(1,8), 8 satellite points each,
    gradient w.r.t. every member's initial F2 and per-step FGEOS with
    checkpoints every `ckpt` steps, against the same ensemble's forward."""
    from ramgpu import inputs
    from ramgpu.capi import RamBatch, RamGpu

    shape = (40, 96, 64, 90)
    log(f"bench: ensemble adjoint, {members} members of 4 species on {'x'.join(map(str, shape))}, "
        f"{nsteps} steps, checkpoint every {ckpt}")
    proto = inputs.make_problem(*shape, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=seed + 31)
    rng = np.random.default_rng(seed + 32)
    kp = rng.uniform(1.0, 8.0, (nsteps, members))
    mem = []
    for m in range(members):
        g = RamGpu.from_problem(proto)
        g.set_kp(kp[0, m])
        mem.append(g)
    pts, w, tgt = [], [], []
    for m in range(members):
        pt, wt = inputs.satellite_points(proto, 8, seed=seed + 40 + m, species=2)
        om = np.log(inputs.satellite_flux(proto.f2, pt, wt))
        pts.append(pt)
        w.append(wt)
        tgt.append(om[None, :] + 0.2 * rng.standard_normal((nsteps, len(pt))))
    cells = members * proto.grid.ns * int(np.prod(shape))
    b = RamBatch(mem)
    b.ram_run_schedule(kp[:1])                 # warm
    for g in mem:
        g.put_f2(proto.f2)
    b.ram_run_schedule(kp)
    fwd_ms = b.last_run_ms()
    for g in mem:
        g.put_f2(proto.f2)
    loss, gf, gb = b.adjoint_run(nsteps, pts, w, tgt, kp, ckpt_every=ckpt)
    adj_ms = b.last_run_ms()
    b.close()
    for g in mem:
        g.close()
    log(f"bench: ensemble adjoint {adj_ms:.1f} ms (forward {fwd_ms:.1f} ms, ratio {adj_ms / fwd_ms:.2f})")
    return {"config": f"{members} members x 4 species on {'x'.join(map(str, shape))}, {nsteps} steps, "
                      f"checkpoint every {ckpt}, 8 satellite points per member, Kp ~ U(1,8)",
            "members": members, "cells": cells, "forward_ms": fwd_ms, "adjoint_ms": adj_ms,
            "ratio": adj_ms / fwd_ms, "forward_cell_updates_per_s": cells * nsteps / (fwd_ms * 1e-3),
            "loss": [float(x) for x in loss],
            "grad_f2_norm": [float(np.linalg.norm(x)) for x in gf]}


def bench_moments(dev, cells_kl: int, reps: int = 50):
    """ANISCH of every species (and with FLUX) on the resident F2: the block
    after ram_run's species loop.  Kernel time from HIP events around each
    launch; algorithmic bytes 8 B per cell read (+ 8 B written for FLUX)."""
    dev.moments(flux=False)
    res = {}
    for tag, flux, n, nbytes in (("anisch", False, reps, 8), ("anisch_flux", True, reps // 5, 16)):
        dev.kernel_timing(True)
        for _ in range(n):
            dev.moments(flux=flux)
        ms, nl = dev.kernel_times()["moments"]
        dev.kernel_timing(False)
        avg = ms / nl * 1e-3
        res[tag] = {"avg_launch_us": avg * 1e6, "bytes_per_launch": nbytes * cells_kl,
                    "achieved_GBps": nbytes * cells_kl / avg / 1e9, "launches": nl}
        log(f"bench: moments {tag:12s} {avg * 1e6:8.2f} us  {nbytes * cells_kl / avg / 1e9:7.0f} GB/s")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=720)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy (counter calibration)")
    ap.add_argument("--members", type=int, default=32,
                    help="ensemble members of the batched measurement (half of each grid)")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--ens-adjoint-members", type=int, default=2,
                    help="members of the high-resolution ensemble adjoint (0: skip)")
    ap.add_argument("--ens-adjoint-steps", type=int, default=120)
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
    # timed region: HIP events on the library stream around the K steps
    f0 = dev.get_f2()
    e0 = dev.energy_state()
    t0 = time.perf_counter()
    dev.ram_run_schedule(kp, bs)     # returns after a stream synchronize
    wall = time.perf_counter() - t0
    dev_ms = dev.last_run_ms()
    log(f"bench: device {dev_ms:.3f} ms, wall {wall * 1e3:.3f} ms for {args.steps} steps")
    # the same K steps from the same state again, every launch bracketed by
    # HIP events (their recording adds idle time between launches, so the
    # throughput above is timed without them): per-kernel durations
    dev.put_f2(f0)
    dev.set_energy_state(**e0)
    dev.kernel_timing(True)
    dev.ram_run_schedule(kp, bs)
    ev_ms = dev.last_run_ms()
    ktimes = dev.kernel_times()
    dev.kernel_timing(False)
    log(f"bench: per-kernel pass {ev_ms:.3f} ms with an event pair around every launch")
    for k, (ms, n) in ktimes.items():
        if n:
            log(f"bench:   {k:9s} {n:6d} launches  {ms / n * 1e3:9.2f} us avg  {ms:9.3f} ms total")

    ms_per_step = dev_ms / args.steps
    value = cells * args.steps / (dev_ms * 1e-3)
    # dominant kernel by total device time
    dom = max(("r", "p", "plane"), key=lambda k: ktimes[k][0])
    per_kernel = {}
    for k in ("r", "p", "plane"):
        ms_tot, nlaunch = ktimes[k]
        if not nlaunch:
            continue
        avg_s = ms_tot / nlaunch * 1e-3
        per_launch = BYTES_PER_CELL_PASS * (cells if k == "r" else cells_kl)
        per_kernel[k] = {"avg_launch_us": avg_s * 1e6, "bytes_per_launch": per_launch,
                         "achieved_GBps": per_launch / avg_s / 1e9, "launches": nlaunch}
    d = per_kernel[dom]
    roofline = {"bound": "hbm", "achieved": d["achieved_GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": d["achieved_GBps"] / HBM_PEAK_GBPS, "traffic": measured_traffic(),
                "kernel": KERNEL_NAMES[dom], "bytes_per_launch": d["bytes_per_launch"],
                "avg_launch_us": d["avg_launch_us"], "per_kernel": per_kernel}
    valu = measured_valu() if dom == "plane" else None
    if valu:
        # the plane kernel is bound by FP64 vector issue, not HBM: its VALU
        # wave instructions per launch over the issue rate, against the launch
        roofline["valu_issue_frac"] = valu / VALU_ISSUE_PER_S / (d["avg_launch_us"] * 1e-6)
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
        "ms_per_step_with_kernel_events": ev_ms / args.steps,
        "copy_peak_GBps": copy_peak,
        "roofline": roofline,
    }
    out["moments"] = bench_moments(dev, cells_kl)
    dev.close()
    if not args.no_batched:
        out["batched"] = bench_batched(args.members, args.batched_steps, args.seed)
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint(args.seed)
        if args.ens_adjoint_members > 0:
            out["batched_adjoint"] = bench_batched_adjoint(args.ens_adjoint_members, args.ens_adjoint_steps, 20,
                                                           args.seed)
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(problem)
        log(f"bench: CPU baseline {out['cpu_baseline']['value']:.3e} cell-updates/s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
