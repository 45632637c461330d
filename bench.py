#!/usr/bin/env python3
"""Benchmark of the FP64 RAM split step on one MI355X.

Headline workload (the batched ensemble): 512 members in one launch per
kernel, 256 of 4 species on 20x25x35x72 and 256 of 4 species on
40x48x48x90 (e-, H+, He+, O+ at 1:0.7:0.05:0.25; the electrons take the
reference's default WAVELO lifetimes), per-member Kp ~ U(1,7) and boundary
flux scale ~ LogNormal(0,0.5), dt = 5 s split steps under a captured HIP
graph with autotuned launch grids.  One step = one ram_run call of every
member (R,P,E,Mu drifts, WAVELO, charge exchange, atmospheric loss twice,
charge exchange, WAVELO, Mu,E,P,R drifts, SUMRC bookkeeping, seam copy and
magnetopause mask).  F2 (78 GB) and the schedules are resident in HBM before
the timed region.

Prints one JSON line.  value = cell-updates/s (sum over members of
NS*NR*NT*NE*NPA cells advanced by one split step, per second of device time
over the timed steps).  roofline: the dominant kernel of the ensemble step,
its launches timed with HIP event pairs on the ensemble's stream over a
further run of the same steps launched eagerly (the graph cannot be split
per kernel); algorithmic bytes per launch as the reference's algorithm
moves them: 16 B (one FP64 read and write of F2) per cell for each of the
8 directional sweeps a step makes, so 16 B per cell for a radial or
azimuthal launch and 64 B per cell (rows I>=2) for the plane launch, which
performs DRIFTE, DRIFTMU, DRIFTMU, DRIFTE with the loss chain between;
traffic from the committed rocprofv3 counter summary of the same workload.
cpu_baseline: the oracle's C restatement (OpenMP over species, as the
reference) on a bounded sample of the multi-species workload.

More objects in the same line:
  "ensemble_vs_sequential": the same members stepped one after another with
    the fused per-member step, and a sample of members stepped by a host loop
    over the per-operator entry points (what a Fortran host running the
    reference's ram_run over the Level-1 drop-in does);
  "multi_species": the plain 4-species 20x25x35x72 run (Kp 2->6 ramp) on one
    context, with its own per-kernel roofline;
  "multi_species_coulomb": the same run for a third of the steps with
    Coulomb collisions on (COULEN and COULMU twice a step in the plane
    kernel);
  "multi_species_all_processes": a sixth of the steps with every optional
    process on (Coulomb, e- WPADIF, H+ EMIC, He+ FLC): the composed step;
  "adjoint": H+ on 20x25x35x72, 60 steps, virtual-satellite misfit at 8
    points, gradient w.r.t. the initial F2 and every step's FGEOS with
    checkpoints every 10 steps, against the 60-step forward alone;
  "batched_adjoint": the same misfit gradient for an ensemble of 4-species
    40x96x64x90 members (Kp ~ U(1,8) per member), 120 steps with checkpoints
    every 20, in one launch per kernel, against the ensemble's forward;
  "moments": ANISCH of every species (and with FLUX) on the resident F2;
  "computehI": the SCB->RAM bounce integrals from 101-sample field lines
    (dipole lines with 3% noise in B, smoothing on) on the 20x25x35x72 and
    40x96x64x90 grids, per call (host lines in, geometry left on the
    device), beside the oracle's time on the smaller grid.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--members M] [--multi-only [--coulomb]]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path[:0] = [str(REPO), str(REPO / "ram-scb_amd")]

import numpy as np  # noqa: E402

HBM_PEAK_GBPS = 8000.0       # MI355X HBM3E, AMD specification
SWEEP_BYTES = 16             # one FP64 read + one FP64 write of F2 per cell and sweep
# algorithmic sweeps per launch: k_r one (DRIFTR), k_p one (DRIFTP, rows
# I>=2), k_plane four (DRIFTE, DRIFTMU, DRIFTMU, DRIFTE on rows I>=2)
SWEEPS_PER_LAUNCH = {"r": 1, "p": 1, "plane": 4}
KERNEL_NAMES = {"r": "k_r (radial sweep, DRIFTR)", "p": "k_p (azimuthal sweep, DRIFTP)",
                "plane": "k_plane (energy + pitch-angle sweeps and loss chain, one (K,L) plane per block)"}
ALGO_BYTES_PER_CELL_UPDATE = 128  # 8 directional sweeps x (read + write) of F2
METRIC = "FP64 RAM split-step cell-updates/s + HBM roofline %; batched & adjoint vs plain"
TRAFFIC_FILE = REPO / "profiles" / "traffic_c3_ensemble.txt"
TRAFFIC_FILE_MULTI = REPO / "profiles" / "traffic_multi_species.txt"
ORACLE_CFLAGS = "gcc -O2 -fPIC -fopenmp -ffp-contract=off (oracle/Makefile)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def schedule(nsteps: int, seed: int, ns: int):
    rng = np.random.default_rng(seed)
    kp = np.array([2.0 + 4.0 * n / max(nsteps - 1, 1) for n in range(nsteps)])
    bscale = np.repeat(np.exp(0.3 * rng.standard_normal(nsteps)), ns)
    return kp, bscale


def measured_traffic(kernel: str, path: Path = TRAFFIC_FILE):
    """HBM bytes per launch of a kernel kind ("plane", "r", "p") of a
    workload, from its committed rocprofv3 counter summary (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated)."""
    if not path.exists():
        return None
    key = f"traffic_bytes_per_launch[{kernel}]="
    for line in path.read_text().splitlines():
        if line.startswith(key):
            return float(line.split("=", 1)[1])
    return None


def host_description():
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "lscpu_model": model or platform.processor() or None,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "compiler": ORACLE_CFLAGS}


def cpu_baseline(problem, target_s: float = 15.0):
    """The oracle (C restatement, OpenMP over species as the reference's
    src/ModRamRun.f90:64) on a bounded sample of the multi-species workload."""
    from oracle.oracle import Oracle
    from ramgpu import inputs

    ns = problem.grid.ns
    cells = ns * problem.grid.nr * problem.grid.nt * problem.grid.ne * problem.grid.npa
    threads = min(ns, os.cpu_count() or 1)
    o = Oracle(problem)
    kp, bs = schedule(720, seed=2, ns=ns)

    def step(n):
        o.set_vt(inputs.volland_stern(problem.grid, kp[n]))
        o.set_wave_kp(kp[n])
        o.set_fgeos(problem.fgeos * bs[n * ns])
        o.ram_run(threads)

    t0 = time.perf_counter()
    step(0)
    one = time.perf_counter() - t0
    n = int(max(1, min(60, target_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    for s in range(1, n + 1):
        step(s)
    dt = time.perf_counter() - t0
    res = {"value": cells * n / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
           "sample": f"{n} split steps of the 4-species 20x25x35x72 Kp-ramp workload with electron WAVELO "
                     f"({dt:.1f} s; oracle C restatement, OpenMP over species as the reference)"}
    res.update(host_description())
    return res


def calibrate_copy(nbytes: int = 1 << 29, reps: int = 5):
    """A plain device copy (k_copy) of known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run, and the measured
    copy bandwidth."""
    from ramgpu.capi import copy_probe

    gbps = copy_probe(nbytes, reps)
    print(json.dumps({"calibration_copy_bytes": nbytes, "reps": reps + 1, "copy_GBps": gbps}))


def per_kernel_roofline(ktimes: dict, cells: int, cells_kl: int, traffic_path: Path = TRAFFIC_FILE):
    """Per-kernel algorithmic GB/s from summed event times (and the measured
    HBM bytes per launch of the workload's counter summary), and the roofline
    object of the dominant kernel (by total device time)."""
    per = {}
    for k in ("r", "p", "plane"):
        ms_tot, n = ktimes[k]
        if not n:
            continue
        avg = ms_tot / n * 1e-3
        nb = SWEEP_BYTES * SWEEPS_PER_LAUNCH[k] * (cells if k == "r" else cells_kl)
        per[k] = {"avg_launch_us": avg * 1e6, "bytes_per_launch": nb, "achieved_GBps": nb / avg / 1e9,
                  "launches": n, "frac": nb / avg / 1e9 / HBM_PEAK_GBPS,
                  "traffic": measured_traffic(k, traffic_path)}
    dom = max(per, key=lambda k: ktimes[k][0])
    d = per[dom]
    roof = {"bound": "hbm", "achieved": d["achieved_GBps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": d["achieved_GBps"] / HBM_PEAK_GBPS, "traffic": d["traffic"],
            "kernel": KERNEL_NAMES[dom], "bytes_per_launch": d["bytes_per_launch"],
            "avg_launch_us": d["avg_launch_us"], "per_kernel": per}
    return roof


def make_members(members: int, seed: int, wave: bool = True):
    """The C3 ensemble: half 20x25x35x72, half 40x48x48x90, 4 species each,
    per-member Kp ~ U(1,7) and boundary scale ~ LogNormal(0,0.5)."""
    from ramgpu import inputs
    from ramgpu.capi import RamGpu

    nsmall = members - members // 2
    shapes = [(20, 25, 35, 72)] * nsmall + [(40, 48, 48, 90)] * (members // 2)
    protos = {sh: inputs.make_problem(*sh, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=seed + 5, wave=wave)
              for sh in set(shapes)}
    rng = np.random.default_rng(seed + 9)
    kpm = rng.uniform(1.0, 7.0, members)
    scm = np.exp(0.5 * rng.standard_normal(members))
    mem = []
    for m, sh in enumerate(shapes):
        g = RamGpu.from_problem(protos[sh])
        g.set_kp(kpm[m])
        if wave:
            g.set_wave_params(kpm[m])
        g.set_boundary_scale(scm[m])
        mem.append(g)
        if (m + 1) % 128 == 0:
            log(f"bench:   {m + 1} members on the device")
    return mem, shapes, protos, kpm, scm


def per_operator_step(g, ns: int, wave: bool):
    """One ram_run step of member g through the per-operator entry points,
    in the order of src/ModRamRun.f90:64-186 (what a Fortran host calling
    the Level-1 drop-in does); species 1 (e-) takes WAVELO, 2-4 charge
    exchange."""
    for s in range(1, ns + 1):
        g.cepara(s)
        g.driftpara(s)
        for op in ("driftr", "driftp", "drifte", "driftmu"):
            getattr(g, op)(s)
        g.sumrc(s)
        if wave and s == 1:
            g.wavelo(s)
            g.sumrc(s)
        if s > 1:
            g.charexchange(s)
            g.sumrc(s)
        g.atmol(s)
        g.sumrc(s)
        g.atmol(s)
        g.sumrc(s)
        if s > 1:
            g.charexchange(s)
            g.sumrc(s)
        if wave and s == 1:
            g.wavelo(s)
            g.sumrc(s)
        for op in ("driftmu", "drifte", "driftp", "driftr"):
            getattr(g, op)(s)
        g.sumrc(s)


def bench_c3(members: int, steps: int, warmup: int, seed: int, seq_steps: int = 2, op_sample: int = 8):
    """The headline: the C3 ensemble in one launch per kernel (graph-captured
    step, autotuned grids), then its per-kernel event pass, then the same
    members one after another (fused step) and a sample by per-operator
    calls."""
    from ramgpu.capi import RamBatch

    log(f"bench: C3 ensemble of {members} members ({members - members // 2} x 4x20x25x35x72, "
        f"{members // 2} x 4x40x48x48x90), warmup {warmup}, timed {steps} steps")
    t0 = time.perf_counter()
    mem, shapes, protos, kpm, scm = make_members(members, seed)
    cells = sum(g.ns * g.nr * g.nt * g.ne * g.npa for g in mem)
    cells_kl = sum(g.ns * (g.nr - 1) * g.nt * g.ne * g.npa for g in mem)
    log(f"bench: members set up in {time.perf_counter() - t0:.1f} s ({cells * 8 / 1e9:.1f} GB of F2)")
    b = RamBatch(mem)
    best = b.autotune(reps=2)
    log(f"bench: autotuned grids {b.shapes()}, best launches (ms) R {best['r']:.2f} P {best['p']:.2f} "
        f"plane {best['plane']:.2f}")
    nt = warmup + steps
    kp = np.tile(kpm, (nt, 1))
    bs = np.repeat(np.tile(scm, (nt, 1)), 4, axis=1)
    if warmup:
        b.ram_run_schedule(kp[:warmup], bs[:warmup])   # also captures the graph
    b.ram_run_schedule(kp[warmup:], bs[warmup:])
    ms = b.last_run_ms()
    log(f"bench: ensemble {ms / steps:.3f} ms per step ({cells * steps / (ms * 1e-3):.3e} cell-updates/s)")
    # per-kernel pass: the same kind of steps launched eagerly with events
    ne = max(1, min(2, steps))
    b.kernel_timing(True)
    b.ram_run_schedule(kp[:ne], bs[:ne])
    ktimes = b.kernel_times()
    b.kernel_timing(False)
    for k, (t, n) in ktimes.items():
        if n:
            log(f"bench:   {k:9s} {n:5d} launches {t / n:10.3f} ms avg")
    roof = per_kernel_roofline(ktimes, cells, cells_kl)
    b.close()
    # the same members one after another: the fused per-member step
    t0 = time.perf_counter()
    dev_ms = 0.0
    for m, g in enumerate(mem):
        g.ram_run_schedule(kp[:seq_steps, m], bs[:seq_steps, 4 * m:4 * m + 4].ravel())
        dev_ms += g.last_run_ms()
    wall = (time.perf_counter() - t0) * 1e3
    seq_dev = dev_ms / seq_steps
    seq_wall = wall / seq_steps
    # a sample of members of each grid through the per-operator entry
    # points, one step each (wall clock: the host calls and synchronisations
    # are part of that way of stepping)
    nsmall = members - members // 2
    per_shape = {}
    for lo, cnt, n_all in ((0, min(op_sample // 2, nsmall), nsmall),
                           (nsmall, min(op_sample // 2, members // 2), members // 2)):
        if cnt == 0:
            continue
        t0 = time.perf_counter()
        for m in range(lo, lo + cnt):
            per_operator_step(mem[m], 4, True)
        per_shape[shapes[lo]] = ((time.perf_counter() - t0) * 1e3 / cnt, cnt, n_all)
    op_step = sum(t * n_all for t, _, n_all in per_shape.values())
    for g in mem:
        g.close()
    ens = {"ensemble_ms_per_step": ms / steps,
           "sequential_fused_device_ms_per_step": seq_dev, "sequential_fused_wall_ms_per_step": seq_wall,
           "speedup_vs_sequential_fused_device": seq_dev / (ms / steps),
           "speedup_vs_sequential_fused_wall": seq_wall / (ms / steps),
           "per_operator_loop_ms_per_step": op_step,
           "per_operator_sample": "; ".join(f"{cnt} members of {'x'.join(map(str, sh))}: {t:.2f} ms per member-step"
                                            for sh, (t, cnt, _) in per_shape.items())
                                  + " (wall clock of the per-operator entry-point calls, one step each, "
                                    "scaled to all members)",
           "speedup_vs_per_operator_loop": op_step / (ms / steps),
           "autotune_best_ms": best, "graph": True}
    log(f"bench: sequential fused {seq_dev:.1f} ms/step device, {seq_wall:.1f} ms wall; per-operator loop "
        f"{op_step:.0f} ms/step (extrapolated)")
    return {"cells": cells, "cells_kl": cells_kl, "ms": ms, "steps": steps, "roofline": roof, "ens": ens,
            "ktimes": ktimes}


def bench_multi_species(steps: int, warmup: int, seed: int, coulomb: bool = False, pitch: bool = False):
    """The plain multi-species run on one context: Kp ramp 2->6, boundary
    scale LogNormal(0,0.3) per step; with coulomb, Coulomb collisions on
    (synthetic plasmasphere of Kp=2, TimeRamElapsed from 0); with pitch also
    WPADIF for e- in place of WAVELO, EMIC WPADIF for H+ and FLC scattering
    for He+ (synthetic coefficients), which takes the composed step."""
    from ramgpu import inputs
    from ramgpu.capi import RamGpu, copy_probe

    problem = inputs.multi_species(seed=seed)
    if coulomb or pitch:
        problem.coulomb = inputs.CoulombInputs(inputs.plasmasphere_density(problem.grid, 2.0), 0.0)
    g = problem.grid
    ns = g.ns
    cells = ns * g.nr * g.nt * g.ne * g.npa
    cells_kl = ns * (g.nr - 1) * g.nt * g.ne * g.npa
    dev = RamGpu.from_problem(problem)
    dev.set_kp(2.0)
    if pitch:
        gr, geo = problem.grid, problem.geom
        dev.set_wpadif(1, inputs.pitch_diffusion_coef(gr, geo, seed=seed + 11),
                       inputs.pitch_diffusion_coef(gr, geo, scale=3e-6, seed=seed + 12), at_wpi=True)
        dev.set_wpadif(2, inputs.pitch_diffusion_coef(gr, geo, scale=2e-6, seed=seed + 13), None,
                       at_wpi=False, at_emic=True)
        dev.set_flc(3, inputs.pitch_diffusion_coef(gr, geo, scale=1e-6, seed=seed + 14), 0.0)
    if warmup:
        kpw, bsw = schedule(warmup, seed + 100, ns)
        dev.ram_run_schedule(kpw, bsw)
    kp, bs = schedule(steps, seed + 1, ns)
    f0 = dev.get_f2()
    e0 = dev.energy_state()
    dev.ram_run_schedule(kp, bs)
    dev_ms = dev.last_run_ms()
    dev.put_f2(f0)
    dev.set_energy_state(**e0)
    dev.kernel_timing(True)
    dev.ram_run_schedule(kp, bs)
    ev_ms = dev.last_run_ms()
    ktimes = dev.kernel_times()
    dev.kernel_timing(False)
    if pitch:  # one launch per operator: device time per step by kernel kind instead of a roofline
        value = cells * steps / (dev_ms * 1e-3)
        per = {k: {"ms_per_step": v[0] / steps, "launches_per_step": v[1] / steps} for k, v in ktimes.items() if v[1]}
        log(f"bench: all processes {dev_ms / steps * 1e3:.1f} us/step, {value:.3e} cell-updates/s")
        res = {"workload": "e-(WPADIF),H+(EMIC),He+(FLC),O+ (1:0.7:0.05:0.25) on 20x25x35x72 with Coulomb "
                           "collisions, dt=5 s, Kp 2->6 ramp, boundary scale LogNormal(0,0.3) per step; the "
                           "composed step (one launch per operator)",
               "cells": cells, "steps": steps, "value": value, "unit": "cell-updates/s",
               "ms_per_step": dev_ms / steps, "ms_per_step_with_kernel_events": ev_ms / steps,
               "kernels": per}
        dev.close()
        return res, problem
    roof = per_kernel_roofline(ktimes, cells, cells_kl, TRAFFIC_FILE_MULTI)
    roof["traffic"] = None  # the committed counters are the headline's
    value = cells * steps / (dev_ms * 1e-3)
    log(f"bench: multi-species {dev_ms / steps * 1e3:.1f} us/step, {value:.3e} cell-updates/s; plane "
        f"{roof['per_kernel']['plane']['avg_launch_us']:.1f} us")
    res = {"workload": "e-(WAVELO),H+,He+,O+ (1:0.7:0.05:0.25) on 20x25x35x72, dt=5 s, Kp 2->6 ramp, "
                       "boundary scale LogNormal(0,0.3) per step" + (", Coulomb collisions" if coulomb else ""),
           "cells": cells, "steps": steps, "value": value, "unit": "cell-updates/s",
           "ms_per_step": dev_ms / steps, "ms_per_step_with_kernel_events": ev_ms / steps,
           "step_algorithmic_GBps": ALGO_BYTES_PER_CELL_UPDATE * value / 1e9,
           "step_frac": ALGO_BYTES_PER_CELL_UPDATE * value / 1e9 / HBM_PEAK_GBPS, "roofline": roof,
           "copy_peak_GBps": copy_probe(1 << 30, 10)}
    if not (coulomb or pitch):
        res["moments"] = bench_moments(dev, cells_kl)
    dev.close()
    return res, problem


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
    40x96x64x90 with per-member Kp ~ U(1,8) (electrons with WAVELO), 8
    satellite points each, gradient w.r.t. every member's initial F2 and
    per-step FGEOS with checkpoints every `ckpt` steps, against the same
    ensemble's forward."""
    from ramgpu import inputs
    from ramgpu.capi import RamBatch, RamGpu

    shape = (40, 96, 64, 90)
    log(f"bench: ensemble adjoint, {members} members of 4 species on {'x'.join(map(str, shape))}, "
        f"{nsteps} steps, checkpoint every {ckpt}")
    proto = inputs.make_problem(*shape, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=seed + 31, wave=True)
    rng = np.random.default_rng(seed + 32)
    kp = rng.uniform(1.0, 8.0, (nsteps, members))
    mem = []
    for m in range(members):
        g = RamGpu.from_problem(proto)
        g.set_kp(kp[0, m])
        g.set_wave_params(kp[0, m])
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
            "forward_step_frac": ALGO_BYTES_PER_CELL_UPDATE * cells * nsteps / (fwd_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
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


def bench_computehI(seed: int, reps: int = 20):
    """ramgpu_computehI per call, wall clock around the call (it uploads the
    lines, runs four kernels and synchronises), on two grids."""
    import numpy as np

    from ramgpu import fieldlines, inputs
    from ramgpu.capi import RamGpu
    res = {}
    for tag, dims in (("20x25x35x72", (20, 25, 35, 72)), ("40x96x64x90", (40, 96, 64, 90))):
        p = inputs.make_problem(*dims, (inputs.SPECIES4[1],), (1.0,), kp=4.0, seed=seed)
        g = p.grid
        lines = fieldlines.dipole_lines(g.lz, g.nt)
        lines.b = lines.b * np.random.default_rng(seed).uniform(0.97, 1.03, lines.b.shape)
        dev = RamGpu.from_problem(p)
        dev.computehI(lines, g.pa, g.pabn, g.lz)
        t0 = time.perf_counter()
        for _ in range(reps):
            dev.computehI(lines, g.pa, g.pabn, g.lz)
        ms = (time.perf_counter() - t0) / reps * 1e3
        dev.close()
        res[tag] = {"ms_per_call": ms, "lines": g.nr * g.nt, "nthe": lines.b.shape[0], "npa": g.npa}
        log(f"bench: computehI {tag} {ms:.3f} ms per call")
    return res


def cpu_computehI(seed: int):
    """The oracle's computehI on the 20x25x35x72 lines of bench_computehI
    (numpy, one thread): its CPU baseline."""
    import numpy as np

    from oracle import bounce as ob
    from ramgpu import fieldlines, inputs
    p = inputs.make_problem(20, 25, 35, 72, (inputs.SPECIES4[1],), (1.0,), kp=4.0, seed=seed)
    g = p.grid
    lines = fieldlines.dipole_lines(g.lz, g.nt)
    lines.b = lines.b * np.random.default_rng(seed).uniform(0.97, 1.03, lines.b.shape)
    prev = {k: getattr(p.geom, k) for k in ("fnhs", "fnis", "bounhs", "bounis", "bnes")}
    t0 = time.perf_counter()
    ob.computehI(lines.x, lines.y, lines.z, lines.b, lines.chi, lines.ieq - 1, g.mu, g.pa, g.pabn, g.lz,
                 np.zeros((g.nr, g.nt), int), np.zeros(g.nt, int), lines.bnormal, 0.0, True, prev)
    return {"ms_per_call": (time.perf_counter() - t0) * 1e3, "kind": "port", "cores": 1,
            "sample": "one call on the 20x25x35x72 lines (numpy restatement)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed ensemble steps of the headline")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--members", type=int, default=512, help="ensemble members (half of each grid)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy (counter calibration)")
    ap.add_argument("--multi-steps", type=int, default=720, help="steps of the multi-species run")
    ap.add_argument("--no-multi", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--ens-adjoint-members", type=int, default=8,
                    help="members of the high-resolution ensemble adjoint (0: skip)")
    ap.add_argument("--ens-adjoint-steps", type=int, default=120)
    ap.add_argument("--ens-adjoint-only", action="store_true",
                    help="only the ensemble adjoint leg (--ens-adjoint-members, --ens-adjoint-steps), its JSON")
    ap.add_argument("--multi-only", action="store_true",
                    help="run only the multi-species leg and print its object (profiling)")
    ap.add_argument("--coulomb", action="store_true", help="with --multi-only: Coulomb collisions on")
    ap.add_argument("--all-processes", action="store_true",
                    help="with --multi-only: every optional process on (the composed step)")
    ap.add_argument("--no-coulomb", action="store_true", help="skip the multi-species leg with Coulomb collisions")
    ap.add_argument("--hi-only", action="store_true", help="only the computehI leg")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "species and ensemble members shard across GPUs without communication, "
              "but no multi-GPU path is built.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.steps < 1 or args.warmup < 0 or args.members < 2:
        ap.error("--steps must be >= 1, --warmup >= 0, --members >= 2")
    if args.ens_adjoint_only:
        r = bench_batched_adjoint(max(args.ens_adjoint_members, 1), args.ens_adjoint_steps, 20, args.seed)
        print(json.dumps({k: v for k, v in r.items() if k != "loss"}), flush=True)
        return
    if args.hi_only:
        print(json.dumps(bench_computehI(args.seed)), flush=True)
        return
    if args.multi_only:
        print(json.dumps(bench_multi_species(args.multi_steps, 20, args.seed, args.coulomb,
                                             args.all_processes)[0]), flush=True)
        return

    c3 = bench_c3(args.members, args.steps, args.warmup, args.seed)
    value = c3["cells"] * args.steps / (c3["ms"] * 1e-3)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": c3["ms"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"C3 batched ensemble: {args.members} members in one launch per kernel, "
                               f"{args.members - args.members // 2} of 4 species on NR=20,NT=25,NE=35,NPA=72 and "
                               f"{args.members // 2} of 4 species on 40x48x48x90 (e-,H+,He+,O+ 1:0.7:0.05:0.25, "
                               "electron WAVELO), per-member Kp ~ U(1,7) and boundary scale ~ LogNormal(0,0.5), "
                               "dt=5 s, graph-captured step, autotuned grids",
                   "members": args.members, "species": 4, "cells": c3["cells"],
                   "grids": [[20, 25, 35, 72], [40, 48, 48, 90]], "parallelism": "single GPU"},
        "step_algorithmic_GBps": ALGO_BYTES_PER_CELL_UPDATE * value / 1e9,
        "step_frac": ALGO_BYTES_PER_CELL_UPDATE * value / 1e9 / HBM_PEAK_GBPS,
        "roofline": c3["roofline"],
        "ensemble_vs_sequential": c3["ens"],
    }
    problem = None

    def leg(name, fn):
        # a leg that fails (e.g. a device allocation that does not fit) is
        # reported as such; the headline line is still printed
        try:
            return fn()
        except Exception as e:  # noqa: BLE001
            log(f"bench: {name} failed: {e}")
            return {"error": f"{type(e).__name__}: {e}"}

    if not args.no_multi:
        r = leg("multi_species", lambda: bench_multi_species(args.multi_steps, 20, args.seed))
        if isinstance(r, tuple):
            out["multi_species"], problem = r
        else:
            out["multi_species"] = r
    if not args.no_coulomb:
        r = leg("multi_species_coulomb", lambda: bench_multi_species(args.multi_steps // 3, 5, args.seed, True))
        out["multi_species_coulomb"] = r[0] if isinstance(r, tuple) else r
        r = leg("multi_species_all_processes",
                lambda: bench_multi_species(args.multi_steps // 6, 5, args.seed, False, True))
        out["multi_species_all_processes"] = r[0] if isinstance(r, tuple) else r
    if not args.no_adjoint:
        out["adjoint"] = leg("adjoint", lambda: bench_adjoint(args.seed))
        if args.ens_adjoint_members > 0:
            out["batched_adjoint"] = leg("batched_adjoint", lambda: bench_batched_adjoint(
                args.ens_adjoint_members, args.ens_adjoint_steps, 20, args.seed))
    out["computehI"] = leg("computehI", lambda: bench_computehI(args.seed))
    if not args.no_cpu_baseline:
        from ramgpu import inputs
        log("bench: timing the CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(problem or inputs.multi_species(seed=args.seed))
        if isinstance(out.get("computehI"), dict) and "error" not in out["computehI"]:
            out["computehI"]["cpu_oracle_20x25x35x72"] = cpu_computehI(args.seed)
        log(f"bench: CPU baseline {out['cpu_baseline']['value']:.3e} cell-updates/s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
