"""GPU runs at the benchmark workloads' own shapes.

Ensemble (members of 4 species on 20x25x35x72 and on 40x48x48x90, the
electrons with WAVELO, per-member Kp ~ U(1,7) and per-species boundary
scale ~ LogNormal(0,0.5), graph-captured steps): every member equals its
run alone bit for bit, and one member of each grid matches the oracle.
Adjoint (H+ on 20x25x35x72, misfit at 8 satellite points over 60 steps):
the one-step VJP passes a dot-product test against central differences of
the oracle, the checkpoint interval does not change the result, and the
loss matches the oracle's 60-step forward.  Ensemble adjoint (members of 4
species on 40x96x64x90): each member's loss and gradients equal its run
alone.
"""
import numpy as np
import pytest

from conftest import gpu_available
from helpers import max_rel_err, rel_err
from oracle.oracle import Oracle
from ramgpu import inputs, species as sp
from ramgpu.capi import RamBatch, RamGpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

RTOL = 1e-12


def _same(a, b, what):
    assert a.shape == b.shape, what
    diff = np.argwhere(a != b)
    assert diff.size == 0, f"{what}: {len(diff)} cells differ, first at {diff[0]}"


def test_ensemble_at_full_member_shapes():
    shapes = [(20, 25, 35, 72), (20, 25, 35, 72), (40, 48, 48, 90), (40, 48, 48, 90)]
    probs = [inputs.make_problem(*sh, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=60 + m, wave=True)
             for m, sh in enumerate(shapes)]
    nsteps = 3
    rng = np.random.default_rng(61)
    kp = rng.uniform(1.0, 7.0, (nsteps, len(probs)))
    bs = np.exp(0.5 * rng.standard_normal((nsteps, 4 * len(probs))))
    alone_f2, alone_out = [], []
    for m, p in enumerate(probs):
        g = RamGpu.from_problem(p)
        g.set_kp(kp[0, m])
        alone_out.append(g.ram_run_schedule(kp[:, m], bs[:, 4 * m:4 * m + 4].ravel()))
        alone_f2.append(g.get_f2())
        g.close()
    mem = []
    for m, p in enumerate(probs):
        g = RamGpu.from_problem(p)
        g.set_kp(kp[0, m])
        mem.append(g)
    b = RamBatch(mem)
    b.use_graph(True)
    outs = b.ram_run_schedule(kp, bs)
    for m, g in enumerate(mem):
        _same(g.get_f2(), alone_f2[m], f"member {m} {shapes[m]} F2")
        for s, (x, y) in enumerate(zip(outs[m], alone_out[m])):
            for k in x:
                assert x[k] == y[k], (m, s, k, x[k], y[k])
    print(f"ensemble of {len(mem)} members at 4x20x25x35x72 and 4x40x48x48x90 == members alone, "
          f"{nsteps} graph-captured steps", flush=True)
    # one member of each grid against the oracle: per step the Volland-Stern
    # field and WAVELO's Kp of the step's Kp, FGEOS times the step's scales
    for m in (0, 2):
        p = probs[m]
        o = Oracle(p)
        for n in range(nsteps):
            o.set_vt(inputs.volland_stern(p.grid, kp[n, m]))
            o.set_wave_kp(kp[n, m])
            o.set_fgeos(p.fgeos * bs[n, 4 * m:4 * m + 4][:, None, None, None])
            o.ram_run(4)
        e1, e2 = rel_err(mem[m].get_f2(), o.f2), max_rel_err(mem[m].get_f2(), o.f2)
        print(f"member {m} {shapes[m]} vs oracle: norm rel {e1:.3e} max rel {e2:.3e}", flush=True)
        assert e1 <= RTOL and e2 <= RTOL
        for s in range(p.grid.ns):
            scale = abs(o.arr("setrc")[s])
            for k in ("setrc", "lsdr", "lscha", "lsatm", "lswae"):
                assert abs(outs[m][s][k] - o.arr(k)[s]) <= RTOL * scale, (m, s, k)
        o.close()
    b.close()
    for g in mem:
        g.close()


def _c4_problem(nsteps, seed=70):
    p = inputs.make_problem(20, 25, 35, 72, (sp.HYDROGEN,), (1.0,), kp=3.0, seed=seed)
    pts, w = inputs.satellite_points(p, 8, seed=seed)
    rng = np.random.default_rng(seed)
    kp = np.linspace(2.0, 5.0, nsteps)
    bs = np.exp(0.3 * rng.standard_normal(nsteps))
    target = np.log(inputs.satellite_flux(p.f2, pts, w))[None, :] + 0.2 * rng.standard_normal((nsteps, len(pts)))
    return p, pts, w, target, kp, bs


def test_adjoint_60_steps_fixed_field_overlapped_segments():
    # C4 as the bench runs it (fixed field and boundary): the segments are
    # recomputed on a second stream while the adjoint of the next runs, and
    # every checkpoint interval gives the bits of the one-segment run
    nsteps = 60
    p, pts, w, target, _, _ = _c4_problem(nsteps, seed=72)
    g = RamGpu.from_problem(p)
    res = {}
    for ck in (60, 10, 7):
        g.put_f2(p.f2)
        res[ck] = g.adjoint_run(nsteps, pts, w, target, ckpt_every=ck)
        print(f"ckpt {ck}: loss {res[ck][0]:.15e}", flush=True)
    for ck in (10, 7):
        assert res[ck][0] == res[60][0], ck
        assert np.array_equal(res[ck][1], res[60][1]), ck
        assert np.array_equal(res[ck][2], res[60][2]), ck
    g.close()


def test_adjoint_one_step_dot_product_at_20x25x35x72():
    # <ybar, J v> by central differences of the oracle's split step against
    # <ram_run_bwd(ybar), v>, normalised by |ybar| |J v|; the step is
    # piecewise linear, so a direction whose interval holds a limiter kink
    # of some cell is checked again at h/10
    p, *_ = _c4_problem(1)
    g = RamGpu.from_problem(p)
    rng = np.random.default_rng(71)

    def step(f2):
        o = Oracle(p)
        o.set_f2(f2)
        o.ram_run(1)
        out = o.f2.copy()
        o.close()
        return out
    for trial in range(2):
        ybar = np.asfortranarray(rng.standard_normal(p.f2.shape))
        v = np.asfortranarray(p.f2 * rng.standard_normal(p.f2.shape))
        g.put_f2(p.f2)
        xbar, _ = g.ram_run_bwd(ybar)
        assert np.array_equal(g.get_f2(), p.f2), "ram_run_bwd must leave F2 unchanged"
        rhs = float(np.sum(xbar * v))
        for h in (1e-6, 1e-7):
            jv = (step(p.f2 + h * v) - step(p.f2 - h * v)) / (2 * h)
            lhs = float(np.sum(ybar * jv))
            err = abs(lhs - rhs) / (np.linalg.norm(ybar) * np.linalg.norm(jv))
            print(f"dot-product trial {trial} h={h:.0e}: fd {lhs:.12e} adjoint {rhs:.12e} rel {err:.2e}", flush=True)
            if err <= 1e-6:
                break
        assert err <= 1e-6
    g.close()


def test_adjoint_60_steps_checkpoints_and_oracle_loss():
    nsteps = 60
    p, pts, w, target, kp, bs = _c4_problem(nsteps)
    g = RamGpu.from_problem(p)
    g.set_kp(kp[0])
    res = {}
    for ck in (10, 60):
        g.put_f2(p.f2)
        res[ck] = g.adjoint_run(nsteps, pts, w, target, ckpt_every=ck, kp=kp, bscale=bs)
    (l10, f10, b10), (l60, f60, b60) = res[10], res[60]
    print(f"loss ckpt 10 {l10:.15e} ckpt 60 {l60:.15e}", flush=True)
    assert abs(l10 - l60) <= RTOL * abs(l60)
    for a, c, what in ((f10, f60, "dL/dF2"), (b10, b60, "dL/dFGEOS")):
        err = np.max(np.abs(a - c)) / max(np.max(np.abs(c)), 1e-300)
        print(f"{what}: ckpt 10 vs 60 max rel {err:.2e}", flush=True)
        assert err <= RTOL, what
    o = Oracle(p)
    ref = 0.0
    for n in range(nsteps):
        o.set_vt(inputs.volland_stern(p.grid, kp[n]))
        o.set_fgeos(p.fgeos * bs[n])
        o.ram_run(1)
        r = np.log(inputs.satellite_flux(o.f2, pts, w)) - target[n]
        ref += float(np.sum(r * r))
    o.close()
    print(f"60-step loss gpu {l10:.15e} oracle {ref:.15e}", flush=True)
    assert abs(l10 - ref) <= 1e-10 * abs(ref)
    g.close()


def test_ensemble_adjoint_at_40x96x64x90():
    shape = (40, 96, 64, 90)
    probs = [inputs.make_problem(*shape, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=80 + m, wave=True)
             for m in range(2)]
    nsteps, ck = 4, 2
    rng = np.random.default_rng(81)
    kp = rng.uniform(1.0, 8.0, (nsteps, 2))
    pts, w, tgt = [], [], []
    for m, p in enumerate(probs):
        pt, wt = inputs.satellite_points(p, 8, seed=82 + m, species=2)
        pts.append(pt)
        w.append(wt)
        tgt.append(np.log(inputs.satellite_flux(p.f2, pt, wt))[None, :] + 0.2 * rng.standard_normal((nsteps, len(pt))))
    mem = []
    for m, p in enumerate(probs):
        g = RamGpu.from_problem(p)
        g.set_kp(kp[0, m])
        mem.append(g)
    b = RamBatch(mem)
    loss, gf, gb = b.adjoint_run(nsteps, pts, w, tgt, kp, ckpt_every=ck)
    b.close()
    for g in mem:
        g.close()
    for m, p in enumerate(probs):
        g = RamGpu.from_problem(p)
        g.set_kp(kp[0, m])
        l1, f1, b1 = g.adjoint_run(nsteps, pts[m], w[m], tgt[m], ckpt_every=ck, kp=kp[:, m])
        g.close()
        print(f"member {m}: loss ensemble {loss[m]:.15e} alone {l1:.15e}", flush=True)
        assert abs(loss[m] - l1) <= RTOL * abs(l1)
        for a, c, what in ((gf[m], f1, "dL/dF2"), (gb[m], b1, "dL/dFGEOS")):
            err = np.max(np.abs(a - c)) / max(np.max(np.abs(c)), 1e-300)
            print(f"member {m} {what}: max rel {err:.2e}", flush=True)
            assert err <= RTOL, (m, what)
