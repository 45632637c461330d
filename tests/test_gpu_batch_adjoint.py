"""GPU reverse mode of an ensemble: ramgpu_batch_adjoint_run over members of
different grid sizes must give each member's loss and gradients as
ramgpu_adjoint_run gives them for that member alone (the same kernels and
arithmetic; only the atomically accumulated misfit and stale-ghost terms may
add in another order, hence a 1e-12 relative bound instead of bit equality),
with whole segments taped and with every step re-taped; one member's
gradient is also checked against central finite differences of the oracle.
"""
import os

import numpy as np
import pytest

from conftest import gpu_available
from oracle.oracle import Oracle
from ramgpu import inputs, species as sp
from ramgpu.capi import RamBatch, RamGpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

NSTEPS = 5
TOL = 1e-12


def _members():
    return [
        inputs.make_problem(nr=8, nt=9, ne=8, npa=10, species=(sp.HYDROGEN,), kp=3.0, seed=41),
        inputs.make_problem(nr=10, nt=8, ne=9, npa=12, species=(sp.HYDROGEN, sp.OXYGEN), composition=(1.0, 0.3),
                            kp=3.0, seed=42),
    ]


def _setup(seed=5):
    probs = _members()
    rng = np.random.default_rng(seed)
    kp = rng.uniform(1.0, 7.0, (NSTEPS, len(probs)))
    sum_ns = sum(p.grid.ns for p in probs)
    bs = np.exp(0.3 * rng.standard_normal((NSTEPS, sum_ns)))
    pts, w, tgt = [], [], []
    for m, p in enumerate(probs):
        pt, wt = inputs.satellite_points(p, 6, seed=seed + m)
        pts.append(pt)
        w.append(wt)
        tgt.append(np.log(inputs.satellite_flux(p.f2, pt, wt))[None, :] + 0.1 * rng.standard_normal((NSTEPS, len(pt))))
    return probs, kp, bs, pts, w, tgt


def _close(a, b, what):
    scale = max(np.max(np.abs(b)), 1e-300)
    err = np.max(np.abs(a - b)) / scale
    print(f"{what}: max rel {err:.2e}", flush=True)
    assert err <= TOL, what


@pytest.mark.parametrize("retape", [False, True])
def test_ensemble_adjoint_equals_members_alone(retape, monkeypatch):
    if retape:
        monkeypatch.setenv("RAMGPU_ADJOINT_RETAPE", "1")
    probs, kp, bs, pts, w, tgt = _setup()
    ck = 2
    mem = [RamGpu.from_problem(p) for p in probs]
    for g, p, k in zip(mem, probs, kp[0]):
        g.set_kp(k)
    b = RamBatch(mem)
    loss, gf, gb = b.adjoint_run(NSTEPS, pts, w, tgt, kp, bs, ckpt_every=ck)
    end = [g.get_f2() for g in mem]
    b.close()
    off = 0
    for m, p in enumerate(probs):
        ns = p.grid.ns
        g = RamGpu.from_problem(p)
        g.set_kp(kp[0, m])
        l1, f1, b1 = g.adjoint_run(NSTEPS, pts[m], w[m], tgt[m], ckpt_every=ck, kp=kp[:, m],
                                   bscale=bs[:, off:off + ns].ravel())
        off += ns
        print(f"member {m} (retape={retape}): loss ensemble {loss[m]:.15e} alone {l1:.15e}", flush=True)
        assert abs(loss[m] - l1) <= TOL * abs(l1)
        _close(gf[m], f1, f"member {m} dL/dF2")
        _close(gb[m], b1, f"member {m} dL/dFGEOS")
        assert np.array_equal(end[m], g.get_f2()), f"member {m}: state after the run"
        g.close()
    for g in mem:
        g.close()


def test_ensemble_adjoint_member_matches_oracle_fd():
    probs, kp, bs, pts, w, tgt = _setup(seed=9)
    mem = [RamGpu.from_problem(p) for p in probs]
    for g, k in zip(mem, kp[0]):
        g.set_kp(k)
    b = RamBatch(mem)
    loss, gf, gb = b.adjoint_run(NSTEPS, pts, w, tgt, kp, bs, ckpt_every=2)
    b.close()
    p = probs[0]  # single species: its boundary scale is column 0
    H = 1e-6

    def oracle_loss(f0, fg):
        o = Oracle(p)
        o.set_f2(f0)
        tot = 0.0
        for n in range(NSTEPS):
            o.set_vt(inputs.volland_stern(p.grid, kp[n, 0]))
            o.set_fgeos(fg[n])
            o.ram_run(1)
            r = np.log(inputs.satellite_flux(o.f2, pts[0], w[0])) - tgt[0][n]
            tot += float(np.sum(r * r))
        return tot

    fg = [p.fgeos * bs[n, 0] for n in range(NSTEPS)]
    ref = oracle_loss(p.f2, fg)
    print(f"loss ensemble {loss[0]:.15e} oracle {ref:.15e}", flush=True)
    assert abs(loss[0] - ref) <= 1e-10 * abs(ref)
    rng = np.random.default_rng(1)
    noise = 100 * np.finfo(float).eps * abs(ref) / H
    for m in range(3):
        v = p.f2 * rng.standard_normal(p.f2.shape)
        fd = (oracle_loss(p.f2 + H * v, fg) - oracle_loss(p.f2 - H * v, fg)) / (2 * H)
        ad = float(np.sum(gf[0] * v))
        err = abs(fd - ad) / (abs(fd) + 1e-30)
        print(f"f0 direction {m}: fd {fd:.10e} adjoint {ad:.10e} rel {err:.2e}", flush=True)
        assert err <= 1e-6 or abs(fd - ad) <= noise
    n = NSTEPS // 2
    d = p.fgeos * rng.standard_normal(p.fgeos.shape)
    fp = [f + (H * d if k == n else 0) for k, f in enumerate(fg)]
    fm = [f - (H * d if k == n else 0) for k, f in enumerate(fg)]
    fd = (oracle_loss(p.f2, fp) - oracle_loss(p.f2, fm)) / (2 * H)
    ad = float(np.sum(gb[0][n] / bs[n, 0] * d))
    err = abs(fd - ad) / (abs(fd) + 1e-30)
    print(f"boundary step {n}: fd {fd:.10e} adjoint {ad:.10e} rel {err:.2e}", flush=True)
    assert err <= 1e-6 or abs(fd - ad) <= noise
    for g in mem:
        g.close()
