"""GPU reverse mode: the backward of every operator, of one ram_run step and
of the checkpointed multi-step misfit gradient, checked against FP64 central
finite differences of the oracle (the CPU restatement of the reference).

Each operator is piecewise linear in (F2, FGEOS), so along a direction v the
central difference (O(x+hv) - O(x-hv))/2h equals J v exactly unless a limiter
kink or a clamp lies within +-h; with relative perturbations h = 1e-6 that is
rare, and the check is the dot-product identity <ybar, J v> = <J^T ybar, v>
to a tolerance of 1e-7 relative (rounding of the difference quotient is
~1e-10).  The multi-step misfit gradient (nonlinear through log) is checked
to 1e-6 relative per direction, or to the rounding floor of the central
difference where the directional derivative is that small, along 64 random
directions of the initial F2 and the FGEOS of three single steps.  A limiter
kink of some cell within +-h of the base point biases the difference quotient
(the map is only piecewise linear); such a direction is checked again at h/10,
where the kink is ten times less likely to lie inside the interval.
"""
import numpy as np
import pytest

from conftest import gpu_available
from helpers import random_problem
from oracle.oracle import Oracle
from ramgpu import inputs, species as sp
from ramgpu.capi import RamGpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

TOL = 1e-7
H = 1e-6


def _oracle_op(p, f2, op, s, fgeos=None):
    o = Oracle(p)
    o.set_f2(f2)
    if fgeos is not None:
        o.set_fgeos(fgeos)
    o.call("cepara", s)
    o.call("driftpara", s)
    o.call(op, s)
    return o.f2


def _state(p, seed):
    """A realistic operator input: the problem's F2 after one oracle step."""
    o = Oracle(p)
    o.ram_run(2)
    return o.f2


def _check(what, lhs, rhs, scale):
    err = abs(lhs - rhs) / scale
    print(f"{what}: <ybar,Jv> {lhs:.12e}  <J^T ybar,v> {rhs:.12e}  rel {err:.2e}", flush=True)
    assert err <= TOL, what


@pytest.mark.parametrize("op", ["driftr", "driftp", "drifte", "driftmu", "charexchange", "atmol"])
@pytest.mark.parametrize("seed", [0, 1])
def test_operator_vjp_matches_oracle_fd(op, seed):
    p = random_problem(seed=30 + seed, species=(sp.HYDROGEN, sp.OXYGEN), spiky=(seed == 1))
    x = _state(p, seed)
    rng = np.random.default_rng(seed)
    for s in (1, 2):
        v = np.asfortranarray(x * rng.standard_normal(x.shape))
        ybar = np.asfortranarray(rng.standard_normal(x.shape))
        yp = _oracle_op(p, x + H * v, op, s)
        ym = _oracle_op(p, x - H * v, op, s)
        jv = (yp - ym) / (2 * H)
        g = RamGpu.from_problem(p)
        g.put_f2(x)
        g.cepara(s)
        g.driftpara(s)
        res = g.op_bwd(op, s, ybar)
        xbar = res[0] if op == "driftr" else res
        lhs, rhs = float(np.sum(ybar * jv)), float(np.sum(xbar * v))
        _check(f"{op} S={s} seed={seed} (F2)", lhs, rhs, np.linalg.norm(ybar) * np.linalg.norm(jv) + 1e-300)
        if op == "driftr":  # boundary direction
            fg = p.fgeos
            d = np.asfortranarray(fg * rng.standard_normal(fg.shape))
            yp = _oracle_op(p, x, op, s, fgeos=fg + H * d)
            ym = _oracle_op(p, x, op, s, fgeos=fg - H * d)
            jd = (yp - ym) / (2 * H)
            lhs, rhs = float(np.sum(ybar * jd)), float(np.sum(res[1] * d))
            if np.linalg.norm(jd) > 0:
                _check(f"driftr S={s} seed={seed} (FGEOS)", lhs, rhs, np.linalg.norm(ybar) * np.linalg.norm(jd))
        g.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_ram_run_step_vjp_matches_oracle_fd(seed):
    _step_vjp(random_problem(seed=40 + seed, species=(sp.HYDROGEN, sp.ELECTRON), spiky=(seed == 1)), seed)


def test_ram_run_step_vjp_with_wavelo():
    # the electrons' WAVELO factors in the loss chain's adjoint
    p = random_problem(seed=48, species=(sp.HYDROGEN, sp.ELECTRON))
    p.wave = inputs.wave_inputs(p.grid, p.kp)
    _step_vjp(p, 0)


def test_ram_run_step_vjp_on_64x90_planes():
    # the ensemble-adjoint members' planes: the plane adjoint kernels hold the
    # forward's plane layout and the adjoint plane in LDS, their largest use
    _step_vjp(random_problem(nr=4, nt=4, ne=64, npa=90, seed=47, species=(sp.HYDROGEN,)), 0)


def _step_vjp(p, seed):
    x = _state(p, seed)
    rng = np.random.default_rng(10 + seed)
    v = np.asfortranarray(x * rng.standard_normal(x.shape))
    d = np.asfortranarray(p.fgeos * rng.standard_normal(p.fgeos.shape))
    ybar = np.asfortranarray(rng.standard_normal(x.shape))

    def step(f2, fgeos):
        o = Oracle(p)
        o.set_f2(f2)
        o.set_fgeos(fgeos)
        o.ram_run(2)
        return o.f2
    g = RamGpu.from_problem(p)
    g.put_f2(x)
    xbar, fbar = g.ram_run_bwd(ybar)
    assert np.array_equal(g.get_f2(), x), "ram_run_bwd must leave F2 unchanged"
    g.close()
    # the step is piecewise linear: a central difference whose interval
    # holds a limiter kink of some cell is off by that cell's jump, so a
    # direction that misses at h is checked again at h/10
    for what, dx, df, bar in (("F2", v, 0 * d, xbar), ("FGEOS", 0 * v, d, fbar)):
        dirn = v if what == "F2" else d
        for h in (H, H / 10):
            jv = (step(x + h * dx, p.fgeos + h * df) - step(x - h * dx, p.fgeos - h * df)) / (2 * h)
            lhs, rhs = float(np.sum(ybar * jv)), float(np.sum(bar * dirn))
            err = abs(lhs - rhs) / (np.linalg.norm(ybar) * np.linalg.norm(jv))
            print(f"ram_run step seed={seed} ({what}) h={h:.0e}: rel {err:.2e}", flush=True)
            if err <= TOL:
                break
        assert err <= TOL, (what, err)


def fd_noise(loss, h=H):
    """Rounding floor of a central difference of the loss with step h: each
    evaluation carries ~eps*|L| per step of accumulated rounding; 100x margin."""
    return 100 * np.finfo(float).eps * abs(loss) / h


def _misfit_setup(nsteps, seed=7):
    p = inputs.make_problem(nr=8, nt=9, ne=8, npa=10, species=(sp.HYDROGEN,), kp=3.0, seed=seed)
    pts, w = inputs.satellite_points(p, 8, seed=seed)
    rng = np.random.default_rng(seed)
    # truth: a run from f0 perturbed by 20%
    truth = Oracle(p)
    truth.set_f2(p.f2 * (1.0 + 0.2 * rng.uniform(-1, 1, p.f2.shape)))
    kp = np.linspace(2.0, 5.0, nsteps)
    bs = np.exp(0.3 * rng.standard_normal(nsteps))
    target = np.zeros((nsteps, len(pts)))
    for n in range(nsteps):
        truth.set_vt(inputs.volland_stern(p.grid, kp[n]))
        truth.set_fgeos(p.fgeos * bs[n])
        truth.ram_run(1)
        target[n] = np.log(inputs.satellite_flux(truth.f2, pts, w))
    return p, pts, w, target, kp, bs


def _oracle_loss(p, f0, fgeos_steps, pts, w, target, kp):
    o = Oracle(p)
    o.set_f2(f0)
    tot = 0.0
    for n in range(len(kp)):
        o.set_vt(inputs.volland_stern(p.grid, kp[n]))
        o.set_fgeos(fgeos_steps[n])
        o.ram_run(1)
        r = np.log(inputs.satellite_flux(o.f2, pts, w)) - target[n]
        tot += float(np.sum(r * r))
    return tot


def test_adjoint_run_gradient_matches_oracle_fd_and_checkpoints():
    nsteps = 6
    p, pts, w, target, kp, bs = _misfit_setup(nsteps)
    g = RamGpu.from_problem(p)
    g.set_kp(kp[0])
    results = {}
    for ck in (1, 2, 6):
        g.put_f2(p.f2)
        results[ck] = g.adjoint_run(nsteps, pts, w, target, ckpt_every=ck, kp=kp, bscale=bs)
    loss, gf, gb = results[2]
    for ck in (1, 6):  # the checkpoint interval does not change the arithmetic
        assert results[ck][0] == loss
        assert np.array_equal(results[ck][1], gf)
        assert np.array_equal(results[ck][2], gb)
    fg_steps = [p.fgeos * bs[n] for n in range(nsteps)]
    ref = _oracle_loss(p, p.f2, fg_steps, pts, w, target, kp)
    print(f"loss gpu {loss:.15e} oracle {ref:.15e}", flush=True)
    assert abs(loss - ref) <= 1e-10 * abs(ref)
    rng = np.random.default_rng(3)
    # 64 random directions in f0 (central differences of the oracle, h = 1e-6 relative)
    worst = 0.0
    for m in range(64):
        v = p.f2 * rng.standard_normal(p.f2.shape)
        ad = float(np.sum(gf * v))
        for h in (H, H / 10):  # a limiter kink within +-h biases the quotient: retry once at h/10
            lp = _oracle_loss(p, p.f2 + h * v, fg_steps, pts, w, target, kp)
            lm = _oracle_loss(p, p.f2 - h * v, fg_steps, pts, w, target, kp)
            fd = (lp - lm) / (2 * h)
            err = abs(fd - ad) / (abs(fd) + 1e-30)
            ok = err <= 1e-6 or abs(fd - ad) <= fd_noise(loss, h)
            if m < 8 or not ok or h != H:
                print(f"f0 direction {m} h={h:.0e}: fd {fd:.10e} adjoint {ad:.10e} rel {err:.2e}", flush=True)
            if ok:
                break
        worst = max(worst, min(err, abs(fd - ad) / fd_noise(loss, h)))
        assert ok
    print(f"64 f0 directions: worst rel error {worst:.2e} (or fraction of the rounding floor)", flush=True)
    # per-step boundary: the FGEOS of one step at a time
    for n in (0, nsteps // 2, nsteps - 1):
        d = p.fgeos * rng.standard_normal(p.fgeos.shape)
        fp = [f + (H * d if m == n else 0) for m, f in enumerate(fg_steps)]
        fm = [f - (H * d if m == n else 0) for m, f in enumerate(fg_steps)]
        fd = (_oracle_loss(p, p.f2, fp, pts, w, target, kp) - _oracle_loss(p, p.f2, fm, pts, w, target, kp)) / (2 * H)
        # the step's applied boundary is FGEOS*bscale: dL/dFGEOS_applied = grad / bscale
        ad = float(np.sum(gb[n] / bs[n] * d))
        err = abs(fd - ad) / (abs(fd) + 1e-30)
        print(f"boundary step {n}: fd {fd:.10e} adjoint {ad:.10e} rel {err:.2e}", flush=True)
        assert err <= 1e-6 or abs(fd - ad) <= fd_noise(loss)
    g.close()


def test_adjoint_run_overlapped_segments_equal_one_segment():
    """Without a Kp or boundary-scale schedule the tables stay fixed over the
    run, and ramgpu_adjoint_run recomputes segment g-1 on a second stream
    while segment g's adjoint runs (two tapes).  The kernels and their data
    are those of the serial order, so every checkpoint interval (odd and even
    segment counts, a short last segment) gives the bits of the one-segment
    run, whose loss is the oracle's."""
    nsteps = 7
    p, pts, w, target, _, _ = _misfit_setup(nsteps, seed=11)
    g = RamGpu.from_problem(p)
    res = {}
    for ck in (7, 1, 2, 3):
        g.put_f2(p.f2)
        res[ck] = g.adjoint_run(nsteps, pts, w, target, ckpt_every=ck)
        print(f"ckpt {ck}: loss {res[ck][0]:.15e}", flush=True)
    loss, gf, gb = res[7]
    for ck in (1, 2, 3):
        assert res[ck][0] == loss, ck
        assert np.array_equal(res[ck][1], gf), ck
        assert np.array_equal(res[ck][2], gb), ck
    # the context is left in the state after the run
    g.put_f2(p.f2)
    g.ram_run(nsteps)
    after = g.get_f2()
    g.put_f2(p.f2)
    g.adjoint_run(nsteps, pts, w, target, ckpt_every=2)
    assert np.array_equal(g.get_f2(), after)
    o = Oracle(p)
    ref = 0.0
    for n in range(nsteps):
        o.ram_run(1)
        r = np.log(inputs.satellite_flux(o.f2, pts, w)) - target[n]
        ref += float(np.sum(r * r))
    o.close()
    print(f"loss gpu {loss:.15e} oracle {ref:.15e}", flush=True)
    assert abs(loss - ref) <= 1e-10 * abs(ref)
    assert np.linalg.norm(gf) > 0 and np.linalg.norm(gb) > 0
    g.close()
