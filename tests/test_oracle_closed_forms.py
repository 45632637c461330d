"""Checks that pin the oracle where the reference's own code cannot be built
here (src/ModRamLoss.f90 and src/ModRamRun.f90 need NetCDF and GSL), plus
textbook special cases of the drift scheme.  Each check computes its answer
a different way from the oracle, so that a dropped term, a wrong index or
range, or a transposed operand fails it.
"""
import math

import numpy as np
import pytest

from helpers import random_problem
from oracle.oracle import Oracle
from ramgpu import inputs, species as sp

# Charge-exchange cross-section fits of CEPARA, log10(sigma [m^2]) as a quartic
# in x = log10(E [keV]) (src/ModRamLoss.f90:46,61,76; the published fits of the
# reference), as (c0, c1, c2, c3, c4).
CEX_FITS = {
    sp.CEX_H: (-18.767, -0.11017, -3.8173e-2, -0.1232, -5.0488e-2),
    sp.CEX_HE: (-20.789, 0.92316, -0.68017, 0.66153, -0.20998),
    sp.CEX_O: (-18.987, -0.10613, -5.4841E-3, -1.6262E-2, -7.0554E-3),
}


def _problem():
    return random_problem(nr=7, nt=8, ne=9, npa=10, species=(sp.HYDROGEN, sp.HELIUM, sp.OXYGEN,
                                                             sp.ELECTRON), seed=11)


def test_cepara_is_the_published_quartic_fit():
    """-ln(CHARGE)/(V*HDNS*DTs) must equal 10**Y(x) with Y the species' quartic
    in x=log10(EKEV), the same for every I,J,L: recover the five coefficients
    by least squares over K and compare to the published fit."""
    p = _problem()
    o = Oracle(p)
    g = p.grid
    for s, spc in enumerate(g.species, start=1):
        o.call("cepara", s)
        ch = o.loss_table(s, "charge")  # (NR,NT,NE,NPA)
        if not spc.cex:
            assert np.all(ch == 1.0)
            continue
        # unchanged rows: I=1, K=1, L=1 (src/ModRamLoss.f90:34,40-43)
        assert np.all(ch[0] == 1.0) and np.all(ch[:, :, 0] == 1.0) and np.all(ch[:, :, :, 0] == 1.0)
        hd = p.geom.hdns[: g.nr]  # (NR,NT,NPA)
        alpha = -np.log(ch[1:, :, 1:, 1:])
        sig = alpha / (g.v[s - 1, 1:][None, None, :, None] * hd[1:, :, None, 1:] * p.dts)
        # one cross-section per K, independent of I, J, L; alpha = -ln(CHARGE)
        # is recovered to ~eps absolute, so the bound scales with eps/alpha
        sig_k = np.array([sig[:, :, k, :].ravel()[np.argmax(alpha[:, :, k, :].ravel())]
                          for k in range(g.ne - 1)])
        bound = 8 * np.finfo(float).eps / alpha + 1e-14
        assert np.all(np.abs(sig / sig_k[None, None, :, None] - 1.0) <= bound)
        x = np.maximum(np.log10(g.ekev[1:]), -2.0)
        coef = np.polynomial.polynomial.polyfit(x, np.log10(sig_k), 4)
        np.testing.assert_allclose(coef, CEX_FITS[spc.cex_kind], rtol=0, atol=1e-6)


def test_atmol_closed_form_and_range():
    """ATMOL multiplies F2 by exp(-DTs*V/(2*RLZ))**(1/FNHS) for K>=2, I>=2 and
    int(UPA(I)) <= L (src/ModRamLoss.f90:162-167,495-503), nothing else."""
    p = _problem()
    o = Oracle(p)
    g = p.grid
    for s in range(1, g.ns + 1):
        o.call("cepara", s)
        before = o.f2.copy()
        o.call("atmol", s)
        after = o.f2
        ratio = after[s - 1] / before[s - 1]
        expect = np.ones_like(ratio)
        for i in range(1, g.nr):
            u = int(g.upa[i])
            for k in range(1, g.ne):
                decay = math.exp(-p.dts * g.v[s - 1, k] / (2.0 * g.rlz[i]))
                for l in range(u - 1, g.npa):
                    expect[i, :, k, l] = decay ** (1.0 / p.geom.fnhs[i, :, l])
        np.testing.assert_allclose(ratio, expect, rtol=4e-15, atol=0)
        other = np.delete(np.arange(g.ns), s - 1)
        assert np.array_equal(after[other], before[other])


def test_charexchange_applies_charge_on_its_range():
    p = _problem()
    o = Oracle(p)
    g = p.grid
    for s, spc in enumerate(g.species, start=1):
        o.call("cepara", s)
        before = o.f2.copy()
        o.call("charexchange", s)
        ch = o.loss_table(s, "charge")
        after = o.f2[s - 1]
        np.testing.assert_array_equal(after[1:, :, 1:, 1:], before[s - 1][1:, :, 1:, 1:] * ch[1:, :, 1:, 1:])
        b = before[s - 1]
        assert np.array_equal(after[0], b[0]) and np.array_equal(after[:, :, 0], b[:, :, 0])
        assert np.array_equal(after[:, :, :, 0], b[:, :, :, 0])


def test_sumrc_is_the_weighted_energy_sum():
    """SETRC = sum over I>=2,K>=2,L>=2,J<=NT-1 of EKEV*F2*WE*WMU
    (src/ModRamRun.f90:246-255) -- compensated sum for comparison."""
    p = _problem()
    o = Oracle(p)
    g = p.grid
    for s in range(1, g.ns + 1):
        o.call("sumrc", s)
        f = o.f2[s - 1][1:, : g.nt - 1, 1:, 1:]
        w = (g.ekev[1:, None] * g.we[1:, None] * g.wmu[None, 1:])
        terms = (f * w[None, None]).ravel()
        exact = math.fsum(terms.tolist())
        assert abs(o.arr("setrc")[s - 1] - exact) <= 1e-13 * abs(exact)
        assert o.arr("elorc")[s - 1] == 0.0 - o.arr("setrc")[s - 1]


def _sweby_periodic(f, nu, beta):
    """One step of the Sweby flux-limited upwind scheme for u_t + a u_x = 0,
    a>0, Courant number nu, limiter phi(r)=max(0,min(beta r,1),min(r,beta)),
    on a periodic grid (LeVeque 2002, Finite Volume Methods for Hyperbolic
    Problems, sec. 6.12)."""
    n = f.size
    out = np.empty(n)
    flux = np.empty(n)  # flux through the right face of cell i
    for i in range(n):
        fm, f0, fp = f[i - 1], f[i], f[(i + 1) % n]
        d = fp - f0
        r = (f0 - fm) / d if d != 0 else 0.0
        phi = max(0.0, min(beta * r, 1.0), min(r, beta))
        flux[i] = nu * (f0 + 0.5 * (1.0 - nu) * phi * d)
    for i in range(n):
        out[i] = f[i] - flux[i] + flux[i - 1]
    return out


def test_driftp_is_the_textbook_limited_scheme():
    """With uniform geometry and no convection the azimuthal Courant number is
    one constant per (I,K,L) line, and DRIFTP must be the Sweby scheme on the
    periodic ring of NT-1 cells (F(1) and F(NT) the same cell)."""
    p = inputs.make_problem(nr=6, nt=13, ne=5, npa=6, kp=0.0, seed=4)
    p.vt[...] = 0.0
    o = Oracle(p)
    o.call("driftpara", 1)
    before = o.f2.copy()
    o.call("driftp", 1)
    c = o.drift_table(1, "cdp")
    nt = p.grid.nt
    checked = 0
    for i in range(1, p.grid.nr):
        for k in range(p.grid.ne):
            for l in range(p.grid.npa):
                nu = c[i, 1, k, l]
                assert np.allclose(c[i, 1:, k, l], nu, rtol=1e-15)
                assert 0 < nu < 1
                ring = before[0, i, 1:, k, l]          # cells J=2..NT (J=NT is J=1)
                ring = np.roll(ring, 1)                 # order J=1(=NT),2,..,NT-1
                ref = _sweby_periodic(ring, nu, p.betalim)
                got = np.roll(o.f2[0, i, 1:, k, l], 1)
                np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
                checked += 1
    assert checked > 0


def test_driftp_conserves_the_ring_total():
    """The flux form telescopes on the periodic ring: sum over J=2..NT of F2 is
    unchanged by DRIFTP when nothing is clamped (src/ModRamDrift.f90:261-272)."""
    p = random_problem(seed=3, noise=0.1)
    o = Oracle(p)
    o.call("driftpara", 1)
    before = o.f2[0][1:, 1:].sum(axis=1)
    o.call("driftp", 1)
    after = o.f2[0][1:, 1:].sum(axis=1)
    np.testing.assert_allclose(after, before, rtol=1e-12)


def test_drifts_leave_uniform_interior_unchanged_when_c_is_zero():
    """C=0 everywhere (DTs-scaled drifts with DTs->0 limit replaced by zero
    fields): no operator changes F2 away from its boundary conventions."""
    p = inputs.make_problem(nr=6, nt=7, ne=6, npa=8, kp=0.0, seed=2)
    p.vt[...] = 0.0
    p.geom.fnis[...] = 2.0 * p.geom.fnhs  # GPA2-type terms vanish with FNIS = 2 FNHS
    o = Oracle(p)
    o.call("driftpara", 1)
    before = o.f2.copy()
    o.call("driftr", 1)   # ExB and gradient-curvature both zero
    np.testing.assert_array_equal(o.f2[0][1:], before[0][1:])


def test_ram_run_is_the_reference_operator_sequence():
    """oracle_ram_run must equal the operators composed in the order of
    src/ModRamRun.f90:66-186,197-201 with the SUMRC accumulators."""
    p = random_problem(nr=7, nt=8, ne=9, npa=10, species=(sp.HYDROGEN, sp.ELECTRON), seed=5)
    whole = Oracle(p)
    whole.ram_run(1)
    parts = Oracle(p)
    g = p.grid
    for s, spc in enumerate(g.species, start=1):
        seq = ["cepara", "driftpara", "driftr", "driftp", "drifte", "driftmu", ("sumrc", "lsdr")]
        if spc.cex:
            seq += ["charexchange", ("sumrc", "lscha")]
        seq += ["atmol", ("sumrc", "lsatm"), "atmol", ("sumrc", "lsatm")]
        if spc.cex:
            seq += ["charexchange", ("sumrc", "lscha")]
        seq += ["driftmu", "drifte", "driftp", "driftr", ("sumrc", "lsdr")]
        for op in seq:
            if isinstance(op, tuple):
                parts.call("sumrc", s)
                parts.arr(op[1])[s - 1] += parts.arr("elorc")[s - 1]
            else:
                parts.call(op, s)
    f = parts.f2
    f[:, :, g.nt - 1] = f[:, :, 0]
    mg = p.geom.outside_mgnp
    for i in range(g.nr):
        for j in range(g.nt):
            if mg[i, j] == 1:
                f[:, i, j] = 1e-31
    assert np.array_equal(whole.f2, f)
    for n in ("setrc", "lsdr", "lscha", "lsatm"):
        assert np.array_equal(whole.arr(n), parts.arr(n)), n
