"""Pins the oracle's whole ram_run step -- the operator order and the SUMRC
bookkeeping of src/ModRamRun.f90:64-205,231-259 -- by what the mathematics
fixes, not by re-composing the oracle's own operators.

Configuration: a geometry that does not depend on local time J (the dipole),
no convection potential (VT = 0) and no induced fields (EIR = EIP = dBdt =
dIdt = dIbndt = 0).  Then, reading src/ModRamDrift.f90:118-146,320-342,
418-433, every term of CDriftR, CDriftE and CDriftMu is a difference of equal
values along J, times finite factors, so they are exactly zero:
  * DRIFTR and DRIFTE change nothing (F2 - 0*FBND + 0*FBND);
  * DRIFTMU changes nothing but its top cell, T: F2(NPA) = F2(NPA-1) *
    FNHS(NPA)*MU(NPA)/FNHS(NPA-1)/MU(NPA-1) (:466);
  * DRIFTP is the textbook Sweby beta-limited upwind scheme S on each
    periodic (I,K,L) ring of NT-1 cells, with one constant Courant number per
    ring (LeVeque 2002, Finite Volume Methods for Hyperbolic Problems, 6.12).
The geocoronal density HDNS enters only CHARGE, so it is given a J- and
L-dependence: the loss factors then do not commute with S or T, and only the
reference's order gives the expected state.  The charge-exchange factor is the
published cross-section fit exp(-10^Y(log10 E) V HDNS DTs) (src/ModRamLoss.f90:
46,61,76) and the atmospheric factor exp(-DTs V/(2 RLZ))^(1/FNHS) (:162-167,
495-503), both formed here from their closed forms.  One step is then
  f1 = T(S(f0)); f2 = C f1; f3 = A f2; f4 = A f3; f5 = C f4; f6 = S(T(f5)),
followed by the seam copy F2(:,:,NT) = F2(:,:,1) and the magnetopause mask,
with SUMRC after each of f1..f6 and the accumulators
  LSDR += E0-E1, LSCHA += E1-E2, LSATM += E2-E3, LSATM += E3-E4,
  LSCHA += E4-E5, LSDR += E5-E6,
where every E is a compensated sum (math.fsum) of the closed-form state.
A dropped or doubled loss operator, a loss placed outside the drift halves, a
drift half in the wrong place relative to the loss (row NPA), a misplaced
accumulator or a wrong seam direction each fail this test.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from ramgpu import inputs, species as sp

# published charge-exchange fits, log10(sigma [m^2]) = c0 + c1 x + .. + c4 x^4,
# x = max(log10(E [keV]), -2)  (src/ModRamLoss.f90:44-46,59-61,74-76)
CEX_FITS = {
    sp.CEX_H: (-18.767, -0.11017, -3.8173e-2, -0.1232, -5.0488e-2),
    sp.CEX_HE: (-20.789, 0.92316, -0.68017, 0.66153, -0.20998),
    sp.CEX_O: (-18.987, -0.10613, -5.4841E-3, -1.6262E-2, -7.0554E-3),
}


def _problem(species, seed, nr=7, nt=9, ne=8, npa=10, masked=((3, 4), (5, 2))):
    p = inputs.make_problem(nr=nr, nt=nt, ne=ne, npa=npa, species=species,
                            composition=tuple(1.0 for _ in species), kp=0.0, seed=seed)
    rng = np.random.default_rng(seed + 100)
    geo = p.geom
    for n in ("fnhs", "fnis", "bounhs", "bounis"):
        a = getattr(geo, n)
        assert np.all(a == a[:, :1, :])          # the dipole is J-independent
    assert np.all(geo.bnes == geo.bnes[:, :1])
    p.vt[...] = 0.0
    for n in ("eir", "eip", "dbdt", "didt", "dibndt"):
        getattr(geo, n)[...] = 0.0
    hd = geo.hdns * rng.uniform(0.3, 3.0, geo.hdns.shape)
    hd[:, nt - 1] = hd[:, 0]      # J=1 and J=NT are one local time
    geo.hdns = np.asfortranarray(hd)
    mg = np.zeros((nr, nt), dtype=np.int32, order="F")
    for i, j in masked:
        mg[i, j] = 1
    geo.outside_mgnp = mg
    # a J-dependent initial state with an exact seam is what any step leaves;
    # the seam cell J=1 is overwritten by DRIFTP anyway
    f2 = p.f2 * np.exp(0.4 * rng.standard_normal(p.f2.shape))
    f2[:, :, nt - 1] = f2[:, :, 0]
    # row I=1 is read by no operator of this configuration; an inexact seam
    # there shows which way the seam copy of src/ModRamRun.f90:186 goes
    f2[:, 0, nt - 1] *= 1.5
    p.f2 = np.asfortranarray(f2)
    return p


def _sweby_ring(f, nu, beta):
    """One step of the Sweby beta-limited scheme, phi(r) = max(0, min(beta r,
    1), min(r, beta)), on a periodic ring, for either sign of the Courant
    number nu (a negative one is the mirrored scheme)."""
    if nu < 0:
        return _sweby_ring(f[::-1], -nu, beta)[::-1]
    n = f.size
    flux = np.empty(n)
    for i in range(n):
        fm, f0, fp = f[i - 1], f[i], f[(i + 1) % n]
        d = fp - f0
        r = (f0 - fm) / d if d != 0 else 0.0
        phi = max(0.0, min(beta * r, 1.0), min(r, beta))
        flux[i] = nu * (f0 + 0.5 * (1.0 - nu) * phi * d)
    return f - flux + np.roll(flux, 1)


def _S(f, cdp, beta):
    """DRIFTP of one species' F2(NR,NT,NE,NPA): rings J=2..NT (J=NT is the
    cell J=1) for I>=2, then F2(I,1) = F2(I,NT)."""
    out = f.copy()
    nr, nt, ne, npa = f.shape
    for i in range(1, nr):
        for k in range(ne):
            for l in range(npa):
                nu = cdp[i, k, l]
                ring = np.roll(f[i, 1:, k, l], 1)      # J = NT(=1), 2, .., NT-1
                new = np.roll(_sweby_ring(ring, nu, beta), -1)
                assert new.min() > 0                   # no clamp on this path
                out[i, 1:, k, l] = new
                out[i, 0, k, l] = out[i, nt - 1, k, l]
    return out


def _sumrc(f, g):
    w = g.ekev[1:, None] * g.we[1:, None] * g.wmu[None, 1:]
    return math.fsum((f[1:, : g.nt - 1, 1:, 1:] * w[None, None]).ravel().tolist())


def _expected_step(p, s, cdp, setrc0):
    g, geo = p.grid, p.geom
    spc = g.species[s - 1]
    nr, nt, ne, npa = g.nr, g.nt, g.ne, g.npa
    beta, dts = p.betalim, p.dts
    fnhs = geo.fnhs[:nr]                                   # (NR,NT,NPA)
    ratio = fnhs[:, :, npa - 1] * g.mu[npa - 1] / fnhs[:, :, npa - 2] / g.mu[npa - 2]

    def T(f):
        out = f.copy()
        out[1:, :, :, npa - 1] = f[1:, :, :, npa - 2] * ratio[1:, :, None]
        return out

    v = g.v[s - 1]
    lc = np.ones((nr, nt, ne, npa))
    if spc.cex:
        c = CEX_FITS[spc.cex_kind]
        x = np.maximum(np.log10(g.ekev), -2.0)
        sig = 10.0 ** (c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3 + c[4] * x ** 4)
        alpha = sig[None, None, :, None] * v[None, None, :, None] * geo.hdns[:nr, :, None, :] * dts
        lc[1:, :, 1:, 1:] = np.exp(-alpha)[1:, :, 1:, 1:]
    la = np.ones((nr, nt, ne, npa))
    for i in range(1, nr):
        u = int(g.upa[i])                                  # L = u..NPA (1-based)
        decay = np.exp(-dts * v[1:] / (2.0 * g.rlz[i]))    # ATLOS(S,I,K), K>=2
        la[i, :, 1:, u - 1:] = decay[None, :, None] ** (1.0 / fnhs[i, :, None, u - 1:])

    f = p.f2[s - 1]
    states = []
    f = T(_S(f, cdp, beta)); states.append(f)
    if spc.cex:
        f = lc * f; states.append(f)
    f = la * f; states.append(f)
    f = la * f; states.append(f)
    if spc.cex:
        f = lc * f; states.append(f)
    f = _S(T(f), cdp, beta); states.append(f)
    e = [setrc0] + [_sumrc(x, g) for x in states]
    if spc.cex:
        ls = {"lsdr": (e[0] - e[1]) + (e[5] - e[6]), "lscha": (e[1] - e[2]) + (e[4] - e[5]),
              "lsatm": (e[2] - e[3]) + (e[3] - e[4])}
    else:
        ls = {"lsdr": (e[0] - e[1]) + (e[3] - e[4]), "lscha": 0.0,
              "lsatm": (e[1] - e[2]) + (e[2] - e[3])}
    return f, e[-1], ls


@pytest.mark.parametrize("species,seed", [
    ((sp.HYDROGEN, sp.ELECTRON), 5),
    ((sp.OXYGEN, sp.HELIUM), 8),
])
def test_ram_run_step_closed_form(species, seed):
    p = _problem(species, seed)
    g = p.grid
    # the Courant numbers of DRIFTP: one per ring, both signs present
    probe = Oracle(p)
    cdps = []
    for s in range(1, g.ns + 1):
        probe.call("driftpara", s)
        probe.call("driftp", s)
        c = probe.drift_table(s, "cdp")                  # (NR,NT,NE,NPA)
        assert np.all(c[1:] == c[1:, 1:2]), "CDriftP must not depend on J here"
        assert np.abs(c[1:]).max() < 1.0
        cdps.append(c[:, 1])
        for n in ("cdr", "cde", "cdmu"):
            assert np.all(probe.drift_table(s, n)[1:] == 0.0), n
    probe.close()
    assert any((c[1:] < 0).any() for c in cdps) or species[0] is not sp.HYDROGEN

    o = Oracle(p)
    setrc0 = [1.0e7 * (s + 1) for s in range(g.ns)]
    o.arr("setrc")[:] = setrc0
    o.ram_run(1)
    expect = np.empty_like(p.f2)
    for s in range(1, g.ns + 1):
        f, setrc, ls = _expected_step(p, s, cdps[s - 1], setrc0[s - 1])
        expect[s - 1] = f
        scale = max(abs(setrc0[s - 1]), abs(setrc))
        assert abs(o.arr("setrc")[s - 1] - setrc) <= 1e-13 * abs(setrc)
        for n, val in ls.items():
            got = o.arr(n)[s - 1]
            assert abs(got - val) <= 1e-12 * scale, (s, n, got, val)
        # the accumulators telescope: their sum is the energy that left
        tot = o.arr("lsdr")[s - 1] + o.arr("lscha")[s - 1] + o.arr("lsatm")[s - 1]
        assert abs(tot - (setrc0[s - 1] - o.arr("setrc")[s - 1])) <= 1e-13 * scale
    expect[:, :, g.nt - 1] = expect[:, :, 0]                   # F2(:,:,NT) = F2(:,:,1)
    mg = p.geom.outside_mgnp
    for i in range(g.nr):
        for j in range(g.nt):
            if mg[i, j] == 1:
                expect[:, i, j] = 1e-31
    got = o.f2
    rel = np.abs(got - expect) / np.abs(expect)
    assert rel.max() <= 1e-13, float(rel.max())
    # the step moved the state: S, T and the losses are all non-trivial here
    assert np.abs(expect - p.f2).max() > 1e-3 * np.abs(p.f2).max()
