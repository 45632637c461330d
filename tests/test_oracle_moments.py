"""Pins the oracle's ANISCH and FLUX (src/ModRamRun.f90:208-222,265-415).

ModRamRun cannot be compiled here: it uses ModRamLoss, which uses ModRamIO
and NetCDF.  The restatement is therefore checked against what the physics
of the moments fixes, on the dipole grids of the bench workloads:

- the trace of the pressure tensor: with the reference's weights,
  2*PPERT + PPART = 2*EDENT for any F2 (the perpendicular and parallel
  pitch-angle weights add up to the number weight, and EPP = ERNH*EKEV);
- a unit isotropic flux at one energy K0 has energy density / number density
  = EKEV(K0), and number density 4*pi/c * ERNH(K0) * (the pitch-angle
  cosine range outside the loss cone);
- an isotropic flux has P_perp / P_par equal to the closed form of the
  truncated pitch-angle integrals, within the quadrature's error;
- FLUX of F2 = j*FFACTOR*FNHS is j;
- the L=1 row of F2 becomes the L=2 row, over I>=2 and K>=2 only.

A plausible slip in the restatement (a dropped factor 2, MU for MU**2, the
perpendicular and parallel weights swapped, an energy or pitch-angle range
off by one, ERNH for EPP) fails at least one of these.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from ramgpu import inputs

RFAC = 4 * math.pi / (2.998e8 * 100)


@pytest.fixture(scope="module")
def plain():
    return inputs.plain_h(seed=3)


def _unit_flux_f2(p, kmask=None):
    """F2 = FFACTOR*FNHS (unit isotropic flux) at the energies kmask."""
    g, geo = p.grid, p.geom
    ns, nr, nt, ne, npa = p.f2.shape
    ff = g.ffactor  # (NS,NR,NE,NPA)
    fn = geo.fnhs[:nr]  # (NR,NT,NPA)
    f2 = ff[:, :, None, :, :] * fn[None, :, :, None, :]
    if kmask is not None:
        f2 = f2 * kmask[None, None, None, :, None]
    return np.asfortranarray(f2)


def test_pressure_trace_identity(plain):
    o = Oracle(plain)
    out = o.anisch(1)
    ppert, ppart, edent = out["ppert"][0, 1:], out["ppart"][0, 1:], out["edent"][0, 1:]
    assert np.all(ppert > 0) and np.all(ppart > 0)
    rel = np.abs(2 * ppert + ppart - 2 * edent) / (2 * edent)
    assert rel.max() < 1e-13, rel.max()
    # I=1 is not computed
    assert np.all(out["ppert"][0, 0] == 0)


def test_mono_energetic_densities(plain):
    g, geo = plain.grid, plain.geom
    ne = g.ne
    i, j = 10, 7
    u = int(g.upa[i] - 1)
    for k0 in range(ne):
        o = Oracle(plain)
        mask = np.zeros(ne)
        mask[k0] = 1.0
        o.set_f2(_unit_flux_f2(plain, mask))
        out = o.anisch(1)
        rn, ed = out["rnhtt"][0, i, j], out["edent"][0, i, j]
        if k0 == 0:  # K=1 is outside the energy sums
            assert rn == 0 and ed == 0
            continue
        assert abs(ed / rn - g.ekev[k0]) <= 1e-14 * g.ekev[k0], k0
        # pitch-angle range: cells 1..u; cell 1 carries F2(L=2), whose flux
        # weight there is FNHS(L=2)/FNHS(L=1) (FFACTOR(L=1) = FFACTOR(L=2))
        mu_range = g.wmu[0] * geo.fnhs[i, j, 1] / geo.fnhs[i, j, 0] + g.wmu[1:u].sum()
        want = RFAC * g.ernh[0, k0] * mu_range
        assert abs(rn - want) <= 1e-13 * want, (k0, rn, want)


def test_isotropic_anisotropy_closed_form(plain):
    g, geo = plain.grid, plain.geom
    o = Oracle(plain)
    o.set_f2(_unit_flux_f2(plain))
    out = o.anisch(1)
    for i in (3, 10, 19):
        j = 5
        u = int(g.upa[i] - 1)
        # continuum: mu over [-a, m], m the upper edge of cell u, a the half
        # width of cell 1 (centred on mu = 0); cell 1 holds F2(L=2), a flux
        # FNHS(L=2)/FNHS(L=1) there, at mu = 0 (perpendicular only)
        m, a = g.mu[u - 1] + 0.5 * g.wmu[u - 1], 0.5 * g.wmu[0]
        r = geo.fnhs[i, j, 1] / geo.fnhs[i, j, 0]
        perp = (m + a) - (m ** 3 + a ** 3) / 3 + (r - 1) * g.wmu[0]
        par = (m ** 3 + a ** 3) / 3
        want = perp / (2 * par)
        got = out["ppert"][0, i, j] / out["ppart"][0, i, j]
        assert abs(got / want - 1) < 3e-3, (i, got, want)


def test_anisch_side_effect_and_ranges(plain):
    o = Oracle(plain)
    f0 = o.f2.copy()
    o.anisch(1)
    f1 = o.f2
    want = f0.copy()
    want[0, 1:, :, 1:, 0] = f0[0, 1:, :, 1:, 1]
    assert np.array_equal(f1, want)


def test_flux_round_trip_and_zeros(plain):
    o = Oracle(plain)
    rng = np.random.default_rng(1)
    j = rng.uniform(0.5, 2.0, plain.f2.shape)
    o.set_f2(np.asfortranarray(_unit_flux_f2(plain) * j))
    fx = o.flux(np.full(plain.f2.shape, np.nan, order="F"))
    inside = fx[:, 1:, :-1, 1:, 1:]
    assert np.max(np.abs(inside / j[:, 1:, :-1, 1:, 1:] - 1)) < 1e-15
    # outside the reference's ranges the array is left alone
    for sl in (np.s_[:, 0], np.s_[:, :, -1], np.s_[:, :, :, 0], np.s_[:, :, :, :, 0]):
        assert np.all(np.isnan(fx[sl]))


def test_windows_cover_energies_once_when_ne_below_30():
    # NE < 30: the windows end at NE (the reference would read past NE)
    p = inputs.make_problem(nr=6, nt=5, ne=12, npa=20, seed=2)
    o = Oracle(p)
    o.set_f2(_unit_flux_f2(p))
    full = o.anisch(1)["rnhtt"][0, 3, 2]
    parts = 0.0
    for k0 in range(1, 12):
        o2 = Oracle(p)
        mask = np.zeros(12)
        mask[k0] = 1.0
        o2.set_f2(_unit_flux_f2(p, mask))
        parts += o2.anisch(1)["rnhtt"][0, 3, 2]
    assert abs(parts - full) <= 1e-13 * full
