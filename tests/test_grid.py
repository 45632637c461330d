"""The host grid (restatement of ARRAYS, src/ModRamInit.f90:364-587) against
closed forms and the dipole field it encodes."""
import math

import numpy as np
import pytest

from ramgpu import grid as G
from ramgpu import inputs, species as sp


@pytest.fixture(scope="module")
def g72():
    return G.build_grid(20, 25, 35, 72, inputs.SPECIES4)


def test_energy_grid_geometric(g72):
    we1, rw = 3e-2, 1.27
    k = np.arange(g72.ne)
    np.testing.assert_allclose(g72.we, we1 * rw ** k, rtol=1e-13)
    ebnd = 0.1 + we1 * (rw ** (k + 1) - 1) / (rw - 1)
    np.testing.assert_allclose(g72.ebnd, ebnd, rtol=1e-13)
    np.testing.assert_allclose(g72.ekev, g72.ebnd - 0.5 * g72.we, rtol=1e-13)
    np.testing.assert_allclose(g72.de[:-1], np.diff(g72.ekev), rtol=1e-12)
    assert abs(g72.ebnd[-1] - 477.0) < 1.0


@pytest.mark.parametrize("ne", [48, 64])
def test_other_ne_keep_the_span(ne):
    g = G.build_grid(8, 9, ne, 10, (sp.HYDROGEN,))
    g35 = G.build_grid(8, 9, 35, 10, (sp.HYDROGEN,))
    assert abs(g.ebnd[-1] / g35.ebnd[-1] - 1) < 1e-9


def test_relativistic_factors(g72):
    # gamma - 1 = E / (m c^2), v = c sqrt(1 - 1/gamma^2)
    for s in range(g72.ns):
        mc2_kev = g72.rmas[s] * G.CS ** 2 / G.Q / 1e3
        np.testing.assert_allclose(g72.grel[s] - 1, g72.ekev / mc2_kev, rtol=1e-12, atol=2.3e-16)
        # the reference forms sqrt(GREL**2-1), which cancels when GREL-1 is
        # tiny: compare with the cancellation-free form to eps/(GREL-1)
        gm1 = g72.ekev / mc2_kev
        exact = G.CS * np.sqrt(gm1 * (gm1 + 2)) / (1 + gm1)
        tol = 4 * np.finfo(float).eps / gm1 + 1e-14
        assert np.all(np.abs(g72.v[s] / exact - 1) <= tol)


def test_pitch_angle_grid_72(g72):
    mu, wmu, dmu = g72.mu, g72.wmu, g72.dmu
    assert mu[0] == 0.0 and mu[-1] == 1.0
    assert np.all(np.diff(mu) > 0)
    assert mu[47] == math.cos(math.radians(18.7)) or abs(mu[47] - math.cos(math.radians(18.7))) < 1e-15
    assert abs(mu[49] - math.cos(math.radians(16.0))) < 1e-15
    np.testing.assert_allclose(wmu[1:47], wmu[0] * 0.98 ** np.arange(1, 47), rtol=1e-13)
    assert wmu[0] == 1 / 32
    np.testing.assert_allclose(dmu[46:71], np.diff(mu)[46:71], rtol=1e-13)


def test_loss_cone_is_the_dipole_mirror_ratio(g72):
    """sin^2 of the loss-cone angle = B_eq / B(atmosphere) on the dipole line
    through L, computed here from the magnetic latitude of the foot point."""
    for i in range(g72.nr):
        r = (G.RE + G.HMIN) / (G.RE * g72.lz[i])   # r_foot / (L RE) = cos^2(lat)
        lat = math.acos(math.sqrt(r))
        ratio = math.sqrt(1 + 3 * math.sin(lat) ** 2) / math.cos(lat) ** 6
        assert abs(math.sin(math.radians(g72.cone[i])) ** 2 - 1 / ratio) < 1e-13
        u = int(g72.upa[i])
        assert g72.pa[u - 1] <= g72.cone[i] and (u == 1 or g72.pa[u - 2] > g72.cone[i])


def test_ffactor_and_conf(g72):
    for s in range(g72.ns):
        beta = g72.v[s] / G.CS
        ff = g72.lz[:g72.nr, None, None] ** 2 / beta[None, :, None] * g72.mu[None, None, :]
        np.testing.assert_allclose(g72.ffactor[s][:, :, 1:], ff[:, :, 1:], rtol=1e-12)
        np.testing.assert_array_equal(g72.ffactor[s][:, :, 0], g72.ffactor[s][:, :, 1])
    assert abs(g72.conf1 - ((6.5 + 0.25) / 6.5) ** 2) < 1e-15
    assert abs(g72.conf2 - ((6.5 + 0.5) / 6.5) ** 2) < 1e-15


def test_ejiri_bounce_integrals_limits():
    """h(mu) and I(mu) of the dipole (Ejiri 1978): h = 1 + ln(2+sqrt3)/(2 sqrt3)
    for field-aligned particles (y=0), I = 0 for equatorially mirroring ones."""
    assert abs(inputs.funt(1.0) - (1 + math.log(2 + math.sqrt(3)) / (2 * math.sqrt(3)))) < 1e-15
    assert inputs.funi(0.0) == 0.0
    assert abs(inputs.funt(0.0) - 0.7405) < 2e-3


def test_volland_stern_closed_form():
    g = G.build_grid(8, 9, 8, 10, (sp.HYDROGEN,))
    vt = inputs.volland_stern(g, 3.0)
    a = 7.05e-6 / (1 - 0.159 * 3 + 0.0093 * 9) ** 3 / G.RE
    expect = a * (g.lz[:, None] * G.RE) ** 2 * np.sin(g.phi[None, :])
    np.testing.assert_allclose(vt, expect, rtol=1e-14, atol=1e-9)


def test_fast_index_division_is_exact():
    """The kernels divide small non-negative integers by NE or NT with a float
    reciprocal and one remainder correction; check it against // for every
    divisor up to 256 and numerators covering each kernel's range."""
    rng = np.random.default_rng(0)
    n = np.concatenate([np.arange(0, 20000), rng.integers(0, 1 << 22, 200000)]).astype(np.int64)
    for d in range(1, 257):
        rd = np.float32(1.0) / np.float32(d)
        q = np.trunc(n.astype(np.float32) * rd).astype(np.int64)
        r = n - q * d
        q = np.where(r < 0, q - 1, np.where(r >= d, q + 1, q))
        assert np.array_equal(q, n // d), d
