"""The oracle's Coulomb operators (COULPARA, COULEN, COULMU of
src/ModRamCoul.f90) against the reference's own Fortran and what the
mathematics fixes.

* Golden vectors (tests/golden/coulomb/, written by the reference's
  ModRamCoul compiled from its sources, tests/golden/make_golden_coulomb.py):
  the six COULPARA tables and F2 after COULEN, COULMU with TimeRamElapsed > 0
  and COULMU with TimeRamElapsed = 0, bit for bit, on a two-species problem
  whose spikes, holes and negative cells make both clamps act.
* Where the reference's library is built (oracle/_ref), operator after
  operator bit for bit on four species, dipole and randomised geometry.
* COULMU is a conservative implicit diffusion in pitch angle with zero-flux
  ends: with the weights MU*WMU*FNHS its couplings are symmetric
  (MU(L)*WMU(L)*FNHS(L)*AN(L) = MU(L+1)*WMU(L+1)*FNHS(L+1)*GN(L+1) =
  COULD(L)*BASCNE(L)*BOUNHS(L)/DMU(L), src/ModRamCoul.f90:97-110,268-271)
  and RL(1) = -1, F2(NPA) = F2(NPA-1) close the ends, so sum over
  L = 2..NPA-1 of WMU(L)*F2(L) of every (I,J,K) line is unchanged.  A wrong
  BASCNE, DMU or MU index, or a dropped FNHS, breaks it.
* ram_run with Coulomb collisions places COULEN, COULMU after the first
  drift half and COULMU, COULEN before the second (src/ModRamRun.f90:81-89,
  160-167) with their SUMRC accumulators LSCOE, LSCSC.
"""
import numpy as np
import pytest

from golden_io import COULOMB_TABLES, coulomb_fixtures, load
from helpers import random_problem
from oracle.oracle import Oracle, Reference
from ramgpu import inputs

needs_ref = pytest.mark.skipif(not Reference.available(),
                               reason="oracle/_ref not built (reference sources absent)")


def _coulomb_problem(seed, random_geometry, time_elapsed=1800.0, negative=False, **kw):
    dims = dict(nr=6, nt=7, ne=12, npa=18)
    dims.update(kw)
    if random_geometry:
        p = random_problem(species=inputs.SPECIES4, composition=inputs.COMPOSITION4, seed=seed,
                           spiky=True, **dims)
    else:
        p = inputs.make_problem(species=inputs.SPECIES4, composition=inputs.COMPOSITION4, kp=4.0,
                                seed=seed, **dims)
    rng = np.random.default_rng(seed + 7)
    if negative:
        m = rng.uniform(0, 1, p.f2.shape)
        p.f2 = np.asfortranarray(np.where(m < 0.03, -p.f2, p.f2))
    necr = inputs.plasmasphere_density(p.grid, 4.0) * rng.uniform(0.5, 1.5, (p.grid.nr, p.grid.nt))
    p.coulomb = inputs.CoulombInputs(np.asfortranarray(necr), time_elapsed)
    return p


@pytest.mark.parametrize("path", coulomb_fixtures(), ids=lambda p: p.stem)
def test_coulomb_golden(path):
    p, exp = load(path)
    ns = p.grid.ns
    for tag, op, time in (("coulen", "coulen", p.coulomb.time_elapsed),
                          ("coulmu", "coulmu", p.coulomb.time_elapsed), ("coulmu_t0", "coulmu", 0.0)):
        o = Oracle(p)
        o.set_coulomb(p.coulomb.necr, time)
        for s in range(1, ns + 1):
            o.call("coulpara", s)
            for n in COULOMB_TABLES:
                assert np.array_equal(o.coul_table(n)[s - 1], exp[f"{n}_s{s}"]), (n, s)
            o.call(op, s)
            assert np.array_equal(o.f2[s - 1], exp[f"f2_s{s}_{tag}"]), (tag, s)
        o.close()
    # the fixture makes both clamps act, and only where TimeRamElapsed > 0 for COULMU
    assert np.any(exp["f2_s1_coulen"] == 1e-15)
    assert np.any(exp["f2_s1_coulmu"] == 1e-15)
    assert np.any(exp["f2_s1_coulmu_t0"] < 0)


@needs_ref
@pytest.mark.parametrize("random_geometry", [False, True])
@pytest.mark.parametrize("time_elapsed", [0.0, 1800.0])
def test_coulomb_reference_bit_exact(random_geometry, time_elapsed):
    p = _coulomb_problem(11, random_geometry, time_elapsed, negative=random_geometry)
    o = Oracle(p)
    r = Reference(p)
    r.setup_coulomb(p.grid.vbnd, p.coulomb.necr, time_elapsed)
    for s in range(1, p.grid.ns + 1):
        for op in ("coulpara", "coulen", "coulmu", "coulmu", "coulen"):
            o.call(op, s)
            r.call(op, s)
            if op == "coulpara":
                tabs = r.coulomb_tables()
                for n in COULOMB_TABLES:
                    assert np.array_equal(o.coul_table(n)[s - 1], tabs[n][s - 1]), (n, s)
            else:
                a, b = o.f2, r.f2
                assert np.array_equal(a, b), f"{op}(S={s}) differs: max |d|={np.abs(a - b).max()}"
    o.close()


def test_coulmu_conserves_pitch_angle_content():
    p = _coulomb_problem(5, False, npa=24)
    g = p.grid
    o = Oracle(p)
    before = o.f2.copy()
    wsum = lambda f: np.einsum("sijkl,l->sijk", f[:, :, :, :, 1:-1], g.wmu[1:-1])  # noqa: E731
    for s in range(1, g.ns + 1):
        o.call("coulpara", s)
        o.call("coulmu", s)
    after = o.f2
    lines = (slice(None), slice(1, None), slice(None), slice(1, None))  # I >= 2, K >= 2
    assert np.all(after[lines][..., 1:] > 0)  # no clamp acted (F2(L=1) = 0 as MU(1) = 0)
    a, b = wsum(before)[lines], wsum(after)[lines]
    assert np.max(np.abs(a - b) / a) < 1e-13
    d = np.abs(after[lines] - before[lines])[..., 1:] / before[lines][..., 1:]
    assert np.max(d) > 1e-3  # and it scattered
    o.close()


def test_ram_run_places_coulomb_and_its_accumulators():
    p = _coulomb_problem(3, False)
    p.wave = inputs.wave_inputs(p.grid, 4.0)
    o = Oracle(p)
    o.ram_run(1)
    # the same step with the Coulomb operators called where ram_run calls them
    m = Oracle(p)
    m.set_coulomb(on=False)
    ls = {n: np.zeros(p.grid.ns) for n in ("lsdr", "lscoe", "lscsc", "lswae", "lscha", "lsatm")}

    def sumrc(s, acc):
        m.call("sumrc", s)
        ls[acc][s - 1] += m.arr("elorc")[s - 1]

    for s in range(1, p.grid.ns + 1):
        wpi = bool(p.wave.wpi[s - 1])
        cex = p.grid.cex_kind[s - 1] != 0
        for op in ("cepara", "driftpara", "coulpara", "driftr", "driftp", "drifte", "driftmu"):
            m.call(op, s)
        sumrc(s, "lsdr")
        m.call("coulen", s)
        sumrc(s, "lscoe")
        m.call("coulmu", s)
        sumrc(s, "lscsc")
        for op, acc, on in (("wavelo", "lswae", wpi), ("charexchange", "lscha", cex), ("atmol", "lsatm", True),
                            ("atmol", "lsatm", True), ("charexchange", "lscha", cex), ("wavelo", "lswae", wpi)):
            if on:
                m.call(op, s)
                sumrc(s, acc)
        m.call("coulmu", s)
        sumrc(s, "lscsc")
        m.call("coulen", s)
        sumrc(s, "lscoe")
        for op in ("driftmu", "drifte", "driftp", "driftr"):
            m.call(op, s)
        sumrc(s, "lsdr")
    f = m.f2.copy()
    f[:, :, -1] = f[:, :, 0]
    f[:, p.geom.outside_mgnp == 1] = 1e-31
    assert np.array_equal(o.f2, f)
    for n, v in ls.items():
        assert np.array_equal(o.arr(n), v), n
    assert np.all(ls["lscoe"] != 0) and np.all(ls["lscsc"] != 0)
    o.close()
    m.close()
