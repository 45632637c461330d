"""GPU parity of the Coulomb operators (src/ModRamCoul.f90) through the C ABI.

COULPARA's tables (formed on the host with the C library's erf/exp/pow, as
the oracle does) and the device's COULEN and COULMU must equal the oracle
bit for bit: both are built of correctly rounded +, -, *, / and sqrt in the
reference's order.  The golden fixture written by the reference's own
Fortran is checked the same way.  ram_run with Coulomb collisions (the
fused plane kernel with COULEN, COULMU and their SUMRC slots) is compared
with the oracle's ram_run to the stated 1e-12 (its loss factors use the
table-driven exponential), with TimeRamElapsed advancing by 2*DTs a step
so that the first step runs COULMU without its clamp and the later ones
with it.  At the largest member shape (40x96x64x90) the operators are
checked bit for bit against the oracle as well.
"""
import numpy as np
import pytest

from conftest import gpu_available
from golden_io import COULOMB_TABLES, coulomb_fixtures, load
from helpers import max_rel_err, random_problem, rel_err
from oracle.oracle import Oracle
from ramgpu import inputs, species as sp
from ramgpu.capi import RamBatch, RamGpu, RamGpuError

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

RTOL = 1e-12


def _same(a, b, what):
    diff = np.argwhere(a != b)
    assert diff.size == 0, f"{what}: {len(diff)} cells differ, first at {diff[0]}"


def _problem(seed, dims=(8, 9, 12, 18), negative=True, time_elapsed=1800.0, species=inputs.SPECIES4,
             composition=inputs.COMPOSITION4):
    nr, nt, ne, npa = dims
    p = random_problem(nr=nr, nt=nt, ne=ne, npa=npa, species=species, composition=composition, seed=seed,
                       spiky=True)
    rng = np.random.default_rng(seed + 7)
    if negative:
        m = rng.uniform(0, 1, p.f2.shape)
        p.f2 = np.asfortranarray(np.where(m < 0.03, -p.f2, p.f2))
    necr = inputs.plasmasphere_density(p.grid, 4.0) * rng.uniform(0.5, 1.5, (nr, nt))
    p.coulomb = inputs.CoulombInputs(np.asfortranarray(necr), time_elapsed)
    return p


def _check_operators(p, species_list, what):
    """COULPARA tables, then COULEN and COULMU alone, with and without the
    clamp, each from the problem's F2, bit for bit against the oracle."""
    g = RamGpu.from_problem(p)
    for time in (p.coulomb.time_elapsed, 0.0):
        for op in ("coulen", "coulmu"):
            o = Oracle(p)
            o.set_coulomb(p.coulomb.necr, time)
            g.put_f2(p.f2)
            g.set_time(time)
            for s in species_list:
                o.call("coulpara", s)
                g.coulpara(s)
                tabs = g.coulomb_tables(s)
                for n in COULOMB_TABLES:
                    _same(tabs[n], o.coul_table(n)[s - 1], f"{what} {n}(S={s})")
                o.call(op, s)
                getattr(g, op)(s)
                _same(g.get_f2(), o.f2, f"{what} {op}(S={s}) time={time}")
            o.close()
        print(f"{what}: tables, COULEN, COULMU bit-exact (TimeRamElapsed={time})", flush=True)
    g.close()


@pytest.mark.parametrize("path", coulomb_fixtures(), ids=lambda p: p.stem)
def test_coulomb_golden_on_device(path):
    p, exp = load(path)
    g = RamGpu.from_problem(p)
    for tag, op, time in (("coulen", "coulen", p.coulomb.time_elapsed),
                          ("coulmu", "coulmu", p.coulomb.time_elapsed), ("coulmu_t0", "coulmu", 0.0)):
        g.put_f2(p.f2)
        g.set_time(time)
        for s in range(1, p.grid.ns + 1):
            g.coulpara(s)
            tabs = g.coulomb_tables(s)
            for n in COULOMB_TABLES:
                _same(tabs[n], exp[f"{n}_s{s}"], f"{n}(S={s})")
            getattr(g, op)(s)
            _same(g.get_f2()[s - 1], exp[f"f2_s{s}_{tag}"], f"{tag}(S={s})")
    print(f"{path.stem}: device tables and operators equal the reference's golden vectors", flush=True)
    g.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_coulomb_operators_bit_exact(seed):
    _check_operators(_problem(seed), [1, 2, 3, 4], f"random seed {seed}")


def test_coulomb_operators_bit_exact_at_40x96x64x90():
    p = _problem(3, dims=(40, 96, 64, 90), negative=True, species=(sp.ELECTRON, sp.OXYGEN),
                 composition=(1.0, 0.25))
    _check_operators(p, [1, 2], "40x96x64x90")


def _ram_run_compare(p, nsteps, what):
    g = RamGpu.from_problem(p)
    o = Oracle(p)
    g.set_time(0.0)
    for n in range(nsteps):
        o.set_coulomb(time_elapsed=2.0 * p.dts * n)
        o.ram_run(4)
        out = g.ram_run(1)
        e1, e2 = rel_err(g.get_f2(), o.f2), max_rel_err(g.get_f2(), o.f2)
        print(f"{what} step {n}: norm rel {e1:.3e} max rel {e2:.3e}", flush=True)
        assert e1 <= RTOL and e2 <= RTOL
        for s in range(p.grid.ns):
            scale = abs(o.arr("setrc")[s])
            for k in ("setrc", "lsdr", "lscha", "lsatm", "lswae", "lscoe", "lscsc"):
                assert abs(out[s][k] - o.arr(k)[s]) <= RTOL * scale, (what, n, s, k, out[s][k], o.arr(k)[s])
    assert g.time() == 2.0 * p.dts * nsteps
    assert all(out[s]["lscoe"] != 0 and out[s]["lscsc"] != 0 for s in range(p.grid.ns))
    o.close()
    g.close()


def test_ram_run_with_coulomb_random_geometry():
    _ram_run_compare(_problem(5, negative=False, time_elapsed=0.0), 3, "random 8x9x12x18")


def test_ram_run_with_coulomb_and_wavelo_multi_species():
    p = inputs.make_problem(20, 25, 35, 72, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=9, wave=True,
                            coulomb=True)
    _ram_run_compare(p, 3, "4 species 20x25x35x72")


def test_coulomb_is_refused_where_not_built():
    p = _problem(6, negative=False)
    g = RamGpu.from_problem(p)
    with pytest.raises(RamGpuError):
        g.ram_run_bwd(np.zeros(p.f2.shape, order="F"))
    with pytest.raises(RamGpuError):
        RamBatch([g])
    g.set_coulomb(on=False)  # and a context without them steps as before
    g.ram_run(1)
    g.close()
