"""GPU parity of the implicit pitch-angle diffusions, WPADIF
(src/ModRamWPI.f90:643-714) and FLCscatter (src/ModRamLoss.f90:513-575),
through the C ABI, and of ram_run with every optional process on.

The line solve is built of correctly rounded operations in the reference's
order (quotients through div_rn and exact reciprocals), so the operators
equal the oracle bit for bit.  ram_run with WPADIF (electrons, at the WPI
positions), EMIC WPADIF (H+), FLC scattering (He+) and Coulomb collisions
runs one launch per operator for all species; it matches the oracle's ram_run to the stated
1e-12 (the loss factors' table-driven exponential), step after step, with
the LS* accumulators to 1e-12 of SETRC.  A context whose FLC never acts
(Dt_bc beyond the run) takes that composed step and must give the fused
step's F2 bit for bit.
"""
import numpy as np
import pytest

from conftest import gpu_available
from helpers import max_rel_err, rel_err
from oracle.oracle import Oracle
from ramgpu import inputs, species as sp
from ramgpu.capi import RamGpu, RamGpuError

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

RTOL = 1e-12


def _same(a, b, what):
    diff = np.argwhere(a != b)
    assert diff.size == 0, f"{what}: {len(diff)} cells differ, first at {diff[0]}"


def _problem(seed, dims=(8, 9, 12, 20)):
    nr, nt, ne, npa = dims
    p = inputs.make_problem(nr, nt, ne, npa, inputs.SPECIES4, inputs.COMPOSITION4, kp=4.0, seed=seed)
    rng = np.random.default_rng(seed + 50)
    p.grid.dmu = p.grid.dmu * rng.uniform(0.7, 1.3, npa)  # uneven widths
    p.grid.wmu = p.grid.wmu * rng.uniform(0.7, 1.3, npa)
    d1 = inputs.pitch_diffusion_coef(p.grid, p.geom, seed=seed + 1)
    d2 = inputs.pitch_diffusion_coef(p.grid, p.geom, scale=3e-6, seed=seed + 2)
    return p, d1, d2


@pytest.mark.parametrize("seed", [0, 1])
def test_wpadif_and_flcscatter_bit_exact(seed):
    p, d1, d2 = _problem(seed)
    g = RamGpu.from_problem(p)
    o = Oracle(p)
    for s in range(1, p.grid.ns + 1):
        g.set_wpadif(s, d1, d2, at_wpi=False, at_emic=False)
        g.wpadif(s)
        o.wpadif(s, d1, d2)
        _same(g.get_f2(), o.f2, f"WPADIF(S={s})")
    g.set_time(500.0)
    o.set_coulomb(time_elapsed=500.0, on=False)
    for s in (2, 3):
        g.set_flc(s, d2, dt_bc=1000.0)
        g.flcscatter(s)                      # TimeRamElapsed < Dt_bc: nothing
        _same(g.get_f2(), o.f2, f"FLCscatter(S={s}) before Dt_bc")
        g.set_flc(s, d2, dt_bc=300.0)
        g.flcscatter(s)
        o.flcscatter(s, d2, 300.0)
        _same(g.get_f2(), o.f2, f"FLCscatter(S={s})")
    print(f"seed {seed}: WPADIF and FLCscatter bit-exact", flush=True)
    o.close()
    g.close()


def test_wpadif_at_40x96x64x90():
    p = inputs.make_problem(40, 96, 64, 90, (sp.ELECTRON,), (1.0,), kp=4.0, seed=7)
    d1 = inputs.pitch_diffusion_coef(p.grid, p.geom, seed=8)
    d2 = inputs.pitch_diffusion_coef(p.grid, p.geom, scale=3e-6, seed=9)
    g = RamGpu.from_problem(p)
    o = Oracle(p)
    g.set_wpadif(1, d1, d2)
    g.wpadif(1)
    o.wpadif(1, d1, d2)
    _same(g.get_f2(), o.f2, "WPADIF 40x96x64x90")
    o.close()
    g.close()


def _all_processes(p, d1, d2, flc, dt_bc):
    """e- WPADIF in place of WAVELO, H+ EMIC, He+ FLC, Coulomb collisions."""
    p.wave = inputs.wave_inputs(p.grid, 4.0)
    necr = inputs.plasmasphere_density(p.grid, 4.0)
    p.coulomb = inputs.CoulombInputs(necr, 0.0)
    g = RamGpu.from_problem(p)
    o = Oracle(p)
    g.set_wpadif(1, d1, d2, at_wpi=True)
    o.set_wpadif(1, d1, d2, at_wpi=True)
    g.set_wpadif(2, d2, d1, at_wpi=False, at_emic=True)
    o.set_wpadif(2, d2, d1, at_wpi=False, at_emic=True)
    g.set_flc(3, flc, dt_bc)
    o.set_flc(3, flc, dt_bc)
    return g, o


def test_ram_run_with_every_optional_process():
    p, d1, d2 = _problem(3)
    flc = inputs.pitch_diffusion_coef(p.grid, p.geom, scale=2e-6, seed=9)
    g, o = _all_processes(p, d1, d2, flc, dt_bc=10.0)   # FLC from the second step on
    nsteps = 3
    for n in range(nsteps):
        o.set_coulomb(time_elapsed=2.0 * p.dts * n)
        o.ram_run(4)
        out = g.ram_run(1)
        e1, e2 = rel_err(g.get_f2(), o.f2), max_rel_err(g.get_f2(), o.f2)
        print(f"step {n}: norm rel {e1:.3e} max rel {e2:.3e}", flush=True)
        assert e1 <= RTOL and e2 <= RTOL
        dt = o.dtdrift()
        for s in range(p.grid.ns):
            scale = abs(o.arr("setrc")[s])
            for k in ("setrc", "lsdr", "lscha", "lsatm", "lswae", "lscoe", "lscsc"):
                assert abs(out[s][k] - o.arr(k)[s]) <= RTOL * scale, (n, s, k, out[s][k], o.arr(k)[s])
            for q, k in enumerate(("dtdrift_r", "dtdrift_p", "dtdrift_e", "dtdrift_mu")):
                assert abs(out[s][k] - dt[s, q]) <= RTOL * abs(dt[s, q]), (n, s, k)
    assert out[1]["lswae"] != 0   # H+ EMIC losses booked
    with pytest.raises(RamGpuError):
        g.ram_run_bwd(np.zeros(p.f2.shape, order="F"))
    o.close()
    g.close()


def test_composed_step_equals_fused_step():
    p, d1, d2 = _problem(4)
    necr = inputs.plasmasphere_density(p.grid, 4.0)
    p.coulomb = inputs.CoulombInputs(necr, 0.0)
    p.wave = inputs.wave_inputs(p.grid, 4.0)
    fused = RamGpu.from_problem(p)
    comp = RamGpu.from_problem(p)
    comp.set_flc(3, d1, dt_bc=1e9)    # takes the composed step; FLC never acts
    for n in range(2):
        a, b = fused.ram_run(1), comp.ram_run(1)
        _same(comp.get_f2(), fused.get_f2(), f"step {n}")
        for s in range(p.grid.ns):
            for k in a[s]:
                assert abs(a[s][k] - b[s][k]) <= RTOL * abs(a[s]["setrc"]), (n, s, k, a[s][k], b[s][k])
    fused.close()
    comp.close()
