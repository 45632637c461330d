"""GPU parity of the moments (ANISCH, FLUX) through the C ABI.

Each device moment is the reference's sum in the reference's order (the
thread of each energy adds its pitch-angle terms row by row; one thread adds
the energies in order, window by window), so PPERT, PPART, EDENT, RNHTT, FLUX and the F2 row
ANISCH rewrites are compared with the oracle bit for bit on the same F2.
"""
import numpy as np
import pytest

from conftest import gpu_available
from helpers import random_problem
from oracle.oracle import Oracle
from ramgpu import inputs
from ramgpu.capi import RamGpu

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="no GPU")]

NAMES = ("ppert", "ppart", "edent", "rnhtt")


def _same(a, b, what):
    nd = int(np.count_nonzero(a != b))
    print(f"{what}: {nd} entries differ", flush=True)
    assert nd == 0, f"{what}: {nd} entries differ"


def _check_anisch(p, what):
    o = Oracle(p)
    g = RamGpu.from_problem(p)
    o_out = {n: np.zeros((p.grid.ns, p.grid.nr, p.grid.nt), order="F") for n in NAMES}
    g_out = {n: np.zeros_like(o_out[n]) for n in NAMES}
    for s in range(1, p.grid.ns + 1):
        o.anisch(s, o_out)
        g.anisch(s, g_out)
    for n in NAMES:
        _same(g_out[n], o_out[n], f"{what} {n}")
    _same(g.get_f2(), o.f2, f"{what} F2 after ANISCH")
    g.close()


def test_anisch_plain_h():
    _check_anisch(inputs.plain_h(seed=2), "H+ 20x25x35x72")


def test_anisch_ne_below_30_and_random_geometry():
    _check_anisch(random_problem(seed=4, spiky=True), "random 8x9x7x10")


def test_anisch_wide_energy_grids():
    # NE = 72: five staged passes of 7 rows; NE = 300: one row per pass and
    # more energies than threads in a block
    p = inputs.make_problem(nr=6, nt=5, ne=72, npa=40, species=inputs.SPECIES4[1:3],
                            composition=(0.7, 0.05), seed=7)
    _check_anisch(p, "H+,He+ 6x5x72x40")
    p = inputs.make_problem(nr=5, nt=4, ne=300, npa=8, species=inputs.SPECIES4[1:2], composition=(0.7,),
                            seed=8)
    _check_anisch(p, "H+ 5x4x300x8")


def test_flux_matches_oracle():
    p = inputs.multi_species(seed=3)
    o = Oracle(p)
    g = RamGpu.from_problem(p)
    fx = g.flux()
    _same(fx, o.flux(), "FLUX e-,H+,He+,O+ 20x25x35x72")
    assert np.all(fx[:, 0] == 0) and np.all(fx[:, :, -1] == 0)
    assert np.all(fx[:, :, :, 0] == 0) and np.all(fx[:, :, :, :, 0] == 0)
    g.close()


def test_moments_after_steps():
    # the block after ram_run's species loop, on the F2 two device steps made
    p = inputs.multi_species(seed=9)
    g = RamGpu.from_problem(p)
    g.ram_run(2)
    f2 = g.get_f2()
    out, fx = g.moments(flux=True)
    o = Oracle(p)
    o.set_f2(f2)
    o_out = {n: np.zeros_like(out[n]) for n in NAMES}
    for s in range(1, p.grid.ns + 1):
        o.anisch(s, o_out)
    for n in NAMES:
        _same(out[n], o_out[n], f"after 2 steps {n}")
    _same(fx, o.flux(), "after 2 steps FLUX")
    _same(g.get_f2(), o.f2, "after 2 steps F2")
    # without FLUX only the moments leave the device; they are the same
    out2, fx2 = g.moments(flux=False)
    assert fx2 is None
    for n in NAMES:
        _same(out2[n], out[n], f"moments without FLUX {n}")
    g.close()
