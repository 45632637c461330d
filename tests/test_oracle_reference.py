"""The oracle's drift operators against the reference's own Fortran.

oracle/_ref/libramref.so is the reference's unmodified src/ModRamDrift.f90
(and src/ModRamFunctions.f90) compiled from its sources with amdflang and the
reference's precision flags (oracle/Makefile).  The C restatement must agree
with it bit for bit, operator by operator, on the dipole workload and on
randomised geometry that reaches every branch (sign changes along a line,
inflow/outflow edges, the stale outer ghost, clamps, the magnetopause mask).

These run in the build container only, where the reference is present; the
committed golden vectors (tests/golden/) carry the same checks elsewhere.
"""
import numpy as np
import pytest

from helpers import random_problem
from oracle.oracle import Oracle, Reference
from ramgpu import inputs, species as sp

pytestmark = pytest.mark.skipif(not Reference.available(),
                                reason="oracle/_ref not built (reference sources absent)")

SEQUENCE = ["driftpara", "driftr", "driftp", "drifte", "driftmu",
            "driftmu", "drifte", "driftp", "driftr"]


def _compare_sequence(problem, species_idx):
    o = Oracle(problem)
    r = Reference(problem)
    for s in species_idx:
        for op in SEQUENCE:
            o.call(op, s)
            r.call(op, s)
            a, b = o.f2, r.f2
            assert np.array_equal(a, b), f"{op}(S={s}) differs: max |d|={np.abs(a - b).max()}"
    np.testing.assert_array_equal(o.dtdrift(), r.dtdrift())


def test_plain_h_bit_exact():
    _compare_sequence(inputs.plain_h(seed=3), [1])


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_geometry_bit_exact(seed):
    p = random_problem(seed=seed, spiky=(seed % 2 == 1))
    _compare_sequence(p, [1, 2])


def test_multi_species_small_bit_exact():
    p = inputs.make_problem(nr=10, nt=13, ne=12, npa=16, species=inputs.SPECIES4,
                            composition=inputs.COMPOSITION4, kp=5.0, seed=7)
    _compare_sequence(p, [1, 2, 3, 4])


def test_stale_ghost_path_is_exercised():
    """The random geometry must contain outflow lines whose face NR-1 is
    upwinded from the ghost (the case the stale F(NR+1) decides)."""
    p = random_problem(seed=1)
    o = Oracle(p)
    o.call("driftpara", 1)
    o.call("driftr", 1)
    c = o.drift_table(1, "cdr")
    nr = p.grid.nr
    outflow = c[nr - 1] >= 0
    upwind_out = c[nr - 2] < 0
    assert np.any(outflow & upwind_out)
    assert np.any(~outflow)


def test_funt_funi_match_reference():
    for x in np.linspace(0.0, 1.0, 41):
        assert inputs.funt(float(x)) == Reference.funt(float(x))
        assert inputs.funi(float(x)) == Reference.funi(float(x))
