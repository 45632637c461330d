"""The oracle against the golden vectors written by the reference's own
Fortran (tests/golden/make_golden.py): every drift operator, every species,
bit for bit, including DtDriftR/P/E/Mu."""
import numpy as np
import pytest

from golden_io import SEQUENCE, fixtures, load
from oracle.oracle import Oracle


@pytest.mark.parametrize("path", fixtures(), ids=lambda p: p.stem)
def test_oracle_matches_reference_vectors(path):
    problem, expected = load(path)
    o = Oracle(problem)
    for s in range(1, problem.grid.ns + 1):
        o.call("driftpara", s)
        for n, op in enumerate(SEQUENCE):
            o.call(op, s)
            ref = expected[f"f2_s{s}_{n}_{op}"]
            got = o.f2[s - 1]
            assert np.array_equal(got, ref), f"{path.stem} {op}(S={s}) step {n}"
        assert np.array_equal(o.dtdrift()[s - 1], expected[f"dtdrift_s{s}"])


def test_fixtures_present():
    assert len(fixtures()) >= 2
