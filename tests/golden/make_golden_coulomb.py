"""Write the golden vectors of the Coulomb operators from the reference's own
Fortran (src/ModRamCoul.f90 in oracle/_ref/libramref.so, built from the
reference sources by oracle/Makefile).  Run in a container that holds the
reference sources:

    make -C oracle ref && python tests/golden/make_golden_coulomb.py

The fixture (tests/golden/coulomb/) stores the full input set of one small
two-species (e-, O+) problem with perturbed geometry, spikes, holes and a few
negative cells (so that both clamps act), VBND, NECR and, for each species:
the COULPARA tables COULE, COULI, ATA, GTA, CEDR, CIDR, and F2 after each of
COULEN, COULMU (TimeRamElapsed > 0) and COULMU with TimeRamElapsed = 0 (no
clamp), each applied alone to the initial F2.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path[:0] = [str(REPO), str(REPO / "ram-scb_amd"), str(REPO / "tests")]

from helpers import random_problem  # noqa: E402
from make_golden import problem_arrays  # noqa: E402
from oracle.oracle import COUL_TABLES, Reference  # noqa: E402
from ramgpu import inputs, species as sp  # noqa: E402

NAME = "coulomb_e_oplus_4x5x8x16_seed4"
TIME = 3600.0


def problem():
    """The fixture's problem: random geometry with spikes and holes, some
    cells negative, NECR of the synthetic plasmasphere at Kp=4."""
    p = random_problem(nr=4, nt=5, ne=8, npa=16, species=(sp.ELECTRON, sp.OXYGEN),
                       composition=(1.0, 0.25), seed=4, spiky=True)
    rng = np.random.default_rng(44)
    m = rng.uniform(0, 1, p.f2.shape)
    p.f2 = np.asfortranarray(np.where(m < 0.03, -p.f2, p.f2))
    necr = inputs.plasmasphere_density(p.grid, 4.0)
    necr = necr * rng.uniform(0.5, 1.5, necr.shape)
    p.coulomb = inputs.CoulombInputs(np.asfortranarray(necr), TIME)
    return p


def main():
    if not Reference.available():
        raise SystemExit("oracle/_ref/libramref.so is not built")
    p = problem()
    g = p.grid
    out = problem_arrays(p)
    out["vbnd"] = g.vbnd
    out["necr"] = p.coulomb.necr
    out["time_elapsed"] = np.array(TIME)
    for tag, op, time in (("coulen", "coulen", TIME), ("coulmu", "coulmu", TIME),
                          ("coulmu_t0", "coulmu", 0.0)):
        r = Reference(p)  # F2 back to the initial one
        r.setup_coulomb(g.vbnd, p.coulomb.necr, time)
        for s in range(1, g.ns + 1):
            r.call("coulpara", s)
            if tag == "coulen":
                tabs = r.coulomb_tables()
                for n in COUL_TABLES:
                    out[f"{n}_s{s}"] = np.asfortranarray(tabs[n][s - 1])
            r.call(op, s)
            out[f"f2_s{s}_{tag}"] = np.asfortranarray(r.f2[s - 1])
    (HERE / "coulomb").mkdir(exist_ok=True)
    np.savez_compressed(HERE / "coulomb" / f"{NAME}.npz", **out)
    print(NAME, sum(a.nbytes for a in out.values()))


if __name__ == "__main__":
    main()
