"""Write the golden vectors of the drift operators from the reference's own
Fortran (oracle/_ref/libramref.so, built from the reference sources by
oracle/Makefile).  Run in a container that holds the reference sources:

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture stores the full input set of one small problem and F2 plus
DtDrift* after every operator of the sequence DRIFTPARA, DRIFTR, DRIFTP,
DRIFTE, DRIFTMU, DRIFTMU, DRIFTE, DRIFTP, DRIFTR for each species.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path[:0] = [str(REPO), str(REPO / "ram-scb_amd"), str(REPO / "tests")]

from helpers import random_problem  # noqa: E402
from oracle.oracle import Reference  # noqa: E402
from ramgpu import inputs, species as sp  # noqa: E402

SEQUENCE = ["driftr", "driftp", "drifte", "driftmu", "driftmu", "drifte", "driftp", "driftr"]

CASES = {
    "random_hplus_oplus_6x7x6x8_seed2": lambda: random_problem(
        nr=6, nt=7, ne=6, npa=8, seed=2, spiky=True),
    "dipole_4species_5x6x5x6_seed3": lambda: inputs.make_problem(
        nr=5, nt=6, ne=5, npa=6, species=inputs.SPECIES4, composition=inputs.COMPOSITION4,
        kp=5.0, seed=3),
}


def problem_arrays(p):
    g, geo = p.grid, p.geom
    d = dict(dims=np.array([g.ns, g.nr, g.nt, g.ne, g.npa]),
             scalars=np.array([p.dts, p.betalim, p.fraccfl, g.mdr, g.dphi, g.conf1, g.conf2]),
             charge=g.charge, cex_kind=g.cex_kind, rmas=g.rmas, lz=g.lz, rlz=g.rlz,
             ekev=g.ekev, ebnd=g.ebnd, we=g.we, de=g.de, grel=g.grel, grbnd=g.grbnd, v=g.v,
             mu=g.mu, wmu=g.wmu, dmu=g.dmu, upa=g.upa, vt=p.vt, fgeos=p.fgeos, f2=p.f2)
    for n in ("fnhs", "fnis", "bounhs", "bounis", "hdns", "didt", "dibndt", "bnes", "dbdt",
              "eir", "eip", "outside_mgnp"):
        d[n] = getattr(geo, n)
    return d


def main():
    if not Reference.available():
        raise SystemExit("oracle/_ref/libramref.so is not built")
    for name, make in CASES.items():
        p = make()
        r = Reference(p)
        out = problem_arrays(p)
        for s in range(1, p.grid.ns + 1):
            r.call("driftpara", s)
            for n, op in enumerate(SEQUENCE):
                r.call(op, s)
                out[f"f2_s{s}_{n}_{op}"] = np.asfortranarray(r.f2[s - 1])
            out[f"dtdrift_s{s}"] = r.dtdrift()[s - 1]
        np.savez_compressed(HERE / f"{name}.npz", **out)
        print(name, sum(a.nbytes for a in out.values()))


if __name__ == "__main__":
    main()
