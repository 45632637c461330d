"""Load a golden fixture (tests/golden/*.npz) into a problem object that the
oracle wrapper and the GPU binding both accept."""
from pathlib import Path
from types import SimpleNamespace

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
COULOMB_TABLES = ("coule", "couli", "ata", "gta", "cedr", "cidr")
SEQUENCE = ["driftr", "driftp", "drifte", "driftmu", "driftmu", "drifte", "driftp", "driftr"]


def fixtures():
    return sorted(GOLDEN.glob("*.npz"))


def coulomb_fixtures():
    return sorted((GOLDEN / "coulomb").glob("*.npz"))


def load(path):
    z = np.load(path, allow_pickle=False)
    ns, nr, nt, ne, npa = (int(v) for v in z["dims"])
    dts, betalim, fraccfl, mdr, dphi, conf1, conf2 = (float(v) for v in z["scalars"])
    charge = z["charge"]
    cex = z["cex_kind"]
    species = tuple(SimpleNamespace(charge=int(charge[s]), cex=bool(cex[s]), cex_kind=int(cex[s]))
                    for s in range(ns))
    grid = SimpleNamespace(
        ns=ns, nr=nr, nt=nt, ne=ne, npa=npa, mdr=mdr, dphi=dphi, conf1=conf1, conf2=conf2,
        charge=charge, cex_kind=cex, species=species, rmas=z["rmas"], lz=z["lz"], rlz=z["rlz"],
        phi=np.array([j * dphi for j in range(nt)]),
        **{n: z[n] for n in ("ekev", "ebnd", "we", "de", "grel", "grbnd", "v", "mu", "wmu",
                             "dmu", "upa")})
    geom = SimpleNamespace(**{n: np.asfortranarray(z[n]) for n in (
        "fnhs", "fnis", "bounhs", "bounis", "hdns", "didt", "dibndt", "bnes", "dbdt", "eir",
        "eip", "outside_mgnp")})
    problem = SimpleNamespace(grid=grid, geom=geom, vt=np.asfortranarray(z["vt"]),
                              fgeos=np.asfortranarray(z["fgeos"]), f2=np.asfortranarray(z["f2"]),
                              dts=dts, betalim=betalim, fraccfl=fraccfl)
    if "vbnd" in z.files:  # a Coulomb fixture
        grid.vbnd = z["vbnd"]
        problem.coulomb = SimpleNamespace(necr=np.asfortranarray(z["necr"]),
                                          time_elapsed=float(z["time_elapsed"]))
    expected = {k: z[k] for k in z.files if k.startswith("f2_s") or k.startswith("dtdrift_s")
                or k.split("_s")[0] in COULOMB_TABLES}
    return problem, expected
