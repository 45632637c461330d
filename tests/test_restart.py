"""Restart files in the reference's layout (src/ModRamRestart.f90:12-600):
the variable and dimension names, the reversed dimension order netCDF gives
a Fortran declaration, the I = 2..NR+1 slices, the text file's fields, and a
bit-exact round trip of the state through the file."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from ramgpu import inputs, restart


def _problem():
    return inputs.make_problem(6, 7, 8, 10, inputs.SPECIES4, inputs.COMPOSITION4, kp=3.0, seed=2, coulomb=True)


def test_restart_layout_and_round_trip(tmp_path):
    p = _problem()
    g, geo = p.grid, p.geom
    geom = {k: getattr(geo, v) for k, v in restart.GEOM_KEYS.items()}
    rng = np.random.default_rng(3)
    ppar = rng.uniform(size=(4, g.nr, g.nt))
    restart.write(tmp_path, f2=p.f2, ekev=g.ekev, pa=g.pa, vt=p.vt, geometry=geom, necr=p.coulomb.necr,
                  ppart=ppar, pperT=2 * ppar, dts_next=3.25, time_elapsed=1234.5, niter=17)
    f = netcdf_file(str(tmp_path / "restart.nc"), "r", mmap=False)
    try:
        v = f.variables
        # Fortran (nR,nT,nE,nPa) is stored nPa, nE, nT, nR; F2(2,...) is FluxH
        assert v["FluxH"].dimensions == ("nPa", "nE", "nT", "nR")
        assert v["FNHS"].dimensions == ("nPa", "nT", "nR")
        assert v["PParT"].dimensions == ("nT", "nR", "nS")
        i, j, k, l = 3, 5, 6, 7
        assert v["FluxH"][l, k, j, i] == p.f2[1, i, j, k, l]
        assert v["FluxO"][0, 1, 2, 3] == p.f2[3, 3, 2, 1, 0]
        # geometry and field rows I = 2..NR+1 (src/ModRamRestart.f90:324-339)
        assert v["FNHS"][l, j, i] == geo.fnhs[i + 1, j, l]
        assert v["VT"][j, i] == p.vt[i + 1, j]
        assert v["necr"][j, i] == p.coulomb.necr[i, j]
        assert float(v["DtNext"].getValue()) == 3.25
        assert set(restart.FLUX_NAMES) <= set(v)
    finally:
        f.close()
    info = (tmp_path / "restart_info.txt").read_text().splitlines()
    assert info[0] == " TIMING:" and info[4] == " GRID:"
    assert info[2] == "Elapsed (seconds)      =       1234.5000"
    assert info[5] == f"nR, nL, nE, nPA        = {g.nr:3d}{g.nt:3d}{g.ne:3d}{g.npa:3d}"
    r = restart.read(tmp_path)
    assert np.array_equal(r["f2"], p.f2)
    for k2, v2 in restart.GEOM_KEYS.items():
        assert np.array_equal(r[k2][1:], getattr(geo, v2)[1:]), k2
    assert np.array_equal(r["VT"][1:], p.vt[1:]) and np.array_equal(r["necr"], p.coulomb.necr)
    assert np.array_equal(r["PParT"], ppar) and np.array_equal(r["PPerT"], 2 * ppar)
    assert r["time_elapsed"] == 1234.5 and r["niter"] == 17 and r["dims"] == (g.nr, g.nt, g.ne, g.npa)


def test_restart_holds_four_species(tmp_path):
    p = _problem()
    f2 = np.concatenate([p.f2, p.f2[:1]])
    with pytest.raises(ValueError):
        restart.write(tmp_path, f2=f2, ekev=p.grid.ekev, pa=p.grid.pa, vt=p.vt,
                      geometry={k: getattr(p.geom, v) for k, v in restart.GEOM_KEYS.items()})
