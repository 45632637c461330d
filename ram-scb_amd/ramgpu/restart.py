"""Restart files of the split step in the reference's layout.

The reference writes ``restart.nc`` and ``restart_info.txt``
(src/ModRamRestart.f90:12-383) and reads them back in read_restart
(:386-600).  This module writes and reads the RAM part of that file: the
same variable names, dimension names and dimension orders (Fortran
``(nR,nT,nE,nPa)`` is stored as ``nPa, nE, nT, nR``, as netCDF lays out a
Fortran declaration), the same slices (the I = 2..NR+1 rows of the
geometry and field arrays), and the text file's fields.  The reference
creates a netCDF-4 (HDF5) file; HDF5 is not available here, so the file is
netCDF classic (CDF-1), which the netCDF library the reference reads with
opens through the same calls.  The SCB half of the file (x, y, z, Bx, ...,
kmax) belongs to the SCB solver and is not written.

Only host memory is touched here: a device context hands its state over with
``RamGpu.get_f2``, ``moments`` and the rest (see ``save``/``load_into``).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
from scipy.io import netcdf_file

# F2(1..4,:,:,:,:) are written as these (src/ModRamRestart.f90:311-314)
FLUX_NAMES = ("FluxE", "FluxH", "FluxHe", "FluxO")
# (NR+1, NT, NPA) arrays written from row I = 2 (src/ModRamRestart.f90:329-339)
GEOM3 = ("FNHS", "BOUNHS", "FNIS", "BOUNIS", "HDNS", "dIdt", "dIbndt")
GEOM2 = ("BNES", "EIR", "EIP", "dBdt")
FIELD2 = ("VTOL", "VTN", "VT")


def _fortran_var(f, name, dims, a):
    """A Fortran-declared (d1, d2, ..) array as netCDF variable with the
    dimensions reversed (the file's row-major order)."""
    v = f.createVariable(name, "d", tuple(reversed(dims)))
    v[...] = np.asarray(a, dtype=np.float64).T
    return v


def _fortran_get(f, name):
    return np.asfortranarray(np.array(f.variables[name][...], dtype=np.float64).T)


def write(path, *, f2, ekev, pa, vt, geometry, necr=None, tau=None, ppart=None, pperT=None,
          dts_next=0.0, told=0.0, vtol=None, vtn=None, time_elapsed=0.0, niter=0,
          start="20000101 000000"):
    """Write ``<path>/restart.nc`` and ``<path>/restart_info.txt``.

    f2 (NS<=4,NR,NT,NE,NPA); ekev (NE); pa (NPA) pitch angles [deg]; vt, vtol,
    vtn (NR+1,NT) (vtol/vtn default to vt: a field held constant); geometry
    a mapping with the GEOM3 arrays (NR+1,NT,NPA) and the GEOM2 arrays
    (NR+1,NT) by the reference's names; necr, tau (NR,NT); ppart, pperT
    (4,NR,NT) (PParT, PPerT); dts_next (DtNext), told (TOld).
    """
    out = Path(path)
    out.mkdir(parents=True, exist_ok=True)
    f2 = np.asarray(f2)
    ns, nr, nt, ne, npa = f2.shape
    if ns > len(FLUX_NAMES):
        raise ValueError(f"the restart file holds {len(FLUX_NAMES)} species, got {ns}")
    with (out / "restart_info.txt").open("w") as t:
        # the fields of src/ModRamRestart.f90:64-74
        t.write(" TIMING:\n")
        t.write(f"Start (YYYYMMDD HHMMSS)= {start}\n")
        t.write(f"Elapsed (seconds)      = {time_elapsed:15.4f}\n")
        t.write(f"Completed Iterations   = {niter:15d}\n")
        t.write(" GRID:\n")
        t.write(f"nR, nL, nE, nPA        = {nr:3d}{nt:3d}{ne:3d}{npa:3d}\n")
    f = netcdf_file(str(out / "restart.nc"), "w", version=1)
    try:
        for name, n in (("nR", nr), ("nT", nt), ("nE", ne), ("nPa", npa), ("nS", 4)):
            f.createDimension(name, n)
        _fortran_var(f, "EnergyGrid", ("nE",), ekev)
        _fortran_var(f, "PitchAngleGrid", ("nPa",), pa)
        for s in range(ns):
            _fortran_var(f, FLUX_NAMES[s], ("nR", "nT", "nE", "nPa"), f2[s])
        for name, a in (("PParT", ppart), ("PPerT", pperT)):
            if a is not None:
                _fortran_var(f, name, ("nS", "nR", "nT"), a)
        vt = np.asarray(vt)
        for name, a in (("VTOL", vtol if vtol is not None else vt), ("VTN", vtn if vtn is not None else vt),
                        ("VT", vt)):
            _fortran_var(f, name, ("nR", "nT"), np.asarray(a)[1:])
        for name in GEOM3:
            _fortran_var(f, name, ("nR", "nT", "nPa"), np.asarray(geometry[name])[1:])
        for name in GEOM2:
            _fortran_var(f, name, ("nR", "nT"), np.asarray(geometry[name])[1:])
        f.createVariable("DtNext", "d", ())[...] = dts_next
        f.createVariable("TOld", "d", ())[...] = told
        _fortran_var(f, "necr", ("nR", "nT"), necr if necr is not None else np.zeros((nr, nt)))
        _fortran_var(f, "tau", ("nR", "nT"), tau if tau is not None else np.zeros((nr, nt)))
    finally:
        f.close()


def read(path) -> dict:
    """The RAM part of a restart written by ``write`` (or by the reference,
    in netCDF classic form): arrays in the reference's Fortran shapes, the
    (NR+1)-row arrays with their I = 1 row zero (the file does not hold it),
    and the text file's fields."""
    p = Path(path)
    out = {}
    f = netcdf_file(str(p / "restart.nc"), "r", mmap=False)
    try:
        names = f.variables.keys()
        out["ekev"] = _fortran_get(f, "EnergyGrid")
        out["pa"] = _fortran_get(f, "PitchAngleGrid")
        fl = [_fortran_get(f, n) for n in FLUX_NAMES if n in names]
        out["f2"] = np.asfortranarray(np.stack(fl))
        for n in ("PParT", "PPerT"):
            if n in names:
                out[n] = _fortran_get(f, n)
        for n in FIELD2 + GEOM2 + GEOM3:
            a = _fortran_get(f, n)
            full = np.zeros((a.shape[0] + 1,) + a.shape[1:], order="F")
            full[1:] = a
            out[n] = full
        out["necr"] = _fortran_get(f, "necr")
        out["tau"] = _fortran_get(f, "tau")
        out["DtNext"] = float(f.variables["DtNext"].getValue())
        out["TOld"] = float(f.variables["TOld"].getValue())
    finally:
        f.close()
    for line in (p / "restart_info.txt").read_text().splitlines():
        if line.startswith("Elapsed"):
            out["time_elapsed"] = float(line.split("=")[1])
        elif line.startswith("Completed"):
            out["niter"] = int(line.split("=")[1])
        elif line.startswith("nR, nL"):
            v = line.split("= ", 1)[1]  # four I3 fields after the label's '= '
            out["dims"] = tuple(int(v[i:i + 3]) for i in range(0, 12, 3))
    return out


GEOM_KEYS = {"FNHS": "fnhs", "BOUNHS": "bounhs", "FNIS": "fnis", "BOUNIS": "bounis", "HDNS": "hdns",
             "dIdt": "didt", "dIbndt": "dibndt", "BNES": "bnes", "EIR": "eir", "EIP": "eip", "dBdt": "dbdt"}


def save(path, dev, problem, niter=0, dts_next=0.0, ppart=None, pperT=None):
    """Write the restart of a device context: its F2, TimeRamElapsed and TOld, the
    problem's grid, field, geometry and plasmasphere density, and the
    caller's DtNext and (NS,NR,NT) PPART/PPERT (from ``RamGpu.moments``,
    which the caller runs where ram_run's ANISCH block runs, as it rewrites
    F2's L=1 row), padded to the file's 4 species."""
    g, geo = problem.grid, problem.geom

    def four(a):
        if a is None:
            return None
        b = np.zeros((4, g.nr, g.nt), order="F")
        b[:g.ns] = a
        return b
    coul = getattr(problem, "coulomb", None)
    write(path, f2=dev.get_f2(), ekev=g.ekev, pa=g.pa, vt=problem.vt,
          geometry={k: getattr(geo, v) for k, v in GEOM_KEYS.items()},
          necr=coul.necr if coul is not None else None, ppart=four(ppart), pperT=four(pperT),
          dts_next=dts_next, told=dev.told(), time_elapsed=dev.time(), niter=niter)


def load_into(path, dev, problem):
    """Resume a device context from a restart: F2, the field and geometry
    rows I = 2..NR+1 (row 1 kept from the problem), NECR when Coulomb
    collisions are on, TimeRamElapsed and TOld."""
    r = read(path)
    geo = problem.geom
    g = {}
    for k, v in GEOM_KEYS.items():
        a = np.array(getattr(geo, v), order="F", copy=True)
        a[1:] = r[k][1:]
        g[v] = a
    for k, v in g.items():
        setattr(geo, k, v)
    dev.set_geometry(geo)
    vt = np.array(problem.vt, order="F", copy=True)
    vt[1:] = r["VT"][1:]
    dev.set_efield(vt)
    dev.put_f2(np.asfortranarray(r["f2"][:problem.grid.ns]))
    if getattr(problem, "coulomb", None) is not None:
        dev.set_coulomb(problem.grid.vbnd, r["necr"], None)
    dev.set_time(r["time_elapsed"])
    dev.set_told(r["TOld"])
    return r
