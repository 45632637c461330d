"""ctypes binding of the C ABI in include/ramgpu.h (libramgpu.so).

``RamGpu`` mirrors the reference's operator interface: one method per
reference subroutine (``driftpara``, ``driftr``, ``driftp``, ``drifte``,
``driftmu``, ``cepara``, ``charexchange``, ``atmol``, ``sumrc``) taking the
1-based species index ``s``, plus the fused ``ram_run``.  Host arrays are
passed in the reference's Fortran layouts.  There is no CPU fallback: if the
HIP library is not built, or no GPU is present, construction raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libramgpu.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

EXPORTS = (
    "ramgpu_last_error", "ramgpu_create", "ramgpu_destroy", "ramgpu_set_species",
    "ramgpu_set_params", "ramgpu_set_grid", "ramgpu_set_geometry", "ramgpu_set_efield",
    "ramgpu_set_kp", "ramgpu_set_boundary", "ramgpu_set_boundary_scale", "ramgpu_put_f2",
    "ramgpu_get_f2", "ramgpu_set_energy_state", "ramgpu_get_energy_state", "ramgpu_driftpara",
    "ramgpu_driftr", "ramgpu_driftp", "ramgpu_drifte", "ramgpu_driftmu", "ramgpu_driftend",
    "ramgpu_cepara", "ramgpu_charexchange", "ramgpu_atmol", "ramgpu_sumrc", "ramgpu_ram_run",
    "ramgpu_ram_run_schedule", "ramgpu_last_run_ms", "ramgpu_kernel_timing",
    "ramgpu_kernel_times", "ramgpu_copy_probe",
)

KERNEL_KINDS = ("prepare", "ij", "kl", "finalize", "vs")


class SpeciesOut(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("setrc", "lsdr", "lscha", "lsatm", "dtdrift_r",
                                          "dtdrift_p", "dtdrift_e", "dtdrift_mu")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class RamGpuError(RuntimeError):
    pass


_lib = None


def load_library(path: Path | None = None):
    """Load libramgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RamGpuError(f"{p} is not built; run __graft_entry__.build() "
                          "or python -m ramgpu.build")
    lib = C.CDLL(str(p))
    lib.ramgpu_last_error.restype = C.c_char_p
    pp = C.POINTER(C.c_void_p)
    v = C.c_void_p
    lib.ramgpu_create.argtypes = [pp] + [C.c_int] * 6
    lib.ramgpu_destroy.argtypes = [v]
    lib.ramgpu_set_species.argtypes = [v, C.c_int, C.c_int, C.c_int, C.c_double]
    lib.ramgpu_set_params.argtypes = [v, C.c_double, C.c_double, C.c_double]
    lib.ramgpu_set_grid.argtypes = [v] + [C.c_double] * 4 + [_dp] * 12
    lib.ramgpu_set_geometry.argtypes = [v] + [_dp] * 11 + [_ip]
    lib.ramgpu_set_efield.argtypes = [v, _dp]
    lib.ramgpu_set_kp.argtypes = [v, C.c_double, _dp, _dp]
    lib.ramgpu_set_boundary.argtypes = [v, _dp]
    lib.ramgpu_set_boundary_scale.argtypes = [v, _dp]
    lib.ramgpu_put_f2.argtypes = [v, _dp]
    lib.ramgpu_get_f2.argtypes = [v, _dp]
    lib.ramgpu_set_energy_state.argtypes = [v, _dp, _dp, _dp, _dp]
    lib.ramgpu_get_energy_state.argtypes = [v, _dp, _dp, _dp, _dp]
    for n in ("driftpara", "cepara", "charexchange", "atmol"):
        getattr(lib, "ramgpu_" + n).argtypes = [v, C.c_int]
    for n in ("driftr", "driftp", "drifte", "driftmu"):
        getattr(lib, "ramgpu_" + n).argtypes = [v, C.c_int, _dp]
    lib.ramgpu_driftend.argtypes = [v]
    lib.ramgpu_sumrc.argtypes = [v, C.c_int, _dp, _dp]
    lib.ramgpu_ram_run.argtypes = [v, C.c_int, C.POINTER(SpeciesOut)]
    lib.ramgpu_ram_run_schedule.argtypes = [v, C.c_int, _dp, _dp, C.POINTER(SpeciesOut)]
    lib.ramgpu_last_run_ms.argtypes = [v, _dp]
    lib.ramgpu_kernel_timing.argtypes = [v, C.c_int]
    lib.ramgpu_kernel_times.argtypes = [v, _dp, C.POINTER(C.c_int)]
    lib.ramgpu_copy_probe.argtypes = [C.c_int, C.c_size_t, C.c_int, _dp]
    for n in EXPORTS:
        if n != "ramgpu_last_error":
            getattr(lib, n).restype = C.c_int
    _lib = lib
    return lib


def _fd(a, n=None):
    a = np.asfortranarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} elements, got {a.size}")
    return a


def _p(a):
    return a.ctypes.data_as(_dp)


def copy_probe(nbytes: int = 1 << 30, reps: int = 20, device: int = 0) -> float:
    """Read+write GB/s of a plain device copy of nbytes (measured HBM peak)."""
    lib = load_library()
    out = C.c_double(0.0)
    rc = lib.ramgpu_copy_probe(device, nbytes, reps, C.byref(out))
    if rc:
        raise RamGpuError(lib.ramgpu_last_error().decode())
    return out.value


class RamGpu:
    """One device-resident RAM problem (a ``ramgpu_ctx``)."""

    def __init__(self, ns, nr, nt, ne, npa, device=0):
        self.lib = load_library()
        self.dims = (ns, nr, nt, ne, npa)
        self.ns, self.nr, self.nt, self.ne, self.npa = self.dims
        h = C.c_void_p()
        self._check(self.lib.ramgpu_create(C.byref(h), ns, nr, nt, ne, npa, device))
        self.h = h

    # -- plumbing -------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = self.lib.ramgpu_last_error().decode()
            raise RamGpuError(f"ramgpu status {rc}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ramgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up -----------------------------------------------------------
    @classmethod
    def from_problem(cls, problem, device=0):
        g = problem.grid
        self = cls(g.ns, g.nr, g.nt, g.ne, g.npa, device)
        for s, spc in enumerate(g.species, start=1):
            self.set_species(s, spc.charge, spc.cex_kind if spc.cex else 0, g.rmas[s - 1])
        self.set_params(problem.dts, problem.betalim, problem.fraccfl)
        self.set_grid(g)
        self.set_geometry(problem.geom)
        self.set_efield(problem.vt)
        self.set_boundary(problem.fgeos)
        self.put_f2(problem.f2)
        return self

    def set_species(self, s, charge, cex_kind, rmas):
        self._check(self.lib.ramgpu_set_species(self.h, s, int(charge), int(cex_kind), float(rmas)))

    def set_params(self, dts, betalim=1.5, fraccfl=0.8):
        self._check(self.lib.ramgpu_set_params(self.h, dts, betalim, fraccfl))

    def set_grid(self, g):
        ns, nr, nt, ne, npa = self.dims
        arrs = [_fd(g.rlz, nr + 1), _fd(g.ekev, ne), _fd(g.ebnd, ne), _fd(g.we, ne), _fd(g.de, ne),
                _fd(g.grel, ns * ne), _fd(g.grbnd, ns * ne), _fd(g.v, ns * ne), _fd(g.mu, npa),
                _fd(g.wmu, npa), _fd(g.dmu, npa), _fd(g.upa, nr)]
        self._keep_grid = arrs
        self._check(self.lib.ramgpu_set_grid(self.h, g.mdr, g.dphi, g.conf1, g.conf2,
                                             *[_p(a) for a in arrs]))
        self._lz = _fd(g.lz, nr + 1)
        self._phi = _fd(g.phi, nt)

    def set_geometry(self, geo):
        ns, nr, nt, ne, npa = self.dims
        n3, n2 = (nr + 1) * nt * npa, (nr + 1) * nt
        arrs = [_fd(geo.fnhs, n3), _fd(geo.fnis, n3), _fd(geo.bounhs, n3), _fd(geo.bounis, n3),
                _fd(geo.hdns, n3), _fd(geo.didt, n3), _fd(geo.dibndt, n3), _fd(geo.bnes, n2),
                _fd(geo.dbdt, n2), _fd(geo.eir, n2), _fd(geo.eip, n2)]
        mg = np.asfortranarray(geo.outside_mgnp, dtype=np.int32)
        if mg.size != nr * nt:
            raise ValueError("outside_mgnp must be (NR,NT)")
        self._check(self.lib.ramgpu_set_geometry(self.h, *[_p(a) for a in arrs],
                                                 mg.ctypes.data_as(_ip)))

    def set_efield(self, vt):
        a = _fd(vt, (self.nr + 1) * self.nt)
        self._check(self.lib.ramgpu_set_efield(self.h, _p(a)))

    def set_kp(self, kp):
        self._check(self.lib.ramgpu_set_kp(self.h, float(kp), _p(self._lz), _p(self._phi)))

    def set_boundary(self, fgeos):
        ns, nr, nt, ne, npa = self.dims
        a = _fd(fgeos, ns * nt * ne * npa)
        self._check(self.lib.ramgpu_set_boundary(self.h, _p(a)))

    def set_boundary_scale(self, scale):
        a = _fd(np.broadcast_to(np.asarray(scale, dtype=np.float64), (self.ns,)), self.ns)
        self._check(self.lib.ramgpu_set_boundary_scale(self.h, _p(a)))

    def put_f2(self, f2):
        a = _fd(f2, int(np.prod(self.dims)))
        self._check(self.lib.ramgpu_put_f2(self.h, _p(a)))

    def get_f2(self) -> np.ndarray:
        out = np.zeros(self.dims, order="F")
        self._check(self.lib.ramgpu_get_f2(self.h, _p(out)))
        return out

    def energy_state(self):
        arrs = [np.zeros(self.ns) for _ in range(4)]
        self._check(self.lib.ramgpu_get_energy_state(self.h, *[_p(a) for a in arrs]))
        return dict(zip(("setrc", "lsdr", "lscha", "lsatm"), arrs))

    def set_energy_state(self, setrc=None, lsdr=None, lscha=None, lsatm=None):
        args = []
        keep = []
        for v in (setrc, lsdr, lscha, lsatm):
            if v is None:
                args.append(None)
            else:
                a = _fd(v, self.ns)
                keep.append(a)
                args.append(_p(a))
        self._check(self.lib.ramgpu_set_energy_state(self.h, *args))

    # -- operators (reference names, 1-based species) -------------------
    def driftpara(self, s):
        self._check(self.lib.ramgpu_driftpara(self.h, s))

    def _drift(self, name, s):
        dt = C.c_double(0.0)
        self._check(getattr(self.lib, "ramgpu_" + name)(self.h, s, C.byref(dt)))
        return dt.value

    def driftr(self, s):
        return self._drift("driftr", s)

    def driftp(self, s):
        return self._drift("driftp", s)

    def drifte(self, s):
        return self._drift("drifte", s)

    def driftmu(self, s):
        return self._drift("driftmu", s)

    def driftend(self):
        self._check(self.lib.ramgpu_driftend(self.h))

    def cepara(self, s):
        self._check(self.lib.ramgpu_cepara(self.h, s))

    def charexchange(self, s):
        self._check(self.lib.ramgpu_charexchange(self.h, s))

    def atmol(self, s):
        self._check(self.lib.ramgpu_atmol(self.h, s))

    def sumrc(self, s):
        setrc, elorc = C.c_double(0.0), C.c_double(0.0)
        self._check(self.lib.ramgpu_sumrc(self.h, s, C.byref(setrc), C.byref(elorc)))
        return setrc.value, elorc.value

    def ram_run(self, nsteps=1):
        out = (SpeciesOut * self.ns)()
        self._check(self.lib.ramgpu_ram_run(self.h, nsteps, out))
        return [o.as_dict() for o in out]

    def ram_run_schedule(self, kp, bscale=None):
        kp = _fd(kp)
        n = kp.size
        bs = None if bscale is None else _fd(bscale, n * self.ns)
        out = (SpeciesOut * self.ns)()
        self._check(self.lib.ramgpu_ram_run_schedule(self.h, n, _p(kp),
                                                     None if bs is None else _p(bs), out))
        return [o.as_dict() for o in out]

    def kernel_timing(self, enable: bool = True):
        self._check(self.lib.ramgpu_kernel_timing(self.h, int(bool(enable))))

    def kernel_times(self) -> dict:
        """{kind: (total_ms, launches)} since the last call (then reset)."""
        ms = (C.c_double * len(KERNEL_KINDS))()
        n = (C.c_int * len(KERNEL_KINDS))()
        self._check(self.lib.ramgpu_kernel_times(self.h, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNEL_KINDS)}

    def last_run_ms(self) -> float:
        ms = C.c_double(0.0)
        self._check(self.lib.ramgpu_last_run_ms(self.h, C.byref(ms)))
        return ms.value
