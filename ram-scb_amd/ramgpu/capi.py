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
    "ramgpu_batch_create", "ramgpu_batch_destroy", "ramgpu_batch_ram_run",
    "ramgpu_batch_ram_run_schedule", "ramgpu_batch_last_run_ms", "ramgpu_batch_use_graph",
    "ramgpu_batch_autotune", "ramgpu_batch_get_shapes", "ramgpu_batch_set_shapes",
    "ramgpu_driftr_bwd", "ramgpu_driftp_bwd", "ramgpu_drifte_bwd", "ramgpu_driftmu_bwd",
    "ramgpu_charexchange_bwd", "ramgpu_atmol_bwd", "ramgpu_ram_run_bwd", "ramgpu_adjoint_run",
    "ramgpu_batch_adjoint_run", "ramgpu_set_plane_shape", "ramgpu_get_plane_shape",
    "ramgpu_set_moment_factors", "ramgpu_anisch", "ramgpu_flux", "ramgpu_moments",
    "ramgpu_ram_run_schedule_noalloc", "ramgpu_set_loss_state", "ramgpu_get_loss_state",
    "ramgpu_set_wavelo", "ramgpu_set_wave_params", "ramgpu_wavelo",
    "ramgpu_batch_kernel_timing", "ramgpu_batch_kernel_times",
    "ramgpu_set_coulomb", "ramgpu_set_necr", "ramgpu_set_time", "ramgpu_get_time", "ramgpu_coulpara",
    "ramgpu_get_coulomb_tables", "ramgpu_coulen", "ramgpu_coulmu",
    "ramgpu_set_wpadif", "ramgpu_wpadif", "ramgpu_set_flc", "ramgpu_flcscatter",
    "ramgpu_get_geometry", "ramgpu_computehI", "ramgpu_set_told", "ramgpu_get_told",
)

KERNEL_KINDS = ("prepare", "r", "plane", "finalize", "vs", "p", "r_bwd", "p_bwd", "plane_bwd",
                "moments", "pitch")


class SpeciesOut(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("setrc", "lsdr", "lscha", "lsatm", "dtdrift_r",
                                          "dtdrift_p", "dtdrift_e", "dtdrift_mu", "lswae",
                                          "lscoe", "lscsc")]

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
    lib.ramgpu_get_geometry.argtypes = [v] + [_dp] * 11 + [_ip]
    lib.ramgpu_set_told.argtypes = [v, C.c_double]
    lib.ramgpu_get_told.argtypes = [v, _dp]
    lib.ramgpu_computehI.argtypes = [v, C.c_int, C.c_int] + [_dp] * 8 + [_ip, _ip, C.c_double, C.c_int]
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
    lib.ramgpu_set_moment_factors.argtypes = [v, _dp, _dp, _dp]
    lib.ramgpu_anisch.argtypes = [v, C.c_int, _dp, _dp, _dp, _dp]
    lib.ramgpu_flux.argtypes = [v, _dp]
    lib.ramgpu_moments.argtypes = [v, _dp, _dp, _dp, _dp, _dp]
    lib.ramgpu_ram_run.argtypes = [v, C.c_int, C.POINTER(SpeciesOut)]
    lib.ramgpu_ram_run_schedule.argtypes = [v, C.c_int, _dp, _dp, C.POINTER(SpeciesOut)]
    lib.ramgpu_ram_run_schedule_noalloc.argtypes = [v, C.c_int, _dp, _dp, C.POINTER(SpeciesOut)]
    lib.ramgpu_set_loss_state.argtypes = [v, _dp, _dp, _dp]
    lib.ramgpu_get_loss_state.argtypes = [v, _dp, _dp, _dp]
    lib.ramgpu_set_wavelo.argtypes = [v, C.c_int, _dp, _dp, _dp]
    lib.ramgpu_set_wave_params.argtypes = [v, C.c_double, C.c_double, _dp, _dp, C.c_int]
    lib.ramgpu_wavelo.argtypes = [v, C.c_int]
    lib.ramgpu_set_coulomb.argtypes = [v, C.c_int, _dp]
    lib.ramgpu_set_necr.argtypes = [v, _dp]
    lib.ramgpu_set_time.argtypes = [v, C.c_double]
    lib.ramgpu_get_time.argtypes = [v, _dp]
    for n in ("coulpara", "coulen", "coulmu"):
        getattr(lib, "ramgpu_" + n).argtypes = [v, C.c_int]
    lib.ramgpu_get_coulomb_tables.argtypes = [v, C.c_int] + [_dp] * 6
    lib.ramgpu_set_wpadif.argtypes = [v, C.c_int, _dp, _dp, C.c_int, C.c_int]
    lib.ramgpu_wpadif.argtypes = [v, C.c_int]
    lib.ramgpu_set_flc.argtypes = [v, C.c_int, _dp, C.c_double]
    lib.ramgpu_flcscatter.argtypes = [v, C.c_int]
    lib.ramgpu_batch_kernel_timing.argtypes = [v, C.c_int]
    lib.ramgpu_batch_kernel_times.argtypes = [v, _dp, C.POINTER(C.c_int)]
    lib.ramgpu_last_run_ms.argtypes = [v, _dp]
    lib.ramgpu_kernel_timing.argtypes = [v, C.c_int]
    lib.ramgpu_kernel_times.argtypes = [v, _dp, C.POINTER(C.c_int)]
    lib.ramgpu_copy_probe.argtypes = [C.c_int, C.c_size_t, C.c_int, _dp]
    lib.ramgpu_batch_create.argtypes = [pp, C.POINTER(C.c_void_p), C.c_int]
    lib.ramgpu_batch_destroy.argtypes = [v]
    lib.ramgpu_batch_ram_run.argtypes = [v, C.c_int, C.POINTER(SpeciesOut)]
    lib.ramgpu_batch_ram_run_schedule.argtypes = [v, C.c_int, _dp, _dp, C.POINTER(SpeciesOut)]
    lib.ramgpu_batch_last_run_ms.argtypes = [v, _dp]
    lib.ramgpu_batch_use_graph.argtypes = [v, C.c_int]
    lib.ramgpu_batch_autotune.argtypes = [v, C.c_int, _dp]
    lib.ramgpu_batch_get_shapes.argtypes = [v, _ip]
    lib.ramgpu_batch_set_shapes.argtypes = [v, _ip]
    lib.ramgpu_driftr_bwd.argtypes = [v, C.c_int, _dp, _dp]
    for n in ("driftp", "drifte", "driftmu", "charexchange", "atmol"):
        getattr(lib, f"ramgpu_{n}_bwd").argtypes = [v, C.c_int, _dp]
    lib.ramgpu_ram_run_bwd.argtypes = [v, _dp, _dp]
    lib.ramgpu_adjoint_run.argtypes = [v, C.c_int, C.c_int, _dp, _dp, C.c_int, _ip, _dp, _dp, _dp,
                                       _dp, _dp]
    lib.ramgpu_set_plane_shape.argtypes = [v, C.c_int, C.c_int]
    lib.ramgpu_get_plane_shape.argtypes = [v, _ip, _ip]
    _pp = C.POINTER(C.c_void_p)
    lib.ramgpu_batch_adjoint_run.argtypes = [v, C.c_int, C.c_int, _dp, _dp, _ip, _pp, _pp, _pp, _dp, _pp, _pp]
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
        if getattr(g, "ernh", None) is not None:
            self.set_moment_factors(g)
        if getattr(problem, "wave", None) is not None:
            self.set_wave(problem.wave)
        if getattr(problem, "coulomb", None) is not None:
            self.set_coulomb(g.vbnd, problem.coulomb.necr, problem.coulomb.time_elapsed)
        return self

    def set_coulomb(self, vbnd=None, necr=None, time_elapsed=None, on=True):
        """Coulomb collisions (DoUseCoulomb) with VBND (NS,NE), the
        plasmasphere density NECR (NR,NT) and TimeRamElapsed."""
        if on:
            self._check(self.lib.ramgpu_set_coulomb(self.h, 1, _p(_fd(vbnd, self.ns * self.ne))))
        else:
            self._check(self.lib.ramgpu_set_coulomb(self.h, 0, None))
        if necr is not None:
            self._check(self.lib.ramgpu_set_necr(self.h, _p(_fd(necr, self.nr * self.nt))))
        if time_elapsed is not None:
            self.set_time(time_elapsed)

    def set_time(self, time_elapsed):
        self._check(self.lib.ramgpu_set_time(self.h, float(time_elapsed)))

    def time(self) -> float:
        t = C.c_double(0.0)
        self._check(self.lib.ramgpu_get_time(self.h, C.byref(t)))
        return t.value

    def coulpara(self, s):
        self._check(self.lib.ramgpu_coulpara(self.h, s))

    def coulomb_tables(self, s) -> dict:
        """COULE, COULI, ATA, GTA, CEDR, CIDR (NE,NPA) of species s."""
        names = ("coule", "couli", "ata", "gta", "cedr", "cidr")
        out = {n: np.zeros((self.ne, self.npa), order="F") for n in names}
        self._check(self.lib.ramgpu_get_coulomb_tables(self.h, s, *[_p(out[n]) for n in names]))
        return out

    def coulen(self, s):
        self._check(self.lib.ramgpu_coulen(self.h, s))

    def set_wpadif(self, s, d1, d2=None, at_wpi=True, at_emic=False):
        """WPADIF coefficients of species s, D = d1 + d2 (NR,NT,NE,NPA);
        d1 None switches WPADIF off for the species."""
        n = self.nr * self.nt * self.ne * self.npa
        a = _fd(d1, n) if d1 is not None else None
        b = _fd(d2, n) if d2 is not None else None
        self._check(self.lib.ramgpu_set_wpadif(self.h, s, _p(a) if a is not None else None,
                                               _p(b) if b is not None else None, int(bool(at_wpi)),
                                               int(bool(at_emic))))

    def wpadif(self, s):
        self._check(self.lib.ramgpu_wpadif(self.h, s))

    def set_flc(self, s, coef, dt_bc=0.0):
        a = _fd(coef, self.nr * self.nt * self.ne * self.npa) if coef is not None else None
        self._check(self.lib.ramgpu_set_flc(self.h, s, _p(a) if a is not None else None, float(dt_bc)))

    def flcscatter(self, s):
        self._check(self.lib.ramgpu_flcscatter(self.h, s))

    def coulmu(self, s):
        self._check(self.lib.ramgpu_coulmu(self.h, s))

    def set_wave(self, wave):
        """WAVELO for the species flagged in wave.wpi (inputs.WaveInputs)."""
        self._check(self.lib.ramgpu_set_wave_params(self.h, float(wave.kp), float(wave.kpmax12),
                                                    _p(self._lz), None, 0))
        for s in range(1, self.ns + 1):
            if wave.wpi[s - 1]:
                w = [_fd(a[s - 1], self.nr * self.ne) for a in (wave.walos1, wave.walos2, wave.walos3)]
                self._check(self.lib.ramgpu_set_wavelo(self.h, s, *[_p(a) for a in w]))
            else:
                self._check(self.lib.ramgpu_set_wavelo(self.h, s, None, None, None))

    def set_wave_params(self, kp, kpmax12=None):
        """WAVELO's KP and plasmapause (5.39 - 0.382 Kpmax12; Kpmax12 = kp
        unless given), no plasmasphere model."""
        km = kp if kpmax12 is None else kpmax12
        self._check(self.lib.ramgpu_set_wave_params(self.h, float(kp), float(km), _p(self._lz), None, 0))

    def wavelo(self, s):
        self._check(self.lib.ramgpu_wavelo(self.h, s))

    def loss_state(self):
        arrs = [np.zeros(self.ns) for _ in range(3)]
        self._check(self.lib.ramgpu_get_loss_state(self.h, *[_p(a) for a in arrs]))
        return dict(zip(("lswae", "lscoe", "lscsc"), arrs))

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

    def get_geometry(self):
        """The context's geometry as a dict of Fortran-ordered arrays in the
        shapes of set_geometry (keys as Geometry's fields)."""
        ns, nr, nt, ne, npa = self.dims
        g3, g2 = (nr + 1, nt, npa), (nr + 1, nt)
        out = {k: np.zeros(g3, order="F") for k in ("fnhs", "fnis", "bounhs", "bounis", "hdns", "didt",
                                                      "dibndt")}
        out.update({k: np.zeros(g2, order="F") for k in ("bnes", "dbdt", "eir", "eip")})
        mg = np.zeros((nr, nt), dtype=np.int32, order="F")
        keys = ("fnhs", "fnis", "bounhs", "bounis", "hdns", "didt", "dibndt", "bnes", "dbdt", "eir", "eip")
        self._check(self.lib.ramgpu_get_geometry(self.h, *[_p(out[k]) for k in keys], mg.ctypes.data_as(_ip)))
        out["outside_mgnp"] = mg
        return out

    def set_told(self, told):
        self._check(self.lib.ramgpu_set_told(self.h, float(told)))

    def told(self) -> float:
        x = C.c_double()
        self._check(self.lib.ramgpu_get_told(self.h, C.byref(x)))
        return x.value

    def computehI(self, lines, pa, pabn, lz, outside_mgnp=None, scale_at=None, smooth=True):
        """computehI from field lines on the RAM grid (``fieldlines.Lines``:
        x, y, z, b as (nthe, NR, NT), chi, ieq 1-based, bnormal): the
        context's geometry from their bounce integrals."""
        ns, nr, nt, ne, npa = self.dims
        n = lines.b.size
        x, y, z, b = (_fd(a, n) for a in (lines.x, lines.y, lines.z, lines.b))
        nthe = lines.b.shape[0]
        mg = np.asfortranarray(np.zeros((nr, nt)) if outside_mgnp is None else outside_mgnp, dtype=np.int32)
        sa = None if scale_at is None else np.ascontiguousarray(scale_at, dtype=np.int32)
        chi, pa_, pb, lz_ = (np.ascontiguousarray(a, dtype=np.float64) for a in (lines.chi, pa, pabn, lz))
        self._check(self.lib.ramgpu_computehI(
            self.h, int(nthe), int(lines.ieq), _p(x), _p(y), _p(z), _p(b), _p(chi), _p(pa_), _p(pb), _p(lz_),
            mg.ctypes.data_as(_ip), None if sa is None else sa.ctypes.data_as(_ip), float(lines.bnormal),
            1 if smooth else 0))

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

    # -- moments (ANISCH, FLUX) ------------------------------------------
    MOMENTS = ("ppert", "ppart", "edent", "rnhtt")

    def set_moment_factors(self, g):
        ns, nr, nt, ne, npa = self.dims
        arrs = [_fd(g.ffactor, ns * nr * ne * npa), _fd(g.ernh, ns * ne), _fd(g.epp, ns * ne)]
        self._check(self.lib.ramgpu_set_moment_factors(self.h, *[_p(a) for a in arrs]))

    def anisch(self, s: int, out=None) -> dict:
        """ANISCH(S): the (S, 2..NR, :) entries of out's (NS,NR,NT) arrays."""
        ns, nr, nt, _, _ = self.dims
        if out is None:
            out = {n: np.zeros((ns, nr, nt), order="F") for n in self.MOMENTS}
        self._check(self.lib.ramgpu_anisch(self.h, s, *[_p(out[n]) for n in self.MOMENTS]))
        return out

    def flux(self) -> np.ndarray:
        out = np.zeros(self.dims, order="F")
        self._check(self.lib.ramgpu_flux(self.h, _p(out)))
        return out

    def moments(self, flux: bool = True, out=None):
        """ANISCH of every species and (flux=True) FLUX, in one launch."""
        ns, nr, nt, _, _ = self.dims
        if out is None:
            out = {n: np.zeros((ns, nr, nt), order="F") for n in self.MOMENTS}
        fx = np.zeros(self.dims, order="F") if flux else None
        self._check(self.lib.ramgpu_moments(self.h, *[_p(out[n]) for n in self.MOMENTS],
                                            _p(fx) if flux else None))
        return out, fx

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

    def ram_run_schedule(self, kp, bscale=None, noalloc=False):
        """nsteps = len(kp) steps with a per-step Kp and boundary scales
        bscale (nsteps*NS); noalloc: the entry point that stages the schedule
        in the library's fixed device array instead of allocating."""
        kp = _fd(kp)
        n = kp.size
        bs = None if bscale is None else _fd(bscale, n * self.ns)
        out = (SpeciesOut * self.ns)()
        fn = self.lib.ramgpu_ram_run_schedule_noalloc if noalloc else self.lib.ramgpu_ram_run_schedule
        self._check(fn(self.h, n, _p(kp), None if bs is None else _p(bs), out))
        return [o.as_dict() for o in out]

    # -- reverse mode ------------------------------------------------------
    def op_bwd(self, op, s, adj, fgeos_adj=None):
        """dL/d(input) of operator `op` (reference name) for species s at the
        current F2, from adj = dL/d(output) (NS,NR,NT,NE,NPA).  DRIFTR also
        returns dL/dFGEOS (NS,NT,NE,NPA) accumulated onto fgeos_adj."""
        a = np.array(adj, dtype=np.float64, order="F", copy=True)
        if a.shape != self.dims:
            raise ValueError("adj must have the shape of F2")
        if op == "driftr":
            g = np.zeros((self.ns, self.nt, self.ne, self.npa), order="F") if fgeos_adj is None else \
                np.array(fgeos_adj, dtype=np.float64, order="F", copy=True)
            self._check(self.lib.ramgpu_driftr_bwd(self.h, s, _p(a), _p(g)))
            return a, g
        self._check(getattr(self.lib, f"ramgpu_{op}_bwd")(self.h, s, _p(a)))
        return a

    def ram_run_bwd(self, adj):
        """(dL/dF2_in, dL/dFGEOS) of one ram_run step at the current F2."""
        a = np.array(adj, dtype=np.float64, order="F", copy=True)
        g = np.zeros((self.ns, self.nt, self.ne, self.npa), order="F")
        self._check(self.lib.ramgpu_ram_run_bwd(self.h, _p(a), _p(g)))
        return a, g

    def adjoint_run(self, nsteps, pts, w, target, ckpt_every=10, kp=None, bscale=None):
        """Loss and gradients of the virtual-satellite misfit over nsteps.
        pts: (npts,4) 1-based S,I,J,K; w: (npts,NPA); target: (nsteps,npts).
        Returns (loss, dL/dF2_0 (NS,NR,NT,NE,NPA), dL/dFGEOS (nsteps,NS,NT,NE,NPA))."""
        pts = np.ascontiguousarray(pts, dtype=np.int32)
        npts = pts.shape[0]
        w = np.ascontiguousarray(w, dtype=np.float64)
        target = np.ascontiguousarray(target, dtype=np.float64)
        if w.shape != (npts, self.npa) or target.shape != (nsteps, npts):
            raise ValueError("w must be (npts,NPA), target (nsteps,npts)")
        kpa = None if kp is None else np.ascontiguousarray(kp, dtype=np.float64)
        bsa = None if bscale is None else np.ascontiguousarray(bscale, dtype=np.float64)
        if kpa is not None and kpa.size != nsteps:
            raise ValueError("kp must have nsteps entries")
        if bsa is not None and bsa.size != nsteps * self.ns:
            raise ValueError("bscale must have nsteps*NS entries")
        loss = C.c_double(0.0)
        gf = np.zeros(self.dims, order="F")
        gb = np.zeros((nsteps, self.ns, self.nt, self.ne, self.npa))
        gbf = np.zeros(nsteps * self.ns * self.nt * self.ne * self.npa)
        self._check(self.lib.ramgpu_adjoint_run(
            self.h, nsteps, ckpt_every, None if kpa is None else _p(kpa), None if bsa is None else _p(bsa),
            npts, pts.ctypes.data_as(_ip), _p(w), _p(target), C.byref(loss), _p(gf), _p(gbf)))
        per = self.ns * self.nt * self.ne * self.npa
        for n in range(nsteps):
            gb[n] = gbf[n * per:(n + 1) * per].reshape((self.ns, self.nt, self.ne, self.npa), order="F")
        return loss.value, gf, gb

    def set_plane_shape(self, threads=0, maxc=0):
        """Block size and row bound of the plane kernels (0, 0: the library's choice)."""
        self._check(self.lib.ramgpu_set_plane_shape(self.h, threads, maxc))

    def plane_shape(self):
        t, m = C.c_int(), C.c_int()
        self._check(self.lib.ramgpu_get_plane_shape(self.h, C.byref(t), C.byref(m)))
        return t.value, m.value

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


class RamBatch:
    """An ensemble of RamGpu members stepped together (``ramgpu_batch``).
    The members stay owned by the caller and must outlive the ensemble."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = C.c_void_p()
        self._check(self.lib.ramgpu_batch_create(C.byref(h), arr, len(self.members)))
        self.h = h
        self.sum_ns = sum(m.ns for m in self.members)

    def _check(self, rc):
        if rc != 0:
            raise RamGpuError(f"ramgpu status {rc}: {self.lib.ramgpu_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ramgpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _split(self, out):
        res, o = [], 0
        for m in self.members:
            res.append([out[o + s].as_dict() for s in range(m.ns)])
            o += m.ns
        return res

    def kernel_timing(self, enable: bool = True):
        """Per-launch HIP events around every kernel of the following
        (eagerly launched) ensemble steps."""
        self._check(self.lib.ramgpu_batch_kernel_timing(self.h, int(bool(enable))))

    def kernel_times(self) -> dict:
        """{kind: (total_ms, launches)} since the last call (then reset)."""
        ms = (C.c_double * len(KERNEL_KINDS))()
        n = (C.c_int * len(KERNEL_KINDS))()
        self._check(self.lib.ramgpu_batch_kernel_times(self.h, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNEL_KINDS)}

    def use_graph(self, enable=True):
        self._check(self.lib.ramgpu_batch_use_graph(self.h, int(bool(enable))))

    def ram_run(self, nsteps=1):
        out = (SpeciesOut * self.sum_ns)()
        self._check(self.lib.ramgpu_batch_ram_run(self.h, nsteps, out))
        return self._split(out)

    def ram_run_schedule(self, kp, bscale=None):
        """kp: (nsteps, M); bscale: (nsteps, sum NS) or None."""
        kp = np.ascontiguousarray(kp, dtype=np.float64)
        n = kp.shape[0]
        if kp.shape != (n, len(self.members)):
            raise ValueError("kp must be (nsteps, members)")
        bs = None
        if bscale is not None:
            bs = np.ascontiguousarray(bscale, dtype=np.float64)
            if bs.shape != (n, self.sum_ns):
                raise ValueError("bscale must be (nsteps, sum of member species)")
        out = (SpeciesOut * self.sum_ns)()
        self._check(self.lib.ramgpu_batch_ram_run_schedule(self.h, n, _p(kp),
                                                           None if bs is None else _p(bs), out))
        return self._split(out)

    def last_run_ms(self) -> float:
        ms = C.c_double(0.0)
        self._check(self.lib.ramgpu_batch_last_run_ms(self.h, C.byref(ms)))
        return ms.value

    def autotune(self, reps=3):
        ms = (C.c_double * 3)()
        self._check(self.lib.ramgpu_batch_autotune(self.h, reps, ms))
        return {"r": ms[0], "p": ms[1], "plane": ms[2]}

    def shapes(self):
        caps = (C.c_int * 3)()
        self._check(self.lib.ramgpu_batch_get_shapes(self.h, caps))
        return {"r": caps[0], "p": caps[1], "plane": caps[2]}

    def set_shapes(self, r=0, p=0, plane=0):
        caps = (C.c_int * 3)(r, p, plane)
        self._check(self.lib.ramgpu_batch_set_shapes(self.h, caps))

    def adjoint_run(self, nsteps, pts, w, target, kp, bscale=None, ckpt_every=10, grad_fgeos=True):
        """Misfit losses and gradients of every member (ramgpu_batch_adjoint_run).
        pts[m]: (npts_m,4) 1-based S,I,J,K; w[m]: (npts_m,NPA_m); target[m]:
        (nsteps,npts_m); kp: (nsteps, M); bscale: (nsteps, sum NS) or None.
        Returns (loss[M], [dL/dF2_0 per member], [dL/dFGEOS (nsteps,NS,NT,NE,NPA) per member or None])."""
        M = len(self.members)
        kp = np.ascontiguousarray(kp, dtype=np.float64)
        if kp.shape != (nsteps, M):
            raise ValueError("kp must be (nsteps, members)")
        bs = None
        if bscale is not None:
            bs = np.ascontiguousarray(bscale, dtype=np.float64)
            if bs.shape != (nsteps, self.sum_ns):
                raise ValueError("bscale must be (nsteps, sum of member species)")
        keep, npts = [], (C.c_int * M)()
        pp, wp, tp, gfp, gbp = ((C.c_void_p * M)() for _ in range(5))
        gfs, gbs = [], []
        for i, m in enumerate(self.members):
            pa = np.ascontiguousarray(pts[i], dtype=np.int32)
            wa = np.ascontiguousarray(w[i], dtype=np.float64)
            ta = np.ascontiguousarray(target[i], dtype=np.float64)
            if pa.ndim != 2 or pa.shape[1] != 4 or wa.shape != (pa.shape[0], m.npa) or \
                    ta.shape != (nsteps, pa.shape[0]):
                raise ValueError(f"member {i}: pts (n,4), w (n,NPA), target (nsteps,n) expected")
            gf = np.zeros(m.dims, order="F")
            gb = np.zeros(nsteps * m.ns * m.nt * m.ne * m.npa) if grad_fgeos else None
            keep += [pa, wa, ta]
            npts[i] = pa.shape[0]
            pp[i], wp[i], tp[i] = pa.ctypes.data, wa.ctypes.data, ta.ctypes.data
            gfp[i] = gf.ctypes.data
            gbp[i] = gb.ctypes.data if gb is not None else None
            gfs.append(gf)
            gbs.append(gb)
        loss = np.zeros(M)
        self._check(self.lib.ramgpu_batch_adjoint_run(
            self.h, nsteps, ckpt_every, _p(kp), None if bs is None else _p(bs), npts, pp, wp, tp, _p(loss),
            gfp, gbp if grad_fgeos else None))
        out_gb = []
        for m, gb in zip(self.members, gbs):
            if gb is None:
                out_gb.append(None)
                continue
            per = m.ns * m.nt * m.ne * m.npa
            out_gb.append(np.stack([gb[n * per:(n + 1) * per].reshape((m.ns, m.nt, m.ne, m.npa), order="F")
                                    for n in range(nsteps)]))
        return loss, gfs, out_gb
