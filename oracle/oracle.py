"""ctypes front end of the oracle libraries (TEST INFRASTRUCTURE ONLY).

``Oracle`` drives oracle/libramoracle.so, the C restatement of the RAM split
step; ``Reference`` drives oracle/_ref/libramref.so, the reference's own
Fortran drift and Coulomb operators behind oracle/ref_harness.f90.  Both take the same
host arrays in the reference's Fortran layout.  Only tests/, the smoke check
and the bench's CPU-baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "libramoracle.so"
REF_SO = HERE / "_ref" / "libramref.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _State(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ns", "nr", "nt", "ne", "npa")] + \
        [(n, C.c_double) for n in ("dts", "betalim", "fraccfl", "mdr", "dphi", "conf1", "conf2")] + \
        [("charge", _ip), ("cex_kind", _ip), ("rmas", _dp), ("rlz", _dp),
         ("ekev", _dp), ("ebnd", _dp), ("we", _dp), ("de", _dp),
         ("grel", _dp), ("grbnd", _dp), ("v", _dp), ("mu", _dp), ("wmu", _dp), ("dmu", _dp),
         ("upa", _dp), ("fnhs", _dp), ("fnis", _dp), ("bounhs", _dp), ("bounis", _dp),
         ("hdns", _dp), ("didt", _dp), ("dibndt", _dp), ("bnes", _dp), ("dbdt", _dp),
         ("vt", _dp), ("eir", _dp), ("eip", _dp), ("outside_mgnp", _ip), ("fgeos", _dp),
         ("f2", _dp), ("dtdrift_r", _dp), ("dtdrift_p", _dp), ("dtdrift_e", _dp),
         ("dtdrift_mu", _dp), ("setrc", _dp), ("elorc", _dp), ("lsdr", _dp), ("lscha", _dp),
         ("lsatm", _dp), ("wpi", _ip), ("walos1", _dp), ("walos2", _dp), ("walos3", _dp),
         ("lz", _dp), ("rlpp", _dp), ("kp", C.c_double), ("lswae", _dp),
         ("vbnd", _dp), ("necr", _dp), ("time_elapsed", C.c_double),
         ("coule", _dp), ("couli", _dp), ("ata", _dp), ("gta", _dp), ("cedr", _dp), ("cidr", _dp),
         ("use_coulomb", C.c_int), ("lscoe", _dp), ("lscsc", _dp),
         ("pd1", C.POINTER(_dp)), ("pd2", C.POINTER(_dp)), ("pd_wpi", _ip), ("pd_emic", _ip),
         ("flc", C.POINTER(_dp)), ("dt_bc", C.c_double)]

COUL_TABLES = ("coule", "couli", "ata", "gta", "cedr", "cidr")


class _Drift(C.Structure):
    _fields_ = [("qs", C.c_int)] + [(n, _dp) for n in
                                   ("vr", "p1", "p2", "edot", "mudot", "cdr", "cdp", "cde", "cdmu")]


class _Loss(C.Structure):
    _fields_ = [("charge", _dp), ("atlos", _dp)]


def _fd(a):
    # always a private copy: the oracle mutates its arrays in place
    return np.array(a, dtype=np.float64, order="F", copy=True)


def _fi(a):
    return np.array(a, dtype=np.int32, order="F", copy=True)


def _ptr(a, typ=_dp):
    return a.ctypes.data_as(typ)


class Oracle:
    """The C restatement, holding its own copies of one problem's arrays."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not ORACLE_SO.exists():
                raise FileNotFoundError(f"{ORACLE_SO} is not built (make -C oracle)")
            lib = C.CDLL(str(ORACLE_SO))
            sp = C.POINTER(_State)
            for name in ("driftpara", "driftr", "driftp", "drifte", "driftmu"):
                f = getattr(lib, "oracle_" + name)
                f.argtypes = [sp, C.c_int, C.POINTER(_Drift)]
                f.restype = None
            lib.oracle_cepara.argtypes = [sp, C.c_int, C.POINTER(_Loss)]
            lib.oracle_charexchange.argtypes = [sp, C.c_int, C.POINTER(_Loss)]
            lib.oracle_atmol.argtypes = [sp, C.c_int, C.POINTER(_Loss)]
            lib.oracle_sumrc.argtypes = [sp, C.c_int]
            lib.oracle_wavelo.argtypes = [sp, C.c_int]
            for name in ("coulpara", "coulen", "coulmu"):
                getattr(lib, "oracle_" + name).argtypes = [sp, C.c_int]
            lib.oracle_wpadif.argtypes = [sp, C.c_int, _dp, _dp]
            lib.oracle_flcscatter.argtypes = [sp, C.c_int, _dp, C.c_double]
            lib.oracle_ram_run.argtypes = [sp, C.c_int, _dp]
            lib.oracle_ram_run.restype = C.c_int
            lib.oracle_drift_alloc.argtypes = [sp, C.POINTER(_Drift)]
            lib.oracle_drift_alloc.restype = C.c_int
            lib.oracle_drift_free.argtypes = [C.POINTER(_Drift)]
            lib.oracle_loss_alloc.argtypes = [sp, C.POINTER(_Loss)]
            lib.oracle_loss_alloc.restype = C.c_int
            lib.oracle_loss_free.argtypes = [C.POINTER(_Loss)]
            lib.oracle_anisch.argtypes = [sp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
            lib.oracle_anisch.restype = None
            lib.oracle_flux.argtypes = [sp, _dp, _dp]
            lib.oracle_flux.restype = None
            cls._lib = lib
        return cls._lib

    def __init__(self, problem):
        g, geo = problem.grid, problem.geom
        self.ns, self.nr, self.nt, self.ne, self.npa = g.ns, g.nr, g.nt, g.ne, g.npa
        keep = {}

        def d(name, a):
            keep[name] = _fd(a)
            return _ptr(keep[name])

        def i(name, a):
            keep[name] = _fi(a)
            return _ptr(keep[name], _ip)

        st = _State()
        st.ns, st.nr, st.nt, st.ne, st.npa = g.ns, g.nr, g.nt, g.ne, g.npa
        st.dts, st.betalim, st.fraccfl = problem.dts, problem.betalim, problem.fraccfl
        st.mdr, st.dphi, st.conf1, st.conf2 = g.mdr, g.dphi, g.conf1, g.conf2
        st.charge = i("charge", g.charge)
        st.cex_kind = i("cex_kind", g.cex_kind)
        st.rmas = d("rmas", g.rmas)
        st.rlz = d("rlz", g.rlz)
        for n in ("ekev", "ebnd", "we", "de", "grel", "grbnd", "v", "mu", "wmu", "dmu", "upa"):
            setattr(st, n, d(n, getattr(g, n)))
        for n in ("fnhs", "fnis", "bounhs", "bounis", "hdns", "didt", "dibndt", "bnes",
                  "dbdt", "eir", "eip"):
            setattr(st, n, d(n, getattr(geo, n)))
        st.outside_mgnp = i("outside_mgnp", geo.outside_mgnp)
        st.vt = d("vt", problem.vt)
        st.fgeos = d("fgeos", problem.fgeos)
        st.f2 = d("f2", problem.f2.copy(order="F"))
        for n in ("dtdrift_r", "dtdrift_p", "dtdrift_e", "dtdrift_mu", "setrc", "elorc",
                  "lsdr", "lscha", "lsatm", "lswae", "lscoe", "lscsc"):
            setattr(st, n, d(n, np.zeros(g.ns)))
        vbnd = getattr(g, "vbnd", None)
        st.vbnd = d("vbnd", np.zeros((g.ns, g.ne)) if vbnd is None else vbnd)
        for n in COUL_TABLES:
            setattr(st, n, d(n, np.zeros((g.ns, g.ne, g.npa))))
        st.necr = d("necr", np.zeros((g.nr, g.nt)))
        # WPADIF / FLC tables, per species (null pointers: none)
        self._pd1 = (_dp * g.ns)()
        self._pd2 = (_dp * g.ns)()
        self._flc = (_dp * g.ns)()
        st.pd1, st.pd2, st.flc = self._pd1, self._pd2, self._flc
        st.pd_wpi = i("pd_wpi", np.zeros(g.ns, dtype=np.int32))
        st.pd_emic = i("pd_emic", np.zeros(g.ns, dtype=np.int32))
        coul = getattr(problem, "coulomb", None)
        if coul is not None:
            keep["necr"][...] = coul.necr
            st.time_elapsed = float(coul.time_elapsed)
            st.use_coulomb = 1
        st.lz = d("lz", g.lz)
        wave = getattr(problem, "wave", None)
        if wave is not None:
            self._set_wave(st, d, i, wave)
        self._keep = keep
        self._grid = g
        self.st = st
        self._drift = {}
        self._loss = {}

    @staticmethod
    def _set_wave(st, d, i, wave):
        """WAVELO inputs: wave.wpi (NS) flags, wave.walos1/2/3 (NS,NR,NE),
        wave.rlpp (NT), wave.kp."""
        st.wpi = i("wpi", wave.wpi)
        st.walos1 = d("walos1", wave.walos1)
        st.walos2 = d("walos2", wave.walos2)
        st.walos3 = d("walos3", wave.walos3)
        st.rlpp = d("rlpp", wave.rlpp)
        st.kp = float(wave.kp)

    def set_wave_kp(self, kp):
        self.st.kp = float(kp)

    def set_coulomb(self, necr=None, time_elapsed=None, on=True):
        """Coulomb collisions in ram_run (DoUseCoulomb), the plasmasphere
        density NECR (NR,NT) and TimeRamElapsed (the COULMU clamp)."""
        if necr is not None:
            self._keep["necr"][...] = necr
        if time_elapsed is not None:
            self.st.time_elapsed = float(time_elapsed)
        self.st.use_coulomb = 1 if on else 0

    def set_wpadif(self, s, d1, d2=None, at_wpi=True, at_emic=False):
        """WPADIF coefficients of species s (D = d1 + d2, (NR,NT,NE,NPA)) and
        where ram_run takes WPADIF: at the WPI positions (in place of WAVELO),
        at the EMIC positions."""
        self._pd1[s - 1] = _ptr(self._store(f"pd1_{s}", d1))
        self._pd2[s - 1] = _ptr(self._store(f"pd2_{s}", d2)) if d2 is not None else _dp()
        self._keep["pd_wpi"][s - 1] = 1 if at_wpi else 0
        self._keep["pd_emic"][s - 1] = 1 if at_emic else 0

    def set_flc(self, s, coef, dt_bc=0.0):
        """FLC_coef(S,...) of species s (NR,NT,NE,NPA) for FLCscatter in
        ram_run, with Dt_bc."""
        self._flc[s - 1] = _ptr(self._store(f"flc_{s}", coef))
        self.st.dt_bc = float(dt_bc)

    def _store(self, name, a):
        self._keep[name] = _fd(a)
        return self._keep[name]

    def wpadif(self, s, d1, d2=None):
        a = self._store("_op_d1", d1)
        b = self._store("_op_d2", d2) if d2 is not None else None
        self.lib().oracle_wpadif(C.byref(self.st), s, _ptr(a), _ptr(b) if b is not None else _dp())

    def flcscatter(self, s, coef, dt_bc=0.0):
        a = self._store("_op_flc", coef)
        self.lib().oracle_flcscatter(C.byref(self.st), s, _ptr(a), float(dt_bc))

    def coul_table(self, name: str) -> np.ndarray:
        """A COULPARA table (NS,NE,NPA) as numpy (copy)."""
        return self._keep[name].copy()

    # -- state access -------------------------------------------------------
    @property
    def f2(self) -> np.ndarray:
        return self._keep["f2"]

    def set_f2(self, f2):
        self._keep["f2"][...] = f2

    def set_vt(self, vt):
        self._keep["vt"][...] = vt

    def set_fgeos(self, fgeos):
        self._keep["fgeos"][...] = fgeos

    def arr(self, name) -> np.ndarray:
        return self._keep[name]

    def dtdrift(self) -> np.ndarray:
        """(NS, 4) DtDriftR, DtDriftP, DtDriftE, DtDriftMu."""
        return np.stack([self._keep[n] for n in ("dtdrift_r", "dtdrift_p", "dtdrift_e",
                                                 "dtdrift_mu")], axis=1)

    # -- operators (species s is 1-based) -----------------------------------
    def _d(self, s):
        if s not in self._drift:
            dr = _Drift()
            if self.lib().oracle_drift_alloc(C.byref(self.st), C.byref(dr)):
                raise MemoryError
            self._drift[s] = dr
        return self._drift[s]

    def _l(self, s):
        if s not in self._loss:
            ls = _Loss()
            if self.lib().oracle_loss_alloc(C.byref(self.st), C.byref(ls)):
                raise MemoryError
            self._loss[s] = ls
        return self._loss[s]

    def call(self, op: str, s: int):
        lib = self.lib()
        if op in ("driftpara", "driftr", "driftp", "drifte", "driftmu"):
            getattr(lib, "oracle_" + op)(C.byref(self.st), s, C.byref(self._d(s)))
        elif op in ("cepara", "charexchange", "atmol"):
            getattr(lib, "oracle_" + op)(C.byref(self.st), s, C.byref(self._l(s)))
        elif op == "sumrc":
            lib.oracle_sumrc(C.byref(self.st), s)
        elif op in ("wavelo", "coulpara", "coulen", "coulmu"):
            getattr(lib, "oracle_" + op)(C.byref(self.st), s)
        else:
            raise ValueError(op)

    def drift_table(self, s: int, name: str) -> np.ndarray:
        """A DRIFTPARA/drift coefficient array of species s as numpy (copy)."""
        dr = self._d(s)
        shapes = {"vr": (self.nr,), "p1": (self.nr,), "p2": (self.nr, self.ne),
                  "edot": (self.nr, self.ne), "mudot": (self.nr, self.npa)}
        shp = shapes.get(name, (self.nr, self.nt, self.ne, self.npa))
        n = int(np.prod(shp))
        return np.ctypeslib.as_array(getattr(dr, name), shape=(n,)).reshape(shp, order="F").copy()

    def loss_table(self, s: int, name: str) -> np.ndarray:
        ls = self._l(s)
        shp = (self.nr, self.nt, self.ne, self.npa) if name == "charge" else (self.nr, self.ne)
        n = int(np.prod(shp))
        return np.ctypeslib.as_array(getattr(ls, name), shape=(n,)).reshape(shp, order="F").copy()

    def _moment_factors(self):
        if "ffactor" not in self._keep:
            g = self._grid
            self._keep["ffactor"] = _fd(g.ffactor)
            self._keep["ernh"] = _fd(g.ernh)
            self._keep["epp"] = _fd(g.epp)
        return self._keep["ffactor"], self._keep["ernh"], self._keep["epp"]

    def anisch(self, s: int, out=None):
        """ANISCH(S): updates F2's L=1 row and the (S, 2..NR, :) entries of
        out = {ppert, ppart, edent, rnhtt} (NS,NR,NT) arrays, zeros when new."""
        ff, ernh, epp = self._moment_factors()
        shp = (self.ns, self.nr, self.nt)
        if out is None:
            out = {n: np.zeros(shp, order="F") for n in ("ppert", "ppart", "edent", "rnhtt")}
        self.lib().oracle_anisch(C.byref(self.st), s, _ptr(ff), _ptr(ernh), _ptr(epp),
                                 *(_ptr(out[n]) for n in ("ppert", "ppart", "edent", "rnhtt")))
        return out

    def flux(self, out=None) -> np.ndarray:
        """FLUX (NS,NR,NT,NE,NPA) of the current F2; entries outside the
        reference's ranges keep out's values (zeros when new)."""
        ff, _, _ = self._moment_factors()
        if out is None:
            out = np.zeros((self.ns, self.nr, self.nt, self.ne, self.npa), order="F")
        self.lib().oracle_flux(C.byref(self.st), _ptr(ff), _ptr(out))
        return out

    def ram_run(self, nthreads: int = 1) -> float:
        out = C.c_double(0.0)
        if self.lib().oracle_ram_run(C.byref(self.st), nthreads, C.byref(out)):
            raise MemoryError("oracle_ram_run failed")
        return out.value

    def close(self):
        lib = self.lib()
        for dr in self._drift.values():
            lib.oracle_drift_free(C.byref(dr))
        for ls in self._loss.values():
            lib.oracle_loss_free(C.byref(ls))
        self._drift.clear()
        self._loss.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Reference:
    """The reference's own DRIFTPARA/DRIFTR/DRIFTP/DRIFTE/DRIFTMU (Fortran)."""

    _lib = None
    OPS = {"driftpara": 0, "driftr": 1, "driftp": 2, "drifte": 3, "driftmu": 4,
           "coulpara": 5, "coulen": 6, "coulmu": 7}

    @classmethod
    def available(cls) -> bool:
        return REF_SO.exists()

    @classmethod
    def lib(cls):
        if cls._lib is None:
            # unresolved MPI entry points are never called: bind lazily
            lib = C.CDLL(str(REF_SO), mode=os.RTLD_LAZY)
            lib.ref_call.argtypes = [C.c_int, C.c_int]
            lib.ref_get_f2.argtypes = [_dp]
            lib.ref_put_f2.argtypes = [_dp]
            lib.ref_get_dtdrift.argtypes = [_dp]
            lib.ref_funt.argtypes = [C.c_double]
            lib.ref_funt.restype = C.c_double
            lib.ref_funi.argtypes = [C.c_double]
            lib.ref_funi.restype = C.c_double
            lib.ref_setup_coulomb.argtypes = [_dp, _dp, C.c_double]
            lib.ref_get_coulomb.argtypes = [_dp] * 6
            cls._lib = lib
        return cls._lib

    def __init__(self, problem):
        g, geo = problem.grid, problem.geom
        self.shape = (g.ns, g.nr, g.nt, g.ne, g.npa)
        self.ns = g.ns
        keep = [_fi(g.charge), _fd(g.rmas), _fd(g.rlz), _fd(g.ekev), _fd(g.ebnd), _fd(g.we),
                _fd(g.de), _fd(g.grel), _fd(g.grbnd), _fd(g.v), _fd(g.mu), _fd(g.wmu),
                _fd(g.dmu), _fd(g.upa), _fd(geo.fnhs), _fd(geo.fnis), _fd(geo.bounhs),
                _fd(geo.bounis), _fd(geo.hdns), _fd(geo.didt), _fd(geo.dibndt), _fd(geo.bnes),
                _fd(geo.dbdt), _fd(problem.vt), _fd(geo.eir), _fd(geo.eip),
                _fi(geo.outside_mgnp), _fd(problem.fgeos), _fd(problem.f2)]
        args = [C.c_int(v) for v in (g.ns, g.nr, g.nt, g.ne, g.npa)]
        args += [C.c_double(v) for v in (problem.dts, problem.betalim, problem.fraccfl, g.mdr,
                                         g.dphi, g.conf1, g.conf2)]
        args += [_ptr(a, _ip if a.dtype == np.int32 else _dp) for a in keep]
        self.lib().ref_setup(*args)

    def call(self, op: str, s: int):
        self.lib().ref_call(self.OPS[op], s)

    def setup_coulomb(self, vbnd, necr, time_elapsed):
        """VBND (NS,NE), NECR (NR,NT), TimeRamElapsed and RAMSpecies."""
        a, b = _fd(vbnd), _fd(necr)
        self.lib().ref_setup_coulomb(_ptr(a), _ptr(b), float(time_elapsed))

    def coulomb_tables(self) -> dict:
        ns, _, _, ne, npa = self.shape
        out = {n: np.zeros((ns, ne, npa), order="F") for n in COUL_TABLES}
        self.lib().ref_get_coulomb(*(_ptr(out[n]) for n in COUL_TABLES))
        return out

    @property
    def f2(self) -> np.ndarray:
        out = np.zeros(self.shape, order="F")
        self.lib().ref_get_f2(_ptr(out))
        return out

    def set_f2(self, f2):
        a = _fd(f2)
        self.lib().ref_put_f2(_ptr(a))

    def dtdrift(self) -> np.ndarray:
        out = np.zeros((self.ns, 4), order="F")
        self.lib().ref_get_dtdrift(_ptr(out))
        return out

    @classmethod
    def funt(cls, x):
        return cls.lib().ref_funt(x)

    @classmethod
    def funi(cls, x):
        return cls.lib().ref_funi(x)
