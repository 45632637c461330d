"""RAM grid construction -- a restatement of ``ARRAYS``
(src/ModRamInit.f90:364-587) for the host side of the step.

The reference builds the radial, local-time, energy and pitch-angle grids,
the relativistic factors and the loss-cone index ``UPA`` once at start-up on
the CPU; this module does the same for the standalone host.  Arithmetic is
written in the reference's order with Python floats (IEEE doubles) so that
the tables match what a Fortran host would hand to the C-ABI.

Arrays are 0-based numpy arrays holding the reference's 1-based entries in
order (``ekev[0]`` is ``EKEV(1)``).  Species-dependent tables have a leading
species axis (``grel[s, k]`` is ``GREL(s+1, k+1)``).

Synthetic-grid choice: the reference fixes ``WE(1)=3E-2``, ``RW=1.27`` for
``EnergyMin=0.1`` keV (src/ModRamInit.f90:438-441).  For NE other than 35 that
ratio pushes the top bin to MeV energies, where the explicit scheme at
dt=5 s is no longer Courant-stable, so ``energy_ratio`` picks the ratio that
keeps the same span (EBND(NE) equal to the NE=35 value, about 477 keV).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .species import Species

# src/ModRamConst.f90:13-22, share/Library/src/ModNumConst.f90:15
RE = 6.371E6
HMIN = 2E5
MP = 1.673E-27
Q = 1.602E-19
CS = 2.998E8
PI = 3.1415926535897932384626433832795
# src/ModRamGrids.f90:36-37
RADIUS_MIN = 1.75
RADIUS_MAX = 6.5


def _acosd(x: float) -> float:  # src/ModRamFunctions.f90:162-174
    return 180.0 / PI * math.acos(x)


def _asind(x: float) -> float:  # src/ModRamFunctions.f90:177-189
    return 180.0 / PI * math.asin(x)


def _cosd(x: float) -> float:  # src/ModRamFunctions.f90:192-204
    return math.cos(PI / 180.0 * x)


def energy_ratio(ne: int, elb: float = 0.1) -> tuple[float, float]:
    """(WE(1), RW) for the energy grid.  NE=35 is the reference's table
    (src/ModRamInit.f90:434-445); other NE keep the NE=35 span."""
    if abs(elb - 0.01) <= 1e-9:
        we1, rw = 2.8E-3, 1.36
    elif abs(elb - 0.1) <= 1e-9:
        we1, rw = 3E-2, 1.27
    elif abs(elb - 1.0) <= 1e-9:
        we1, rw = 0.31, 1.16
    else:
        raise ValueError(f"EnergyMin={elb} has no reference energy table")
    if ne == 35:
        return we1, rw
    target = we1 * (rw ** 35 - 1.0) / (rw - 1.0)  # sum of the NE=35 widths
    lo, hi = 1.0 + 1e-9, 3.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if we1 * (mid ** ne - 1.0) / (mid - 1.0) < target:
            lo = mid
        else:
            hi = mid
    return we1, 0.5 * (lo + hi)


@dataclass
class RamGrid:
    nr: int
    nt: int
    ne: int
    npa: int
    species: tuple[Species, ...]
    dl1: float = 0.0
    mdr: float = 0.0
    dphi: float = 0.0
    conf1: float = 0.0
    conf2: float = 0.0
    rw: float = 0.0
    lz: np.ndarray = field(default=None)       # (NR+1)
    rlz: np.ndarray = field(default=None)      # (NR+1)
    phi: np.ndarray = field(default=None)      # (NT)
    mlt: np.ndarray = field(default=None)      # (NT)
    we: np.ndarray = field(default=None)       # (NE)
    de: np.ndarray = field(default=None)
    ekev: np.ndarray = field(default=None)
    ebnd: np.ndarray = field(default=None)
    rmas: np.ndarray = field(default=None)     # (NS)
    grel: np.ndarray = field(default=None)     # (NS,NE)
    grbnd: np.ndarray = field(default=None)
    v: np.ndarray = field(default=None)
    vbnd: np.ndarray = field(default=None)
    pa: np.ndarray = field(default=None)       # (NPA)
    mu: np.ndarray = field(default=None)
    dmu: np.ndarray = field(default=None)
    wmu: np.ndarray = field(default=None)
    pabn: np.ndarray = field(default=None)
    cone: np.ndarray = field(default=None)     # (NR+4)
    upa: np.ndarray = field(default=None)      # (NR) real-valued index
    ffactor: np.ndarray = field(default=None)  # (NS,NR,NE,NPA)
    ernh: np.ndarray = field(default=None)     # (NS,NE)
    epp: np.ndarray = field(default=None)      # (NS,NE)

    @property
    def ns(self) -> int:
        return len(self.species)

    @property
    def charge(self) -> np.ndarray:
        return np.array([sp.charge for sp in self.species], dtype=np.int32)

    @property
    def cex_kind(self) -> np.ndarray:
        return np.array([sp.cex_kind if sp.cex else 0 for sp in self.species], dtype=np.int32)


def build_grid(nr: int, nt: int, ne: int, npa: int, species, elb: float = 0.1) -> RamGrid:
    """ARRAYS (src/ModRamInit.f90:364-587) for every species."""
    species = tuple(species)
    ns = len(species)
    if nr < 4 or nt < 4 or ne < 4 or npa < 4:
        raise ValueError("RAM grids need at least 4 points per axis")
    g = RamGrid(nr, nt, ne, npa, species)

    # radial and local-time grids, :392-412
    dl1 = (RADIUS_MAX - RADIUS_MIN) / (nr - 1)
    g.dl1 = dl1
    g.mdr = dl1 * RE
    lz = np.empty(nr + 1)
    rlz = np.empty(nr + 1)
    for i in range(nr + 1):
        lz[i] = RADIUS_MIN + i * dl1
        rlz[i] = RE * lz[i]
    g.lz, g.rlz = lz, rlz
    g.dphi = 2.0 * PI / (nt - 1)
    g.phi = np.array([j * g.dphi for j in range(nt)])
    g.mlt = np.array([p * 12.0 / PI for p in g.phi])

    # energy grid and relativistic factors, :414-459
    g.rmas = np.array([MP * sp.mass for sp in species])
    we1, rw = energy_ratio(ne, elb)
    g.rw = rw
    we = np.zeros(ne)
    ebnd = np.zeros(ne)
    de = np.zeros(ne)
    ekev = np.zeros(ne)
    we[0] = we1
    ekev[0] = elb + 0.5 * we[0]
    ebnd[0] = elb + we[0]
    for k in range(ne - 1):
        we[k + 1] = we[k] * rw
        ebnd[k + 1] = ebnd[k] + we[k + 1]
        de[k] = 0.5 * (we[k] + we[k + 1])
        ekev[k + 1] = ekev[k] + de[k]
    de[ne - 1] = 0.5 * we[ne - 1] * (1.0 + rw)
    g.we, g.ebnd, g.de, g.ekev = we, ebnd, de, ekev
    grel = np.zeros((ns, ne))
    grbnd = np.zeros((ns, ne))
    v = np.zeros((ns, ne))
    vbnd = np.zeros((ns, ne))
    for s in range(ns):
        rm = g.rmas[s]
        for k in range(ne):
            grel[s, k] = 1.0 + ekev[k] * 1000.0 * Q / rm / CS / CS
            v[s, k] = CS * math.sqrt(grel[s, k] * grel[s, k] - 1.0) / grel[s, k]
            grbnd[s, k] = 1.0 + ebnd[k] * 1000.0 * Q / rm / CS / CS
            vbnd[s, k] = CS * math.sqrt(grbnd[s, k] * grbnd[s, k] - 1.0) / grbnd[s, k]
    g.grel, g.grbnd, g.v, g.vbnd = grel, grbnd, v, vbnd

    # dipole loss cone, :461-474
    cone = np.zeros(nr + 4)
    for i in range(nr):
        clc = (RE + HMIN) / rlz[i]
        cone[i] = _asind(math.sqrt(clc * clc * clc / math.sqrt(4.0 - 3.0 * clc)))
    cone[nr], cone[nr + 1], cone[nr + 2], cone[nr + 3] = 2.5, 1.5, 1.0, 0.0
    g.cone = cone

    # pitch-angle grid, :476-535
    pa = np.zeros(npa)
    mu = np.zeros(npa)
    wmu = np.zeros(npa)
    dmu = np.zeros(npa)
    pabn = np.zeros(npa)
    if npa != 72:
        # the reference's uniform branch (written for NPA=90), used here for
        # any NPA other than 72
        pa[0], mu[0], pa[npa - 1], mu[npa - 1] = 90.0, 0.0, 0.0, 1.0
        wmu[0] = (mu[npa - 1] - mu[0]) / npa
        for l in range(npa - 1):
            wmu[l + 1] = wmu[l]
            dmu[l] = 0.5 * (wmu[l] + wmu[l + 1])
            mu[l + 1] = mu[l] + dmu[l]
            pa[l + 1] = _acosd(mu[l + 1])
            pabn[l] = _acosd(mu[l] + 0.5 * wmu[l])
        dmu[npa - 1] = 0.5 * (wmu[npa - 1] + wmu[npa - 2])
        pabn[npa - 1] = 0.0
    else:
        if nr != 20:
            raise ValueError("the non-uniform NPA=72 pitch-angle grid of ARRAYS is "
                             "defined for NR=20")
        rwu = 0.98
        pa[0], mu[0], pa[npa - 1], mu[npa - 1] = 90.0, 0.0, 0.0, 1.0
        wmu[0] = (mu[npa - 1] - mu[0]) / 32
        for l in range(46):  # reference L=1..46
            wmu[l + 1] = wmu[l] * rwu
            dmu[l] = 0.5 * (wmu[l] + wmu[l + 1])
            mu[l + 1] = mu[l] + dmu[l]
            pa[l + 1] = _acosd(mu[l + 1])
        pa[47] = 18.7
        mu[47] = _cosd(pa[47])
        dmu[46] = mu[47] - mu[46]
        ic = 2
        for L in range(48, npa):  # reference L=48..NPA-1
            pa[L] = cone[ic - 1]
            if L == 49:
                pa[49] = 16.0
            else:
                ic = ic + (nr - 1) // 19 if ic < nr else ic + 1
            mu[L] = _cosd(pa[L])
            dmu[L - 1] = mu[L] - mu[L - 1]
            wmu[L - 1] = 2.0 * (dmu[L - 2] - 0.5 * wmu[L - 2])
            if L > 55:
                wmu[L - 1] = 0.5 * (dmu[L - 1] + dmu[L - 2])
        dmu[npa - 1] = dmu[npa - 2]
        wmu[npa - 1] = dmu[npa - 2]
        for l in range(npa - 1):
            pabn[l] = _acosd(mu[l] + 0.5 * wmu[l])
        pabn[npa - 1] = 0.0
    g.pa, g.mu, g.wmu, g.dmu, g.pabn = pa, mu, wmu, dmu, pabn

    # loss-cone index, :539-544
    upa = np.zeros(nr)
    for i in range(nr):
        u = float(npa)
        for L in range(npa, 0, -1):
            if pa[L - 1] <= cone[i]:
                u = float(L)
        upa[i] = u
    g.upa = upa

    # FFACTOR, :560-570
    ff = np.zeros((ns, nr, ne, npa))
    for s in range(ns):
        for i in range(nr):
            for k in range(ne):
                gk = grel[s, k]
                base = lz[i] * lz[i] * gk / math.sqrt(gk * gk - 1.0)
                for l in range(1, npa):
                    ff[s, i, k, l] = base * mu[l]
                ff[s, i, k, 0] = ff[s, i, k, 1]
    g.ffactor = ff

    # energy factors of the moments, :573-575
    ernh = np.zeros((ns, ne))
    epp = np.zeros((ns, ne))
    for s in range(ns):
        for k in range(ne):
            gk = grel[s, k]
            ernh[s, k] = we[k] * gk / math.sqrt((gk - 1.0) * (gk + 1.0))
            epp[s, k] = ernh[s, k] * ekev[k]
    g.ernh, g.epp = ernh, epp

    # ghost-cell factors, :580-581
    g.conf1 = ((lz[nr - 1] + dl1) / lz[nr - 1]) ** 2
    g.conf2 = ((lz[nr - 1] + 2.0 * dl1) / lz[nr - 1]) ** 2
    return g
