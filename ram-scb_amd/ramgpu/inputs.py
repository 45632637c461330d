"""Synthetic inputs for the RAM split step: dipole geometry, Volland-Stern
potential, kappa initial and boundary distributions, and the named workload
presets.

In the reference these arrays come from the SCB/field-line producer
(``computehI``, src/ModRamScb.f90:144-643), the electric-field module and the
boundary module.  For a dipole field the reference fills its bounce integrals
from the Ejiri closed forms ``funt``/``funi`` (src/ModRamFunctions.f90:90-143),
exactly as it does for the I=1 row (src/ModRamIO.f90:457-471); this module
does the same for every row.  Everything here is host-side set-up, not part of
the GPU step.

Array layouts follow the reference (Fortran order, 1-based semantics kept in
0-based numpy arrays):
  FNHS, FNIS, BOUNHS, BOUNIS, HDNS, dIdt, dIbndt : (NR+1, NT, NPA)
  BNES, dBdt, VT, EIR, EIP                       : (NR+1, NT)
  outsideMGNP                                    : (NR, NT) int32
  FGEOS                                          : (NS, NT, NE, NPA)
  F2                                             : (NS, NR, NT, NE, NPA)
All are allocated with ``order='F'`` so their memory matches a Fortran host.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import species as sp
from .grid import PI, Q, RE, RamGrid, build_grid


def funt(x: float) -> float:
    """h(mu), Ejiri 1978 (src/ModRamFunctions.f90:90-114)."""
    y = math.sqrt(1 - x * x)
    alpha = 1.0 + math.log(2.0 + math.sqrt(3.0)) / 2.0 / math.sqrt(3.0)
    beta = alpha / 2.0 - PI * math.sqrt(2.0) / 12.0
    a1, a2, a3, a4 = 0.055, -0.037, -0.074, 0.056
    return (alpha - beta * (y + math.sqrt(y)) + a1 * y ** (1.0 / 3.0)
            + a2 * y ** (2.0 / 3.0) + a3 * y + a4 * y ** (4.0 / 3.0))


def funi(x: float) -> float:
    """I(mu), Ejiri 1978 (src/ModRamFunctions.f90:117-143)."""
    ylog = 0.0
    y = math.sqrt(1 - x * x)
    if y > 0:
        ylog = math.log(y)
    alpha = 1.0 + math.log(2.0 + math.sqrt(3.0)) / 2.0 / math.sqrt(3.0)
    beta = alpha / 2.0 - PI * math.sqrt(2.0) / 12.0
    a1, a2, a3, a4 = 0.055, -0.037, -0.074, 0.056
    return (2.0 * alpha * (1.0 - y) + 2.0 * beta * y * ylog + 4.0 * beta * (y - math.sqrt(y))
            + 3.0 * a1 * (y ** (1.0 / 3.0) - y) + 6.0 * a2 * (y ** (2.0 / 3.0) - y)
            + 6.0 * a4 * (y - y ** (4.0 / 3.0)) - 2.0 * a3 * y * ylog)


def rairden_density(r: float) -> float:
    """Geocoronal hydrogen density [m^-3] at geocentric distance r [RE]
    (the RAIRDEN profile, src/ModRamScb.f90:365-370)."""
    return 10 ** (13.326 - 3.6908 * r + 1.1362 * r ** 2 - 0.16984 * r ** 3 + 0.009553 * r ** 4)


def volland_stern(grid: RamGrid, kp: float) -> np.ndarray:
    """VT(NR+1,NT) in volts for the given Kp (src/ModRamRun.f90:45-54,
    PHIOFS=0 as set in src/ModRamDrift.f90:61)."""
    d = 1.0 - 0.159 * kp + 0.0093 * kp ** 2
    avs = 7.05E-6 / (d * d * d) / RE
    vt = np.zeros((grid.nr + 1, grid.nt), order="F")
    for i in range(grid.nr + 1):
        lre = grid.lz[i] * RE
        for j in range(grid.nt):
            vt[i, j] = avs * (lre * lre) * math.sin(grid.phi[j] - 0.0)
    return vt


@dataclass
class Geometry:
    """Drift-velocity and geometry inputs of the step (SCB-side products)."""
    fnhs: np.ndarray
    fnis: np.ndarray
    bounhs: np.ndarray
    bounis: np.ndarray
    hdns: np.ndarray
    didt: np.ndarray
    dibndt: np.ndarray
    bnes: np.ndarray
    dbdt: np.ndarray
    eir: np.ndarray
    eip: np.ndarray
    outside_mgnp: np.ndarray


def dipole_geometry(grid: RamGrid) -> Geometry:
    """Static dipole: bounce integrals from funt/funi at the cell and boundary
    pitch angles, BNES = 0.32e-4/L^3 T, a Rairden geocorona weighted towards
    small pitch angles (a synthetic bounce average), zero induced fields."""
    nr1, nt, npa = grid.nr + 1, grid.nt, grid.npa
    shape3 = (nr1, nt, npa)
    h = np.array([funt(m) for m in grid.mu])
    ii = np.array([funi(m) for m in grid.mu])
    cpb = [math.cos(PI / 180.0 * p) for p in grid.pabn]
    hb = np.array([funt(c) for c in cpb])
    ib = np.array([funi(c) for c in cpb])
    fnhs = np.asfortranarray(np.broadcast_to(h[None, None, :], shape3))
    fnis = np.asfortranarray(np.broadcast_to(ii[None, None, :], shape3))
    bounhs = np.asfortranarray(np.broadcast_to(hb[None, None, :], shape3))
    bounis = np.asfortranarray(np.broadcast_to(ib[None, None, :], shape3))
    nh = np.array([rairden_density(l) for l in grid.lz])
    hdns = np.asfortranarray(nh[:, None, None] * (1.0 + grid.mu[None, None, :] ** 2)
                             * np.ones(shape3))
    bnes = np.zeros((nr1, nt), order="F")
    for i in range(nr1):
        bnes[i, :] = 0.32 / grid.lz[i] ** 3 / 1.e4
    z3 = np.zeros(shape3, order="F")
    z2 = np.zeros((nr1, nt), order="F")
    return Geometry(fnhs, fnis, bounhs, bounis, hdns, z3.copy(), z3.copy(), bnes,
                    z2.copy(), z2.copy(), z2.copy(),
                    np.zeros((grid.nr, nt), dtype=np.int32, order="F"))


def kappa_flux(e_kev: np.ndarray, rmas: float, n_cm3: float, kt_kev: float = 10.0,
               kappa: float = 3.0) -> np.ndarray:
    """Differential flux j(E) [cm^-2 s^-1 sr^-1 keV^-1] of an isotropic kappa
    distribution with density n, temperature kT and index kappa
    (non-relativistic): j = v^2 f(v) / m."""
    e_j = np.asarray(e_kev) * 1e3 * Q
    theta2 = (2.0 * kappa - 3.0) / kappa * kt_kev * 1e3 * Q / rmas
    norm = (n_cm3 * 1e6) / (PI * kappa * theta2) ** 1.5 * math.exp(
        math.lgamma(kappa + 1.0) - math.lgamma(kappa - 0.5))
    v2 = 2.0 * e_j / rmas
    f = norm * (1.0 + v2 / (kappa * theta2)) ** (-(kappa + 1.0))
    j_si = v2 * f / rmas                     # per m^2 s sr J
    return j_si * 1e-4 * (1e3 * Q)           # per cm^2 s sr keV


def initial_f2(grid: RamGrid, geom: Geometry, composition, rng: np.random.Generator,
               density_cm3: float = 1.0, noise_sigma: float = 0.05) -> np.ndarray:
    """F2 = j_kappa(E) * sin^2(alpha) * LogNormal(0, sigma) * FFACTOR * FNHS."""
    ns, nr, nt, ne, npa = grid.ns, grid.nr, grid.nt, grid.ne, grid.npa
    sin2 = 1.0 - grid.mu ** 2
    f2 = np.zeros((ns, nr, nt, ne, npa), order="F")
    for s in range(ns):
        j = kappa_flux(grid.ekev, grid.rmas[s], density_cm3 * composition[s])
        base = (j[None, None, :, None] * sin2[None, None, None, :]
                * grid.ffactor[s][:, None, :, :]
                * geom.fnhs[:nr, :, None, :])
        noise = np.exp(noise_sigma * rng.standard_normal((nr, nt, ne, npa)))
        f2[s] = base * noise
    # the periodic local-time seam holds one value, as after any RAM step
    f2[:, :, nt - 1] = f2[:, :, 0]
    return np.asfortranarray(f2)


def boundary_fgeos(grid: RamGrid, composition, scale: float = 1.0,
                   density_cm3: float = 1.0) -> np.ndarray:
    """FGEOS(S,J,K,L) = j_kappa * scale * FFACTOR(S,NR,K,L) for
    L=2..int(UPA(NR)-1), zero elsewhere (the fill of
    src/ModRamBoundary.f90:288-295 with an isotropic kappa spectrum)."""
    ns, nt, ne, npa = grid.ns, grid.nt, grid.ne, grid.npa
    fg = np.zeros((ns, nt, ne, npa), order="F")
    u = int(grid.upa[grid.nr - 1] - 1)
    for s in range(ns):
        j = kappa_flux(grid.ekev, grid.rmas[s], density_cm3 * composition[s]) * scale
        for l in range(1, u):  # reference L=2..u
            fg[s, :, :, l] = j[None, :] * grid.ffactor[s, grid.nr - 1, :, l][None, :]
    return fg


@dataclass
class WaveInputs:
    """Inputs of WAVELO (src/ModRamWPI.f90:580-636) for the species whose
    species%WPI is set (electrons by default, src/ModRamSpecies.f90:49):
    per-species flags, the WAVEPARA1/2 lifetime tables, the plasmapause
    RLpp(J) and Kp."""
    wpi: np.ndarray       # (NS) int32
    walos1: np.ndarray    # (NS,NR,NE) [s]
    walos2: np.ndarray
    walos3: np.ndarray
    rlpp: np.ndarray      # (NT)
    kp: float
    kpmax12: float


def wave_inputs(grid: RamGrid, kp: float, kpmax12: float | None = None, wpi=None) -> WaveInputs:
    """WAVELO's inputs on this grid: species%WPI as the reference sets it
    (the electron species) unless wpi (NS flags) is given, WALOS1/2/3 of
    WAVEPARA1/2 for those species, the plasmapause 5.39 - 0.382 Kpmax12 at
    every local time (no plasmasphere model), Kpmax12 = Kp unless given."""
    from . import wpi as W
    ns, nr, ne = grid.ns, grid.nr, grid.ne
    if wpi is None:
        wpi = [1 if spc.wpi else 0 for spc in grid.species]
    wpi = np.asarray(wpi, dtype=np.int32)
    w1 = np.zeros((ns, nr, ne), order="F")
    w2 = np.zeros((ns, nr, ne), order="F")
    w3 = np.zeros((ns, nr, ne), order="F")
    t1 = W.wavepara1(grid)
    for s in range(ns):
        if wpi[s]:
            w1[s] = t1
            w2[s], w3[s] = W.wavepara2(grid, s + 1)
    km = kp if kpmax12 is None else kpmax12
    return WaveInputs(wpi, w1, w2, w3, W.plasmapause(grid, km), float(kp), float(km))


@dataclass
class CoulombInputs:
    """Inputs of the Coulomb operators (src/ModRamCoul.f90), in ram_run
    with DoUseCoulomb: the plasmasphere electron density NECR(NR,NT) [cm^-3]
    and the elapsed time TimeRamElapsed [s] (COULMU clamps negative cells
    only when it is positive)."""
    necr: np.ndarray      # (NR,NT)
    time_elapsed: float


def plasmasphere_density(grid: RamGrid, kp: float) -> np.ndarray:
    """Synthetic NECR(NR,NT) [cm^-3] shaped like the reference's Carpenter
    model (src/ModRamPlasmasphere.f90:286-299), which is not part of this
    library: the saturated plasmasphere 10**(-0.3145 L + 3.9043) inside the
    plasmapause L = 5.6 - 0.46 Kp (bulging by 0.8 at dusk), the trough
    n_pp (L/L_pp)**-4.5 + 1 - exp((2-L)/10) outside it."""
    necr = np.zeros((grid.nr, grid.nt))
    for j in range(grid.nt):
        mlt = 24.0 * j / (grid.nt - 1)
        lpp = 5.6 - 0.46 * kp + 0.8 * math.exp(-((mlt - 18.0) / 3.0) ** 2)
        npp = 10.0 ** (-0.3145 * lpp + 3.9043)
        for i in range(grid.nr):
            lv = grid.lz[i]
            if lv <= lpp:
                necr[i, j] = 10.0 ** (-0.3145 * lv + 3.9043)
            else:
                necr[i, j] = npp * (lv / lpp) ** -4.5 + 1.0 - math.exp((2.0 - lv) / 10.0)
    return necr


def pitch_diffusion_coef(grid: RamGrid, geom, scale: float = 1e-5, seed: int = 0) -> np.ndarray:
    """A synthetic bounce-averaged pitch-angle diffusion coefficient
    D(NR,NT,NE,NPA) of the form WPADIF and FLCscatter take (ATAW, ATAC,
    ATAW_emic_*, FLC_coef; the reference reads or derives them from wave and
    field-line-curvature models that are not part of this library):
    scale * (1 - MU^2) * FNHS * MU * E^-1/2 (E in keV), with a seeded
    log-normal texture, so that DTs*D/(DMU*FNHS*MU*WMU) is of order 0.1-1."""
    rng = np.random.default_rng(seed)
    mu = grid.mu[None, None, None, :]
    fh = geom.fnhs[:grid.nr, :, None, :]
    e = grid.ekev[None, None, :, None]
    d = scale * (1.0 - mu * mu) * fh * mu / np.sqrt(e) * np.exp(0.3 * rng.standard_normal(
        (grid.nr, grid.nt, grid.ne, grid.npa)))
    return np.asfortranarray(d)


@dataclass
class Problem:
    """One member's full input set for the split step."""
    grid: RamGrid
    geom: Geometry
    vt: np.ndarray
    fgeos: np.ndarray
    f2: np.ndarray
    kp: float
    composition: tuple
    dts: float = 5.0
    betalim: float = 1.5
    fraccfl: float = 0.8
    wave: WaveInputs | None = None   # WAVELO inputs; None: no species takes it
    coulomb: CoulombInputs | None = None  # Coulomb inputs; None: DoUseCoulomb off


SPECIES4 = (sp.ELECTRON, sp.HYDROGEN, sp.HELIUM, sp.OXYGEN)
COMPOSITION4 = (1.0, 0.7, 0.05, 0.25)


def make_problem(nr=20, nt=25, ne=35, npa=72, species=(sp.HYDROGEN,), composition=(1.0,),
                 kp=3.0, seed=0, boundary_scale=1.0, dts=5.0, wave=False, coulomb=False,
                 time_elapsed=0.0) -> Problem:
    """wave: WAVELO for the species whose species%WPI is set (the electron
    species, as the reference does by default); coulomb: Coulomb collisions
    on, with the synthetic plasmasphere of Kp and TimeRamElapsed =
    time_elapsed."""
    grid = build_grid(nr, nt, ne, npa, species)
    geom = dipole_geometry(grid)
    rng = np.random.default_rng(seed)
    f2 = initial_f2(grid, geom, composition, rng)
    fgeos = boundary_fgeos(grid, composition, boundary_scale)
    w = wave_inputs(grid, kp) if wave else None
    c = CoulombInputs(plasmasphere_density(grid, kp), float(time_elapsed)) if coulomb else None
    return Problem(grid, geom, volland_stern(grid, kp), fgeos, f2, kp, tuple(composition), dts, wave=w,
                   coulomb=c)


def plain_h(seed=0) -> Problem:
    """H+ on 20x25x35x72, Kp=3, one dt=5 s split step."""
    return make_problem(seed=seed)


def multi_species(seed=0, nr=20, nt=25, ne=35, npa=72, wave=True) -> Problem:
    """e-, H+, He+, O+ (1:0.7:0.05:0.25) on 20x25x35x72, Kp=2 at step 0, with
    the reference's default electron lifetimes (WAVELO) unless wave=False."""
    return make_problem(nr, nt, ne, npa, SPECIES4, COMPOSITION4, kp=2.0, seed=seed, wave=wave)


def multi_species_schedule(nsteps: int, seed: int = 0):
    """Per-step (Kp, boundary scale) of the multi-species run: Kp ramps
    linearly 2 -> 6 over the run, the boundary scale is LogNormal(0, 0.3)."""
    rng = np.random.default_rng(seed + 1)
    kps = [2.0 + 4.0 * n / max(nsteps - 1, 1) for n in range(nsteps)]
    scales = np.exp(0.3 * rng.standard_normal(nsteps))
    return list(zip(kps, scales.tolist()))


def satellite_points(problem: Problem, npts: int, seed: int = 0, species: int = 1):
    """Seeded virtual-satellite points (S,I,J,K), 1-based, away from the
    boundary rows, and their omnidirectional-flux weights
    w[p, L] = 4*pi*WMU(L) / (FFACTOR(S,I,K,L) * FNHS(I,J,L)), so that
    omega_p = sum_L w[p,L] * F2(S,I,J,K,L) is the flux of src/ModRamRun.f90:214
    (as src/ModRamSats.f90:661 samples it)."""
    g = problem.grid
    rng = np.random.default_rng(seed)
    pts = np.zeros((npts, 4), dtype=np.int32)
    w = np.zeros((npts, g.npa))
    for p in range(npts):
        i = int(rng.integers(2, g.nr))          # I = 3..NR (0-based 2..NR-1)
        j = int(rng.integers(1, g.nt - 1))      # J = 2..NT-1
        k = int(rng.integers(1, g.ne - 1))      # K = 2..NE-1
        pts[p] = (species, i + 1, j + 1, k + 1)
        ff = g.ffactor[species - 1, i, k, :]
        w[p] = 4.0 * np.pi * g.wmu / (ff * problem.geom.fnhs[i, j, :])
    return pts, w


def satellite_flux(f2: np.ndarray, pts: np.ndarray, w: np.ndarray) -> np.ndarray:
    """omega_p of satellite_points for a Fortran-layout F2 (NS,NR,NT,NE,NPA)."""
    return np.array([np.dot(w[p], f2[s - 1, i - 1, j - 1, k - 1, :]) for p, (s, i, j, k) in enumerate(pts)])
