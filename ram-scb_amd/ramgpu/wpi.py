"""Electron wave-particle lifetime tables of WAVEPARA1 and WAVEPARA2
(src/ModRamWPI.f90:18-126,132-182), computed on the host once per grid as the
reference does at initialisation, and the plasmapause of WAVELO
(:596-603).

WALOS1(I,K): lifetime [s] inside the plasmasphere, a log-log interpolation
in (L, E) of the published table of lifetimes in days at E = 0.2..2 MeV and
L = 5..1.65 (extrapolated linearly in log beyond the table).  At L = 5
exactly (I = 14 of the NR = 20 grid) the reference's L interpolation assigns
nothing and keeps the previous iteration's values; so does this one.
WALOS2(I,K): strong-diffusion-limited lifetime [s] outside it,
1/min(0.08 E^-1.32, 0.4*10^(2L-6+0.4 log10(29 E))) days, E in MeV.
WALOS3(I,K): the strong-diffusion lifetime 64 L RE / 35 / (1-0.25) /
sin^2(cone) / V [s] at the loss cone of the grid.
All (NR, NE) arrays; rows I=1 and columns K=1 stay 0 as in the reference.
"""
from __future__ import annotations

import math

import numpy as np

# published lifetimes [days] (src/ModRamWPI.f90:29-69): rows E = rEa, columns L = rL
R_EA = (0.2, 0.5, 1.0, 1.5, 2.0)
R_L = (5.0, 4.5, 4.0, 3.5, 3.0, 2.5, 2.0, 1.65)
R_LIFE = (
    (6.80, 16.44, 13.75, 17.38, 53.08, 187.06, 93.72, 101571.57),
    (23.38, 55.98, 43.43, 31.75, 38.20, 104.90, 164.86, 185.67),
    (343.16, 475.15, 99.87, 62.46, 98.82, 134.95, 171.96, 73.63),
    (619.62, 356.89, 139.64, 130.32, 210.25, 283.46, 359.03, 159.19),
    (1062.13, 381.88, 210.37, 231.97, 370.61, 498.14, 638.07, 473.75),
)


def _clife_at(xl: float, prev=None) -> list:
    """Lifetimes [days] at the five table energies, interpolated in L
    (log10 of the lifetime linear in L), src/ModRamWPI.f90:75-99.  At L = 5
    exactly no branch of the reference assigns them, so its values from the
    previous (K, I) iteration stand: prev."""
    rl, lf = R_L, R_LIFE
    cl = list(prev) if prev is not None else [0.0] * 5
    if 1.65 <= xl <= 5.0:
        for i in range(7, 0, -1):      # reference i = 8..2
            if rl[i] <= xl < rl[i - 1]:
                for j in range(5):
                    c = (math.log10(lf[j][i - 1]) - math.log10(lf[j][i])) / (rl[i - 1] - rl[i]) * (xl - rl[i]) \
                        + math.log10(lf[j][i])
                    cl[j] = 10.0 ** c
                break
    elif xl > 5.0:
        for j in range(5):
            c = (math.log10(lf[j][0]) - math.log10(lf[j][1])) / (rl[0] - rl[1]) * (xl - rl[0]) + math.log10(lf[j][0])
            cl[j] = 10.0 ** c
    else:
        for j in range(5):
            c = (math.log10(lf[j][6]) - math.log10(lf[j][7])) / (rl[6] - rl[7]) * (xl - rl[7]) + math.log10(lf[j][7])
            cl[j] = 10.0 ** c
    return cl


def _xlife(xe: float, cl: list) -> float:
    """Lifetime [days] at xe MeV, log-log interpolation over the table
    energies, src/ModRamWPI.f90:101-119."""
    ea = R_EA
    if 0.2 <= xe < 2.0:
        for i in range(4):
            if ea[i] <= xe < ea[i + 1]:
                x = (math.log10(cl[i + 1]) - math.log10(cl[i])) / (math.log10(ea[i + 1]) - math.log10(ea[i])) \
                    * (math.log10(xe) - math.log10(ea[i])) + math.log10(cl[i])
                return 10.0 ** x
        raise AssertionError("unreachable")
    if xe < 0.2:
        x = (math.log10(cl[1]) - math.log10(cl[0])) / (math.log10(ea[1]) - math.log10(ea[0])) \
            * (math.log10(xe) - math.log10(ea[0])) + math.log10(cl[0])
        return 10.0 ** x
    x = (math.log10(cl[4]) - math.log10(cl[3])) / (math.log10(ea[4]) - math.log10(ea[3])) \
        * (math.log10(xe) - math.log10(ea[4])) + math.log10(cl[4])
    return 10.0 ** x


def wavepara1(grid) -> np.ndarray:
    """WALOS1(NR,NE) [s], src/ModRamWPI.f90:71-124."""
    w = np.zeros((grid.nr, grid.ne), order="F")
    cl = None
    for k in range(1, grid.ne):
        for i in range(1, grid.nr):
            cl = _clife_at(float(grid.lz[i]), cl)
            w[i, k] = _xlife(grid.ekev[k] / 1000.0, cl) * 60.0 * 60.0 * 24.0
    return w


def wavepara2(grid, s: int):
    """WALOS2, WALOS3 (NR,NE) [s] for species s (1-based; WALOS3 uses its
    speeds V(S,K)), src/ModRamWPI.f90:148-179."""
    from .grid import PI, RE
    nr, ne = grid.nr, grid.ne
    w2 = np.zeros((nr, ne), order="F")
    w3 = np.zeros((nr, ne), order="F")
    for k in range(1, ne):
        for i in range(1, nr):
            emev = grid.ekev[k] * 0.001
            r1 = 0.08 * emev ** (-1.32)
            r2 = 0.4 * 10.0 ** (2.0 * grid.lz[i] - 6.0 + 0.4 * math.log10(29.0 * emev))
            tau = 1.0 / min(r1, r2)
            w2[i, k] = tau * 60.0 * 60.0 * 24.0
    for i in range(1, nr):
        sc = math.sin(grid.cone[i] * PI / 180.0)
        for k in range(1, ne):
            w3[i, k] = 64.0 * grid.lz[i] * RE / 35.0 / (1 - 0.25) / sc / sc / grid.v[s - 1, k]
    return w2, w3


def plasmapause(grid, kpmax12: float, necr=None, use_plasmasphere: bool = False) -> np.ndarray:
    """RLpp(NT): the plasmapause L of WAVELO, 5.39 - 0.382 Kpmax12 (Moldwin
    et al. 2002), or with a plasmasphere model the largest L (I=2..NR) whose
    density NECR(I,J) exceeds 50 cm^-3 (src/ModRamWPI.f90:598-603)."""
    rl = np.full(grid.nt, 5.39 - 0.382 * kpmax12)
    if use_plasmasphere and necr is not None:
        for j in range(grid.nt):
            for i in range(1, grid.nr):
                if necr[i, j] > 50.0:
                    rl[j] = grid.lz[i]
    return rl
