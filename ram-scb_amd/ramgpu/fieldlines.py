"""Field lines on the RAM grid, the input of computehI.

computehI (src/ModRamScb.f90:144-643) first maps field lines onto the RAM
grid (xRAM, yRAM, zRAM, bRAM, each (nthe, NR, NT)) and then integrates along
them.  With the dipole boundary ('DIPL', the reference's default) the lines
come from get_dipole_lines (src/ModScbFunctions.f90:13-55) with RAM=.true.,
B scaled by b0dip/bnormal (src/ModRamScb.f90:244-247), and the line
coordinate is chiVal = thetaVal + constTheta*sin(2*thetaVal)
(src/ModScbIO.f90:72-80, thetaVal from src/ModScbInit.f90:278-280).  Other
boundaries trace the lines through the SCB equilibrium or Geopack on the
host; they reach the library in the same form.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

B0DIP = 30570.0                              # src/ModRamConst.f90:23
XZERO = 6.6                                  # src/ModScbInit.f90:246
BNORMAL = 0.31 / XZERO ** 3 * 1.E5           # src/ModScbInit.f90:264
NTHE = 101                                   # src/ModScbGrids.f90:14
CONST_THETA = 0.2                            # src/ModScbParams.f90:40


@dataclass
class Lines:
    x: np.ndarray      # (nthe, NR, NT) [RE]
    y: np.ndarray
    z: np.ndarray
    b: np.ndarray      # [units of bnormal]
    chi: np.ndarray    # (nthe) line coordinate
    ieq: int           # equator sample, 1-based (nThetaEquator)
    bnormal: float = BNORMAL


def dipole_lines(lz, nt: int, nthe: int = NTHE, const_theta: float = CONST_THETA) -> Lines:
    """get_dipole_lines(radiusMin, radiusMax, constTheta, nthe, NR, NT, ...,
    RAM=.true.) with the RAM rows LZ(2..NR+1) as its radii, B in units of
    bnormal."""
    lz = np.asarray(lz, dtype=np.float64)
    nr = len(lz) - 1
    rmin, rmax = lz[1], lz[nr]
    shape = (nthe, nr, nt)
    x, y, z, b = (np.zeros(shape, order="F") for _ in range(4))
    pi = math.pi
    for i in range(nr):
        r0 = rmin + float(i) / float(nr - 1) * (rmax - rmin)
        t1 = math.asin(math.sqrt(1.0 / r0))
        t0 = pi - t1
        for j in range(1, nt):
            zt = 2.0 * pi * float(j) / float(nt - 1) - pi
            for k in range(nthe):
                tt = t0 + float(k) / float(nthe - 1) * (t1 - t0)
                tt = tt + const_theta * math.sin(2.0 * tt)
                rt = r0 * math.sin(tt) ** 2
                x[k, i, j] = rt * math.cos(zt) * math.sin(tt)
                y[k, i, j] = rt * math.sin(zt) * math.sin(tt)
                z[k, i, j] = rt * math.cos(tt)
                b[k, i, j] = math.sqrt(1 + 3 * math.cos(tt) ** 2) / rt ** 3
    for a in (x, y, z, b):
        a[:, :, 0] = a[:, :, nt - 1]
    b *= B0DIP / BNORMAL
    theta = np.array([pi * float(k) / float(nthe - 1) for k in range(nthe)])
    chi = theta + const_theta * np.sin(2.0 * theta)
    return Lines(x, y, z, b, chi, (nthe + 1) // 2)
