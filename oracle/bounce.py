"""CPU restatement of the SCB->RAM bounce integrals, computehI
(src/ModRamScb.f90:144-643) from field lines already on the RAM grid
(TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the device path).

The reference evaluates its integrals with GSL (not vendored; any GSL 2.x):
``gsl_integration_cquad`` at 1e-3 relative and absolute tolerance of
integrands built on ``gsl_interp_linear`` (src/RamGSL.c:295-478), and
``gsl_interp_steffen`` for the pitch-angle interpolation (src/RamGSL.c:111-175
through src/ModRamGSL.f90:241-312).  This restatement integrates the same
integrands exactly: on a segment where B and the bounce-averaged variable are
linear in the field-line coordinate, sqrt(Bm-B), 1/sqrt(Bm-B) and
v/sqrt(Bm-B) have closed-form integrals, summed over the segments in order.
The reference's quadrature approximates these values to its 1e-3 tolerance.
The Steffen interpolant follows Steffen (1990, A&A 239:443), eqs. 2-11,
with the one-sided end slopes of its section 2.2, as GSL implements it.

Array conventions: field lines (NTHE, NR, NT) as the reference's
xRAM(k,i,j) with i = 1..NR the RAM rows I = 2..NR+1; results in the
reference's (NR+1, NT[, NPA]) geometry shapes.  All indices below are
0-based.
"""
from __future__ import annotations

import math

import numpy as np

# Rairden geocorona (src/ModRamScb.f90:363-370)
RAIRDEN = (13.326, -3.6908, 1.1362, -0.16984, 0.009553)


def rairden(d):
    """Geocoronal H density [m^-3] at geocentric distance d [RE]; powers of
    d as products, 10**x through pow."""
    d2 = d * d
    e = (((RAIRDEN[0] + RAIRDEN[1] * d) + RAIRDEN[2] * d2) + RAIRDEN[3] * (d2 * d)) + RAIRDEN[4] * (d2 * d2)
    return np.power(10.0, e)


def segment_integrals(c0, c1, b0, b1, v0, v1, m):
    """Exact integrals over [c0, c1] of sqrt(m-B), 1/sqrt(m-B) and
    v/sqrt(m-B) (zero where B >= m) for B and v linear from (b0, v0) to
    (b1, v1); elementwise over arrays.  With s, e the square roots of m-B at
    the ends of the part where B < m, of length dl:
      int 1/sqrt = 2 dl/(s+e),  int sqrt = 2 dl (s^2+se+e^2)/(3(s+e)),
      int v/sqrt = vs int 1/sqrt + (ve-vs) 2 dl (2s+e)/(3(s+e)^2)."""
    c0, c1, b0, b1, v0, v1, m = np.broadcast_arrays(*(np.asarray(a, dtype=np.float64)
                                                      for a in (c0, c1, b0, b1, v0, v1, m)))
    u0 = m - b0
    u1 = m - b1
    d = c1 - c0
    both = (u0 > 0) & (u1 > 0)
    lo = (u0 > 0) & ~(u1 > 0)
    hi = ~(u0 > 0) & (u1 > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        t_lo = u0 / (u0 - u1)
        t_hi = u1 / (u1 - u0)
        dl = np.where(both, d, np.where(lo, d * t_lo, d * t_hi))
        us = np.where(hi, 0.0, u0)
        ue = np.where(lo, 0.0, u1)
        vs = np.where(hi, v1 + (v0 - v1) * t_hi, v0)
        ve = np.where(lo, v0 + (v1 - v0) * t_lo, v1)
        s = np.sqrt(us)
        e = np.sqrt(ue)
        den = s + e
        jh = (2.0 * dl) / den
        ji = ((2.0 * dl) * ((us + s * e) + ue)) / (3.0 * den)
        jd = vs * jh + (ve - vs) * (((2.0 * dl) * (2.0 * s + e)) / ((3.0 * den) * den))
    any_ = both | lo | hi
    z = np.zeros_like(d)
    return np.where(any_, ji, z), np.where(any_, jh, z), np.where(any_, jd, z)


def line_integrals(chi, bf, var, mirror):
    """integrator_c and bounceaverage_c (src/RamGSL.c:480-602) of one line:
    yI, yH, yD per pitch angle and the mirror fields as the reference leaves
    them (pitch angles whose mirror point falls off the line take the last
    one's).  chi, bf, var (NTHE); mirror (NPA)."""
    n, npa = len(chi), len(mirror)
    m = np.array(mirror, dtype=np.float64)
    ia = np.zeros(npa, dtype=np.int64)  # bracket [a, b] as knot indices
    ib = np.zeros(npa, dtype=np.int64)
    code = np.zeros(npa, dtype=np.int64)  # 0 integrate, 1 last pitch angle's values, 2 the next one's
    for L in range(1, npa - 1):
        a = b = 0.0
        lh = rh = 0
        for i in range(1, n - 1):  # :550-556
            if m[L] <= bf[i - 1] and m[L] >= bf[i]:
                a, lh = chi[i - 1], i - 1
                break
        for i in range(n - 2, 0, -1):  # :557-563
            if m[L] >= bf[i - 1] and m[L] <= bf[i]:
                b, rh = chi[i], i
                break
        if m[L] >= bf[1] or a == 0:  # :573-574
            a, lh = chi[0], 0
        if m[L] >= bf[n - 1] or b == 0:
            b, rh = chi[n - 1], n - 1
        if a <= chi[0] or b >= chi[n - 1]:
            code[L] = 1
        elif rh - lh <= 4:
            code[L] = 2
        ia[L], ib[L] = lh, rh
    ia[npa - 1], ib[npa - 1] = 0, n - 1
    # the integrals, segments added in order along the line
    yI = np.zeros(npa)
    yH = np.zeros(npa)
    yD = np.zeros(npa)
    todo = (code == 0)
    todo[0] = False
    todo[npa - 1] = True
    for k in range(n - 1):
        on = todo & (ia <= k) & (k < ib)
        if not on.any():
            continue
        ji, jh, jd = segment_integrals(chi[k], chi[k + 1], bf[k], bf[k + 1], var[k], var[k + 1], m)
        yI = np.where(on, yI + ji, yI)
        yH = np.where(on, yH + jh, yH)
        yD = np.where(on, yD + jd, yD)
    for L in range(npa - 2, 0, -1):  # :566-594
        if code[L] == 1:
            m[L] = m[npa - 1]
            yI[L], yH[L], yD[L] = yI[npa - 1], yH[npa - 1], yD[npa - 1]
        elif code[L] == 2:
            yI[L], yH[L], yD[L] = yI[L + 1], yH[L + 1], yD[L + 1]
        else:
            if yI[L] <= 0:
                yI[L] = yI[L + 1]
            if yH[L] <= 0:
                yH[L] = yH[L + 1]
            if yD[L] <= 0:
                yD[L] = yD[L + 1]
    yI[0] = 0.0
    yH[0] = yH[1]
    yD[0] = yD[1]
    return yI, yH, yD, m


def steffen(xa, fa, xb):
    """Steffen's monotone cubic through (xa, fa) at xb, with the reference's
    linear extrapolation outside [xa[0], xa[-1]] (src/RamGSL.c:158-166).
    Slopes: eq. 11 inside, the end secants at the ends; coefficients eqs. 2-5;
    Horner evaluation on the interval xa[i] <= x < xa[i+1]."""
    n = len(xa)
    yp = np.zeros(n)
    yp[0] = (fa[1] - fa[0]) / (xa[1] - xa[0])
    for i in range(1, n - 1):
        hi, him1 = xa[i + 1] - xa[i], xa[i] - xa[i - 1]
        si, sim1 = (fa[i + 1] - fa[i]) / hi, (fa[i] - fa[i - 1]) / him1
        p = (sim1 * hi + si * him1) / (him1 + hi)
        sg = (-1.0 if sim1 < 0 else 1.0) + (-1.0 if si < 0 else 1.0)
        yp[i] = sg * min(abs(sim1), min(abs(si), 0.5 * abs(p)))
    yp[n - 1] = (fa[n - 1] - fa[n - 2]) / (xa[n - 1] - xa[n - 2])
    out = np.zeros(len(xb))
    for q, x in enumerate(xb):
        if x <= xa[0]:
            out[q] = fa[0] + (x - xa[0]) / (xa[1] - xa[0]) * (fa[1] - fa[0])
        elif x >= xa[n - 1]:
            out[q] = fa[n - 1] + (x - xa[n - 1]) / (xa[n - 2] - xa[n - 1]) * (fa[n - 2] - fa[n - 1])
        else:
            i = int(np.searchsorted(xa, x, side="right")) - 1
            h = xa[i + 1] - xa[i]
            s = (fa[i + 1] - fa[i]) / h
            a = (yp[i] + yp[i + 1] - 2.0 * s) / h / h
            b = (3.0 * s - 2.0 * yp[i] - yp[i + 1]) / h
            dx = x - xa[i]
            out[q] = fa[i] + dx * (yp[i] + dx * (b + dx * a))
    return out


def gaussian_weights(sigma=1.0, truncate=4.0):
    """gaussian_kernel (srcExternal/gaussian_filter.f90:18-56): radius
    int(truncate*sigma+0.5), w = 2 exp(-0.5 (x^2+y^2)/sigma^2), normalised
    by its sum (column-major order).  w[di+r, dj+r], di the row offset."""
    r = int(int(truncate) * sigma + 0.5)
    s = sigma * sigma
    w = np.zeros((2 * r + 1, 2 * r + 1))
    for dj in range(-r, r + 1):
        for di in range(-r, r + 1):
            w[di + r, dj + r] = 2.0 * math.exp((-0.5 * float(di * di + dj * dj)) / s)
    tot = 0.0
    for dj in range(2 * r + 1):
        for di in range(2 * r + 1):
            tot += w[di, dj]
    return w / tot


def smooth2d(a, w):
    """convolve (srcExternal/gaussian_filter.f90:104-178) without a mask:
    edges by whole-array reflection (d c b a | a b c d), each output the sum
    of weight times input over the window in column-major order."""
    rows, cols = a.shape
    r = w.shape[0] // 2
    if w.shape[0] >= rows + 1 or w.shape[1] >= cols + 1:
        raise ValueError("Input size too small for weights matrix")

    def refl(p, n):
        return np.where(p < 0, -1 - p, np.where(p >= n, 2 * n - 1 - p, p))
    ii = np.arange(rows)[:, None]
    jj = np.arange(cols)[None, :]
    acc = np.zeros_like(a)
    for dj in range(-r, r + 1):
        for di in range(-r, r + 1):
            acc = acc + w[di + r, dj + r] * a[refl(ii + di, rows), refl(jj + dj, cols)]
    return acc


def computehI(x, y, z, b, chi, ieq, mu, pa, pabn, lz, outside, scale_at, bnormal, dthi, smooth, prev):
    """computehI after its field-line step (src/ModRamScb.f90:358-629):
    integrals of every line inside the magnetopause, the scaling of rows
    beyond the SCB domain (scale_at[j] = first such row, 1-based as the
    reference's ScaleAt; 0 none), MLT continuity, the near-90-degree and
    monotonicity corrections, Steffen interpolation to the boundary pitch
    angles, optional Gaussian smoothing, and the RAM variables with their
    time derivatives.  prev holds the geometry before the call (fnhs, fnis,
    bounhs, bounis, bnes); returns the new geometry and the line values."""
    nthe, nr, nt = b.shape
    npa = len(mu)
    b = np.array(b, dtype=np.float64, order="F", copy=True)
    hC = np.zeros((nr, nt, npa))
    iC = np.zeros((nr, nt, npa))
    dC = np.zeros((nr, nt, npa))
    bz = np.zeros((nr, nt))
    for j in range(nt):
        for i in range(nr):
            if outside[i, j] != 0:
                continue
            xl, yl, zl, bl = x[:, i, j], y[:, i, j], z[:, i, j], b[:, i, j]
            dist = np.sqrt((xl * xl + yl * yl) + zl * zl)
            dens = rairden(dist)
            length = 0.0
            for k in range(1, nthe):  # :379-384
                dx, dy, dz = xl[k] - xl[k - 1], yl[k] - yl[k - 1], zl[k] - zl[k - 1]
                length += math.sqrt((dx * dx + dy * dy) + dz * dz)
            r0 = math.sqrt(xl[ieq] * xl[ieq] + yl[ieq] * yl[ieq])
            bmin = bl.min()
            if abs(bl[ieq] - bmin) > 1e-9:  # :386-392
                bl[ieq] = 2.0 * bmin - bl[ieq] if 2.0 * bmin - bl[ieq] > 0 else bmin - 0.01
            mirror = np.empty(npa)
            mirror[:npa - 1] = bl[ieq] / (1.0 - mu[:npa - 1] * mu[:npa - 1])
            mirror[npa - 1] = bl[nthe - 1]
            yI, yH, yD, m = line_integrals(chi, bl, dens, mirror)
            sm = np.sqrt(m)
            iC[i, j] = (length / (math.pi * r0)) * yI / sm
            hC[i, j] = (length / (math.pi * 2.0 * r0)) * yH * sm
            dC[i, j] = yD / yH
            bz[i, j] = bl[ieq] * bnormal
    # rows beyond the SCB domain (:413-461); lz is LZ(1..NR+1), 0-based
    for j in range(1, nt):
        ii = int(scale_at[j])
        if ii == 0:
            continue
        if ii < 3:
            raise ValueError("scale_at must be 0 or at least 3")
        a0, a1, a2 = ii - 1, ii - 2, ii - 3  # cart rows ii, ii-1, ii-2 (0-based)
        fr = (lz[ii - 1] - lz[ii - 2]) / (lz[ii - 3] - lz[ii - 2])  # (Lz(ii)-Lz(ii-1))/(Lz(ii-2)-Lz(ii-1))
        for L in range(1, npa):
            sc = []
            for X in (iC, hC, dC):
                t = X[a1, j, L] + fr * (X[a2, j, L] - X[a1, j, L])
                sc.append(X[a1, j, L] / X[a0, j, L] if t <= 0 else t / X[a0, j, L])
            for i in range(a0, nr):
                if outside[i, j] == 0:
                    iC[i, j, L] *= sc[0]
                    hC[i, j, L] *= sc[1]
                    dC[i, j, L] *= sc[2]
                else:
                    iC[i, j, L] = iC[i - 1, j, L]
                    hC[i, j, L] = hC[i - 1, j, L]
                    dC[i, j, L] = dC[i - 1, j, L]
                bz[i, j] = bz[i - 1, j]
    for X in (iC, hC, dC):  # :464-467
        X[:, 0, :] = X[:, nt - 1, :]
    bz[:, 0] = bz[:, nt - 1]
    iC[:, :, 2] = 0.50 * iC[:, :, 3]  # :470-478
    iC[:, :, 1] = 0.20 * iC[:, :, 2]
    iC[:, :, 0] = 0.0
    for X, f in ((hC, 0.99), (dC, 0.999)):
        X[:, :, 2] = f * X[:, :, 3]
        X[:, :, 1] = f * X[:, :, 2]
        X[:, :, 0] = f * X[:, :, 1]
    for i in range(1, nr):  # :481-500 (only negative cells change)
        for X in (hC, iC, dC):
            neg = X[i] < 0
            X[i][neg] = X[i - 1][neg]
    for L in range(npa - 2, -1, -1):  # :502-510
        for X, f in ((iC, 0.99), (hC, 0.99), (dC, 0.999)):
            bad = X[:, :, L] > X[:, :, L + 1]
            X[:, :, L] = np.where(bad, f * X[:, :, L + 1], X[:, :, L])
    hB = np.zeros_like(hC)
    iB = np.zeros_like(iC)
    xa = pa[::-1]
    xb = pabn[npa - 2:0:-1]  # PAbn(NPA-1:2:-1)
    for j in range(nt):  # :520-533
        for i in range(nr):
            for X, Y in ((hC, hB), (iC, iB)):
                Y[i, j, npa - 2:0:-1] = steffen(xa, X[i, j, ::-1], xb)
                Y[i, j, npa - 1] = Y[i, j, npa - 2]
                Y[i, j, 0] = Y[i, j, 1]
    if smooth:  # :539-562
        w = gaussian_weights(1.0)
        for L in range(1, npa):
            for X in (hC, iC, hB, iB, dC):
                X[:, :, L] = smooth2d(X[:, :, L], w)
    out = {"h": hC, "i": iC, "hb": hB, "ib": iB, "d": dC, "bz": bz}
    g3 = (nr + 1, nt, npa)
    fnhs, fnis, bounhs, bounis, hdns = (np.zeros(g3, order="F") for _ in range(5))
    didt, dibndt = np.zeros(g3, order="F"), np.zeros(g3, order="F")
    bnes, dbdt = np.zeros((nr + 1, nt), order="F"), np.zeros((nr + 1, nt), order="F")
    fnhs[1:], fnis[1:], bounhs[1:], bounis[1:], hdns[1:] = hC, iC, hB, iB, dC
    bnes[1:] = bz
    if abs(dthi) > 1e-9:  # :580-588
        didt[1:] = (fnis[1:] - prev["fnis"][1:]) / dthi
        dibndt[1:] = (bounis[1:] - prev["bounis"][1:]) / dthi
    bnes[1:] = bnes[1:] / 1e9
    if abs(dthi) > 1e-9:
        dbdt[1:] = (bnes[1:] - prev["bnes"][1:]) / dthi
    bnes[0, :] = 0.32 / (lz[0] * lz[0] * lz[0]) / 1.e4  # :598-611
    for X in (fnhs, fnis, bounhs, bounis, hdns):
        X[0] = X[1]
    for i in range(1, nr + 1):  # :614-626
        for X in (fnis, fnhs, bounis, bounhs, hdns):
            nan = np.isnan(X[i])
            X[i][nan] = X[i - 1][nan]
        for X in (didt, dibndt):
            X[i][np.isnan(X[i])] = 0.0
    out.update(fnhs=fnhs, fnis=fnis, bounhs=bounhs, bounis=bounis, hdns=hdns, didt=didt, dibndt=dibndt,
               bnes=bnes, dbdt=dbdt)
    return out
