// bounce.hip -- the SCB->RAM bounce integrals, computehI
// (src/ModRamScb.f90:144-643), from field lines already on the RAM grid:
// the h and I integrals and the bounce-averaged geocoronal density of every
// line, the reference's corrections, the interpolation to the boundary
// pitch angles, the optional Gaussian smoothing, and the RAM geometry with
// its time derivatives, written straight into the context's geometry.
//
// The reference integrates with GSL's cquad at 1e-3 tolerance an integrand
// built on GSL's linear interpolation of B (src/RamGSL.c:326-602).  Here the
// same integrands are integrated exactly: on a segment between two samples
// B and the density are linear in the line coordinate, so sqrt(Bm-B),
// 1/sqrt(Bm-B) and v/sqrt(Bm-B) have closed-form integrals, summed segment
// by segment along the line.  The pitch-angle interpolation is Steffen's
// monotone cubic (GSL's gsl_interp_steffen; Steffen 1990, A&A 239:443).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ctx.h"

using namespace ramgpu_host;

namespace {

constexpr int kHiThreads = 64;      // one wave per field line
constexpr int kHiLineThreads = 128; // two waves per line for the integrals (one pitch angle each up to 128)
constexpr int kSmoothR = 4;        // Gaussian kernel of sigma 1, truncated at 4 sigma
constexpr int kSmoothW = 2 * kSmoothR + 1;

struct HiArgs {
  int nthe, nr, nt, npa, ieq;  // ieq 0-based
  const double *x, *y, *z, *b;  // (NTHE, NR, NT)
  const double *chi, *mu, *pa, *pabn, *lz, *w;
  const int *mgnp, *scale_at;
  double bnormal, dthi;
  // line values (NR, NT, NPA) at [i + NR*(j + NT*L)], bz (NR, NT)
  double *hC, *iC, *dC, *bz, *hB, *iB;
  // their smoothed copies
  double *sh, *si, *shb, *sib, *sd;
  // the context's geometry, (NR+1, NT[, NPA]) Fortran layout
  double *fnhs, *fnis, *bounhs, *bounis, *hdns, *didt, *dibndt, *bnes, *dbdt, *eir, *eip;
  int *mgnp_dst;
  int smooth;
};

// Geocoronal H density [m^-3] at d [RE], the Rairden profile
// (src/ModRamScb.f90:363-370)
__device__ inline double rairden(double d) {
  const double d2 = d * d;
  const double e = (((13.326 + -3.6908 * d) + 1.1362 * d2) + -0.16984 * (d2 * d)) + 0.009553 * (d2 * d2);
  return pow(10.0, e);
}

// Integrals over one segment [c0, c1] of sqrt(m-B), 1/sqrt(m-B) and
// v/sqrt(m-B) (zero where B >= m), B and v linear between the samples.
// With s, e the square roots of m-B at the ends of the part below m, of
// length dl: 2dl/(s+e), 2dl(s^2+se+e^2)/(3(s+e)), and
// vs*2dl/(s+e) + (ve-vs)*2dl(2s+e)/(3(s+e)^2).  False if no part is below m.
__device__ inline bool segment(double c0, double c1, double b0, double b1, double v0, double v1, double m,
                               double &ji, double &jh, double &jd) {
  const double u0 = m - b0, u1 = m - b1, d = c1 - c0;
  const bool p0 = u0 > 0, p1 = u1 > 0;
  if (!p0 && !p1) return false;
  double dl, us, ue, vs, ve;
  if (p0 && p1) {
    dl = d; us = u0; ue = u1; vs = v0; ve = v1;
  } else if (p0) {  // B reaches m inside the segment
    const double t = u0 / (u0 - u1);
    dl = d * t; us = u0; ue = 0.0; vs = v0; ve = v0 + (v1 - v0) * t;
  } else {
    const double t = u1 / (u1 - u0);
    dl = d * t; us = 0.0; ue = u1; vs = v1 + (v0 - v1) * t; ve = v1;
  }
  const double s = sqrt(us), e = sqrt(ue), den = s + e;
  jh = (2.0 * dl) / den;
  ji = ((2.0 * dl) * ((us + s * e) + ue)) / (3.0 * den);
  jd = vs * jh + (ve - vs) * (((2.0 * dl) * (2.0 * s + e)) / ((3.0 * den) * den));
  return true;
}

// One field line per block of two waves: integrator_c and bounceaverage_c
// (src/RamGSL.c:480-602) and the line's I, h, density
// (src/ModRamScb.f90:376-404).
__global__ void __launch_bounds__(kHiLineThreads) k_hi_lines(HiArgs a) {
  extern __shared__ double hs[];
  const int nthe = a.nthe, npa = a.npa, tid = threadIdx.x;
  const int line = blockIdx.x, i = line % a.nr, j = line / a.nr;
  const size_t cell0 = (size_t)i + (size_t)a.nr * j, plane = (size_t)a.nr * a.nt;
  if (a.mgnp[cell0] != 0) {  // outside the magnetopause: no integrals
    for (int L = tid; L < npa; L += blockDim.x) {
      a.hC[cell0 + plane * L] = 0.0;
      a.iC[cell0 + plane * L] = 0.0;
      a.dC[cell0 + plane * L] = 0.0;
    }
    if (tid == 0) a.bz[cell0] = 0.0;
    return;
  }
  double *bs = hs, *cs = bs + nthe, *vs = cs + nthe, *ms = vs + nthe;
  double *yI = ms + npa, *yH = yI + npa, *yD = yH + npa, *sc = yD + npa;  // sc: length, r0
  double *sg = sc + 2;  // segment lengths
  int *code = (int *)(sg + nthe);
  const size_t base = (size_t)nthe * cell0;
  for (int k = tid; k < nthe; k += blockDim.x) {
    const double x = a.x[base + k], y = a.y[base + k], z = a.z[base + k];
    bs[k] = a.b[base + k];
    cs[k] = a.chi[k];
    vs[k] = rairden(sqrt((x * x + y * y) + z * z));
    if (k >= 1) {
      const double dx = x - a.x[base + k - 1], dy = y - a.y[base + k - 1], dz = z - a.z[base + k - 1];
      sg[k] = sqrt((dx * dx + dy * dy) + dz * dz);
    }
  }
  __syncthreads();
  if (tid == 0) {  // the line's length, summed along it
    double length = 0.0, bmin = bs[0];
    for (int k = 1; k < nthe; ++k) {
      length += sg[k];
      bmin = fmin(bmin, bs[k]);
    }
    const double xe = a.x[base + a.ieq], ye = a.y[base + a.ieq];
    sc[0] = length;
    sc[1] = sqrt(xe * xe + ye * ye);
    const double be = bs[a.ieq];
    if (fabs(be - bmin) > 1e-9) bs[a.ieq] = (2.0 * bmin - be > 0) ? 2.0 * bmin - be : bmin - 0.01;
  }
  __syncthreads();
  const double beq = bs[a.ieq];
  for (int L = tid; L < npa; L += blockDim.x) {
    const double m = (L < npa - 1) ? beq / (1.0 - a.mu[L] * a.mu[L]) : bs[nthe - 1];
    ms[L] = m;
    int cd = 0, lh = 0, rh = nthe - 1;
    if (L == 0) {
      cd = 3;  // 90 degrees: set from L=1 below
    } else if (L < npa - 1) {
      double av = 0.0, bv = 0.0;
      lh = rh = 0;
      for (int k = 1; k < nthe - 1; ++k)
        if (m <= bs[k - 1] && m >= bs[k]) {
          av = cs[k - 1];
          lh = k - 1;
          break;
        }
      for (int k = nthe - 2; k > 0; --k)
        if (m >= bs[k - 1] && m <= bs[k]) {
          bv = cs[k];
          rh = k;
          break;
        }
      if (m >= bs[1] || av == 0) av = cs[0];
      if (m >= bs[nthe - 1] || bv == 0) bv = cs[nthe - 1];
      if (av <= cs[0] || bv >= cs[nthe - 1])
        cd = 1;  // mirrors off the line: the last pitch angle's values
      else if (rh - lh <= 4)
        cd = 2;  // too few samples: the next pitch angle's values
    }
    double si = 0.0, sh = 0.0, sd = 0.0;
    if (cd == 0)
      for (int k = lh; k < rh; ++k) {
        double ji, jh, jd;
        if (segment(cs[k], cs[k + 1], bs[k], bs[k + 1], vs[k], vs[k + 1], m, ji, jh, jd)) {
          si += ji;
          sh += jh;
          sd += jd;
        }
      }
    yI[L] = si;
    yH[L] = sh;
    yD[L] = sd;
    code[L] = cd;
  }
  __syncthreads();
  if (tid == 0) {  // the reference's fallbacks, from the loss cone down (src/RamGSL.c:566-594)
    const int t = npa - 1;
    for (int L = npa - 2; L > 0; --L) {
      if (code[L] == 1) {
        ms[L] = ms[t];
        yI[L] = yI[t]; yH[L] = yH[t]; yD[L] = yD[t];
      } else if (code[L] == 2) {
        yI[L] = yI[L + 1]; yH[L] = yH[L + 1]; yD[L] = yD[L + 1];
      } else {
        if (yI[L] <= 0) yI[L] = yI[L + 1];
        if (yH[L] <= 0) yH[L] = yH[L + 1];
        if (yD[L] <= 0) yD[L] = yD[L + 1];
      }
    }
    yI[0] = 0.0;
    yH[0] = yH[1];
    yD[0] = yD[1];
  }
  __syncthreads();
  const double length = sc[0], r0 = sc[1];
  for (int L = tid; L < npa; L += blockDim.x) {
    const double sm = sqrt(ms[L]);
    a.iC[cell0 + plane * L] = (length / (M_PI * r0)) * yI[L] / sm;
    a.hC[cell0 + plane * L] = (length / (M_PI * 2.0 * r0)) * yH[L] * sm;
    a.dC[cell0 + plane * L] = yD[L] / yH[L];
  }
  if (tid == 0) a.bz[cell0] = beq * a.bnormal;
}

// Steffen's interpolant at x of one line's f (in L order) over the pitch
// angles pa.  The reference interpolates over PA(NPA:1:-1)
// (src/ModRamScb.f90:520-533), so xa and fa are pa and f read reversed;
// slopes by Steffen's eq. 11 inside and the end secants at the ends,
// coefficients by eqs. 2-5, Horner evaluation on xa(r) <= x < xa(r+1), and
// linear extrapolation outside (src/RamGSL.c:158-166).
__device__ double steffen_at(const double *pa, const double *f, int n, double x) {
  auto XA = [&](int r) { return pa[n - 1 - r]; };
  auto FA = [&](int r) { return f[n - 1 - r]; };
  auto slope = [&](int r) -> double {
    if (r == 0) return (FA(1) - FA(0)) / (XA(1) - XA(0));
    if (r == n - 1) return (FA(n - 1) - FA(n - 2)) / (XA(n - 1) - XA(n - 2));
    const double hi = XA(r + 1) - XA(r), him1 = XA(r) - XA(r - 1);
    const double si = (FA(r + 1) - FA(r)) / hi, sim1 = (FA(r) - FA(r - 1)) / him1;
    const double p = (sim1 * hi + si * him1) / (him1 + hi);
    const double sg = (sim1 < 0 ? -1.0 : 1.0) + (si < 0 ? -1.0 : 1.0);
    return sg * fmin(fabs(sim1), fmin(fabs(si), 0.5 * fabs(p)));
  };
  if (x <= XA(0)) return FA(0) + (x - XA(0)) / (XA(1) - XA(0)) * (FA(1) - FA(0));
  if (x >= XA(n - 1)) return FA(n - 1) + (x - XA(n - 1)) / (XA(n - 2) - XA(n - 1)) * (FA(n - 2) - FA(n - 1));
  int r = 0;
  while (r + 1 < n - 1 && XA(r + 1) <= x) ++r;
  const double y0 = slope(r), y1 = slope(r + 1);
  const double h = XA(r + 1) - XA(r);
  const double s = (FA(r + 1) - FA(r)) / h;
  const double ca = (y0 + y1 - 2.0 * s) / h / h;
  const double cb = (3.0 * s - 2.0 * y0 - y1) / h;
  const double dx = x - XA(r);
  return FA(r) + dx * (y0 + dx * (cb + dx * ca));
}

// The corrections of src/ModRamScb.f90:406-533, in the reference's order,
// one launch per stage: rows beyond the SCB domain, MLT continuity with the
// near-90-degree values, negative values, then monotonicity in pitch angle
// with the interpolation to PAbn.

// Rows from ScaleAt(j) on (:413-461): one thread per (j, L), j = 2..NT,
// L = 2..NPA, walking the rows.
__global__ void k_hi_scale(HiArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nt * a.npa) return;
  const int nr = a.nr, j = t % a.nt, L = t / a.nt, ii = a.scale_at[j];
  if (j == 0 || L == 0 || ii == 0) return;
  const size_t plane = (size_t)nr * a.nt;
  double *const XS[3] = {a.iC, a.hC, a.dC};
  auto at = [&](int i) { return (size_t)i + (size_t)nr * j + plane * L; };
  const int a0 = ii - 1, a1 = ii - 2, a2 = ii - 3;
  const double fr = (a.lz[ii - 1] - a.lz[ii - 2]) / (a.lz[ii - 3] - a.lz[ii - 2]);
  double s[3];
  for (int q = 0; q < 3; ++q) {
    const double *X = XS[q];
    const double tmp = X[at(a1)] + fr * (X[at(a2)] - X[at(a1)]);
    s[q] = (tmp <= 0) ? X[at(a1)] / X[at(a0)] : tmp / X[at(a0)];
  }
  for (int i = a0; i < nr; ++i) {
    for (int q = 0; q < 3; ++q) {
      double *X = XS[q];
      X[at(i)] = (a.mgnp[i + (size_t)nr * j] == 0) ? X[at(i)] * s[q] : X[at(i - 1)];
    }
    if (L == 1) a.bz[i + (size_t)nr * j] = a.bz[i - 1 + (size_t)nr * j];
  }
}

// Continuity across MLT 0 (:464-467) and the near-90-degree values
// (:470-478): one thread per line.  Column 1 takes column NT's L >= 4 (the
// corrections rewrite L = 1..3 of every line from its L = 4, so copying
// first and correcting after equals the reference's order).
__global__ void k_hi_near90(HiArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, nr = a.nr, nt = a.nt;
  if (t >= nr * nt) return;
  const size_t plane = (size_t)nr * nt, c = t;
  double *I = a.iC, *H = a.hC, *D = a.dC;
  if (t < nr) {
    const size_t src = c + (size_t)nr * (nt - 1);
    for (int L = 3; L < a.npa; ++L) {
      I[c + plane * L] = I[src + plane * L];
      H[c + plane * L] = H[src + plane * L];
      D[c + plane * L] = D[src + plane * L];
    }
    a.bz[c] = a.bz[src];
  }
  I[c + 2 * plane] = 0.50 * I[c + 3 * plane];
  I[c + plane] = 0.20 * I[c + 2 * plane];
  I[c] = 0.0;
  H[c + 2 * plane] = 0.99 * H[c + 3 * plane];
  H[c + plane] = 0.99 * H[c + 2 * plane];
  H[c] = 0.99 * H[c + plane];
  D[c + 2 * plane] = 0.999 * D[c + 3 * plane];
  D[c + plane] = 0.999 * D[c + 2 * plane];
  D[c] = 0.999 * D[c + plane];
}

// Negative values take the row below's (:481-500): one thread per (j, L).
__global__ void k_hi_negative(HiArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nt * a.npa) return;
  const int nr = a.nr, j = t % a.nt, L = t / a.nt;
  const size_t b0 = (size_t)nr * j + (size_t)nr * a.nt * L;
  double *const XS[3] = {a.hC, a.iC, a.dC};
  for (int i = 1; i < nr; ++i)
    for (int q = 0; q < 3; ++q) {
      double *X = XS[q] + b0;
      if (X[i] < 0) X[i] = X[i - 1];
    }
}

// No value above the next pitch angle's (:502-510), then the interpolation
// to PAbn (:520-533): one wave per line, the line's tables in LDS; one lane
// per table walks the pitch angles down, then the lanes interpolate one L
// each, the ends copied.
__global__ void __launch_bounds__(kHiThreads) k_hi_pitch(HiArgs a) {
  extern __shared__ double ls[];
  const int n = a.npa, tid = threadIdx.x;
  const size_t plane = (size_t)a.nr * a.nt, c = blockIdx.x;
  double *I = ls, *H = I + n, *D = H + n, *P = D + n;
  for (int L = tid; L < n; L += blockDim.x) {
    I[L] = a.iC[c + plane * L];
    H[L] = a.hC[c + plane * L];
    D[L] = a.dC[c + plane * L];
    P[L] = a.pa[L];
  }
  __syncthreads();
  if (tid < 3) {
    double *X = (tid == 0) ? I : (tid == 1 ? H : D);
    const double f = (tid == 2) ? 0.999 : 0.99;
    for (int L = n - 2; L >= 0; --L)
      if (X[L] > X[L + 1]) X[L] = f * X[L + 1];
  }
  __syncthreads();
  for (int L = tid; L < n; L += blockDim.x) {
    a.iC[c + plane * L] = I[L];
    a.hC[c + plane * L] = H[L];
    a.dC[c + plane * L] = D[L];
  }
  for (int L = 1 + tid; L <= n - 2; L += blockDim.x) {
    const double x = a.pabn[L];
    const double hb = steffen_at(P, H, n, x), ib = steffen_at(P, I, n, x);
    a.hB[c + plane * L] = hb;
    a.iB[c + plane * L] = ib;
    if (L == n - 2) {
      a.hB[c + plane * (n - 1)] = hb;
      a.iB[c + plane * (n - 1)] = ib;
    }
    if (L == 1) {
      a.hB[c] = hb;
      a.iB[c] = ib;
    }
  }
}

__device__ inline int reflect(int p, int n) { return p < 0 ? -1 - p : (p >= n ? 2 * n - 1 - p : p); }

// Gaussian smoothing of the five tables over (I,J) at each L >= 2
// (src/ModRamScb.f90:539-562 with srcExternal/gaussian_filter.f90:104-178):
// edges by reflection, the window summed in column-major order.  One block
// per (L, table) with the (I,J) plane in LDS; L = 1 is copied.
__global__ void __launch_bounds__(256) k_hi_smooth(HiArgs a) {
  extern __shared__ double pl[];
  const int nr = a.nr, nt = a.nt, L = blockIdx.x, q = blockIdx.y;
  const size_t plane = (size_t)nr * nt;
  const double *src[5] = {a.hC, a.iC, a.hB, a.iB, a.dC};
  double *dst[5] = {a.sh, a.si, a.shb, a.sib, a.sd};
  const double *X = src[q] + plane * L;
  double *Y = dst[q] + plane * L;
  if (L == 0) {
    for (size_t c = threadIdx.x; c < plane; c += blockDim.x) Y[c] = X[c];
    return;
  }
  double *w = pl + plane;
  for (size_t c = threadIdx.x; c < plane; c += blockDim.x) pl[c] = X[c];
  for (int c = threadIdx.x; c < kSmoothW * kSmoothW; c += blockDim.x) w[c] = a.w[c];
  __syncthreads();
  for (size_t c = threadIdx.x; c < plane; c += blockDim.x) {
    const int j = (int)(c / nr), i = (int)(c - (size_t)j * nr);
    double acc = 0.0;
    for (int dj = -kSmoothR; dj <= kSmoothR; ++dj) {
      const double *col = pl + (size_t)nr * reflect(j + dj, nt);
      for (int di = -kSmoothR; di <= kSmoothR; ++di)
        acc = acc + w[(di + kSmoothR) + kSmoothW * (dj + kSmoothR)] * col[reflect(i + di, nr)];
    }
    Y[c] = acc;
  }
}

// The RAM variables and their time derivatives (src/ModRamScb.f90:564-596):
// one thread per cell of the rows I = 2..NR+1.
__global__ void k_hi_fill(HiArgs a) {
  const int nr = a.nr, nt = a.nt;
  const size_t plane = (size_t)nr * nt, t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= plane * a.npa) return;
  const int L = (int)(t / plane), ij = (int)(t - plane * L), J = ij / nr, i = ij - J * nr + 1;
  const bool sm = a.smooth != 0, dt = fabs(a.dthi) > 1e-9;
  const size_t d = (size_t)i + (size_t)(nr + 1) * J + (size_t)(nr + 1) * nt * L;
  const double fi = (sm ? a.si : a.iC)[t], bi = (sm ? a.sib : a.iB)[t];
  a.didt[d] = dt ? (fi - a.fnis[d]) / a.dthi : 0.0;
  a.dibndt[d] = dt ? (bi - a.bounis[d]) / a.dthi : 0.0;
  a.fnhs[d] = (sm ? a.sh : a.hC)[t];
  a.fnis[d] = fi;
  a.hdns[d] = (sm ? a.sd : a.dC)[t];
  a.bounhs[d] = (sm ? a.shb : a.hB)[t];
  a.bounis[d] = bi;
  if (L == 0) {
    const size_t e = (size_t)i + (size_t)(nr + 1) * J;
    const double bn = a.bz[ij] / 1e9;
    a.dbdt[e] = dt ? (bn - a.bnes[e]) / a.dthi : 0.0;
    a.bnes[e] = bn;
  }
}

// Row I = 1 (src/ModRamScb.f90:598-611) and the NaN repair (:614-626): one
// thread per (J, L) walks I, so a repaired cell takes the finished row below.
__global__ void k_hi_rows(HiArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nt * a.npa) return;
  const int J = t % a.nt, L = t / a.nt, nr = a.nr, nr1 = nr + 1;
  auto g = [&](int i) { return (size_t)i + (size_t)nr1 * J + (size_t)nr1 * a.nt * L; };
  if (L == 0) {
    const double l1 = a.lz[0];
    const size_t d = (size_t)nr1 * J;
    a.bnes[d] = 0.32 / (l1 * l1 * l1) / 1.e4;
    a.dbdt[d] = 0.0;
    a.eir[d] = 0.0;
    a.eip[d] = 0.0;
  }
  {
    const size_t d0 = g(0), d1 = g(1);
    a.fnhs[d0] = a.fnhs[d1];
    a.fnis[d0] = a.fnis[d1];
    a.bounhs[d0] = a.bounhs[d1];
    a.bounis[d0] = a.bounis[d1];
    a.hdns[d0] = a.hdns[d1];
    a.didt[d0] = 0.0;
    a.dibndt[d0] = 0.0;
  }
  for (int i = 1; i <= nr; ++i) {
    const size_t d = g(i), p = g(i - 1);
    if (isnan(a.fnis[d])) a.fnis[d] = a.fnis[p];
    if (isnan(a.fnhs[d])) a.fnhs[d] = a.fnhs[p];
    if (isnan(a.bounis[d])) a.bounis[d] = a.bounis[p];
    if (isnan(a.bounhs[d])) a.bounhs[d] = a.bounhs[p];
    if (isnan(a.hdns[d])) a.hdns[d] = a.hdns[p];
    if (isnan(a.didt[d])) a.didt[d] = 0.0;
    if (isnan(a.dibndt[d])) a.dibndt[d] = 0.0;
  }
  for (int i = t; i < nr * a.nt; i += a.nt * a.npa) a.mgnp_dst[i] = a.mgnp[i];
}

size_t hi_lds_bytes(int nthe, int npa) { return (4 * (size_t)nthe + 4 * (size_t)npa + 2) * 8 + (size_t)npa * 4; }
size_t hi_smooth_lds_bytes(int nr, int nt) { return ((size_t)nr * nt + kSmoothW * kSmoothW) * 8; }

}  // namespace

extern "C" {

int ramgpu_computehI(ramgpu_ctx *c, int nthe, int ieq, const double *x, const double *y, const double *z,
                     const double *b, const double *chi, const double *pa, const double *pabn, const double *lz,
                     const int *outside_mgnp, const int *scale_at, double bnormal, int smooth) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_grid) return fail(RAMGPU_ESTATE, "ramgpu_set_grid has not been called (MU is needed)");
  const int nr = c->nr, nt = c->nt, npa = c->npa;
  if (!x || !y || !z || !b || !chi || !pa || !pabn || !lz || !outside_mgnp)
    return fail(RAMGPU_EINVAL, "computehI: a required array is null");
  if (nthe < 3) return fail(RAMGPU_EINVAL, "computehI: NTHE=%d, at least 3 samples per line", nthe);
  if (ieq < 1 || ieq > nthe) return fail(RAMGPU_EINVAL, "computehI: equator index %d outside 1..%d", ieq, nthe);
  if (npa < 4) return fail(RAMGPU_EINVAL, "computehI: NPA=%d, the corrections need NPA >= 4", npa);
  if (hi_lds_bytes(nthe, npa) > 65536) return fail(RAMGPU_EINVAL, "computehI: NTHE=%d too large", nthe);
  if (smooth && (nr < kSmoothW || nt < kSmoothW))
    return fail(RAMGPU_EINVAL, "computehI: smoothing needs NR and NT of at least %d (grid %dx%d)", kSmoothW, nr, nt);
  if (smooth && hi_smooth_lds_bytes(nr, nt) > 65536)
    return fail(RAMGPU_EINVAL, "computehI: smoothing of a %dx%d plane exceeds the LDS of one block", nr, nt);
  for (int l = 1; l < npa; ++l)
    if (!(pa[l] < pa[l - 1])) return fail(RAMGPU_EINVAL, "computehI: PA must decrease strictly");
  for (int k = 1; k < nthe; ++k)
    if (!(chi[k] > chi[k - 1])) return fail(RAMGPU_EINVAL, "computehI: the line coordinate must increase strictly");
  if (scale_at)
    for (int j = 1; j < nt; ++j)
      if (scale_at[j] != 0 && (scale_at[j] < 3 || scale_at[j] > nr))
        return fail(RAMGPU_EINVAL, "computehI: ScaleAt(%d)=%d outside 3..NR", j + 1, scale_at[j]);
  // Gaussian weights on the host (src: srcExternal/gaussian_filter.f90:18-56)
  double w[kSmoothW * kSmoothW], tot = 0.0;
  for (int dj = -kSmoothR; dj <= kSmoothR; ++dj)
    for (int di = -kSmoothR; di <= kSmoothR; ++di)
      w[(di + kSmoothR) + kSmoothW * (dj + kSmoothR)] = 2.0 * std::exp((-0.5 * (double)(di * di + dj * dj)) / 1.0);
  for (int q = 0; q < kSmoothW * kSmoothW; ++q) tot += w[q];
  for (int q = 0; q < kSmoothW * kSmoothW; ++q) w[q] = w[q] / tot;
  // one workspace for the call: inputs, line values, smoothed copies
  const size_t nl = (size_t)nthe * nr * nt, n3 = (size_t)nr * nt * npa, n2 = (size_t)nr * nt;
  const size_t nd = 4 * nl + nthe + 2 * (size_t)npa + (nr + 1) + kSmoothW * kSmoothW + 10 * n3 + n2;
  const size_t ni = n2 + nt;
  char *ws = nullptr;
  if (hipMalloc(&ws, nd * sizeof(double) + ni * sizeof(int)) != hipSuccess)
    return fail(RAMGPU_ENOMEM, "computehI: workspace allocation failed");
  struct Free {
    char *p;
    ramgpu_ctx *c;
    ~Free() {
      (void)hipStreamSynchronize(c->stream);
      (void)hipFree(p);
    }
  } guard{ws, c};
  double *d = (double *)ws;
  HiArgs a{};
  a.nthe = nthe; a.nr = nr; a.nt = nt; a.npa = npa; a.ieq = ieq - 1;
  double *p = d;
  auto take = [&](size_t n) { double *q = p; p += n; return q; };
  double *dx = take(nl), *dy = take(nl), *dz = take(nl), *db = take(nl), *dchi = take(nthe);
  double *dpa = take(npa), *dpabn = take(npa), *dlz = take(nr + 1), *dw = take(kSmoothW * kSmoothW);
  a.hC = take(n3); a.iC = take(n3); a.dC = take(n3); a.hB = take(n3); a.iB = take(n3);
  a.sh = take(n3); a.si = take(n3); a.shb = take(n3); a.sib = take(n3); a.sd = take(n3);
  a.bz = take(n2);
  int *im = (int *)p, *isa = im + n2;
  std::vector<int> sa(nt, 0);
  if (scale_at)
    for (int j = 1; j < nt; ++j) sa[j] = scale_at[j];
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(dx, x, nl * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dy, y, nl * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dz, z, nl * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(db, b, nl * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dchi, chi, (size_t)nthe * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dpa, pa, (size_t)npa * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dpabn, pabn, (size_t)npa * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dlz, lz, (size_t)(nr + 1) * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dw, w, sizeof(w), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(im, outside_mgnp, n2 * sizeof(int), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(isa, sa.data(), (size_t)nt * sizeof(int), hipMemcpyHostToDevice, s));
  a.x = dx; a.y = dy; a.z = dz; a.b = db; a.chi = dchi; a.mu = c->d_mu; a.pa = dpa; a.pabn = dpabn;
  a.lz = dlz; a.w = dw; a.mgnp = im; a.scale_at = isa;
  a.bnormal = bnormal;
  a.dthi = c->time_elapsed - c->told;
  a.fnhs = c->d_fnhs; a.fnis = c->d_fnis; a.bounhs = c->d_bounhs; a.bounis = c->d_bounis; a.hdns = c->d_hdnsf;
  a.didt = c->d_didt; a.dibndt = c->d_dibndt; a.bnes = c->d_bnes; a.dbdt = c->d_dbdt; a.eir = c->d_eir;
  a.eip = c->d_eip; a.mgnp_dst = c->d_mgnp;
  a.smooth = smooth ? 1 : 0;
  hipLaunchKernelGGL(k_hi_lines, dim3((unsigned)n2), dim3(kHiLineThreads), hi_lds_bytes(nthe, npa), s, a);
  if ((rc = launch_check("k_hi_lines"))) return rc;
  const unsigned gjl = (unsigned)((nt * npa + 255) / 256), gij = (unsigned)((n2 + 255) / 256);
  if (scale_at) {
    hipLaunchKernelGGL(k_hi_scale, dim3(gjl), dim3(256), 0, s, a);
    if ((rc = launch_check("k_hi_scale"))) return rc;
  }
  hipLaunchKernelGGL(k_hi_near90, dim3(gij), dim3(256), 0, s, a);
  if ((rc = launch_check("k_hi_near90"))) return rc;
  hipLaunchKernelGGL(k_hi_negative, dim3(gjl), dim3(256), 0, s, a);
  if ((rc = launch_check("k_hi_negative"))) return rc;
  hipLaunchKernelGGL(k_hi_pitch, dim3((unsigned)n2), dim3(kHiThreads), 4 * (size_t)npa * sizeof(double), s, a);
  if ((rc = launch_check("k_hi_pitch"))) return rc;
  if (smooth) {
    hipLaunchKernelGGL(k_hi_smooth, dim3((unsigned)npa, 5), dim3(256), hi_smooth_lds_bytes(nr, nt), s, a);
    if ((rc = launch_check("k_hi_smooth"))) return rc;
  }
  hipLaunchKernelGGL(k_hi_fill, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, a);
  if ((rc = launch_check("k_hi_fill"))) return rc;
  hipLaunchKernelGGL(k_hi_rows, dim3(gjl), dim3(256), 0, s, a);
  if ((rc = launch_check("k_hi_rows"))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  c->told = c->time_elapsed;  // TOld = TimeRamElapsed (src/ModRamScb.f90:612)
  c->have_geom = true;
  c->tables_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_told(ramgpu_ctx *c, double told) {
  int rc = check_ctx(c);
  if (rc) return rc;
  c->told = told;
  return RAMGPU_OK;
}

int ramgpu_get_told(ramgpu_ctx *c, double *told) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!told) return fail(RAMGPU_EINVAL, "told is null");
  *told = c->told;
  return RAMGPU_OK;
}

int ramgpu_get_geometry(ramgpu_ctx *c, double *fnhs, double *fnis, double *bounhs, double *bounis, double *hdns,
                        double *didt, double *dibndt, double *bnes, double *dbdt, double *eir, double *eip,
                        int *outside_mgnp) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const size_t g3 = (size_t)(c->nr + 1) * c->nt * c->npa, g2 = (size_t)(c->nr + 1) * c->nt;
  const struct {
    double *h;
    const double *d;
    size_t n;
  } v[] = {{fnhs, c->d_fnhs, g3}, {fnis, c->d_fnis, g3}, {bounhs, c->d_bounhs, g3}, {bounis, c->d_bounis, g3},
           {hdns, c->d_hdnsf, g3}, {didt, c->d_didt, g3}, {dibndt, c->d_dibndt, g3}, {bnes, c->d_bnes, g2},
           {dbdt, c->d_dbdt, g2}, {eir, c->d_eir, g2}, {eip, c->d_eip, g2}};
  for (const auto &e : v)
    if (e.h) HIP_TRY(hipMemcpyAsync(e.h, e.d, e.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (outside_mgnp)
    HIP_TRY(hipMemcpyAsync(outside_mgnp, c->d_mgnp, (size_t)c->nr * c->nt * sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

}  // extern "C"
