// kernels.hip -- gfx950 kernels of the RAM kinetic split step.
//
// Reference: src/ModRamDrift.f90 (DRIFTPARA/DRIFTR/DRIFTP/DRIFTE/DRIFTMU),
// src/ModRamLoss.f90 (CEPARA/CHAREXCHANGE/ATMOL), src/ModRamRun.f90
// (ram_run sequence, SUMRC).  Index names in comments are the reference's
// 1-based ones; code indices are 0-based.
//
// Kernel map
//   k_species   per-species 1-D tables of DRIFTPARA/CEPARA, grid reciprocals
//   k_coef      per-(I,J,L) / per-(I,J) / per-I intermediate terms of the
//               four Courant numbers
//   k_ghost     DRIFTR outer-ghost carry between lines (see k_ghost)
//   k_ij        R/P sweeps on (I,J) planes x KW consecutive (L,K) positions
//   k_kl        E/Mu sweeps and loss operators on whole (K,L) planes
//   k_finalize  DtDrift* from the CFL partial maxima; SUMRC bookkeeping
//   k_vs        Volland-Stern VT from Kp
//
// Arithmetic: FP64 throughout, compiled with -ffp-contract=off.  Every
// Courant number, limited face value and flux update is evaluated with the
// reference's own sequence of roundings (left-to-right Fortran evaluation):
// intermediate terms that do not depend on K (or on K and L) are tabulated
// exactly as the reference computes them, and a division by a tabulated
// divisor d is done as div_rn(a, d, 1/d) below, which returns the correctly
// rounded quotient, i.e. the same double as a/d.  The drift operators are
// therefore bit-identical to the reference on identical inputs; the loss
// operators differ only through the libm/ocml exp, pow and log10.
#include "ramgpu_dev.h"

#include <hip/hip_runtime.h>

namespace ramgpu {

static constexpr double kOME = 7.3E-5;      // src/ModRamDrift.f90:224
static constexpr double kQ = 1.602E-19;     // src/ModRamConst.f90:19
static constexpr double kCS = 2.998E8;      // src/ModRamConst.f90:20
static constexpr double kRE = 6.371E6;      // src/ModRamConst.f90:14

// geometry accessors, 1-based as the reference
#define GEO3(A, I, J, L) (A)[((I)-1) + (size_t)(t.nr + 1) * (((J)-1) + (size_t)t.nt * ((L)-1))]
#define GEO2(A, I, J) (A)[((I)-1) + (size_t)(t.nr + 1) * ((J)-1)]
#define MGNP(I, J) t.mgnp[((I)-1) + (size_t)t.nr * ((J)-1)]

__device__ __forceinline__ size_t ij_of(const Tab &t, int i, int j) { return (size_t)i * t.nt + j; }
__device__ __forceinline__ size_t lij_of(const Tab &t, int l, int i, int j) {
  return ((size_t)l * t.nr + i) * t.nt + j;
}
__device__ __forceinline__ size_t ijl_of(const Tab &t, int i, int j, int l) {
  return ((size_t)i * t.nt + j) * t.npa + l;
}

// n/d for 0 <= n < 2^22, d >= 1, from rd = RN_float(1/d): the float
// quotient is within one of the true one, and one remainder test fixes it.
__device__ __forceinline__ int fast_div(int n, int d, float rd) {
  int q = __float2int_rz((float)n * rd);
  const int r = n - q * d;
  if (r < 0) --q;
  else if (r >= d) ++q;
  return q;
}

// a/d, correctly rounded, from rd = RN(1/d): q = RN(a*rd) lies within one ulp
// of a/d, the residual a - d*q is exact under FMA, and one correction step
// rounds to RN(a/d) (Markstein, IBM J. Res. Dev. 34(1), 1990).
__device__ __forceinline__ double div_rn(double a, double d, double rd) {
  const double q = a * rd;
  const double r = fma(-q, d, a);
  return fma(r, rd, q);
}

// Radial Courant number CDriftR(I,J,K,L) = CR(I,J) + CGR with
// CGR = CGR3/(FNHS+FNHS')*P4/2./(BNES+BNES')/(RLZ+0.5*MDR)
// (src/ModRamDrift.f90:140-146).  One function for the sweep, the CFL and the
// ghost carry, so the inflow/outflow decision is the same everywhere.
__device__ __forceinline__ double c_radial(const Tab &t, int s, int i, int j, int k, int l) {
  const size_t ij = ij_of(t, i, j);
  const double x = (t.q1[lij_of(t, l, i, j)] * t.p4[s * t.ne + k]) * 0.5;
  return t.cr[ij] + div_rn(div_rn(x, t.bsr[ij], t.rbsr[ij]), t.rsum[i], t.rrsum[i]);
}

// Azimuthal Courant number (src/ModRamDrift.f90:236-239).
__device__ __forceinline__ double c_azim(const Tab &t, int s, int i, int j, int k, int l) {
  const size_t ij = ij_of(t, i, j), q = lij_of(t, l, i, j);
  const double u = div_rn(t.p2[((size_t)s * t.nr + i) * t.ne + k] * t.gpa[q], t.fhs[q], t.rfhs[q]);
  const double num = t.pa[ij] - u - t.pce[ij];
  return div_rn(num, t.pbs[ij], t.rpbs[ij]) + t.omet;
}

// Limited face value of the reference's upwind scheme (src/ModRamDrift.f90:
// 169-183) for the face between F(I) = f0 and F(I+1) = fp1.  fn and fnm1 are
// F(N) and F(N-1) with N = I+1-sgn as the reference forms it (including its
// periodic wrap in DRIFTP); cn is the Courant number as the limiter's CORR
// term sees it (C, C/DE(K) or C/DMU(L)).
__device__ __forceinline__ double face_value(double f0, double fp1, double fn, double fnm1,
                                             double c, double cn, double beta) {
  const double sg = (c < 0) ? -1.0 : 1.0;
  const double x = fp1 - f0;
  const double fup = 0.5 * (f0 + fp1 - sg * x);
  if (fabs(x) <= 1.E-27) return fup;
  const double r = (fn - fnm1) / x;
  if (r <= 0) return fup;
  const double lim = fmax(fmin(beta * r, 1.0), fmin(r, beta));
  const double corr = -0.5 * (cn - sg) * x;
  return fup + lim * corr;
}

// ---------------------------------------------------------------------------
// Block reductions (fixed order: deterministic for a given launch shape).
__device__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double tot = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) tot += red[w];
  }
  return tot;  // valid in thread 0
}

__device__ double block_max(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) m = fmax(m, red[w]);
  }
  return m;  // valid in thread 0
}

__device__ double block_min(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double m = v;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) m = fmin(m, red[w]);
  }
  return m;  // valid in thread 0
}

__device__ __forceinline__ void lds_max(unsigned long long *slot, double v) {
  atomicMax(slot, (unsigned long long)__double_as_longlong(v));  // v >= 0
}

// ---------------------------------------------------------------------------
// Per-species tables.  DRIFTPARA (src/ModRamDrift.f90:64-85): P2, EDOT; the
// K-only factor P4 of DRIFTR (:131); the K-only prefixes of DRIFTE's EDT1
// (:337) and DRIFTMU's EDT (:424); CEPARA's cross-section x velocity
// (src/ModRamLoss.f90:39-81) and ATLOS (:162-167); DRIFTE's relativistic
// ghost factors (:310-311,334-335).  Block 0 also forms the grid reciprocals.
__global__ void k_species(Tab t) {
  const int s = blockIdx.x;
  const double qs = (double)t.charge[s];
  const double dts = t.dts, dphi = t.dphi, mdr = t.mdr;
  for (int n = threadIdx.x; n < t.nr * t.ne; n += blockDim.x) {
    const int i = n / t.ne, k = n % t.ne;
    const double rlz = t.rlz[i];
    const double g = t.grel[s + t.ns * k], gb = t.grbnd[s + t.ns * k];
    const double ekev = t.ekev[k], ebnd = t.ebnd[k];
    const size_t sik = ((size_t)s * t.nr + i) * t.ne + k;
    t.p2[sik] = dts * ekev * 1000.0 * (g + 1.0) / g / (rlz * rlz) / dphi / qs;
    t.edot[sik] = ebnd * dts / rlz * (gb + 1.0) / gb / 2.0;
    double atl = 1.0;  // ATLOS(S,I,K)=EXP(-DTs/TAUB), TAUB=2*RLZ(I)/V(S,K), I,K>=2
    if (i >= 1 && k >= 1) {
      const double taub = 2.0 * rlz / t.v[s + t.ns * k];
      atl = exp(-dts / taub);
    }
    t.atlos[sik] = atl;
    t.latl[sik] = log(atl);
    if (i == 0) {
      t.p4[s * t.ne + k] = dts * ekev * 1000.0 * (g + 1.0) / g / dphi / mdr / qs;
      t.ekp[s * t.ne + k] = ebnd * 1e3 * (gb + 1.0) / 2.0 / gb;
      t.mkp[s * t.ne + k] = ekev * 1e3 * (g + 1.0) / 2.0 / g;
      double sv = 0.0;
      const int kind = t.cex[s];
      if (kind != 0 && k >= 1) {
        double x = log10(ekev);
        if (x < -2.0) x = -2.0;
        const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
        double y;
        if (kind == 1)
          y = -18.767 - 0.11017 * x - 3.8173e-2 * x2 - 0.1232 * x3 - 5.0488e-2 * x4;
        else if (kind == 2)
          y = -20.789 + 0.92316 * x - 0.68017 * x2 + 0.66153 * x3 - 0.20998 * x4;
        else
          y = -18.987 - 0.10613 * x - 5.4841E-3 * x2 - 1.6262E-2 * x3 - 7.0554E-3 * x4;
        sv = pow(10.0, y) * t.v[s + t.ns * k];
      }
      t.sigv[s * t.ne + k] = sv;
    }
  }
  if (threadIdx.x == 0) {
    const double g1 = t.grel[s], g2 = t.grel[s + t.ns];
    const double ezero = t.ekev[0] - t.we[0];
    const double grzero = 1.0 + ezero * 1000.0 * kQ / t.rmas[s] / kCS / kCS;
    double *rel = t.rel + s * 8;
    rel[0] = g1;
    rel[1] = g2;
    rel[2] = sqrt((g2 * g2 - 1.0) / (g1 * g1 - 1.0));
    rel[3] = grzero;
    rel[4] = sqrt((g1 * g1 - 1.0) / (grzero * grzero - 1.0));
  }
  if (s == 0) {
    for (int k = threadIdx.x; k < t.ne; k += blockDim.x) {
      t.rde[k] = 1.0 / t.de[k];
      t.rwe[k] = 1.0 / t.we[k];
    }
    for (int l = threadIdx.x; l < t.npa; l += blockDim.x) {
      t.rdmu[l] = 1.0 / t.dmu[l];
      t.rwmu[l] = 1.0 / t.wmu[l];
    }
  }
}

// ---------------------------------------------------------------------------
// Species-independent intermediate terms, each written with the reference's
// expression: k_coef_ij the per-(I,J) and per-I terms (the ones the
// convection potential VT enters), k_coef one block per (I,J), one thread per
// L, the per-(I,J,L) terms:
//   DRIFTR  src/ModRamDrift.f90:118-144   CR, CGR3/(FNHS+FNHS')
//   DRIFTP  :232-239                       GPA1+GPA2, FNHS sums, VT and EIR terms
//   DRIFTE  :320-341                       DRD1, DPD1, GPR*, GPP*, DRD2, DPD2, dBdt1, dIdt1
//   DRIFTMU :415-432                       DRM1, DPM1, CMUDOT, GMR*, GMP*, DRM2, DPM2, dBdt2, dIbndt2
//   ATMOL   src/ModRamLoss.f90:501         1/FNHS
__global__ void k_coef_ij(Tab t) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= t.nr * t.nt) return;
  const int i = n / t.nt, j = n % t.nt;
  const int I = i + 1, J = j + 1;
  const int NT = t.nt;
  const double dts = t.dts, mdr = t.mdr, dphi = t.dphi;
  const double rlz = t.rlz[i];
  const size_t ij = ij_of(t, i, j);
  if (j == 0) {
    t.rrlz[i] = 1.0 / rlz;
    const double rs = rlz + 0.5 * mdr;
    t.rsum[i] = rs;
    t.rrsum[i] = 1.0 / rs;
  }
  {
    const int J0 = (J == 1) ? NT - 1 : J - 1;
    const int J1 = (J == NT) ? 2 : J + 1;
    const double vr = dts / mdr / (rlz + 0.5 * mdr) / 2.0 / dphi;  // :67
    const double bsum = GEO2(t.bnes, I, J) + GEO2(t.bnes, I + 1, J);
    const double dv = GEO2(t.vt, I, J0) + GEO2(t.vt, I + 1, J0) - GEO2(t.vt, I, J1) - GEO2(t.vt, I + 1, J1);
    t.cr[ij] = vr * dv / bsum + (GEO2(t.eip, I, J) + GEO2(t.eip, I + 1, J)) / bsum * dts / mdr;
    t.bsr[ij] = bsum;
    t.rbsr[ij] = 1.0 / bsum;
  }
  double pa = 0.0, pce = 0.0, pbs = 1.0;
  double drd1 = 0.0, dpd1 = 0.0, drm1 = 0.0, dpm1 = 0.0, db2 = 0.0;
  const double bij = GEO2(t.bnes, I, J);
  if (I >= 2) {
    if (J >= 2) {
      const int J1 = (J == NT) ? 2 : J + 1;
      const double p1 = dts / dphi / 2.0 / mdr / rlz;  // :68
      pa = (GEO2(t.vt, I + 1, J) + GEO2(t.vt, I + 1, J1) - GEO2(t.vt, I - 1, J) - GEO2(t.vt, I - 1, J1)) * p1;
      pce = (GEO2(t.eir, I, J1) + GEO2(t.eir, I, J)) / rlz * dts / dphi;
      pbs = GEO2(t.bnes, I, J) + GEO2(t.bnes, I, J1);
    }
    const int J0 = (J == 1) ? NT - 1 : J - 1;
    const int J2 = (J == NT) ? 2 : J + 1;
    drd1 = (GEO2(t.eip, I, J) * rlz - (GEO2(t.vt, I, J2) - GEO2(t.vt, I, J0)) / 2.0 / dphi) / bij;
    dpd1 = kOME * rlz + ((GEO2(t.vt, I + 1, J) - GEO2(t.vt, I - 1, J)) / 2.0 / mdr - GEO2(t.eir, I, J)) / bij;
    // DRIFTMU's DRM1/DPM1 (:415-416) are the same expressions as DRD1/DPD1
    drm1 = (GEO2(t.eip, I, J) * rlz - (GEO2(t.vt, I, J2) - GEO2(t.vt, I, J0)) / 2.0 / dphi) / bij;
    dpm1 = kOME * rlz + ((GEO2(t.vt, I + 1, J) - GEO2(t.vt, I - 1, J)) / 2.0 / mdr - GEO2(t.eir, I, J)) / bij;
    db2 = GEO2(t.dbdt, I, J) / 2.0 / bij * rlz;  // :431
  }
  t.pa[ij] = pa;
  t.pce[ij] = pce;
  t.pbs[ij] = pbs;
  t.rpbs[ij] = 1.0 / pbs;
  t.drd1[ij] = drd1;
  t.dpd1[ij] = dpd1;
  t.bij[ij] = bij;
  t.rbij[ij] = 1.0 / bij;
  t.drm1[ij] = drm1;
  t.dpm1[ij] = dpm1;
  t.db2[ij] = db2;
}

__global__ void k_coef(Tab t) {
  const int b = blockIdx.x;
  const int i = b / t.nt, j = b % t.nt;
  const int I = i + 1, J = j + 1;
  const int NT = t.nt, NPA = t.npa;
  const int l = threadIdx.x;
  const int L = l + 1;
  const double dts = t.dts, mdr = t.mdr, dphi = t.dphi;
  const double rlz = t.rlz[i];
  if (l >= NPA) return;

  // ---- DRIFTR, all I (:140-144)
  {
    const int J0 = (J == 1) ? NT - 1 : J - 1;
    const int J1 = (J == NT) ? 2 : J + 1;
    const double cgr1 = GEO3(t.fnis, I + 1, J1, L) + GEO3(t.fnis, I, J1, L) -
                        GEO3(t.fnis, I + 1, J0, L) - GEO3(t.fnis, I, J0, L);
    const double cgr2 = GEO2(t.bnes, I + 1, J1) + GEO2(t.bnes, I, J1) - GEO2(t.bnes, I + 1, J0) -
                        GEO2(t.bnes, I, J0);
    const double cgr3 = cgr1 + (GEO3(t.fnis, I + 1, J, L) + GEO3(t.fnis, I, J, L) -
                                2.0 * GEO3(t.fnhs, I + 1, J, L) - 2.0 * GEO3(t.fnhs, I, J, L)) *
                                   cgr2 / 2.0 / (GEO2(t.bnes, I + 1, J) + GEO2(t.bnes, I, J));
    t.q1[lij_of(t, l, i, j)] = cgr3 / (GEO3(t.fnhs, I, J, L) + GEO3(t.fnhs, I + 1, J, L));
  }
  // ---- DRIFTP, I>=2, J>=2 (:232-235)
  {
    double gpa = 0.0, fhs = 1.0;
    if (I >= 2 && J >= 2) {
      const int J1 = (J == NT) ? 2 : J + 1;
      const double gpa1 = GEO3(t.fnis, I, J, L) + GEO3(t.fnis, I, J1, L) +
                          (GEO3(t.fnis, I + 1, J1, L) + GEO3(t.fnis, I + 1, J, L) - GEO3(t.fnis, I - 1, J, L) -
                           GEO3(t.fnis, I - 1, J1, L)) * rlz / 2.0 / mdr;
      const double gpa2 = rlz / 4.0 / mdr *
                          (GEO3(t.fnis, I, J, L) + GEO3(t.fnis, I, J1, L) - 2.0 * GEO3(t.fnhs, I, J, L) -
                           2.0 * GEO3(t.fnhs, I, J1, L)) *
                          (GEO2(t.bnes, I + 1, J1) + GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J) -
                           GEO2(t.bnes, I - 1, J1)) /
                          (GEO2(t.bnes, I, J) + GEO2(t.bnes, I, J1));
      gpa = gpa1 + gpa2;
      fhs = GEO3(t.fnhs, I, J, L) + GEO3(t.fnhs, I, J1, L);
    }
    const size_t q = lij_of(t, l, i, j);
    t.gpa[q] = gpa;
    t.fhs[q] = fhs;
    t.rfhs[q] = 1.0 / fhs;
  }
  const size_t q = ijl_of(t, i, j, l);
  // ---- DRIFTE, I>=2 (:323-341)
  {
    double gr = 0, gp = 0, drd2 = 0, dpd2 = 0, db1 = 0, di1 = 0;
    const double fnhs = GEO3(t.fnhs, I, J, L);
    if (I >= 2) {
      const int J0 = (J == 1) ? NT - 1 : J - 1;
      const int J2 = (J == NT) ? 2 : J + 1;
      const double bij = GEO2(t.bnes, I, J);
      const double fnis = GEO3(t.fnis, I, J, L);
      const double gpa = (1.0 - fnis / 2.0 / fnhs) / bij;
      const double gpr1 = gpa * (GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J)) / 2.0 / mdr;
      const double gpr2 = -fnis / fnhs / rlz;
      const double gpr3 = -(GEO3(t.fnis, I + 1, J, L) - GEO3(t.fnis, I - 1, J, L)) / 2.0 / mdr / fnhs;
      const double gpp1 = gpa * (GEO2(t.bnes, I, J2) - GEO2(t.bnes, I, J0)) / 2.0 / dphi;
      const double gpp2 = -(GEO3(t.fnis, I, J2, L) - GEO3(t.fnis, I, J0, L)) / 2.0 / dphi / fnhs;
      drd2 = (GEO3(t.fnis, I, J2, L) - GEO3(t.fnis, I, J0, L)) / 2.0 / dphi +
             (fnis - 2.0 * fnhs) * (GEO2(t.bnes, I, J2) - GEO2(t.bnes, I, J0)) / 4.0 / bij / dphi;
      dpd2 = fnis + (GEO3(t.fnis, I + 1, J, L) - GEO3(t.fnis, I - 1, J, L)) * rlz / 2.0 / mdr +
             rlz * (fnis - 2.0 * fnhs) / 4.0 / mdr * (GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J)) / bij;
      db1 = GEO2(t.dbdt, I, J) * (1.0 - fnis / 2.0 / fnhs) * rlz / bij;
      di1 = -GEO3(t.didt, I, J, L) * rlz / fnhs;
      gr = gpr1 + gpr2 + gpr3;
      gp = gpp1 + gpp2;
    }
    t.gr[q] = gr;
    t.gp[q] = gp;
    t.drd2[q] = drd2;
    t.dpd2[q] = dpd2;
    t.db1[q] = db1;
    t.di1[q] = di1;
    t.hs[q] = fnhs;
    t.rhs[q] = 1.0 / fnhs;
  }
  // ---- DRIFTMU, I>=2, L>=2 (:418-432; MUDOT from DRIFTPARA :77-81)
  {
    double cmudot = 0, gmr = 0, gmp = 0, drm2 = 0, dpm2 = 0, di2 = 0;
    const double bhs = GEO3(t.bounhs, I, J, L);
    if (I >= 2 && L >= 2) {
      const int J0 = (J == 1) ? NT - 1 : J - 1;
      const int J1 = (J == NT) ? 2 : J + 1;
      const double bij = GEO2(t.bnes, I, J);
      double mudot = 0.0;
      if (L < NPA) {
        const double muboun = t.mu[l] + 0.5 * t.wmu[l];
        mudot = (1.0 - muboun * muboun) * dts / 2.0 / muboun / rlz;
      }
      const double bis = GEO3(t.bounis, I, J, L);
      cmudot = mudot * bis / bhs;
      const double gmr1 = (GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J)) / 4.0 / mdr / bij;
      const double gmr2 = 1.0 / rlz;
      const double gmr3 = (GEO3(t.bounis, I + 1, J, L) - GEO3(t.bounis, I - 1, J, L)) / 2.0 / mdr / bis;
      const double gmp1 = (GEO2(t.bnes, I, J1) - GEO2(t.bnes, I, J0)) / 4.0 / dphi / bij;
      const double gmp2 = (GEO3(t.bounis, I, J1, L) - GEO3(t.bounis, I, J0, L)) / 2.0 / dphi / bis;
      drm2 = (GEO3(t.bounis, I, J1, L) - GEO3(t.bounis, I, J0, L)) / 2.0 / dphi +
             (bis - 2.0 * bhs) * (GEO2(t.bnes, I, J1) - GEO2(t.bnes, I, J0)) / 4.0 / bij / dphi;
      dpm2 = bis + (GEO3(t.bounis, I + 1, J, L) - GEO3(t.bounis, I - 1, J, L)) * rlz / 2.0 / mdr +
             (bis - 2.0 * bhs) * rlz / 4.0 / mdr * (GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J)) / bij;
      di2 = GEO3(t.dibndt, I, J, L) * rlz / bis;
      gmr = gmr1 + gmr2 + gmr3;
      gmp = gmp1 + gmp2;
    }
    t.cmudot[q] = cmudot;
    t.gmr[q] = gmr;
    t.gmp[q] = gmp;
    t.drm2[q] = drm2;
    t.dpm2[q] = dpm2;
    t.di2[q] = di2;
    t.bhs[q] = bhs;
    t.rbhs[q] = 1.0 / bhs;
  }
  t.hdns[q] = GEO3(t.hdns_f, I, J, L);
  t.invfnhs[q] = 1.0 / GEO3(t.fnhs, I, J, L);
}

// ---------------------------------------------------------------------------
// DRIFTR's outer ghost F(NR+1) is a per-call buffer that is not reset between
// lines (src/ModRamDrift.f90:112-113,159-167): an outflow line (C(NR)>=0)
// whose face NR-1 is upwinded from outside (C(NR-1)<0, :174-175) reads the
// ghost left by the most recent inflow line in the reference's K,L,J loop
// order, or 0 when no inflow line came before.  Whether a line is inflow
// depends on geometry, fields and grid only, so the carry is precomputed:
// carry[s][p] is the index p'*NT+(J'-1) of the last inflow line before plane
// p=(K-1)*NPA+(L-1), or -1.  One block per species; block-wide max-scan.
__device__ __forceinline__ bool line_inflow(const Tab &t, int s, int j, int k, int l) {
  return c_radial(t, s, t.nr - 1, j, k, l) < 0;
}

__global__ void k_ghost(Tab t) {
  __shared__ int s_scan[1024];
  const int s = blockIdx.x;
  const int P = t.ne * t.npa;
  const int nthr = blockDim.x;
  const int chunk = (P + nthr - 1) / nthr;
  const int p0 = threadIdx.x * chunk, p1 = min(P, p0 + chunk);
  int last = -1;
  for (int p = p0; p < p1; ++p) {
    const int k = p / t.npa, l = p % t.npa;
    for (int j = t.nt - 1; j >= 0; --j)
      if (line_inflow(t, s, j, k, l)) {
        last = p * t.nt + j;
        break;
      }
  }
  s_scan[threadIdx.x] = last;
  __syncthreads();
  for (int off = 1; off < nthr; off <<= 1) {  // inclusive max-scan
    const int v = (threadIdx.x >= off) ? s_scan[threadIdx.x - off] : -1;
    __syncthreads();
    s_scan[threadIdx.x] = max(s_scan[threadIdx.x], v);
    __syncthreads();
  }
  int run = (threadIdx.x > 0) ? s_scan[threadIdx.x - 1] : -1;
  for (int p = p0; p < p1; ++p) {
    t.carry[(size_t)s * P + p] = run;
    const int k = p / t.npa, l = p % t.npa;
    for (int j = t.nt - 1; j >= 0; --j)
      if (line_inflow(t, s, j, k, l)) {
        run = p * t.nt + j;
        break;
      }
  }
}

// F(NR+1) (conf=CONF1) or F(NR+2) (conf=CONF2) of inflow line (J,K,L):
// 0 outside the magnetopause, else FGEOS*CONF*FNHS(NR,J,L) (:161-167).
__device__ __forceinline__ double ghost_value(const Tab &t, int s, int j, int k, int l, double conf) {
  const int NR = t.nr;
  if (t.mgnp[(NR - 1) + (size_t)NR * j] == 1) return 0.0;
  const double fg = t.fgeos[s + (size_t)t.ns * (j + (size_t)t.nt * (k + (size_t)t.ne * l))] * t.bscale[s];
  return fg * conf * t.fnhs[(NR - 1) + (size_t)(NR + 1) * (j + (size_t)t.nt * l)];
}

__device__ double stale_ghost(const Tab &t, int s, int j, int k, int l) {
  for (int jj = j - 1; jj >= 0; --jj)
    if (line_inflow(t, s, jj, k, l)) return ghost_value(t, s, jj, k, l, t.conf1);
  const int p = k * t.npa + l;
  const int c = t.carry[(size_t)s * t.ne * t.npa + p];
  if (c < 0) return 0.0;
  const int pp = c / t.nt, jj = c % t.nt;
  return ghost_value(t, s, jj, pp / t.npa, pp % t.npa, t.conf1);
}

// ---------------------------------------------------------------------------
// IJ kernel: one block per (species, tile of KW = 2^kws consecutive (L,K)
// positions), holding F[(I,J)][w] and the face fluxes G = C*FBND in LDS, plus
// the inflow flag of every radial line.  Runs a program of OP_R / OP_P /
// OP_SUM / OP_SEAM.
__device__ __forceinline__ void atomic_max_pos(unsigned long long *slot, double v) {
  atomicMax(slot, (unsigned long long)__double_as_longlong(v));  // v >= 0
}

__global__ void __launch_bounds__(256) k_ij(Tab t, Program prog, int s0, int ntiles, int kws) {
  extern __shared__ double sm[];
  const int kw = 1 << kws;
  const int nij = t.nr * t.nt;
  double *F = sm;
  double *G = sm + (size_t)nij * kw;
  double *red = G + (size_t)nij * kw;  // 16 doubles
  int *inflow = (int *)(red + 16);      // NT*kw line flags
  const int s = s0 + blockIdx.x / ntiles;
  const int tile = blockIdx.x % ntiles;
  const int P = t.ne * t.npa;
  const int q0 = tile * kw;
  const int nw = min(kw, P - q0);
  const int ncell = nij * kw;
  double *gF = t.f2 + (size_t)s * f2_species_stride(t);
  const int NR = t.nr, NT = t.nt, NE = t.ne;
  const double beta = t.betalim;

  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int ij = c >> kws, w = c & (kw - 1);
    F[c] = (w < nw) ? gF[(size_t)ij * P + q0 + w] : 0.0;
  }
  __syncthreads();

  for (int o = 0; o < prog.n; ++o) {
    const int op = prog.op[o];
    if (op == OP_R) {
      // DRIFTR, src/ModRamDrift.f90:139-191: face I of line (J,w) per cell
      for (int n = threadIdx.x; n < NT * kw; n += blockDim.x) {  // sgn(NR) per line (:154)
        const int j = n >> kws, w = n & (kw - 1);
        int f = 0;
        if (w < nw) {
          const int q = q0 + w, l = fast_div(q, NE, t.rne), k = q - l * NE;
          f = line_inflow(t, s, j, k, l) ? 1 : 0;
        }
        inflow[n] = f;
      }
      __syncthreads();
      double cmax = 0.0;
      for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        const int ij = c >> kws, w = c & (kw - 1);
        if (w >= nw) continue;
        const int i = fast_div(ij, NT, t.rnt), j = ij - i * NT;
        const int q = q0 + w, l = fast_div(q, NE, t.rne), k = q - l * NE;
        const int I = i + 1;
        const double ci = c_radial(t, s, i, j, k, l);
        if (MGNP(I, j + 1) == 0) cmax = fmax(cmax, fmax(fabs(ci), 1E-10));
        const bool in = inflow[(j << kws) + w] != 0;
        const int base = (j << kws) + w;  // F(I') at base + (I'-1)*NT*kw
#define FR(II) F[base + (((II)-1) * NT << kws)]
        double fb;
        if (I == 1) {
          fb = in ? FR(2) : 0.0;
        } else if (!in && I == NR) {
          fb = FR(NR);
        } else {
          const double f0 = FR(I);
          double fp1, fn, fnm1;
          if (I == NR) {  // inflow only: ghosts F(NR+1), F(NR+2)
            fp1 = ghost_value(t, s, j, k, l, t.conf1);
            if (ci < 0) {
              fn = ghost_value(t, s, j, k, l, t.conf2);
              fnm1 = fp1;
            } else {
              fn = f0;
              fnm1 = FR(I - 1);
            }
          } else {
            fp1 = FR(I + 1);
            if (ci < 0) {
              fnm1 = fp1;
              if (I + 2 <= NR)
                fn = FR(I + 2);
              else  // I == NR-1: the ghost F(NR+1)
                fn = in ? ghost_value(t, s, j, k, l, t.conf1) : stale_ghost(t, s, j, k, l);
            } else {
              fn = f0;
              fnm1 = FR(I - 1);
            }
          }
          fb = face_value(f0, fp1, fn, fnm1, ci, ci, beta);
        }
#undef FR
        G[c] = ci * fb;
      }
      if (prog.arg[o]) {
        const double m = block_max(cmax, red);
        if (threadIdx.x == 0 && m > 0) atomic_max_pos(&t.cflmax[(size_t)s * (2 + NE + t.npa) + 0], m);
      }
      __syncthreads();
      for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        const int ij = c >> kws, w = c & (kw - 1);
        if (w >= nw || ij < NT) continue;  // I>=2
        double v = F[c] - G[c] + G[c - (NT << kws)];
        if (v < 0) v = 1E-15;
        F[c] = v;
      }
      __syncthreads();
    } else if (op == OP_P) {
      // DRIFTP, src/ModRamDrift.f90:229-272 (periodic in J)
      double cmax = 0.0;
      for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        const int ij = c >> kws, w = c & (kw - 1);
        if (w >= nw) continue;
        const int i = fast_div(ij, NT, t.rnt), j = ij - i * NT;
        const int J = j + 1;
        if (i < 1 || J < 2) continue;
        const int q = q0 + w, l = fast_div(q, NE, t.rne), k = q - l * NE;
        const double cp = c_azim(t, s, i, j, k, l);
        if (MGNP(i + 1, J) == 0) cmax = fmax(cmax, fmax(fabs(cp), 1E-10));
        const int base = ((i * NT) << kws) + w;  // F(J') at base + (J'-1)*kw
#define FP(JJ) F[base + (((JJ)-1) << kws)]
        const int J1 = (J == NT) ? 2 : J + 1;
        const double f0 = FP(J), fp1 = FP(J1);
        double fn, fnm1;
        if (cp < 0) {
          int N = J + 2;
          if (N > NT) N = N - NT + 1;  // :251
          fn = FP(N);
          fnm1 = FP(N - 1);
        } else {
          fn = f0;
          fnm1 = FP(J - 1);
        }
#undef FP
        G[c] = cp * face_value(f0, fp1, fn, fnm1, cp, cp, beta);
      }
      if (prog.arg[o]) {
        const double m = block_max(cmax, red);
        if (threadIdx.x == 0 && m > 0) atomic_max_pos(&t.cflmax[(size_t)s * (2 + NE + t.npa) + 1], m);
      }
      __syncthreads();
      for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        const int ij = c >> kws, w = c & (kw - 1);
        if (w >= nw) continue;
        const int i = fast_div(ij, NT, t.rnt), j = ij - i * NT;
        if (i < 1 || j < 1) continue;
        // FBND(1)=FBND(NT), CDriftP(I,1)=CDriftP(I,NT)  (:261-262)
        const double gm = (j == 1) ? G[((i * NT + NT - 1) << kws) + w] : G[c - kw];
        double v = F[c] - G[c] + gm;
        if (v < 0) v = 1E-15;
        F[c] = v;
      }
      __syncthreads();
      for (int n = threadIdx.x; n < (NR - 1) * kw; n += blockDim.x) {  // F2(S,I,1)=F2(S,I,NT)
        const int i = 1 + (n >> kws), w = n & (kw - 1);
        if (w < nw) F[((i * NT) << kws) + w] = F[((i * NT + NT - 1) << kws) + w];
      }
      __syncthreads();
    } else if (op == OP_SUM) {
      // SUMRC over I>=2, K>=2, L>=2, J<=NT-1 (src/ModRamRun.f90:246-255)
      double acc = 0.0;
      for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        const int ij = c >> kws, w = c & (kw - 1);
        if (w >= nw) continue;
        const int i = fast_div(ij, NT, t.rnt), j = ij - i * NT;
        const int q = q0 + w, l = fast_div(q, NE, t.rne), k = q - l * NE;
        if (i < 1 || k < 1 || l < 1 || j > NT - 2) continue;
        acc += t.ekev[k] * (F[c] * t.we[k] * t.wmu[l]);
      }
      const double tot = block_sum(acc, red);
      if (threadIdx.x == 0) t.partials[((size_t)prog.arg[o] * t.ns + s) * t.maxblk + tile] = tot;
      __syncthreads();
    } else if (op == OP_SEAM) {
      // F2(:,:,NT,:,:)=F2(:,:,1,:,:) then outside-magnetopause cells to 1e-31
      // (src/ModRamRun.f90:186,197-201)
      for (int n = threadIdx.x; n < NR * kw; n += blockDim.x) {
        const int i = n >> kws, w = n & (kw - 1);
        if (w < nw) F[((i * NT + NT - 1) << kws) + w] = F[((i * NT) << kws) + w];
      }
      __syncthreads();
      for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        const int ij = c >> kws, w = c & (kw - 1);
        if (w >= nw) continue;
        const int i = fast_div(ij, NT, t.rnt), j = ij - i * NT;
        if (MGNP(i + 1, j + 1) == 1) F[c] = 1e-31;
      }
      __syncthreads();
    }
  }

  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int ij = c >> kws, w = c & (kw - 1);
    if (w < nw) gF[(size_t)ij * P + q0 + w] = F[c];
  }
}

// ---------------------------------------------------------------------------
// KL kernel: one block per (species, I>=2, J), holding the whole (K,L) plane
// F[l*NE+k] in LDS.  The reference divides both face fluxes of cell K by that
// cell's width, CDriftE(K)/WE(K)*FBND(K) and CDriftE(K-1)/WE(K)*FBND(K-1)
// (src/ModRamDrift.f90:364; :460 for Mu), so the owner of face K writes the
// cell's own partial update O = F(K) - C(K)/WE(K)*FBND(K) and, in H, the term
// C(K)/WE(K+1)*FBND(K) its upper neighbour adds.
__device__ __forceinline__ double cex_factor(const Tab &t, int s, int k, size_t q) {
  // CHARGE=EXP(-ALPHA), ALPHA=10**Y*V*HDNS*DTs (src/ModRamLoss.f90:47-49)
  return exp(-(t.sigv[s * t.ne + k] * t.hdns[q] * t.dts));
}
__device__ __forceinline__ double atm_factor(const Tab &t, int s, int i, int k, size_t q) {
  // ATLOS(S,I,K)**(1/FNHS(I,J,L)) (src/ModRamLoss.f90:501) as
  // exp(log(ATLOS)*(1/FNHS))
  return exp(t.latl[((size_t)s * t.nr + i) * t.ne + k] * t.invfnhs[q]);
}

__global__ void __launch_bounds__(512) k_kl(Tab t, Program prog, int s0) {
  extern __shared__ double sm[];
  const int NE = t.ne, NPA = t.npa, NT = t.nt;
  const int P = NE * NPA;
  double *F = sm;
  double *O = F + P;
  double *H = O + P;
  double *red = H + P;  // 16
  unsigned long long *cmx = (unsigned long long *)(red + 16);  // NE + NPA
  const int nplane = (t.nr - 1) * NT;
  const int s = s0 + blockIdx.x / nplane;
  const int pl = blockIdx.x % nplane;
  const int i = 1 + pl / NT, j = pl % NT;  // I=i+1>=2
  double *gF = t.f2 + (size_t)s * f2_species_stride(t) + ((size_t)i * NT + j) * P;
  const double beta = t.betalim;
  const size_t qb = ijl_of(t, i, j, 0);
  const size_t ij = ij_of(t, i, j);
  const size_t sib = ((size_t)s * t.nr + i) * NE;
  const bool cex = t.cex[s] != 0;
  const int upa = t.upa[i];  // int(UPA(I))
  const bool sum_row = j <= NT - 2;
  const bool inside = MGNP(i + 1, j + 1) == 0;
  const double qs = (double)t.charge[s];  // x/QS == x*QS for QS = +-1
  const double rlz = t.rlz[i], rrlz = t.rrlz[i];
  const double bij = t.bij[ij], rbij = t.rbij[ij];
  const float rne = t.rne;
  unsigned long long *gcfl = t.cflmax + (size_t)s * (2 + NE + NPA);

  for (int c = threadIdx.x; c < P; c += blockDim.x) F[c] = gF[c];
  __syncthreads();

  for (int o = 0; o < prog.n; ++o) {
    const int op = prog.op[o];
    const bool rec = prog.arg[o] != 0 && inside;
    if (op == OP_E) {
      // DRIFTE, src/ModRamDrift.f90:333-369.  F(1),F(0) relativistic ghosts
      // from F(2); F(NE+1)=F(NE+2)=0.
      if (rec) {
        for (int k = threadIdx.x; k < NE; k += blockDim.x) cmx[k] = 0ull;
        __syncthreads();
      }
      const double *rel = t.rel + s * 8;
      const double drd1 = t.drd1[ij], dpd1 = t.dpd1[ij];
      for (int c = threadIdx.x; c < P; c += blockDim.x) {
        const int l = fast_div(c, NE, rne), k = c - l * NE;
        const int K = k + 1;
        const size_t q = qb + l;
        // CDriftE(I,J,K,L), :337-342
        const double edt1 = div_rn(div_rn(div_rn(t.ekp[s * NE + k], t.hs[q], t.rhs[q]), rlz, rrlz), bij, rbij) * qs;
        const double drdt = drd1 + edt1 * t.drd2[q] * rlz;
        const double dpdt = dpd1 - edt1 * t.dpd2[q];
        const double ce = t.edot[sib + k] * (t.gr[q] * drdt + t.gp[q] * dpdt + t.db1[q] + t.di1[q]);
        if (rec) lds_max(&cmx[k], fmax(fabs(ce), 1E-10));
        const double *row = F + l * NE;
        const double f2v = row[1];
        const double g1 = f2v * rel[0] / rel[1] * rel[2];
        const double g0 = g1 * rel[3] / rel[0] * rel[4];
#define FE(KK) ((KK) == 0 ? g0 : (KK) == 1 ? g1 : (KK) <= NE ? row[(KK)-1] : 0.0)
        const double f0 = FE(K), fp1 = FE(K + 1);
        const double fn = (ce < 0) ? FE(K + 2) : f0;
        const double fnm1 = (ce < 0) ? fp1 : FE(K - 1);
#undef FE
        const double fb = face_value(f0, fp1, fn, fnm1, ce, div_rn(ce, t.de[k], t.rde[k]), beta);
        O[c] = row[k] - div_rn(ce, t.we[k], t.rwe[k]) * fb;
        H[c] = (k + 1 < NE) ? div_rn(ce, t.we[k + 1], t.rwe[k + 1]) * fb : 0.0;
      }
      __syncthreads();
      for (int c = threadIdx.x; c < P; c += blockDim.x) {
        const int l = fast_div(c, NE, rne), k = c - l * NE;
        if (k >= 1) {
          double v = O[c] + H[c - 1];
          if (v < 0) v = 1E-15;
          F[c] = v;
        }
      }
      if (rec) {
        __syncthreads();
        for (int k = threadIdx.x; k < NE; k += blockDim.x)
          if (cmx[k]) atomicMax(&gcfl[2 + k], cmx[k]);
      }
      __syncthreads();
    } else if (op == OP_MU) {
      // DRIFTMU, src/ModRamDrift.f90:413-466.  F(1)=F(2) for the stencil;
      // limiter faces L=2..NPA-2; CDriftMu(1)=0, FBND(1)=0, FBND(NPA-1)=F(NPA).
      if (rec) {
        for (int l = threadIdx.x; l < NPA; l += blockDim.x) cmx[NE + l] = 0ull;
        __syncthreads();
      }
      const double drm1 = t.drm1[ij], dpm1 = t.dpm1[ij], db2 = t.db2[ij];
      for (int c = threadIdx.x; c < P; c += blockDim.x) {
        const int l = fast_div(c, NE, rne), k = c - l * NE;
        const int L = l + 1;
        double on = 0.0, h = 0.0;
        if (L >= 2) {
          const size_t q = qb + l;
          // CDriftMu(I,J,K,L), :418-433
          const double edt = div_rn(div_rn(div_rn(t.mkp[s * NE + k], t.bhs[q], t.rbhs[q]), rlz, rrlz), bij, rbij) * qs;
          const double drdm = drm1 + edt * t.drm2[q] * rlz;
          const double dpdm = dpm1 - edt * t.dpm2[q];
          const double cm = -t.cmudot[q] * (t.gmr[q] * drdm + t.gmp[q] * dpdm + db2 + t.di2[q]);
          if (rec) lds_max(&cmx[NE + l], fmax(fabs(cm), 1E-32));
          if (L <= NPA - 1) {
#define FM(LL) F[(((LL) == 1 ? 2 : (LL)) - 1) * NE + k]
            double fb;
            if (L == NPA - 1) {
              fb = FM(NPA);
            } else {
              const double f0 = FM(L), fp1 = FM(L + 1);
              const double fn = (cm < 0) ? FM(L + 2) : f0;
              const double fnm1 = (cm < 0) ? fp1 : FM(L - 1);
              fb = face_value(f0, fp1, fn, fnm1, cm, div_rn(cm, t.dmu[l], t.rdmu[l]), beta);
            }
#undef FM
            on = div_rn(cm, t.wmu[l], t.rwmu[l]) * fb;
            h = div_rn(cm, t.wmu[l + 1], t.rwmu[l + 1]) * fb;
          }
        }
        // face 1: CDriftMu(1)=0, FBND(1)=0, so its term is 0/WMU(2)*0 = 0
        O[c] = F[c] - on;
        H[c] = h;
      }
      __syncthreads();
      for (int c = threadIdx.x; c < P; c += blockDim.x) {
        const int l = fast_div(c, NE, rne);
        if (l >= 1 && l <= NPA - 2) {
          double v = O[c] + H[c - NE];
          if (v < 0) v = 1E-15;
          F[c] = v;
        }
      }
      __syncthreads();
      // F2(NPA)=F2(NPA-1)*FNHS(NPA)*MU(NPA)/FNHS(NPA-1)/MU(NPA-1)  (:466)
      const double hN = GEO3(t.fnhs, i + 1, j + 1, NPA), hN1 = GEO3(t.fnhs, i + 1, j + 1, NPA - 1);
      for (int k = threadIdx.x; k < NE; k += blockDim.x)
        F[(NPA - 1) * NE + k] = F[(NPA - 2) * NE + k] * hN * t.mu[NPA - 1] / hN1 / t.mu[NPA - 2];
      if (rec)
        for (int l = threadIdx.x; l < NPA; l += blockDim.x)
          if (cmx[NE + l]) atomicMax(&gcfl[2 + NE + l], cmx[NE + l]);
      __syncthreads();
    } else if (op == OP_CEX) {
      if (cex) {  // CHAREXCHANGE, src/ModRamLoss.f90:467-475: K>=2, L>=2
        for (int c = threadIdx.x; c < P; c += blockDim.x) {
          const int l = fast_div(c, NE, rne), k = c - l * NE;
          if (k >= 1 && l >= 1) F[c] = F[c] * cex_factor(t, s, k, qb + l);
        }
      }
      __syncthreads();
    } else if (op == OP_ATM) {
      // ATMOL, src/ModRamLoss.f90:495-503: K>=2, L>=int(UPA(I))
      for (int c = threadIdx.x; c < P; c += blockDim.x) {
        const int l = fast_div(c, NE, rne), k = c - l * NE;
        if (k >= 1 && l + 1 >= upa) F[c] = F[c] * atm_factor(t, s, i, k, qb + l);
      }
      __syncthreads();
    } else if (op == OP_SUM) {
      double acc = 0.0;
      if (sum_row)
        for (int c = threadIdx.x; c < P; c += blockDim.x) {
          const int l = fast_div(c, NE, rne), k = c - l * NE;
          if (k >= 1 && l >= 1) acc += t.ekev[k] * (F[c] * t.we[k] * t.wmu[l]);
        }
      const double tot = block_sum(acc, red);
      if (threadIdx.x == 0) t.partials[((size_t)prog.arg[o] * t.ns + s) * t.maxblk + pl] = tot;
      __syncthreads();
    } else if (op == OP_LOSS) {
      // ram_run :77-141 between the two drift halves: SUMRC, [CEX, SUMRC],
      // ATMOL, SUMRC, ATMOL, SUMRC, [CEX, SUMRC].  Slots arg..arg+4.
      double a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0;
      for (int c = threadIdx.x; c < P; c += blockDim.x) {
        const int l = fast_div(c, NE, rne), k = c - l * NE;
        const bool w_on = sum_row && k >= 1 && l >= 1;
        const bool cx = cex && k >= 1 && l >= 1;
        const bool atm = k >= 1 && l + 1 >= upa;
        double f = F[c];
        const double cxf = cx ? cex_factor(t, s, k, qb + l) : 1.0;
        const double af = atm ? atm_factor(t, s, i, k, qb + l) : 1.0;
        const double wk = t.we[k], wl = t.wmu[l], ek = t.ekev[k];
        if (w_on) a1 += ek * (f * wk * wl);
        if (cx) f = f * cxf;
        if (w_on) a2 += ek * (f * wk * wl);
        if (atm) f = f * af;
        if (w_on) a3 += ek * (f * wk * wl);
        if (atm) f = f * af;
        if (w_on) a4 += ek * (f * wk * wl);
        if (cx) f = f * cxf;
        if (w_on) a5 += ek * (f * wk * wl);
        F[c] = f;
      }
      const double acc[5] = {a1, a2, a3, a4, a5};
      for (int n = 0; n < 5; ++n) {
        const double tot = block_sum(acc[n], red);
        if (threadIdx.x == 0) t.partials[((size_t)(prog.arg[o] + n) * t.ns + s) * t.maxblk + pl] = tot;
      }
      __syncthreads();
    }
  }

  for (int c = threadIdx.x; c < P; c += blockDim.x) gF[c] = F[c];
}

// ---------------------------------------------------------------------------
// DtDrift*(S) from the CFL maxima (cflmask bits 1 R, 2 P, 4 E, 8 Mu): min
// over cells of FracCFL*DTs*w/ctemp equals FracCFL*DTs*w / max(ctemp) because
// correctly rounded division is monotone; initial values 1e5 (R,P) and 1e4
// (E,Mu) as the reference's (:115,223,308,404).  The maxima are cleared after
// use for the next recording.
// SUMRC bookkeeping (src/ModRamRun.f90:77-175,239-257), summode 1: one SUMRC
// from slot 0 (SETRC:=S, ELORC=old-S); summode 2: the full ram_run chain from
// slots 0..5 (slot 5 from the IJ tiles).
__global__ void k_finalize(Tab t, int cflmask, int summode, int s0, int nblk_kl, int nblk_ij,
                           double *elorc_out) {
  __shared__ double red[16];
  __shared__ double S[6];
  const int s = s0 + blockIdx.x;
  const int NE = t.ne, NPA = t.npa;
  const double a = t.fraccfl * t.dts;
  unsigned long long *m = t.cflmax + (size_t)s * (2 + NE + NPA);
  if (cflmask) {
    double de = 10000.0, dm = 10000.0;
    if (cflmask & 4)
      for (int k = threadIdx.x; k < NE; k += blockDim.x) {
        const double v = __longlong_as_double((long long)m[2 + k]);
        if (v > 0) de = fmin(de, a * t.de[k] / v);
      }
    if (cflmask & 8)
      for (int l = 1 + threadIdx.x; l < NPA; l += blockDim.x) {
        const double v = __longlong_as_double((long long)m[2 + NE + l]);
        if (v > 0) dm = fmin(dm, a * t.dmu[l] / v);
      }
    de = block_min(de, red);
    __syncthreads();
    dm = block_min(dm, red);
    if (threadIdx.x == 0) {
      if (cflmask & 1) {
        const double v = __longlong_as_double((long long)m[0]);
        t.dtdrift[s * 4 + 0] = (v > 0) ? fmin(100000.0, a / v) : 100000.0;
      }
      if (cflmask & 2) {
        const double v = __longlong_as_double((long long)m[1]);
        t.dtdrift[s * 4 + 1] = (v > 0) ? fmin(100000.0, a / v) : 100000.0;
      }
      if (cflmask & 4) t.dtdrift[s * 4 + 2] = de;
      if (cflmask & 8) t.dtdrift[s * 4 + 3] = dm;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 2 + NE + NPA; e += blockDim.x) m[e] = 0ull;
  }
  if (summode == 0) return;
  const int nslot = (summode == 1) ? 1 : 6;
  for (int sl = 0; sl < nslot; ++sl) {
    const int nb = (summode == 2 && sl == 5) ? nblk_ij : nblk_kl;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += t.partials[((size_t)sl * t.ns + s) * t.maxblk + b];
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) S[sl] = tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double *e = t.energy + s * 4;
    const double old = e[0];
    if (summode == 1) {
      e[0] = S[0];
      if (elorc_out) elorc_out[s] = old - S[0];
    } else {
      const bool cex = t.cex[s] != 0;
      double prev = old;
      e[1] += prev - S[0];  // LSDR
      prev = S[0];
      if (cex) {
        e[2] += prev - S[1];  // LSCHA
        prev = S[1];
      }
      e[3] += prev - S[2];  // LSATM
      prev = S[2];
      e[3] += prev - S[3];  // LSATM
      prev = S[3];
      if (cex) {
        e[2] += prev - S[4];  // LSCHA
        prev = S[4];
      }
      e[1] += prev - S[5];  // LSDR
      e[0] = S[5];
    }
  }
}

// VT(I,J)=AVS*(LZ(I)*RE)**2*SIN(PHI(J)-PHIOFS), src/ModRamRun.f90:45-54
__global__ void k_vs(double *vt, const double *lz, const double *phi, const double *kp_ptr,
                     double kp_val, int nr1, int nt) {
  const double kp = kp_ptr ? *kp_ptr : kp_val;
  const double d = 1.0 - 0.159 * kp + 0.0093 * (kp * kp);
  const double avs = 7.05E-6 / (d * d * d) / kRE;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nr1 * nt; n += gridDim.x * blockDim.x) {
    const int i = n % nr1, j = n / nr1;
    const double lre = lz[i] * kRE;
    vt[n] = avs * (lre * lre) * sin(phi[j] - 0.0);
  }
}

// Plain streaming copy (16-byte accesses), the bandwidth and counter
// calibration probe.
__global__ void k_copy(double2 *__restrict__ dst, const double2 *__restrict__ src, size_t n2) {
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n2; x += (size_t)gridDim.x * blockDim.x)
    dst[x] = src[x];
}

// Fortran F2(NS,NR,NT,NE,NPA) <-> device [S][I][J][L][K]
__global__ void k_f2_in(double *dst, const double *src, int ns, int nr, int nt, int ne, int npa) {
  const size_t n = (size_t)ns * nr * nt * ne * npa;
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
    size_t r = x;  // device index: k fastest, then l, j, i, s
    const int k = r % ne;
    r /= ne;
    const int l = r % npa;
    r /= npa;
    const int j = r % nt;
    r /= nt;
    const int i = r % nr;
    r /= nr;
    const int s = (int)r;
    dst[x] = src[s + (size_t)ns * (i + (size_t)nr * (j + (size_t)nt * (k + (size_t)ne * l)))];
  }
}

__global__ void k_f2_out(double *dst, const double *src, int ns, int nr, int nt, int ne, int npa) {
  const size_t n = (size_t)ns * nr * nt * ne * npa;
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
    size_t r = x;
    const int k = r % ne;
    r /= ne;
    const int l = r % npa;
    r /= npa;
    const int j = r % nt;
    r /= nt;
    const int i = r % nr;
    r /= nr;
    const int s = (int)r;
    dst[s + (size_t)ns * (i + (size_t)nr * (j + (size_t)nt * (k + (size_t)ne * l)))] = src[x];
  }
}

}  // namespace ramgpu
