// sweep_common.h -- device helpers shared by the forward kernels
// (kernels.hip) and their adjoints (adjoint.hip): exact-rounding arithmetic,
// the limited face value, launch units, block reductions, the DRIFTR
// Courant number and ghosts, and the LDS layout of the plane kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "ramgpu_dev.h"

namespace ramgpu {

static constexpr double kOME = 7.3E-5;      // src/ModRamDrift.f90:224
static constexpr double kQ = 1.602E-19;     // src/ModRamConst.f90:19
static constexpr double kCS = 2.998E8;      // src/ModRamConst.f90:20
static constexpr double kRE = 6.371E6;      // src/ModRamConst.f90:14

// geometry accessors, 1-based as the reference
#define GEO3(A, I, J, L) (A)[((I)-1) + (size_t)(t.nr + 1) * (((J)-1) + (size_t)t.nt * ((L)-1))]
#define GEO2(A, I, J) (A)[((I)-1) + (size_t)(t.nr + 1) * ((J)-1)]
#define MGNP(I, J) t.mgnp[((I)-1) + (size_t)t.nr * ((J)-1)]

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

// exp(x) by Tang's table-driven method (ACM TOMS 15(2):144-157, 1989):
// x = (64m + j) ln2/64 + r with |r| <= ln2/128, exp(x) = 2^m 2^(j/64) (1 + p(r)),
// p the degree-5 Taylor polynomial of e^r - 1 (truncation below 4e-17; Horner
// form with fused multiply-adds), tab[j]
// = 2^(j/64) correctly rounded.  Within about 1 ulp of exp; it serves the loss
// factors, whose bar is 1e-12 relative.  Any NaN or out-of-range argument
// still indexes tab within 0..63.
__host__ __device__ inline double exp_tab(double x, const double *tab) {
  const double inv = 92.33248261689366;                           // 64/ln2
  const double l1 = 0.010830424696249145, l2 = 3.623510646634843e-19;  // ln2/64 = l1 + l2
  double n = rint(x * inv);
  n = fmin(fmax(n, -131072.0), 131072.0);
  const double r = fma(-n, l2, fma(-n, l1, x));
  const int ni = (int)n;
  const int j = ni & 63, m = (ni - j) / 64;
  const double p = r * fma(r, fma(r, fma(r, fma(r, 1.0 / 120.0, 1.0 / 24.0), 1.0 / 6.0), 0.5), 1.0);
  return ldexp(fma(tab[j], p, tab[j]), m);
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

// face_value without branches: every lane forms the ratio and the limited
// value, then keeps FUP where the reference returns early (|X| <= 1e-27 or
// R <= 0); the arithmetic of the value kept is face_value's.
__device__ __forceinline__ double face_value_sel(double f0, double fp1, double fn, double fnm1,
                                                 double c, double cn, double beta) {
  const double sg = (c < 0) ? -1.0 : 1.0;
  const double x = fp1 - f0;
  const double fup = 0.5 * (f0 + fp1 - sg * x);
  const double r = (fn - fnm1) / x;
  const double lim = fmax(fmin(beta * r, 1.0), fmin(r, beta));
  const double corr = -0.5 * (cn - sg) * x;
  const double v = fup + lim * corr;
  return (fabs(x) <= 1.E-27 || r <= 0) ? fup : v;
}

// The same face value with its partial derivatives with respect to the four
// stencil values (reverse mode).  fb = FUP + lim(R)*CORR is piecewise linear:
// on each piece lim = a + b*R, so with kap = -0.5*(cn-sg), X = fp1-f0,
//   dfb/dfn = kap*b,  dfb/dfnm1 = -kap*b,  dfb/dX = kap*a,
//   dfb/df0 = (1+sg)/2 - kap*a,  dfb/dfp1 = (1-sg)/2 + kap*a.
// Subgradient at the kinks of lim = max(min(beta*R,1), min(R,beta)): a tie
// of min or max takes the derivative of its first argument; R <= 0 and
// |X| <= 1e-27 take the upwind branch (lim = 0), as the forward does.
struct FaceD {
  double fb, d0, d1, dn, dm;
};

// Every case is formed and the one that applies is selected (no divergent
// branches in the adjoint sweeps), with the values of the case analysis:
// the upwind face when |X| <= 1e-27 or R <= 0, else the limited face, its
// derivatives from the active branch of the limiter (a, b).
__device__ __forceinline__ FaceD face_d(double f0, double fp1, double fn, double fnm1, double c, double cn,
                                        double beta) {
  const double sg = (c < 0) ? -1.0 : 1.0;
  const double x = fp1 - f0;
  const double fup = 0.5 * (f0 + fp1 - sg * x);
  const double u0 = 0.5 * (1.0 + sg), u1 = 0.5 * (1.0 - sg);
  const double rr = (fn - fnm1) / x;
  const bool lin = !(fabs(x) <= 1.E-27) && !(rr <= 0);
  const double A = fmin(beta * rr, 1.0), B = fmin(rr, beta);
  const double lim = fmax(A, B);
  const bool ab = A >= B;
  const bool lo = ab ? (beta * rr <= 1.0) : (rr <= beta);
  const double a = lo ? 0.0 : (ab ? 1.0 : beta);
  const double b = lo ? (ab ? beta : 1.0) : 0.0;
  const double kap = -0.5 * (cn - sg);
  FaceD r;
  r.fb = lin ? fup + lim * (kap * x) : fup;
  r.d0 = lin ? u0 - kap * a : u0;
  r.d1 = lin ? u1 + kap * a : u1;
  r.dn = lin ? kap * b : 0.0;
  r.dm = lin ? -kap * b : 0.0;
  return r;
}

// Global-memory views of pointers: loads and stores through them are
// global (not flat) instructions, so a wait for one load does not also wait
// for LDS traffic or for every earlier flat access.
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<T> as_global(T *p) {
  return (gptr<T>)p;
}
// element n of a double4 / double2 table read through a global view
__device__ __forceinline__ double4 ld_d4(gptr<const double> p, size_t n) {
  const gptr<const double> q = p + 4 * n;
  return make_double4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ double2 ld_d2(gptr<const double> p, size_t n) {
  const gptr<const double> q = p + 2 * n;
  return make_double2(q[0], q[1]);
}

// ---------------------------------------------------------------------------
// Launch units
struct Unit {
  const Tab *t;
  int s, b;
};

__device__ __forceinline__ Unit unit_of(const Launch &L, int vb) {
  if (!L.vmap) {
    if (L.ilv) {
      const int nsp = L.nv / L.nb, span = 8 * nsp;
      const int g = vb / span, r = vb - g * span;
      const int rem = min(8, L.nb - 8 * g);
      const int s = r / rem;
      return {L.tabs, L.s0 + s, 8 * g + (r - s * rem)};
    }
    const int n = vb / L.nb;
    return {L.tabs, L.s0 + n, vb - n * L.nb};
  }
  const int e = L.vmap[vb];
  const Job jb = L.jobs[e >> kVmapShift];
  return {L.tabs + jb.tab, jb.s, e & ((1 << kVmapShift) - 1)};
}

// ---------------------------------------------------------------------------
// Face terms of the radial lines of one block (kLineThreads consecutive
// lines from line0), staged in LDS by the whole block: q1 of the block's
// (J,L) pairs [pair][I], rij of its J values [j][I][4], rsi [I][2] and the
// flags rij(I,NT).w [I] (radial_lds_bytes).  The sweeps then read their face
// terms without memory latency.
// source of element e of the tables
__device__ __forceinline__ gptr<const double> radial_term_src(const Tab &t, int e, int npair, int nj, int pair0,
                                                              int j0) {
  const int NR = t.nr, NT = t.nt;
  const int nq = npair * NR, nr4 = nj * NR * 4;
  if (e < nq) {
    const int pp = e / NR, i = e - pp * NR, pair = pair0 + pp;
    const int jj = pair / t.npa, l = pair - jj * t.npa;
    return as_global(t.q1) + ((size_t)l * NR + i) * NT + jj;
  }
  const gptr<const double> rij = as_global((const double *)t.rij);
  e -= nq;
  if (e < nr4) {
    const int jl = e / (4 * NR), rem = e - jl * 4 * NR, i = rem >> 2;
    return rij + ((size_t)i * NT + j0 + jl) * 4 + (rem & 3);
  }
  e -= nr4;
  if (e < 2 * NR) return as_global((const double *)t.rsi) + e;
  e -= 2 * NR;
  return rij + ((size_t)e * NT + NT - 1) * 4 + 3;
}

// Face terms of the azimuthal rings of one block (lines line0..lastl, line
// = (I-2)*NE*NPA + (L-1)*NE + K-1): pl of the block's (I,L) pairs [pair][J][4]
// and pij of its I values [i][J][4] (azim_lds_bytes), staged by the whole
// block.
struct AzimTerms {
  const double *pl, *pij;
  int pair0, i0, nt;
  __device__ const double *pl_row(int i1, int l, int npa) const { return pl + (size_t)(i1 * npa + l - pair0) * nt * 4; }
  __device__ const double *pij_row(int i1) const { return pij + (size_t)(i1 - i0) * nt * 4; }
};
__device__ inline AzimTerms stage_azim_terms(const Tab &t, int line0, int lastl, double *sterm) {
  const int NR = t.nr, NT = t.nt, NE = t.ne, NPA = t.npa, P = NE * NPA;
  const int pair0 = line0 / NE, npair = lastl / NE - pair0 + 1;
  const int i0 = line0 / P, ni = lastl / P - i0 + 1;
  const int n4 = NT * 4, ntot = (npair + ni) * n4;
  const gptr<const double> plg = as_global((const double *)t.pl), pijg = as_global((const double *)t.pij);
  for (int e0 = threadIdx.x; e0 < ntot; e0 += 4 * blockDim.x) {
    double v[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int e = e0 + m * blockDim.x;
      v[m] = 0.0;
      if (e < npair * n4) {
        const int pp = e / n4, rem = e - pp * n4, pair = pair0 + pp;
        const int i1 = pair / NPA, l = pair - i1 * NPA;
        v[m] = plg[((size_t)l * NR + i1 + 1) * n4 + rem];
      } else if (e < ntot) {
        const int f = e - npair * n4, ii = f / n4, rem = f - ii * n4;
        v[m] = pijg[(size_t)(i0 + ii + 1) * n4 + rem];
      }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
      if (e0 + m * blockDim.x < ntot) sterm[e0 + m * blockDim.x] = v[m];
  }
  __syncthreads();
  AzimTerms at;
  at.pl = sterm;
  at.pij = sterm + (size_t)npair * n4;
  at.pair0 = pair0;
  at.i0 = i0;
  at.nt = NT;
  return at;
}

struct RadialTerms {
  const double *q, *r, *rs, *mg;
  int pair0, j0, nr;
  __device__ const double *q_row(int j, int l, int npa) const { return q + (size_t)(j * npa + l - pair0) * nr; }
  __device__ const double *r_row(int j) const { return r + (size_t)(j - j0) * nr * 4; }
};
__device__ inline RadialTerms stage_radial_terms(const Tab &t, int line0, double *sterm) {
  const int NR = t.nr, NE = t.ne, P = NE * t.npa;
  const int lastl = min(line0 + kLineThreads, t.nt * P) - 1;
  const int pair0 = line0 / NE, npair = lastl / NE - pair0 + 1;
  const int j0 = line0 / P, nj = lastl / P - j0 + 1;
  const int ntot = (npair + 4 * nj + 3) * NR;
  for (int e0 = threadIdx.x; e0 < ntot; e0 += 4 * kLineThreads) {
    double v[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int e = e0 + m * kLineThreads;
      v[m] = (e < ntot) ? *radial_term_src(t, e, npair, nj, pair0, j0) : 0.0;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
      if (e0 + m * kLineThreads < ntot) sterm[e0 + m * kLineThreads] = v[m];
  }
  __syncthreads();
  RadialTerms rt;
  rt.q = sterm;
  rt.r = rt.q + (size_t)npair * NR;
  rt.rs = rt.r + (size_t)nj * NR * 4;
  rt.mg = rt.rs + 2 * NR;
  rt.pair0 = pair0;
  rt.j0 = j0;
  rt.nr = NR;
  return rt;
}

// ---------------------------------------------------------------------------
// Block reductions (fixed order: deterministic for a given launch shape).
template <int N>
__device__ inline void block_sum_n(double (&v)[N], double *red, double (&out)[N]) {
  for (int m = 0; m < N; ++m)
    for (int off = 32; off > 0; off >>= 1) v[m] += __shfl_down(v[m], off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
    for (int m = 0; m < N; ++m) red[wid * N + m] = v[m];
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  for (int m = 0; m < N; ++m) {
    double tot = 0.0;
    for (int w = 0; w < nw; ++w) tot += red[w * N + m];
    out[m] = tot;
  }
  __syncthreads();
}

__device__ inline double block_max(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double m = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int w = 0; w < nw; ++w) m = fmax(m, red[w]);
  __syncthreads();
  return m;
}

__device__ inline double block_min(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double m = v;
  const int nw = (blockDim.x + 63) >> 6;
  for (int w = 0; w < nw; ++w) m = fmin(m, red[w]);
  __syncthreads();
  return m;
}

__device__ __forceinline__ void atomic_max_pos(unsigned long long *slot, double v) {
  atomicMax(slot, (unsigned long long)__double_as_longlong(v));  // v >= 0
}

// Radial Courant number CDriftR(I,J,K,L) = CR(I,J) + CGR with
// CGR = CGR3/(FNHS+FNHS')*P4/2./(BNES+BNES')/(RLZ+0.5*MDR)
// (src/ModRamDrift.f90:140-146).  One function for the sweep, the CFL and the
// ghost carry, so the inflow/outflow decision is the same everywhere.
// from the face's terms: r = rij(I,J), rs = rsi(I), q1 = q1(L,I,J)
__device__ __forceinline__ double c_radial_of(double4 r, double2 rs, double q1, double p4k) {
  const double x = (q1 * p4k) * 0.5;
  return r.x + div_rn(div_rn(x, r.y, r.z), rs.x, rs.y);
}
__device__ __forceinline__ double c_radial(const Tab &t, double p4k, int i, int j, int l) {
  return c_radial_of(t.rij[(size_t)i * t.nt + j], t.rsi[i], t.q1[((size_t)l * t.nr + i) * t.nt + j], p4k);
}

// CDriftE(I,J,K,L) (src/ModRamDrift.f90:337-342) from the per-(I,J,L)
// terms e[KleField], the per-(I,J) DRD1, DPD1, BNES and the per-(S,K)
// prefix ekpq = ekp*QS and EDOT.  The reference's EDT1 = ekp/FNHS/RLZ/BNES/QS
// with QS = +-1 equals the same chain of correctly rounded quotients started
// from ekp*QS: every step is odd in its dividend, so the sign passes through
// exactly.
__device__ __forceinline__ double c_energy(const double *e, double ekpq, double edot, double drd1, double dpd1,
                                           double bij, double rbij, double rlz, double rrlz) {
  const double edt1 = div_rn(div_rn(div_rn(ekpq, e[KLE_HS], e[KLE_RHS]), rlz, rrlz), bij, rbij);
  const double drdt = drd1 + edt1 * e[KLE_DRD2] * rlz;
  const double dpdt = dpd1 - edt1 * e[KLE_DPD2];
  return edot * (e[KLE_GR] * drdt + e[KLE_GP] * dpdt + e[KLE_DB1] + e[KLE_DI1]);
}

// CDriftMu(I,J,K,L) (src/ModRamDrift.f90:418-433) from the per-(I,J,L)
// terms m[KlmField], the per-(I,J) DRM1, DPM1, dBdt2, BNES and the per-(S,K)
// prefix mkpq = mkp*QS (QS folded in as in c_energy).
__device__ __forceinline__ double c_mu(const double *m, double mkpq, double drm1, double dpm1, double db2,
                                       double bij, double rbij, double rlz, double rrlz) {
  const double edt = div_rn(div_rn(div_rn(mkpq, m[KLM_BHS], m[KLM_RBHS]), rlz, rrlz), bij, rbij);
  const double drdm = drm1 + edt * m[KLM_DRM2] * rlz;
  const double dpdm = dpm1 - edt * m[KLM_DPM2];
  return -m[KLM_CMUDOT] * (m[KLM_GMR] * drdm + m[KLM_GMP] * dpdm + db2 + m[KLM_DI2]);
}

__device__ __forceinline__ bool line_inflow(const Tab &t, int s, int j, int k, int l) {
  return c_radial(t, t.p4[s * t.ne + k], t.nr - 1, j, l) < 0;
}

// F(NR+1) (conf=CONF1) or F(NR+2) (conf=CONF2) of inflow line (J,K,L):
// 0 outside the magnetopause, else FGEOS*CONF*FNHS(NR,J,L) (:161-167).
__device__ __forceinline__ double ghost_value(const Tab &t, int s, int j, int k, int l, double conf) {
  const int NR = t.nr;
  if (t.mgnp[(NR - 1) + (size_t)NR * j] == 1) return 0.0;
  const double fg = t.fgeos[s + (size_t)t.ns * (j + (size_t)t.nt * (k + (size_t)t.ne * l))] * t.bscale[s];
  return fg * conf * t.fnhs[(NR - 1) + (size_t)(NR + 1) * (j + (size_t)t.nt * l)];
}

__device__ inline double stale_ghost(const Tab &t, int s, int j, int k, int l) {
  for (int jj = j - 1; jj >= 0; --jj)
    if (line_inflow(t, s, jj, k, l)) return ghost_value(t, s, jj, k, l, t.conf1);
  const int p = k * t.npa + l;
  const int c = t.carry[(size_t)s * t.ne * t.npa + p];
  if (c < 0) return 0.0;
  const int pp = c / t.nt, jj = c % t.nt;
  return ghost_value(t, s, jj, pp / t.npa, pp % t.npa, t.conf1);
}

// The line (J',K',L') whose F(NR+1) stale_ghost reads, as its FGEOS index
// s + NS*(J' + NT*(K' + NE*L')), or -1 when it reads 0.
__device__ inline long stale_ghost_source(const Tab &t, int s, int j, int k, int l) {
  for (int jj = j - 1; jj >= 0; --jj)
    if (line_inflow(t, s, jj, k, l)) return s + (long)t.ns * (jj + (long)t.nt * (k + (long)t.ne * l));
  const int c = t.carry[(size_t)s * t.ne * t.npa + k * t.npa + l];
  if (c < 0) return -1;
  const int pp = c / t.nt, jj = c % t.nt;
  return s + (long)t.ns * (jj + (long)t.nt * (pp / t.npa + (long)t.ne * (pp % t.npa)));
}

// d F(NR+1) / d FGEOS(S,J,K,L) of an inflow line (conf = CONF1 or CONF2)
__device__ __forceinline__ double ghost_dfgeos(const Tab &t, int s, int j, int l, double conf) {
  const int NR = t.nr;
  if (t.mgnp[(NR - 1) + (size_t)NR * j] == 1) return 0.0;
  return t.bscale[s] * conf * t.fnhs[(NR - 1) + (size_t)(NR + 1) * (j + (size_t)t.nt * l)];
}

enum PlaneSeq : int {
  SEQ_E = 0,     // DRIFTE
  SEQ_MU = 1,    // DRIFTMU
  SEQ_CEX = 2,   // CHAREXCHANGE
  SEQ_ATM = 3,   // ATMOL
  SEQ_SUM = 4,   // SUMRC partial sum, slot 0
  SEQ_STEP = 5,  // the middle of ram_run: DRIFTE, DRIFTMU, loss chain, DRIFTMU, DRIFTE
  SEQ_LOSS = 6,  // the loss chain alone (WAVELO, CEX, ATMOL, ATMOL, CEX, WAVELO with seven SUMRC partials)
  SEQ_TAPE = 7,  // SEQ_STEP out of place, also writing the input of each later phase to a tape
  SEQ_WAVE = 8,  // WAVELO
  SEQ_COULEN = 9,   // COULEN
  SEQ_COULMU = 10,  // COULMU
  SEQ_STEPC = 11,   // SEQ_STEP with Coulomb collisions: DRIFTE, DRIFTMU, COULEN, COULMU, loss chain,
                    // COULMU, COULEN, DRIFTMU, DRIFTE with their SUMRC partials
};
// plane kernels whose LDS holds the Coulomb operators' tables
__host__ __device__ constexpr bool seq_coul(int seq) {
  return seq == SEQ_COULEN || seq == SEQ_COULMU || seq == SEQ_STEPC;
}

// Elements of one F2-sized buffer (all species) of the problem t.
__device__ __forceinline__ size_t f2_elems(const Tab &t) { return (size_t)t.ns * f2_species_stride(t); }

// LDS layout of the plane kernel (doubles)
struct PlaneLds {
  double *F, *H, *cf, *cm, *cl, *kt, *lt, *red, *ex, *pc, *gh;
  unsigned long long *cmx;
};
// per-K terms of a plane; KT_WAVE is WAVELO's decay factor exp(-DTs/TAU) of
// the plane's (I,J) (1 for K=1 and for species without WAVELO)
enum KtField { KT_EKP, KT_MKP, KT_SIGV, KT_EDOT, KT_LATL, KT_DE, KT_RDE, KT_WE, KT_RWE, KT_EKEV, KT_WAVE, KT_N };
enum LtField { LT_DMU, LT_RDMU, LT_WMU, LT_RWMU, LT_N };
// Scalars of one plane, staged in LDS with the plane so that the phases read
// them as broadcasts instead of making dependent global loads after each
// barrier: the (I,J) row of kij, the species' relativistic factors, the
// top-cell FNHS(NPA), FNHS(NPA-1), MU(NPA), MU(NPA-1) of DRIFTMU, and DTs.
enum PcField { PC_KIJ = 0, PC_REL = 8, PC_HN = 16, PC_HN1, PC_MUN, PC_MUN1, PC_DTS, PC_N = 24 };

__device__ __forceinline__ PlaneLds plane_lds(double *sm, int ne, int npa) {
  const int P = ne * npa;
  PlaneLds d;
  d.F = sm;
  d.H = d.F + P;
  d.cf = d.H + P;            // 8*NPA: per-L DRIFTE terms (KleField)
  d.cm = d.cf + 8 * npa;     // 8*NPA: per-L DRIFTMU terms (KlmField)
  d.cl = d.cm + 8 * npa;     // 2*NPA: HDNS, 1/FNHS
  d.kt = d.cl + 2 * npa;     // KT_N*NE
  d.lt = d.kt + KT_N * ne;   // LT_N*NPA
  d.red = d.lt + LT_N * npa; // 128: block reductions (up to 16 waves x 7 sums: 112)
  d.ex = d.red + 128;        // 64: 2^(j/64) of exp_tab
  d.pc = d.ex + 64;          // PC_N: the plane's scalars (PcField)
  d.gh = d.pc + PC_N;        // 2*NPA: DRIFTE's relativistic ghosts F(1), F(0) of each row
  d.cmx = (unsigned long long *)(d.gh + 2 * npa);  // NE + kMuSpread*NPA
  return d;
}

struct PlaneCtx {
  int s, i, j, NE, NPA, NT, P, nmine, nmax;  // nmax: the most cells any thread of the block owns
  int cex, upa;  // the species' charge-exchange kind, UPA(I)
  size_t ij;
  double rlz, rrlz, qs, beta;
  float rne;
  // row-wise ownership of the forward plane kernels: this thread owns K = ke
  // of rows L = le0, le0+lestep, .. with lestep = blockDim/NE
  int ke, le0, lestep;
  bool eact;
};

// (K,L) of the n-th cell owned by this thread of a BT-thread block (plane adjoints): cell threadIdx.x + n*BT
#define CELL_KL_T(n, BT)                                    \
  const int cc = threadIdx.x + (n) * (BT);                   \
  const int l = fast_div(cc, c.NE, c.rne), k = cc - l * c.NE; \
  (void)k

// Block copy of n doubles between global memory and LDS (either way).  A
// strided loop of plain assignments waits for each load before its store
// (flat loads are not moved past stores that may alias them), so each pass
// costs a memory latency; here every thread has up to U loads in flight
// before its first store.
template <int U>
__device__ __forceinline__ void block_copy(double *dst, const double *src, int n) {
  const int T = blockDim.x;
  for (int i0 = threadIdx.x; i0 < n; i0 += U * T) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = (i0 + u * T < n) ? src[i0 + u * T] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * T < n) dst[i0 + u * T] = v[u];
  }
}

// WAVELO's decay factor exp(-DTs/TAU_LIF) of cell row (I,J) at energy K of
// species s (src/ModRamWPI.f90:596-630) at Kp kp, 0-based i, j, k with k >= 1
// (tabulated per step by k_coef_ij, read by the plane kernels): the
// lifetime WALOS1 scaled by the wave amplitude Bw (30 pT, 100 above Kp 4)
// inside the plasmapause; outside, WALOS2 with the strong-diffusion WALOS3
// and the reference's corrections below 5 keV, or 5 days / KP above 1 MeV.
__device__ inline double wave_factor(const Tab &t, int s, int i, int j, int k, double kp) {
  const size_t nik = (size_t)t.nr * t.ne, ik = (size_t)i * t.ne + k;
  const double lz = t.lz[i], rl = t.rlpp[j], ek = t.ekev[k];
  const double bw = (kp >= 4.0) ? 100. : 30.;
  double tau = 0.0;
  if (lz <= rl) {
    const double b = 10. / bw;
    tau = t.wl[(size_t)s * nik + ik] * (b * b);
  } else if (lz > rl) {
    if (ek <= 1000.) {
      const double w2 = t.wl[((size_t)t.ns + s) * nik + ik], w3 = t.wl[((size_t)2 * t.ns + s) * nik + ik];
      tau = w2 * (1 + w3 / w2);
      if (ek <= 1.1)
        tau = tau * 37.5813 * exp(-1.81255 * ek);
      else if (ek > 1.1 && ek <= 5.)
        tau = tau * (7.5 - 1.15 * ek);
    } else if (ek > 1000.) {
      tau = 5. * 3600 * 24 / kp;
    }
  }
  return exp(-t.dts / tau);
}

// LDS copies of the tables of plane (I,J) = (i,j) of species s, for any
// plane layout D with the table pointers of PlaneLds: the plane's scalars
// (PcField), the per-K and per-L tables, with need_loss HDNS, 1/FNHS and the
// exponential's table, and with need_kl the per-L DRIFTE and DRIFTMU terms.
// The loads go in two batches, each issued whole before its LDS stores.
template <class D>
__device__ inline void plane_tables(const Tab &t, int s, int i, int j, const D &d, bool need_loss, bool need_kl) {
  const int NE = t.ne, NPA = t.npa, NT = t.nt, NR = t.nr;
  const int tid = threadIdx.x, T = blockDim.x;
  const size_t ij = (size_t)i * NT + j;
  const int nkl = 8 * NPA;
  const double *se = t.kle + ij * nkl, *sm8 = t.klm + ij * nkl;

  // batch A: per-L terms and scalars
  double vc[2] = {0.0, 0.0}, vm[2] = {0.0, 0.0};
  if (need_kl) {
#pragma unroll
    for (int m = 0; m < 2; ++m)
      if (tid + m * T < nkl) {
        vc[m] = se[tid + m * T];
        vm[m] = sm8[tid + m * T];
      }
  }
  double vp = 0.0;
  if (tid < PC_N) {
    const double *a;
    if (tid < PC_REL) a = t.kij + ij * 8 + tid;
    else if (tid < PC_HN) a = t.rel + s * 8 + (tid - PC_REL);
    else if (tid == PC_HN) a = &GEO3(t.fnhs, i + 1, j + 1, NPA);
    else if (tid == PC_HN1) a = &GEO3(t.fnhs, i + 1, j + 1, NPA - 1);
    else if (tid == PC_MUN) a = t.mu + NPA - 1;
    else if (tid == PC_MUN1) a = t.mu + NPA - 2;
    else a = &t.dts;
    vp = *a;
  }
  if (need_kl) {
#pragma unroll
    for (int m = 0; m < 2; ++m)
      if (tid + m * T < nkl) {
        d.cf[tid + m * T] = vc[m];
        d.cm[tid + m * T] = vm[m];
      }
  }
  if (tid < PC_N) d.pc[tid] = vp;

  // batch B: per-K and per-L tables
  const size_t sk = (size_t)s * NE, sik = ((size_t)s * NR + i) * NE;
  const double *wav = t.wavef + (((size_t)s * NR + i) * NT + j) * NE;
  const bool wave = need_loss && t.wpi[s];
  double vk[KT_N];
  if (tid < NE) {
    vk[KT_EKP] = t.ekp[sk + tid];
    vk[KT_MKP] = t.mkp[sk + tid];
    vk[KT_SIGV] = t.sigv[sk + tid];
    vk[KT_EDOT] = t.edot[sik + tid];
    vk[KT_LATL] = t.latl[sik + tid];
    vk[KT_DE] = t.de[tid];
    vk[KT_RDE] = t.rde[tid];
    vk[KT_WE] = t.we[tid];
    vk[KT_RWE] = t.rwe[tid];
    vk[KT_EKEV] = t.ekev[tid];
    // WAVELO's factor of the step (1 for K=1 and without WAVELO)
    vk[KT_WAVE] = (wave && tid >= 1) ? wav[tid] : 1.0;
  }
  double vl[LT_N];
  double2 hv = make_double2(0.0, 0.0);
  if (tid < NPA) {
    vl[LT_DMU] = t.dmu[tid];
    vl[LT_RDMU] = t.rdmu[tid];
    vl[LT_WMU] = t.wmu[tid];
    vl[LT_RWMU] = t.rwmu[tid];
    if (need_loss) hv = t.kll[ij * NPA + tid];
  }
  const double vx = (need_loss && tid < 64) ? t.ex64[tid] : 0.0;

  if (tid < NE)
#pragma unroll
    for (int f = 0; f < KT_N; ++f) d.kt[f * NE + tid] = vk[f];
  if (tid < NPA) {
#pragma unroll
    for (int f = 0; f < LT_N; ++f) d.lt[f * NPA + tid] = vl[f];
    if (need_loss) {
      d.cl[2 * tid] = hv.x;
      d.cl[2 * tid + 1] = hv.y;
    }
  }
  if (need_loss && tid < 64) d.ex[tid] = vx;

  // what one pass did not cover (small blocks)
  if (need_kl && nkl > 2 * T) {
    block_copy<2>(d.cf + 2 * T, se + 2 * T, nkl - 2 * T);
    block_copy<2>(d.cm + 2 * T, sm8 + 2 * T, nkl - 2 * T);
  }
  for (int k = tid + T; k < NE; k += T) {
    d.kt[KT_EKP * NE + k] = t.ekp[sk + k];
    d.kt[KT_MKP * NE + k] = t.mkp[sk + k];
    d.kt[KT_SIGV * NE + k] = t.sigv[sk + k];
    d.kt[KT_EDOT * NE + k] = t.edot[sik + k];
    d.kt[KT_LATL * NE + k] = t.latl[sik + k];
    d.kt[KT_DE * NE + k] = t.de[k];
    d.kt[KT_RDE * NE + k] = t.rde[k];
    d.kt[KT_WE * NE + k] = t.we[k];
    d.kt[KT_RWE * NE + k] = t.rwe[k];
    d.kt[KT_EKEV * NE + k] = t.ekev[k];
    d.kt[KT_WAVE * NE + k] = wave ? wav[k] : 1.0;
  }
  for (int l = tid + T; l < NPA; l += T) {
    d.lt[LT_DMU * NPA + l] = t.dmu[l];
    d.lt[LT_RDMU * NPA + l] = t.rdmu[l];
    d.lt[LT_WMU * NPA + l] = t.wmu[l];
    d.lt[LT_RWMU * NPA + l] = t.rwmu[l];
    if (need_loss) {
      const double2 h = t.kll[ij * NPA + l];
      d.cl[2 * l] = h.x;
      d.cl[2 * l + 1] = h.y;
    }
  }
}

// Block set-up of the plane adjoint kernels (linear plane layout of
// PlaneLds): context of plane b of species s (I>=2), the LDS copy of the
// plane (from src, device layout of one species) and its tables
// (plane_tables).  The plane's first KS cells per thread are in flight
// while the tables load; the rest follow in a further pass.
template <int KS>
__device__ inline void plane_setup(const Tab &t, const Unit &u, double *sm, const double *src, bool need_loss,
                                   bool need_kl, PlaneCtx &c, PlaneLds &d) {
  const int NE = t.ne, NPA = t.npa, NT = t.nt;
  const int tid = threadIdx.x, T = blockDim.x;
  c.s = u.s;
  c.i = 1 + u.b / NT;  // I = i+1 >= 2
  c.j = u.b - (c.i - 1) * NT;
  c.NE = NE;
  c.NPA = NPA;
  c.NT = NT;
  c.P = NE * NPA;
  c.ij = (size_t)c.i * NT + c.j;
  c.rlz = t.rlz[c.i];
  c.rrlz = t.rrlz[c.i];
  c.qs = (double)t.charge[c.s];  // x/QS == x*QS for QS = +-1
  c.cex = t.cex[c.s];
  c.upa = t.upa[c.i];
  c.beta = t.betalim;
  c.rne = t.rne;
  c.nmine = (c.P - tid + T - 1) / T;  // cells tid, tid+blockDim, .. < P
  c.nmax = (c.P + T - 1) / T;
  c.lestep = T / NE;
  c.le0 = fast_div(tid, NE, c.rne);
  c.ke = tid - c.le0 * NE;
  c.eact = tid < c.lestep * NE;
  d = plane_lds(sm, NE, NPA);
  const int P = c.P;
  const double *gF = src + (size_t)c.s * f2_species_stride(t) + c.ij * P;
  double vf[KS];
#pragma unroll
  for (int m = 0; m < KS; ++m) vf[m] = (tid + m * T < P) ? gF[tid + m * T] : 0.0;
  plane_tables(t, c.s, c.i, c.j, d, need_loss, need_kl);
#pragma unroll
  for (int m = 0; m < KS; ++m)
    if (tid + m * T < P) d.F[tid + m * T] = vf[m];
  if (P > KS * T) block_copy<4>(d.F + KS * T, gF + KS * T, P - KS * T);
  __syncthreads();
}

// CHARGE=EXP(-ALPHA), ALPHA=10**Y*V*HDNS*DTs (src/ModRamLoss.f90:47-49)
template <class D>
__device__ __forceinline__ double cex_factor(const D &d, int NE, int k, int l, double dts) {
  return exp_tab(-(d.kt[KT_SIGV * NE + k] * d.cl[2 * l] * dts), d.ex);
}
// ATLOS(S,I,K)**(1/FNHS(I,J,L)) (src/ModRamLoss.f90:501) as exp(log(ATLOS)*(1/FNHS))
template <class D>
__device__ __forceinline__ double atm_factor(const D &d, int NE, int k, int l) {
  return exp_tab(d.kt[KT_LATL * NE + k] * d.cl[2 * l + 1], d.ex);
}

}  // namespace ramgpu
