// adjoint.hip -- reverse-mode (vector-Jacobian product) kernels of the split
// step operators.
//
// Every drift Courant number, CHARGE and ATLOS factor depends on fields,
// geometry and grid only, never on F2, so each operator is a piecewise
// linear map of (F2, FGEOS): the limiter branch, the upwind direction and
// the clamp F2<0 -> 1e-15 are fixed by the operator's input.  The adjoint of
// an operator therefore needs that input x (kept on a tape by the caller)
// and maps the adjoint of its output (adj, in place) to the adjoint of x;
// DRIFTR also accumulates the adjoint of FGEOS through its outer ghosts.
// Conventions (subgradients): a clamped cell passes no derivative (a cell
// exactly at 0 is not clamped, as in the forward); limiter kinks follow
// face_d (ties of min/max take the first argument).
//
// Each kernel recomputes, from x, the forward quantities of its operator in
// the forward's arithmetic (the same Courant numbers, face values and clamp
// decisions as kernels.hip), then propagates the adjoint.
#include "sweep_common.h"

#include <hip/hip_runtime.h>

namespace ramgpu {

// ---------------------------------------------------------------------------
// Line adjoints.  Face I of a line carries G_I = C_I*FBND_I; cell I is
// updated as F(I) - G_I + G_(I-1), so the adjoint of G_I is -w_I + w_(I+1),
// with w the output adjoint of each unclamped updated cell.  FBND_I depends
// on the stencil F(I), F(I+1) and, by the upwind side, on F(I+2), F(I+1)
// (C<0) or F(I), F(I-1) (C>=0): fixed offsets from I, so a backward walk
// keeps the adjoint of cells I-1..I+2 in four registers (acc[0..3]) and
// writes cell I+2 back when the window moves on.  The stencil values are a
// register window of the input, one load per cell, issued one step ahead.
struct LFace {
  double G, p0, p1, pn, pm;  // G and C * dFBND / d(F(I), F(I+1), F(N), F(N-1))
  bool neg;                  // C < 0: N = I+2, else N = I
};

__device__ __forceinline__ LFace line_face(double c, double xm1, double x0, double x1, double x2, double beta) {
  const bool neg = c < 0;
  const FaceD D = face_d(x0, x1, neg ? x2 : x0, neg ? x1 : xm1, c, c, beta);
  return {c * D.fb, c * D.d0, c * D.d1, c * D.dn, c * D.dm, neg};
}

// adjoint gb of face f's value into the window of cells I-1..I+2
__device__ __forceinline__ void line_route(const LFace &f, double gb, double (&acc)[4]) {
  const double a = f.pn * gb, b = f.pm * gb;
  acc[1] += f.p0 * gb;
  acc[2] += f.p1 * gb;
  if (f.neg) {
    acc[3] += a;
    acc[2] += b;
  } else {
    acc[1] += a;
    acc[0] += b;
  }
}

// DRIFTR adjoint: one thread per radial line (J,K,L), walking I downward.
// The outer ghosts are cells NR+1 and NR+2 of the window: F(NR+1) =
// FGEOS*CONF1*FNHS and F(NR+2) = FGEOS*CONF2*FNHS on inflow lines; on an
// outflow line F(NR+1) is the stale ghost that face NR-1 reads when its
// Courant number is negative (another line's F(NR+1)), and face NR is
// FBND(NR) = F(NR).  Face 1 is FBND(1) = F(2) on inflow lines, else 0.
// The adjoints of the ghosts go to FGEOS.
// FINAL: the last sweep of ram_run, whose output then went through
// F2(:,:,NT)=F2(:,:,1) and the magnetopause mask; its J=NT line output was
// discarded (zero adjoint), and line J=1 also receives line NT's adjoint.
// The output adjoint of cell I in the final sweep: 0 where the mask set the
// cell, plus line J=NT's adjoint folded into line J=1 (the seam copy).
__device__ __forceinline__ double ybar_of(double a, double b, bool masked, bool fold) {
  double y = masked ? 0.0 : a;
  if (fold) y += b;
  return y;
}

template <int FINAL>
__global__ void __launch_bounds__(kLineThreads, FINAL ? 3 : 4) k_r_bwd(Launch L, const double *xin, double *adj, double *bbar) {
  extern __shared__ double sterm[];  // the block's face terms (radial_lds_bytes)
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int s = u.s;
    const int NR = t.nr, NT = t.nt, NE = t.ne;
    const int P = NE * t.npa;
    const int line0 = u.b * kLineThreads;
    const int line = line0 + threadIdx.x;
    const int j = fast_div(line, P, t.rp);
    const RadialTerms rt = stage_radial_terms(t, line0, sterm);
    if (line < NT * P && !(FINAL && j == NT - 1)) {
      const int q = line - j * P;
      const int l = fast_div(q, NE, t.rne), k = q - l * NE;
      const double p4k = t.p4[s * NE + k];
      const double beta = t.betalim;
      const size_t di = (size_t)NT * P;
      const size_t off = (size_t)s * f2_species_stride(t) + (size_t)j * P + q;
      const gptr<const double> xl = as_global((xin ? xin : launch_src(L, t)) + off);
      const gptr<double> al = as_global((adj ? adj : t.lam) + off);
      const gptr<double> as = al + (size_t)(NT - 1 - j) * P;  // line J=NT
      double *bb = bbar ? bbar : (t.bbar ? t.bbar + (size_t)L.bstep * t.ns * NT * NE * t.npa : nullptr);
      const double *const qrow = rt.q_row(j, l, t.npa), *const rrow = rt.r_row(j), *const rst = rt.rs, *const mgt = rt.mg;
#define RTERM(II) make_double4(rrow[4 * (II)], rrow[4 * (II) + 1], rrow[4 * (II) + 2], rrow[4 * (II) + 3])
#define RSTERM(II) make_double2(rst[2 * (II)], rst[2 * (II) + 1])
#define CFACE(II) c_radial_of(RTERM(II), RSTERM(II), qrow[II], p4k)  // CDriftR at face II+1
#define MASKED(II) (FINAL && rrow[4 * ((II)-1) + 3] == 1.0)
#define FOLD(II) (FINAL && j == 0 && mgt[(II)-1] != 1.0)
      const double cnr = CFACE(NR - 1);
      const bool in = cnr < 0;
      double gh1 = 0.0, gh2 = 0.0;
      if (in) {
        gh1 = ghost_value(t, s, j, k, l, t.conf1);
        gh2 = ghost_value(t, s, j, k, l, t.conf2);
      }
      // face NR, then the window X(NR-2..NR+1) for the first step
      double w0 = xl[(size_t)(NR - 3) * di], w1 = xl[(size_t)(NR - 2) * di], w2 = xl[(size_t)(NR - 1) * di],
             w3 = gh1;
      LFace fc;
      if (in) fc = line_face(cnr, w1, w2, gh1, gh2, beta);
      else fc = {cnr * w2, cnr, 0.0, 0.0, 0.0, false};  // FBND(NR) = F(NR)
      // F(NR+1) read by face NR-1 in the first step: the stale ghost when it
      // is upwinded from outside an outflow line (formed once, here)
      const double x2nr = (!in && CFACE(NR - 2) < 0) ? stale_ghost(t, s, j, k, l) : w3;
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      double wnext = 0.0, g1 = 0.0, g2 = 0.0;
      // step I reads X(max(I-3,1)) (the cell entering the window), the
      // output adjoint of cell I and, for the seam, line J=NT's; they come
      // from two rings of registers loaded four to eight steps ahead
#define BLOAD(IV, XR, YR, SR)                                             \
  {                                                                       \
    const int I_ = (IV);                                                  \
    if (I_ >= 2) {                                                        \
      XR = xl[(size_t)(max(I_ - 3, 1) - 1) * di];                         \
      YR = al[(size_t)(I_ - 1) * di];                                     \
      SR = (FINAL && j == 0) ? as[(size_t)(I_ - 1) * di] : 0.0;           \
    }                                                                     \
  }
#define BSTEP(IV, XN, YA, SB)                                             \
  {                                                                       \
    const int I = (IV);                                                   \
    const double yI = ybar_of(YA, SB, MASKED(I), FOLD(I)), xn = (XN);     \
    LFace fp;                                                             \
    const double c = CFACE(I - 2); /* face I-1 from the window X(I-2..I+1) */ \
    if (I >= 3) {                                                         \
      fp = line_face(c, w0, w1, w2, (I == NR) ? x2nr : w3, beta);         \
    } else {                                                              \
      fp = {c * (in ? w2 : 0.0), 0.0, in ? c : 0.0, 0.0, 0.0, false}; /* FBND(1) */ \
    }                                                                     \
    const double v = w2 - fc.G + fp.G;                                    \
    const double w = (v < 0) ? 0.0 : yI;                                  \
    acc[1] += w; /* cell I's own term */                                  \
    line_route(fc, -w + wnext, acc);                                      \
    if (I + 2 <= NR) al[(size_t)(I + 1) * di] = acc[3];                   \
    else if (I + 2 == NR + 1) g1 = acc[3];                                \
    else g2 = acc[3];                                                     \
    acc[3] = acc[2];                                                      \
    acc[2] = acc[1];                                                      \
    acc[1] = acc[0];                                                      \
    acc[0] = 0.0;                                                         \
    fc = fp;                                                              \
    wnext = w;                                                            \
    w3 = w2;                                                              \
    w2 = w1;                                                              \
    w1 = w0;                                                              \
    w0 = xn;                                                              \
  }
      double xa[4], ya[4], sa[4], xb[4], yb[4], sb[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) BLOAD(NR - m, xa[m], ya[m], sa[m]);
      for (int I0 = NR; I0 >= 2; I0 -= 8) {
#pragma unroll
        for (int m = 0; m < 4; ++m) BLOAD(I0 - 4 - m, xb[m], yb[m], sb[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (I0 - m >= 2) BSTEP(I0 - m, xa[m], ya[m], sa[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m) BLOAD(I0 - 8 - m, xa[m], ya[m], sa[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (I0 - 4 - m >= 2) BSTEP(I0 - 4 - m, xb[m], yb[m], sb[m]);
      }
      // face 1 (window: cells 0..3); cell 1 is not updated: its output adjoint passes
      line_route(fc, wnext, acc);
      acc[1] += ybar_of(al[0], (FINAL && j == 0) ? as[0] : 0.0, MASKED(1), FOLD(1));
      al[2 * di] = acc[3];
      al[di] = acc[2];
      al[0] = acc[1];
#undef BSTEP
#undef BLOAD
#undef FOLD
#undef MASKED
#undef CFACE
#undef RSTERM
#undef RTERM
      if (FINAL && j == 0)  // line NT's adjoint has been folded into line 1
        for (int I = 1; I <= NR; ++I) as[(size_t)(I - 1) * di] = 0.0;
      if (bb) {
        if (in) {
          if (g1 != 0.0 || g2 != 0.0) {
            const long own = s + (long)t.ns * (j + (long)NT * (k + (long)NE * l));
            atomicAdd(&bb[own], g1 * ghost_dfgeos(t, s, j, l, t.conf1) + g2 * ghost_dfgeos(t, s, j, l, t.conf2));
          }
        } else if (g1 != 0.0) {  // the stale ghost: F(NR+1) of an earlier inflow line
          const long src = stale_ghost_source(t, s, j, k, l);
          if (src >= 0) {  // src = s + NS*(J' + NT*(K' + NE*L'))
            const long r = src / t.ns;
            const int jj = (int)(r % NT), ll = (int)(r / ((long)NT * NE));
            atomicAdd(&bb[src], g1 * ghost_dfgeos(t, s, jj, ll, t.conf1));
          }
        }
      }
    }
    __syncthreads();  // the block's face terms are read
  }
}

// DRIFTP adjoint: one thread per ring (I>=2,K,L), walking J downward.  The
// ring is periodic through cells NT+1 = 2 and NT+2 = 3 (face NT reads
// F(2), F(3) as its F(NT+1), F(NT+2)); face NT-1 with C<0 reads F(2), F(1)
// as its F(N), F(N-1) (the reference wraps N, not N-1).  Cell J>=2 is
// updated from faces J and J-1, face 1 being face NT; the output F2(I,1) is
// a copy of F2(I,NT), so its adjoint joins cell NT's, and the input F(1)
// only feeds stencils.  The adjoints of cells 1..3, which the wrapped faces
// reach first, collect in wr[] and join the window's at the end.
__global__ void __launch_bounds__(kLineThreads) k_p_bwd(Launch L, const double *xin, double *adj) {
  extern __shared__ double sterm[];  // the block's face terms (azim_lds_bytes)
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int s = u.s;
    const int NR = t.nr, NT = t.nt, NE = t.ne;
    const int P = NE * t.npa;
    const int line0 = u.b * kLineThreads;
    const int line = line0 + threadIdx.x;
    const int nring = (NR - 1) * P;
    if (line0 >= nring) continue;  // (uniform over the block)
    const AzimTerms at = stage_azim_terms(t, line0, min(line0 + kLineThreads, nring) - 1, sterm);
    if (line < nring) {
      const int i1 = fast_div(line, P, t.rp);
      const int q = line - i1 * P;
      const int i = i1 + 1;
      const int l = fast_div(q, NE, t.rne), k = q - l * NE;
      const double p2k = t.p2[((size_t)s * NR + i) * NE + k];
      const double beta = t.betalim, omet = t.omet;
      const size_t off = (size_t)s * f2_species_stride(t) + (size_t)i * NT * P + q;
      const gptr<const double> xl = as_global((xin ? xin : launch_src(L, t)) + off);
      const gptr<double> al = as_global((adj ? adj : t.lam) + off);
      const double *const ap = at.pij_row(i1), *const gp = at.pl_row(i1, l, t.npa);
#define CFACE(JJ) (div_rn(ap[4 * ((JJ)-1)] - div_rn(p2k * gp[4 * ((JJ)-1)], gp[4 * ((JJ)-1) + 1], gp[4 * ((JJ)-1) + 2]) - \
                          ap[4 * ((JJ)-1) + 1],                                                                  \
                          ap[4 * ((JJ)-1) + 2], ap[4 * ((JJ)-1) + 3]) +                                          \
                   omet)
#define XJ(JJ) xl[(size_t)((JJ)-1) * P]
      const double x1 = XJ(1), x2 = XJ(2), x3 = XJ(3), x4 = XJ(4);
      const double yb1 = al[0], yb2 = al[P];
      // face NT (stencil NT-1, NT, 2, 3) and face 2
      double w0 = XJ(NT - 2), w1 = XJ(NT - 1), w2 = XJ(NT);
      const LFace fN = line_face(CFACE(NT), w1, w2, x2, x3, beta);
      const LFace f2 = line_face(CFACE(2), x1, x2, x3, x4, beta);
      const double v2 = x2 - f2.G + fN.G;
      const double w2c = (v2 < 0) ? 0.0 : yb2;
      double wr[4] = {0.0, 0.0, 0.0, 0.0};  // cells 1..3 (index 0 unused)
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      LFace fc = fN;
      double wnext = w2c;  // face NT feeds cell 2 as face 1
      double w3 = x2;      // window X(J-2..J+1), X(NT+1) = X(2)
      // step J reads X(max(J-3,1)) (the cell entering the window) and the
      // output adjoint of cell J (plus cell 1's at J = NT: F2(I,1) =
      // F2(I,NT)), from two rings of registers loaded four to eight steps
      // ahead
#define PLOAD(JV, XR, YR)                        \
  {                                              \
    const int J_ = (JV);                         \
    if (J_ >= 3) {                               \
      XR = xl[(size_t)(max(J_ - 3, 1) - 1) * P]; \
      YR = al[(size_t)(J_ - 1) * P];             \
    }                                            \
  }
#define PSTEPB(JV, XN, YR)                                                          \
  {                                                                                 \
    const int J = (JV);                                                             \
    const double yJ = (J == NT) ? (YR) + yb1 : (YR), xn = (XN);                     \
    LFace fp;                                                                       \
    if (J - 1 == 2) {                                                               \
      fp = f2;                                                                      \
    } else {                                                                        \
      const double c = CFACE(J - 1);                                                \
      /* face NT-1 with C<0 reads F(N-1) = F(1), not F(NT) */                       \
      if (J == NT && c < 0) {                                                       \
        const FaceD D = face_d(w1, w2, x2, x1, c, c, beta);                         \
        fp = {c * D.fb, c * D.d0, c * D.d1, c * D.dn, c * D.dm, true};             \
      } else {                                                                      \
        fp = line_face(c, w0, w1, w2, w3, beta);                                    \
      }                                                                             \
    }                                                                               \
    const double v = w2 - fc.G + fp.G;                                              \
    const double w = (v < 0) ? 0.0 : yJ;                                            \
    acc[1] += w;                                                                    \
    const double gb = -w + wnext;                                                   \
    if (J == NT - 1 && fc.neg) { /* face NT-1, C<0: F(N) -> cell 2 (as NT+1), F(N-1) -> cell 1 */ \
      acc[1] += fc.p0 * gb;                                                         \
      acc[2] += fc.p1 * gb;                                                         \
      acc[3] += fc.pn * gb;                                                         \
      wr[1] += fc.pm * gb;                                                          \
    } else {                                                                        \
      line_route(fc, gb, acc);                                                      \
    }                                                                               \
    if (J + 2 <= NT) al[(size_t)(J + 1) * P] = acc[3];                              \
    else if (J + 2 == NT + 1) wr[2] += acc[3];                                      \
    else wr[3] += acc[3];                                                           \
    acc[3] = acc[2];                                                                \
    acc[2] = acc[1];                                                                \
    acc[1] = acc[0];                                                                \
    acc[0] = 0.0;                                                                   \
    fc = fp;                                                                        \
    wnext = w;                                                                      \
    w3 = w2;                                                                        \
    w2 = w1;                                                                        \
    w1 = w0;                                                                        \
    w0 = xn;                                                                        \
  }
      double xa[4], ya[4], xb[4], yb[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) PLOAD(NT - m, xa[m], ya[m]);
      for (int J0 = NT; J0 >= 3; J0 -= 8) {
#pragma unroll
        for (int m = 0; m < 4; ++m) PLOAD(J0 - 4 - m, xb[m], yb[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (J0 - m >= 3) PSTEPB(J0 - m, xa[m], ya[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m) PLOAD(J0 - 8 - m, xa[m], ya[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (J0 - 4 - m >= 3) PSTEPB(J0 - 4 - m, xb[m], yb[m]);
      }
#undef PSTEPB
#undef PLOAD
#undef XJ
#undef CFACE
      {  // J = 2
        acc[1] += w2c;
        line_route(f2, -w2c + wnext, acc);
        al[3 * (size_t)P] = acc[3];  // cell 4 (NT >= 4)
      }
      al[0] = wr[1] + acc[0];
      al[P] = wr[2] + acc[1];
      al[2 * (size_t)P] = wr[3] + acc[2];
    }
    __syncthreads();  // the block's face terms are read
  }
}

// ---------------------------------------------------------------------------
// Plane adjoints.  LDS: the linear plane layout of PlaneLds (F = the
// operator's input x, H) plus one more plane Y (the output adjoint, then the
// flux weights w).  Thread n of a BT-thread block owns cells n, n+BT, ..
// (CELL_KL_T).  The adjoint of the input is assembled in H and written back
// over adj.

// The partial scatter of the plane adjoints.  Each face sends partials to
// the cells at offsets -1..+2 along its sweep.  The faces go in four passes
// by their index mod 4 along the sweep, with a barrier after each: two faces
// of a pass are four or more cells apart, so their targets are disjoint and
// each adds into H without atomics, and every cell takes its partials in the
// order of the passes (the same on every run and for every launch shape).

// DRIFTE adjoint (forward: plane_e).  Slots of face K (0-based k): f0 = F(K)
// or the ghost F(1) (from F(2)), fp1 = F(K+1) or 0, fn = F(K+2) or 0 (C<0)
// or f0, fnm1 = fp1 (C<0) or F(K-1) / ghosts F(0), F(1).
// TAPED: the forward's clamp decisions of the plane's cells are in cb (one
// byte per cell, [L][K], written by k_plane<SEQ_TAPE>), so the faces are
// formed once, with their derivatives; otherwise they are formed first for
// the decisions.
template <int BT, bool TAPED = false>
__device__ void plane_e_bwd(const Tab &t, const PlaneCtx &c, const PlaneLds &d, double *Y,
                            const unsigned char *cb = nullptr) {
  const int NE = c.NE, NPA = c.NPA;
  __syncthreads();
  const double *k8 = d.pc + PC_KIJ;  // the per-L terms are in d.cf (plane_setup)
  const double drd1 = k8[KIJ_DRD1], dpd1 = k8[KIJ_DPD1], bij = k8[KIJ_BIJ], rbij = k8[KIJ_RBIJ];
  const double *rel = d.pc + PC_REL;
  const double r0 = rel[0], r1 = rel[1], r2 = rel[2], r3 = rel[3], r4 = rel[4];
  const double dg1 = r0 / r1 * r2, dg0 = dg1 * r3 / r0 * r4;  // dF(1)/dF(2), dF(0)/dF(2)
  const double rlz = c.rlz, rrlz = c.rrlz, qs = c.qs, beta = c.beta;
  // forward face of cell (k,l): Courant number, limited value and slots
  auto eface = [&](int k, int l, double &ce, FaceD &D, int (&tg)[4], double (&gf)[4]) {
    ce = c_energy(d.cf + l * 8, d.kt[KT_EKP * NE + k] * qs, d.kt[KT_EDOT * NE + k], drd1, dpd1, bij, rbij, rlz,
                  rrlz);
    const double *row = d.F + l * NE;
    const double g1 = row[1] * r0 / r1 * r2;
    const double g0 = g1 * r3 / r0 * r4;
    // slot targets are cells of row l (-1: none); gf: factor into that cell
    const double f0 = (k == 0) ? g1 : row[k];
    tg[0] = (k == 0) ? 1 : k;
    gf[0] = (k == 0) ? dg1 : 1.0;
    const double fp1 = (k + 1 < NE) ? row[k + 1] : 0.0;
    tg[1] = (k + 1 < NE) ? k + 1 : -1;
    gf[1] = 1.0;
    double fn, fnm1;
    if (ce < 0) {
      fn = (k + 2 < NE) ? row[k + 2] : 0.0;
      tg[2] = (k + 2 < NE) ? k + 2 : -1;
      gf[2] = 1.0;
      fnm1 = fp1;
      tg[3] = tg[1];
      gf[3] = 1.0;
    } else {
      fn = f0;
      tg[2] = tg[0];
      gf[2] = gf[0];
      if (k == 0) {
        fnm1 = g0;
        tg[3] = 1;
        gf[3] = dg0;
      } else if (k == 1) {
        fnm1 = g1;
        tg[3] = 1;
        gf[3] = dg1;
      } else {
        fnm1 = row[k - 1];
        tg[3] = k - 1;
        gf[3] = 1.0;
      }
    }
    D = face_d(f0, fp1, fn, fnm1, ce, div_rn(ce, d.kt[KT_DE * NE + k], d.kt[KT_RDE * NE + k]), beta);
  };
  if constexpr (TAPED) {
    // flux weights w (unclamped updated cells) into Y, pass-through terms into H
#pragma unroll
    for (int n = 0; n < 8; ++n)
      if (n < c.nmine) {
        CELL_KL_T(n, BT);
        const double yb = Y[cc];
        if (k >= 1) {
          const double w = cb[cc] ? 0.0 : yb;
          Y[cc] = w;
          d.H[cc] = w;
        } else {
          Y[cc] = 0.0;
          d.H[cc] = yb;  // cell 1 is not updated
        }
      }
    __syncthreads();
  } else {
  double O[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    if (n < c.nmine) {
      CELL_KL_T(n, BT);
      double ce;
      FaceD D;
      int tg[4];
      double gf[4];
      eface(k, l, ce, D, tg, gf);
      O[n] = d.F[cc] - div_rn(ce, d.kt[KT_WE * NE + k], d.kt[KT_RWE * NE + k]) * D.fb;
      d.H[cc] = (k + 1 < NE) ? div_rn(ce, d.kt[KT_WE * NE + k + 1], d.kt[KT_RWE * NE + k + 1]) * D.fb : 0.0;
    }
  }
  __syncthreads();
  // flux weights w (unclamped updated cells) into Y; pass-through terms in O
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    if (n < c.nmine) {
      CELL_KL_T(n, BT);
      const double yb = Y[cc];
      if (k >= 1) {
        const double v = O[n] + d.H[cc - 1];
        const double w = (v < 0) ? 0.0 : yb;
        Y[cc] = w;
        O[n] = w;
      } else {
        Y[cc] = 0.0;
        O[n] = yb;  // cell 1 is not updated
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < 8; ++n)
    if (n < c.nmine) d.H[threadIdx.x + n * BT] = O[n];
  __syncthreads();
  }
  // face K's partials go to cells K-1..K+2 of its row (the ghosts' to cell
  // 2).  The faces go in four passes by K mod 4, each ending in a barrier:
  // the targets of two faces of a pass lie in disjoint spans of four cells,
  // so no two faces add to one cell at once, and every cell takes its
  // partials in the order of the passes (no atomics; the same bits for any
  // launch shape)
#pragma unroll 1
  for (int r = 0; r < 4; ++r) {
    const int nk = (NE - r + 3) / 4;  // K = r, r+4, .. < NE (0-based)
    const int cnt = NPA * nk;
    const float rnk = 1.0f / (float)max(nk, 1);
#pragma unroll 1
    for (int j = threadIdx.x; j < cnt; j += BT) {
      const int l = fast_div(j, nk, rnk), k = r + 4 * (j - l * nk);
      const int cc = l * NE + k;
      double ce;
      FaceD D;
      int tg[4];
      double gf[4];
      eface(k, l, ce, D, tg, gf);
      double fbb = (k >= 1) ? -div_rn(ce, d.kt[KT_WE * NE + k], d.kt[KT_RWE * NE + k]) * Y[cc] : 0.0;
      if (k + 1 < NE) fbb += div_rn(ce, d.kt[KT_WE * NE + k + 1], d.kt[KT_RWE * NE + k + 1]) * Y[cc + 1];
      const double pd[4] = {D.d0, D.d1, D.dn, D.dm};
#pragma unroll
      for (int m = 0; m < 4; ++m)
        if (tg[m] >= 0) d.H[l * NE + tg[m]] += pd[m] * gf[m] * fbb;
    }
    __syncthreads();
  }
}

// DRIFTMU adjoint (forward: plane_mu).  Updated cells L=2..NPA-1 from faces
// L and L-1 (face 1 carries nothing); the output F2(NPA) is F2(NPA-1) times
// FNHS(NPA)*MU(NPA)/FNHS(NPA-1)/MU(NPA-1), so its adjoint joins cell NPA-1's;
// the input F(NPA) only feeds FBND(NPA-1) = F(NPA) and stencils.
// TAPED: as plane_e_bwd's.
template <int BT, bool TAPED = false>
__device__ void plane_mu_bwd(const Tab &t, const PlaneCtx &c, const PlaneLds &d, double *Y,
                             const unsigned char *cb = nullptr) {
  const int NE = c.NE, NPA = c.NPA;
  const double hN = d.pc[PC_HN], hN1 = d.pc[PC_HN1];
  const double ratio = hN * d.pc[PC_MUN] / hN1 / d.pc[PC_MUN1];
  __syncthreads();
  const double *k8 = d.pc + PC_KIJ;  // the per-L terms are in d.cm (plane_setup)
  const double drm1 = k8[KIJ_DRM1], dpm1 = k8[KIJ_DPM1], db2 = k8[KIJ_DB2];
  const double bij = k8[KIJ_BIJ], rbij = k8[KIJ_RBIJ];
  const double rlz = c.rlz, rrlz = c.rrlz, qs = c.qs, beta = c.beta;
  // face L = l+1 (2..NPA-1) of column k: Courant number, value, slots (rows)
  auto mface = [&](int k, int l, double &cm, FaceD &D, int (&tg)[4]) {
    cm = c_mu(d.cm + l * 8, d.kt[KT_MKP * NE + k] * qs, drm1, dpm1, db2, bij, rbij, rlz, rrlz);
    const int L1 = l + 1, cc = l * NE + k;
    if (L1 == NPA - 1) {  // FBND(NPA-1) = F(NPA)
      D.fb = d.F[cc + NE];
      D.d0 = 1.0;
      D.d1 = D.dn = D.dm = 0.0;
      tg[0] = l + 1;
      tg[1] = tg[2] = tg[3] = -1;
      return;
    }
    const double f0 = d.F[cc], fp1 = d.F[cc + NE];
    double fn, fnm1;
    tg[0] = l;
    tg[1] = l + 1;
    if (cm < 0) {
      fn = d.F[cc + 2 * NE];
      tg[2] = l + 2;
      fnm1 = fp1;
      tg[3] = l + 1;
    } else {
      fn = f0;
      tg[2] = l;
      fnm1 = (L1 == 2) ? f0 : d.F[cc - NE];  // F(1) = F(2)
      tg[3] = (L1 == 2) ? l : l - 1;
    }
    D = face_d(f0, fp1, fn, fnm1, cm, div_rn(cm, d.lt[LT_DMU * NPA + l], d.lt[LT_RDMU * NPA + l]), beta);
  };
  if constexpr (TAPED) {
    // the top cell's output adjoint joins row NPA-1's, before any row's Y changes
#pragma unroll
    for (int n = 0; n < 8; ++n)
      if (n < c.nmine) {
        CELL_KL_T(n, BT);
        if (l == NPA - 2) Y[cc] = Y[cc] + ratio * Y[cc + NE];
      }
    __syncthreads();
#pragma unroll
    for (int n = 0; n < 8; ++n)
      if (n < c.nmine) {
        CELL_KL_T(n, BT);
        const double yb = Y[cc];
        if (l >= 1 && l <= NPA - 2) {
          const double w = cb[cc] ? 0.0 : yb;
          Y[cc] = w;
          d.H[cc] = w;
        } else {
          Y[cc] = 0.0;
          d.H[cc] = (l == 0) ? yb : 0.0;  // L=1 passes; the input F(NPA) is overwritten
        }
      }
    __syncthreads();
  } else {
  double O[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    if (n < c.nmine) {
      CELL_KL_T(n, BT);
      double on = 0.0, h = 0.0;
      if (l >= 1 && l <= NPA - 2) {
        double cm;
        FaceD D;
        int tg[4];
        mface(k, l, cm, D, tg);
        on = div_rn(cm, d.lt[LT_WMU * NPA + l], d.lt[LT_RWMU * NPA + l]) * D.fb;
        h = div_rn(cm, d.lt[LT_WMU * NPA + l + 1], d.lt[LT_RWMU * NPA + l + 1]) * D.fb;
      }
      O[n] = d.F[cc] - on;
      d.H[cc] = h;
      if (l == NPA - 2) Y[cc] = Y[cc] + ratio * Y[cc + NE];  // the top cell's output adjoint
    }
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    if (n < c.nmine) {
      CELL_KL_T(n, BT);
      const double yb = Y[cc];
      if (l >= 1 && l <= NPA - 2) {
        const double v = O[n] + d.H[cc - NE];
        const double w = (v < 0) ? 0.0 : yb;
        Y[cc] = w;
        O[n] = w;
      } else {
        Y[cc] = 0.0;
        O[n] = (l == 0) ? yb : 0.0;  // L=1 passes; the input F(NPA) is overwritten
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < 8; ++n)
    if (n < c.nmine) d.H[threadIdx.x + n * BT] = O[n];
  __syncthreads();
  }
  // face L's partials go to rows L-1..L+2 of its column, in four passes by
  // L mod 4 as DRIFTE's
#pragma unroll 1
  for (int r = 0; r < 4; ++r) {
    const int nl = (NPA - r + 3) / 4;  // rows r, r+4, .. < NPA (0-based)
    const int cnt = nl * NE;
#pragma unroll 1
    for (int j = threadIdx.x; j < cnt; j += BT) {
      const int ll = fast_div(j, NE, c.rne), k = j - ll * NE, l = r + 4 * ll;
      if (l < 1 || l > NPA - 2) continue;
      const int cc = l * NE + k;
      double cm;
      FaceD D;
      int tg[4];
      mface(k, l, cm, D, tg);
      double fbb = -div_rn(cm, d.lt[LT_WMU * NPA + l], d.lt[LT_RWMU * NPA + l]) * Y[cc];
      if (l + 1 <= NPA - 2) fbb += div_rn(cm, d.lt[LT_WMU * NPA + l + 1], d.lt[LT_RWMU * NPA + l + 1]) * Y[cc + NE];
      const double pd[4] = {D.d0, D.d1, D.dn, D.dm};
#pragma unroll
      for (int m = 0; m < 4; ++m)
        if (tg[m] >= 0) d.H[tg[m] * NE + k] += pd[m] * fbb;
    }
    __syncthreads();
  }
}

// Loss operators: F2 *= factor on their cell ranges, so the adjoint is
// multiplied by the same factors.  which: 0 CHAREXCHANGE, 1 ATMOL, 2 the
// chain CEX, ATMOL, ATMOL, CEX of ram_run.
template <int BT>
__device__ void plane_loss_bwd(const Tab &t, const PlaneCtx &c, const PlaneLds &d, double *Y, int which) {
  const int NE = c.NE;
  const bool cex = c.cex != 0;
  const int upa = c.upa;
#pragma unroll
  for (int n = 0; n < 8; ++n)
    if (n < c.nmine) {
      CELL_KL_T(n, BT);
      const bool cx = cex && k >= 1 && l >= 1 && which != 1;
      const bool atm = k >= 1 && l + 1 >= upa && which != 0;
      double y = Y[cc];
      const double cxf = cx ? cex_factor(d, NE, k, l, d.pc[PC_DTS]) : 1.0;
      const double af = atm ? atm_factor(d, NE, k, l) : 1.0;
      // WAVELO's factor (1 for K=1 and species without it), on L>=2
      const double wf = (l >= 1) ? d.kt[KT_WAVE * NE + k] : 1.0;
      if (which == 2) {  // the chain WAVELO, CEX, ATMOL, ATMOL, CEX, WAVELO, reversed
        y = y * wf;
        if (cx) y = y * cxf;
        if (atm) y = y * af;
        if (atm) y = y * af;
        if (cx) y = y * cxf;
        y = y * wf;
      } else {
        if (cx) y = y * cxf;
        if (atm) y = y * af;
      }
      d.H[cc] = y;
    }
  __syncthreads();
}

// SEQ_STEP: the adjoint of the whole middle of ram_run (DRIFTE, DRIFTMU, the
// loss chain, DRIFTMU, DRIFTE) in one pass over the plane.  xin is tape slot
// 0; slots 0..4 hold the inputs of the five phases (as k_plane<SEQ_TAPE>
// writes them).  The adjoint stays in LDS between the phases: each phase
// leaves the adjoint of its input in H, which becomes the next phase's Y.
template <int SEQ, int BT>
__global__ void __launch_bounds__(BT, 4) k_plane_bwd(Launch L, const double *xin,
                                                                                 double *adj) {
  extern __shared__ double sm[];
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    constexpr bool need_loss = SEQ == SEQ_CEX || SEQ == SEQ_ATM || SEQ == SEQ_LOSS || SEQ == SEQ_STEP;
    const size_t po = (size_t)u.s * f2_species_stride(t), ns = f2_elems(t);
    const double *x0 = xin ? xin : launch_src(L, t);
    PlaneCtx c;
    PlaneLds d;
    constexpr bool need_kl = SEQ == SEQ_E || SEQ == SEQ_MU || SEQ == SEQ_STEP;
    plane_setup<4>(t, u, sm, (SEQ == SEQ_STEP) ? x0 + 4 * ns : x0, need_loss, need_kl, c, d);
    double *Y = (double *)d.cmx;  // in place of the CFL slots (plane_bwd_lds_bytes)
    double *ga = (adj ? adj : t.lam) + po + c.ij * c.P;
    block_copy<4>(Y, ga, c.P);
    __syncthreads();
    if constexpr (SEQ == SEQ_E) plane_e_bwd<BT>(t, c, d, Y);
    else if constexpr (SEQ == SEQ_MU) plane_mu_bwd<BT>(t, c, d, Y);
    else if constexpr (SEQ == SEQ_CEX) plane_loss_bwd<BT>(t, c, d, Y, 0);
    else if constexpr (SEQ == SEQ_ATM) plane_loss_bwd<BT>(t, c, d, Y, 1);
    else if constexpr (SEQ == SEQ_LOSS) plane_loss_bwd<BT>(t, c, d, Y, 2);
    else {
      // phases alternate the two adjoint planes: A holds the output adjoint
      // of a phase, B receives the adjoint of its input
      const double *x = x0 + po + c.ij * c.P;
      // the clamp decisions of E1, MU1, MU2, E2: byte planes of the slot 7
      // after slot 0 (k_plane<SEQ_TAPE>)
      const unsigned char *cb = reinterpret_cast<const unsigned char *>(x0 + 7 * ns) + po + c.ij * c.P;
      double *const A = Y, *const B = d.H;
      PlaneLds d2 = d;
      d2.H = A;
      plane_e_bwd<BT, true>(t, c, d, A, cb + 3 * ns);  // E2: A -> B
      block_copy<4>(d.F, x + 3 * ns, c.P);
      __syncthreads();
      plane_mu_bwd<BT, true>(t, c, d2, B, cb + 2 * ns);  // MU2: B -> A
      plane_loss_bwd<BT>(t, c, d, A, 2);  // loss chain: A -> B (needs only its factors)
      block_copy<4>(d.F, x + ns, c.P);
      __syncthreads();
      plane_mu_bwd<BT, true>(t, c, d2, B, cb + ns);  // MU1: B -> A
      block_copy<4>(d.F, x, c.P);
      __syncthreads();
      plane_e_bwd<BT, true>(t, c, d, A, cb);  // E1: A -> B
    }
    block_copy<4>(ga, d.H, c.P);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Flux misfit at virtual-satellite points (loss of the adjoint driver):
//   omega_p = sum_L w[p][L] * F2(S_p,I_p,J_p,K_p,L),
//   loss_p  = (log(omega_p) - target_p)^2,
// one 64-thread block per point; with lam != NULL also lam += d loss / d F2
// (= 2*(log(omega)-target)/omega * w, atomically: points may share cells).
// pts[p] = {S,I,J,K} 1-based.  The products w*F2 are formed by all lanes at
// once and summed by lane 0 in the order of L (so omega is the sequential
// sum, whatever the launch shape).
constexpr int kMisfitLds = 1025;  // NPA <= 1024 products and the adjoint factor
__device__ inline void misfit_point(int p, const double *f2, int nr, int nt, int ne, int npa, const int *pts,
                                    const double *w, const double *target, double *loss_out, double *lam,
                                    double *sh) {
  const int s = pts[4 * p] - 1, i = pts[4 * p + 1] - 1, j = pts[4 * p + 2] - 1, k = pts[4 * p + 3] - 1;
  const size_t base = ((((size_t)s * nr + i) * nt + j) * npa) * ne + k;  // + l*NE
  const double *wp = w + (size_t)p * npa;
  for (int l = threadIdx.x; l < npa; l += blockDim.x) sh[l] = wp[l] * f2[base + (size_t)l * ne];
  __syncthreads();
  if (threadIdx.x == 0) {
    double om = 0.0;
    for (int l = 0; l < npa; ++l) om += sh[l];
    const double r = log(om) - target[p];
    if (loss_out) loss_out[p] = r * r;
    sh[npa] = 2.0 * r / om;
  }
  __syncthreads();
  if (lam) {
    const double g = sh[npa];
    for (int l = threadIdx.x; l < npa; l += blockDim.x) atomicAdd(&lam[base + (size_t)l * ne], g * wp[l]);
  }
  __syncthreads();  // sh is reused by the block's next point
}

__global__ void k_misfit(const double *f2, int nr, int nt, int ne, int npa, int npts, const int *pts,
                         const double *w, const double *target, double *loss_out, double *lam) {
  __shared__ double sh[kMisfitLds];
  for (int p = blockIdx.x; p < npts; p += gridDim.x)
    misfit_point(p, f2, nr, nt, ne, npa, pts, w, target, loss_out, lam, sh);
}

// The same for every problem of an ensemble launch (one unit per problem
// along x, points along y): the state after step `step` is the launch's
// source (a tape slot or F2); losses go to Tab::mloss[step], and with
// adjoint != 0 dL/dF2 to Tab::lam.
__global__ void k_misfit_ens(Launch L, int step, int adjoint) {
  __shared__ double sh[kMisfitLds];
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const size_t np = (size_t)t.mnpts;
    for (int p = blockIdx.y; p < t.mnpts; p += gridDim.y)
      misfit_point(p, launch_src(L, t), t.nr, t.nt, t.ne, t.npa, t.mpts, t.mw, t.mtarget + (size_t)step * np,
                   adjoint ? nullptr : t.mloss + (size_t)step * np, adjoint ? t.lam : nullptr, sh);
  }
}

// y += a*x (n doubles)
__global__ void k_axpy(double *y, const double *x, double a, size_t n) {
  for (size_t m = blockIdx.x * (size_t)blockDim.x + threadIdx.x; m < n; m += (size_t)gridDim.x * blockDim.x)
    y[m] += a * x[m];
}

// fill (the device-side clear used where a graph may capture it)
__global__ void k_fill(double *y, double v, size_t n) {
  for (size_t m = blockIdx.x * (size_t)blockDim.x + threadIdx.x; m < n; m += (size_t)gridDim.x * blockDim.x)
    y[m] = v;
}

template __global__ void k_r_bwd<0>(Launch, const double *, double *, double *);
template __global__ void k_r_bwd<1>(Launch, const double *, double *, double *);

template __global__ void k_plane_bwd<SEQ_E, 512>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_E, 1024>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_MU, 512>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_MU, 1024>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_CEX, 512>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_CEX, 1024>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_ATM, 512>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_ATM, 1024>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_LOSS, 512>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_LOSS, 1024>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_STEP, 512>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_STEP, 1024>(Launch, const double *, double *);

}  // namespace ramgpu
