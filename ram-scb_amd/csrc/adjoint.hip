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
// DRIFTR adjoint: one thread per radial line (J,K,L), walking I downward.
// Face I (value G_I = C_I*FBND_I) feeds cells I (-G_I) and I+1 (+G_I); the
// adjoint of G_I is -w_I + w_(I+1) with w the output adjoint of each
// unclamped updated cell.  The partials of FBND_I go to stencil cells
// I-1..I+2 (a 4-register window, written back once complete) or to FGEOS
// (ghosts F(NR+1), F(NR+2), and the stale ghost of another line).
// FINAL: the last sweep of ram_run, whose output then went through
// F2(:,:,NT)=F2(:,:,1) and the magnetopause mask; its J=NT line output was
// discarded (zero adjoint), and line J=1 also receives line NT's adjoint.
struct RFace {
  double G;
  double p[4];  // C * dFBND/d(slot)
  int tgt[4];   // slot cell I' (1..NR), 0 none, -1 F(NR+1), -2 F(NR+2), -3 stale ghost
};

template <int FINAL>
__global__ void __launch_bounds__(kLineThreads) k_r_bwd(Launch L, const double *xin, double *adj, double *bbar) {
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int s = u.s;
    const int NR = t.nr, NT = t.nt, NE = t.ne;
    const int P = NE * t.npa;
    const int line = u.b * kLineThreads + threadIdx.x;
    const int j = fast_div(line, P, t.rp);
    if (line >= NT * P || (FINAL && j == NT - 1)) continue;
    const int q = line - j * P;
    const int l = fast_div(q, NE, t.rne), k = q - l * NE;
    const double p4k = t.p4[s * NE + k];
    const double beta = t.betalim;
    const size_t di = (size_t)NT * P;
    const size_t off = (size_t)s * f2_species_stride(t) + (size_t)j * P + q;
    const double *xl = xin + off;
    double *al = adj + off;
    double *as = al + (size_t)(NT - 1 - j) * P;  // line J=NT
    const bool in = c_radial(t, p4k, NR - 1, j, l) < 0;
    double gh1 = 0.0, gh2 = 0.0, dg1 = 0.0, dg2 = 0.0;
    if (in) {
      gh1 = ghost_value(t, s, j, k, l, t.conf1);
      gh2 = ghost_value(t, s, j, k, l, t.conf2);
      dg1 = ghost_dfgeos(t, s, j, l, t.conf1);
      dg2 = ghost_dfgeos(t, s, j, l, t.conf2);
    }
    const long own = s + (long)t.ns * (j + (long)NT * (k + (long)NE * l));
    double gacc1 = 0.0, gacc2 = 0.0;

    auto X = [&](int I) -> double { return xl[(size_t)(I - 1) * di]; };
    auto ybar = [&](int I) -> double {
      double y = al[(size_t)(I - 1) * di];
      if (FINAL) {
        if (t.rij[(size_t)(I - 1) * NT + j].w == 1.0) y = 0.0;
        if (j == 0 && t.rij[(size_t)(I - 1) * NT + NT - 1].w != 1.0) y += as[(size_t)(I - 1) * di];
      }
      return y;
    };
    auto face = [&](int I) -> RFace {
      RFace f;
      const int i = I - 1;
      const double ci = c_radial(t, p4k, i, j, l);
      f.tgt[0] = f.tgt[1] = f.tgt[2] = f.tgt[3] = 0;
      f.p[0] = f.p[1] = f.p[2] = f.p[3] = 0.0;
      if (I == 1) {  // FBND(1) = F(2) on inflow lines, else 0
        f.G = ci * (in ? X(2) : 0.0);
        if (in) {
          f.tgt[0] = 2;
          f.p[0] = ci;
        }
        return f;
      }
      if (I == NR && !in) {  // FBND(NR) = F(NR)
        f.G = ci * X(NR);
        f.tgt[0] = NR;
        f.p[0] = ci;
        return f;
      }
      const double f0 = X(I);
      const double fp1 = (I == NR) ? gh1 : X(I + 1);
      const int tp1 = (I == NR) ? -1 : I + 1;
      double fn, fnm1;
      int tn, tm;
      if (ci < 0) {
        fnm1 = fp1;
        tm = tp1;
        if (I == NR) {
          fn = gh2;
          tn = -2;
        } else if (I + 2 <= NR) {
          fn = X(I + 2);
          tn = I + 2;
        } else if (in) {
          fn = gh1;
          tn = -1;
        } else {
          fn = stale_ghost(t, s, j, k, l);
          tn = -3;
        }
      } else {
        fn = f0;
        tn = I;
        fnm1 = X(I - 1);
        tm = I - 1;
      }
      const FaceD D = face_d(f0, fp1, fn, fnm1, ci, ci, beta);
      f.G = ci * D.fb;
      f.tgt[0] = I;
      f.p[0] = ci * D.d0;
      f.tgt[1] = tp1;
      f.p[1] = ci * D.d1;
      f.tgt[2] = tn;
      f.p[2] = ci * D.dn;
      f.tgt[3] = tm;
      f.p[3] = ci * D.dm;
      return f;
    };
    // window acc[0..3] holds the adjoint of cells I-1..I+2 for face I
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    auto route = [&](int I, int tgt, double v) {
      if (tgt > 0) {
        const int m = tgt - (I - 1);
        if (m == 0) acc[0] += v;
        else if (m == 1) acc[1] += v;
        else if (m == 2) acc[2] += v;
        else acc[3] += v;
      } else if (tgt == -1) {
        gacc1 += v;
      } else if (tgt == -2) {
        gacc2 += v;
      } else if (tgt == -3 && bbar) {
        const long src = stale_ghost_source(t, s, j, k, l);
        if (src >= 0) {  // src = s + NS*(J' + NT*(K' + NE*L'))
          const long r = src / t.ns;
          const int jj = (int)(r % NT), ll = (int)(r / ((long)NT * NE));
          atomicAdd(&bbar[src], v * ghost_dfgeos(t, s, jj, ll, t.conf1));
        }
      }
    };
    RFace fc = face(NR);
    double wnext = 0.0;
    for (int I = NR; I >= 2; --I) {
      const RFace fp = face(I - 1);
      double v = X(I) - fc.G + fp.G;
      const double w = (v < 0) ? 0.0 : ybar(I);
      acc[1] += w;  // cell I's own term
      const double gb = -w + wnext;
      for (int m = 0; m < 4; ++m)
        if (fc.tgt[m] != 0) route(I, fc.tgt[m], fc.p[m] * gb);
      if (I + 2 <= NR) al[(size_t)(I + 1) * di] = acc[3];
      acc[3] = acc[2];
      acc[2] = acc[1];
      acc[1] = acc[0];
      acc[0] = 0.0;
      fc = fp;
      wnext = w;
    }
    // face 1 (window: cells 0..3); cell 1 is not updated: its output adjoint passes
    for (int m = 0; m < 4; ++m)
      if (fc.tgt[m] != 0) route(1, fc.tgt[m], fc.p[m] * wnext);
    acc[1] += ybar(1);
    if (NR >= 3) al[2 * di] = acc[3];
    al[di] = acc[2];
    al[0] = acc[1];
    if (FINAL && j == 0)  // line NT's adjoint has been folded into line 1
      for (int I = 1; I <= NR; ++I) as[(size_t)(I - 1) * di] = 0.0;
    if (bbar && (gacc1 != 0.0 || gacc2 != 0.0)) atomicAdd(&bbar[own], gacc1 * dg1 + gacc2 * dg2);
  }
}

// ---------------------------------------------------------------------------
// DRIFTP adjoint: one thread per ring (I>=2,K,L), walking J downward.  Cell
// J>=2 is updated from faces J and J-1 (face 1 is face NT), and the output
// F2(I,1) is a copy of F2(I,NT), so its adjoint joins cell NT's.  The input
// F(1) only feeds stencils.  Stencil partials go to a 4-register window for
// cells >= 4 and to three accumulators for cells 1..3, which the periodic
// wrap reaches from faces NT-1 and NT.
struct PFace {
  double G;
  double p[4];
  int tgt[4];
};

__global__ void __launch_bounds__(kLineThreads) k_p_bwd(Launch L, const double *xin, double *adj) {
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int s = u.s;
    const int NR = t.nr, NT = t.nt, NE = t.ne;
    const int P = NE * t.npa;
    const int line = u.b * kLineThreads + threadIdx.x;
    if (line >= (NR - 1) * P) continue;
    const int i1 = fast_div(line, P, t.rp);
    const int q = line - i1 * P;
    const int i = i1 + 1;
    const int l = fast_div(q, NE, t.rne), k = q - l * NE;
    const double p2k = t.p2[((size_t)s * NR + i) * NE + k];
    const double beta = t.betalim, omet = t.omet;
    const size_t off = (size_t)s * f2_species_stride(t) + (size_t)i * NT * P + q;
    const double *xl = xin + off;
    double *al = adj + off;
    const double4 *pij = t.pij + (size_t)i * NT;
    const double4 *pl = t.pl + ((size_t)l * NR + i) * NT;
    auto X = [&](int J) -> double { return xl[(size_t)(J - 1) * P]; };
    auto face = [&](int J) -> PFace {
      const double4 a = pij[J - 1];
      const double4 g = pl[J - 1];
      const double uu = div_rn(p2k * g.x, g.y, g.z);
      const double num = a.x - uu - a.y;
      const double cp = div_rn(num, a.z, a.w) + omet;
      int t0 = J, t1, tn, tm;
      if (J == NT) {
        t1 = 2;
        if (cp < 0) {
          tn = 3;
          tm = 2;
        } else {
          tn = NT;
          tm = NT - 1;
        }
      } else {
        t1 = J + 1;
        if (cp < 0) {
          if (J + 2 <= NT) {
            tn = J + 2;
            tm = J + 1;
          } else {
            tn = 2;
            tm = 1;
          }
        } else {
          tn = J;
          tm = J - 1;
        }
      }
      const FaceD D = face_d(X(t0), X(t1), X(tn), X(tm), cp, cp, beta);
      PFace f;
      f.G = cp * D.fb;
      f.tgt[0] = t0;
      f.p[0] = cp * D.d0;
      f.tgt[1] = t1;
      f.p[1] = cp * D.d1;
      f.tgt[2] = tn;
      f.p[2] = cp * D.dn;
      f.tgt[3] = tm;
      f.p[3] = cp * D.dm;
      return f;
    };
    const double yb1 = al[0], ybN = al[(size_t)(NT - 1) * P];
    const PFace fN = face(NT), f2 = face(2);
    double v2 = X(2) - f2.G + fN.G;
    const double w2 = (v2 < 0) ? 0.0 : al[P];
    double wr[4] = {0.0, 0.0, 0.0, 0.0};  // cells 1..3 (index 0 unused)
    double acc[4] = {0.0, 0.0, 0.0, 0.0};  // cells J-1..J+2 (>= 4) for face J
    auto route = [&](int J, int tgt, double v) {
      if (tgt <= 3) {
        if (tgt == 1) wr[1] += v;
        else if (tgt == 2) wr[2] += v;
        else wr[3] += v;
      } else {
        const int m = tgt - (J - 1);
        if (m == 0) acc[0] += v;
        else if (m == 1) acc[1] += v;
        else if (m == 2) acc[2] += v;
        else acc[3] += v;
      }
    };
    PFace fc = fN;
    double wnext = w2;  // face NT feeds cell 2 as face 1
    for (int J = NT; J >= 3; --J) {
      const PFace fp = (J - 1 == 2) ? f2 : face(J - 1);
      double v = X(J) - fc.G + fp.G;
      const double yb = (J == NT) ? ybN + yb1 : al[(size_t)(J - 1) * P];
      const double w = (v < 0) ? 0.0 : yb;
      route(J, J, w);
      const double gb = -w + wnext;
      for (int m = 0; m < 4; ++m) route(J, fc.tgt[m], fc.p[m] * gb);
      if (J + 2 <= NT && J + 2 >= 4) al[(size_t)(J + 1) * P] = acc[3];
      acc[3] = acc[2];
      acc[2] = acc[1];
      acc[1] = acc[0];
      acc[0] = 0.0;
      fc = fp;
      wnext = w;
    }
    {  // J = 2
      route(2, 2, w2);
      const double gb = -w2 + wnext;
      for (int m = 0; m < 4; ++m) route(2, f2.tgt[m], f2.p[m] * gb);
      al[3 * (size_t)P] = acc[3];  // cell 4 (NT >= 4)
    }
    al[0] = wr[1];
    al[P] = wr[2];
    al[2 * (size_t)P] = wr[3];
  }
}

// ---------------------------------------------------------------------------
// Plane adjoints.  LDS: the forward layout (F = the operator's input x, H)
// plus one more plane Y (the output adjoint, then the flux weights w).
// Cells of the plane are owned as in k_plane.  The adjoint of the input is
// assembled in H and written back over adj.

// DRIFTE adjoint (forward: plane_e).  Slots of face K (0-based k): f0 = F(K)
// or the ghost F(1) (from F(2)), fp1 = F(K+1) or 0, fn = F(K+2) or 0 (C<0)
// or f0, fnm1 = fp1 (C<0) or F(K-1) / ghosts F(0), F(1).
__device__ void plane_e_bwd(const Tab &t, const PlaneCtx &c, const PlaneLds &d, double *Y) {
  const int NE = c.NE, NPA = c.NPA;
  {
    const double *src = t.kle + c.ij * NPA * 8;
    for (int n = threadIdx.x; n < NPA * 8; n += blockDim.x) d.cf[n] = src[n];
  }
  __syncthreads();
  const double *k8 = t.kij + c.ij * 8;
  const double drd1 = k8[KIJ_DRD1], dpd1 = k8[KIJ_DPD1], bij = k8[KIJ_BIJ], rbij = k8[KIJ_RBIJ];
  const double *rel = t.rel + c.s * 8;
  const double r0 = rel[0], r1 = rel[1], r2 = rel[2], r3 = rel[3], r4 = rel[4];
  const double dg1 = r0 / r1 * r2, dg0 = dg1 * r3 / r0 * r4;  // dF(1)/dF(2), dF(0)/dF(2)
  const double rlz = c.rlz, rrlz = c.rrlz, qs = c.qs, beta = c.beta;
  // forward face of cell (k,l): Courant number, limited value and slots
  auto eface = [&](int k, int l, double &ce, FaceD &D, int (&tg)[4], double (&gf)[4]) {
    ce = c_energy(d.cf + l * 8, d.kt[KT_EKP * NE + k], d.kt[KT_EDOT * NE + k], drd1, dpd1, bij, rbij, rlz, rrlz,
                  qs);
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
  double O[kMaxCellsPerThread];
#pragma unroll
  for (int n = 0; n < kMaxCellsPerThread; ++n) {
    if (n < c.nmine) {
      CELL_KL(n);
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
  for (int n = 0; n < kMaxCellsPerThread; ++n) {
    if (n < c.nmine) {
      CELL_KL(n);
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
  for (int n = 0; n < kMaxCellsPerThread; ++n)
    if (n < c.nmine) d.H[threadIdx.x + n * kPlaneThreads] = O[n];
  __syncthreads();
#pragma unroll
  for (int n = 0; n < kMaxCellsPerThread; ++n) {
    if (n < c.nmine) {
      CELL_KL(n);
      double ce;
      FaceD D;
      int tg[4];
      double gf[4];
      eface(k, l, ce, D, tg, gf);
      double fbb = (k >= 1) ? -div_rn(ce, d.kt[KT_WE * NE + k], d.kt[KT_RWE * NE + k]) * Y[cc] : 0.0;
      if (k + 1 < NE) fbb += div_rn(ce, d.kt[KT_WE * NE + k + 1], d.kt[KT_RWE * NE + k + 1]) * Y[cc + 1];
      if (fbb != 0.0) {
        const double pd[4] = {D.d0, D.d1, D.dn, D.dm};
        for (int m = 0; m < 4; ++m)
          if (tg[m] >= 0 && pd[m] != 0.0) atomicAdd(&d.H[l * NE + tg[m]], pd[m] * gf[m] * fbb);
      }
    }
  }
  __syncthreads();
}

// DRIFTMU adjoint (forward: plane_mu).  Updated cells L=2..NPA-1 from faces
// L and L-1 (face 1 carries nothing); the output F2(NPA) is F2(NPA-1) times
// FNHS(NPA)*MU(NPA)/FNHS(NPA-1)/MU(NPA-1), so its adjoint joins cell NPA-1's;
// the input F(NPA) only feeds FBND(NPA-1) = F(NPA) and stencils.
__device__ void plane_mu_bwd(const Tab &t, const PlaneCtx &c, const PlaneLds &d, double *Y) {
  const int NE = c.NE, NPA = c.NPA;
  {
    const double *src = t.klm + c.ij * NPA * 8;
    for (int n = threadIdx.x; n < NPA * 8; n += blockDim.x) d.cf[n] = src[n];
  }
  const double hN = GEO3(t.fnhs, c.i + 1, c.j + 1, NPA), hN1 = GEO3(t.fnhs, c.i + 1, c.j + 1, NPA - 1);
  const double ratio = hN * t.mu[NPA - 1] / hN1 / t.mu[NPA - 2];
  __syncthreads();
  const double *k8 = t.kij + c.ij * 8;
  const double drm1 = k8[KIJ_DRM1], dpm1 = k8[KIJ_DPM1], db2 = k8[KIJ_DB2];
  const double bij = k8[KIJ_BIJ], rbij = k8[KIJ_RBIJ];
  const double rlz = c.rlz, rrlz = c.rrlz, qs = c.qs, beta = c.beta;
  // face L = l+1 (2..NPA-1) of column k: Courant number, value, slots (rows)
  auto mface = [&](int k, int l, double &cm, FaceD &D, int (&tg)[4]) {
    cm = c_mu(d.cf + l * 8, d.kt[KT_MKP * NE + k], drm1, dpm1, db2, bij, rbij, rlz, rrlz, qs);
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
  double O[kMaxCellsPerThread];
#pragma unroll
  for (int n = 0; n < kMaxCellsPerThread; ++n) {
    if (n < c.nmine) {
      CELL_KL(n);
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
  for (int n = 0; n < kMaxCellsPerThread; ++n) {
    if (n < c.nmine) {
      CELL_KL(n);
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
  for (int n = 0; n < kMaxCellsPerThread; ++n)
    if (n < c.nmine) d.H[threadIdx.x + n * kPlaneThreads] = O[n];
  __syncthreads();
#pragma unroll
  for (int n = 0; n < kMaxCellsPerThread; ++n) {
    if (n < c.nmine) {
      CELL_KL(n);
      if (l >= 1 && l <= NPA - 2) {
        double cm;
        FaceD D;
        int tg[4];
        mface(k, l, cm, D, tg);
        double fbb = -div_rn(cm, d.lt[LT_WMU * NPA + l], d.lt[LT_RWMU * NPA + l]) * Y[cc];
        if (l + 1 <= NPA - 2) fbb += div_rn(cm, d.lt[LT_WMU * NPA + l + 1], d.lt[LT_RWMU * NPA + l + 1]) * Y[cc + NE];
        if (fbb != 0.0) {
          const double pd[4] = {D.d0, D.d1, D.dn, D.dm};
          for (int m = 0; m < 4; ++m)
            if (tg[m] >= 0 && pd[m] != 0.0) atomicAdd(&d.H[tg[m] * NE + k], pd[m] * fbb);
        }
      }
    }
  }
  __syncthreads();
}

// Loss operators: F2 *= factor on their cell ranges, so the adjoint is
// multiplied by the same factors.  which: 0 CHAREXCHANGE, 1 ATMOL, 2 the
// chain CEX, ATMOL, ATMOL, CEX of ram_run.
__device__ void plane_loss_bwd(const Tab &t, const PlaneCtx &c, const PlaneLds &d, double *Y, int which) {
  const int NE = c.NE;
  const bool cex = t.cex[c.s] != 0;
  const int upa = t.upa[c.i];
#pragma unroll
  for (int n = 0; n < kMaxCellsPerThread; ++n)
    if (n < c.nmine) {
      CELL_KL(n);
      const bool cx = cex && k >= 1 && l >= 1 && which != 1;
      const bool atm = k >= 1 && l + 1 >= upa && which != 0;
      double y = Y[cc];
      const double cxf = cx ? cex_factor(d, NE, k, l, t.dts) : 1.0;
      const double af = atm ? atm_factor(d, NE, k, l) : 1.0;
      if (which == 2) {
        if (cx) y = y * cxf;
        if (atm) y = y * af;
        if (atm) y = y * af;
        if (cx) y = y * cxf;
      } else {
        if (cx) y = y * cxf;
        if (atm) y = y * af;
      }
      d.H[cc] = y;
    }
  __syncthreads();
}

template <int SEQ>
__global__ void __launch_bounds__(kPlaneThreads, 4) k_plane_bwd(Launch L, const double *xin, double *adj) {
  extern __shared__ double sm[];
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    constexpr bool need_loss = SEQ == SEQ_CEX || SEQ == SEQ_ATM || SEQ == SEQ_LOSS;
    PlaneCtx c;
    PlaneLds d;
    plane_setup(t, u, sm, xin, need_loss, c, d);
    double *Y = (double *)(d.cmx + c.NE + c.NPA);
    double *ga = adj + (size_t)c.s * f2_species_stride(t) + c.ij * c.P;
    for (int n = threadIdx.x; n < c.P; n += blockDim.x) Y[n] = ga[n];
    __syncthreads();
    if constexpr (SEQ == SEQ_E) plane_e_bwd(t, c, d, Y);
    else if constexpr (SEQ == SEQ_MU) plane_mu_bwd(t, c, d, Y);
    else if constexpr (SEQ == SEQ_CEX) plane_loss_bwd(t, c, d, Y, 0);
    else if constexpr (SEQ == SEQ_ATM) plane_loss_bwd(t, c, d, Y, 1);
    else plane_loss_bwd(t, c, d, Y, 2);
    for (int n = threadIdx.x; n < c.P; n += blockDim.x) ga[n] = d.H[n];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Flux misfit at virtual-satellite points (loss of the adjoint driver):
//   omega_p = sum_L w[p][L] * F2(S_p,I_p,J_p,K_p,L),
//   loss_p  = (log(omega_p) - target_p)^2,
// one thread per point; with lam != NULL also lam += d loss / d F2
// (= 2*(log(omega)-target)/omega * w, atomically: points may share cells).
// pts[p] = {S,I,J,K} 1-based.
__global__ void k_misfit(const double *f2, int nr, int nt, int ne, int npa, int npts, const int *pts,
                         const double *w, const double *target, double *loss_out, double *lam) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const int s = pts[4 * p] - 1, i = pts[4 * p + 1] - 1, j = pts[4 * p + 2] - 1, k = pts[4 * p + 3] - 1;
  const size_t base = ((((size_t)s * nr + i) * nt + j) * npa) * ne + k;  // + l*NE
  const double *wp = w + (size_t)p * npa;
  double om = 0.0;
  for (int l = 0; l < npa; ++l) om += wp[l] * f2[base + (size_t)l * ne];
  const double r = log(om) - target[p];
  loss_out[p] = r * r;
  if (lam) {
    const double g = 2.0 * r / om;
    for (int l = 0; l < npa; ++l) atomicAdd(&lam[base + (size_t)l * ne], g * wp[l]);
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
template __global__ void k_plane_bwd<SEQ_E>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_MU>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_CEX>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_ATM>(Launch, const double *, double *);
template __global__ void k_plane_bwd<SEQ_LOSS>(Launch, const double *, double *);

}  // namespace ramgpu
