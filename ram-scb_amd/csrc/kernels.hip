// kernels.hip -- gfx950 kernels of the RAM kinetic split step.
//
// Reference: src/ModRamDrift.f90 (DRIFTPARA/DRIFTR/DRIFTP/DRIFTE/DRIFTMU),
// src/ModRamLoss.f90 (CEPARA/CHAREXCHANGE/ATMOL), src/ModRamRun.f90
// (ram_run sequence, SUMRC).  Index names in comments are the reference's
// 1-based ones; code indices are 0-based.
//
// Kernel map
//   k_species   per-species 1-D tables of DRIFTPARA/CEPARA, grid reciprocals
//   k_coef      per-(I,J,L) intermediate terms of the four Courant numbers
//   k_coef_ij   per-(I,J) / per-I terms (the ones VT enters)
//   k_ghost     DRIFTR outer-ghost carry between lines (see k_ghost)
//   k_r         DRIFTR, one thread per radial line (J,K,L)
//   k_p         DRIFTP, one thread per azimuthal ring (I,K,L)
//   k_plane     DRIFTE / DRIFTMU sweeps, loss operators and SUMRC on whole
//               (K,L) planes in LDS
//   k_finalize  DtDrift* from the CFL partial maxima; SUMRC bookkeeping
//   k_vs        Volland-Stern VT from Kp
//
// Every kernel takes a Launch: a list of (problem, species) units, so one
// launch can cover all species of one problem or all members of an
// ensemble.
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
#include "sweep_common.h"

#include <hip/hip_runtime.h>

namespace ramgpu {

// ---------------------------------------------------------------------------
// Per-species tables.  DRIFTPARA (src/ModRamDrift.f90:64-85): P2, EDOT; the
// K-only factor P4 of DRIFTR (:131); the K-only prefixes of DRIFTE's EDT1
// (:337) and DRIFTMU's EDT (:424); CEPARA's cross-section x velocity
// (src/ModRamLoss.f90:39-81) and ATLOS (:162-167); DRIFTE's relativistic
// ghost factors (:310-311,334-335).  Species 0's block also forms the grid
// reciprocals.  One block per (problem, species).
__global__ void k_species(Launch L) {
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int s = u.s;
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
// One unit per problem; k_coef_ij: one thread per (I,J), k_coef: one block per
// (I,J).
// Volland-Stern potential VT = AVS*(LZ*RE)^2*sin(PHI-PHIOFS) with AVS =
// 7.05e-6/(1-0.159Kp+0.0093Kp^2)^3/RE (src/ModRamRun.f90:45-54): one
// expression for every kernel that forms it, so they agree bit for bit.
__device__ __forceinline__ double vs_avs(double kp) {
  const double d = 1.0 - 0.159 * kp + 0.0093 * (kp * kp);
  return 7.05E-6 / (d * d * d) / kRE;
}
__device__ __forceinline__ double vs_vt(double avs, double lz, double phi) {
  const double lre = lz * kRE;
  return avs * (lre * lre) * sin(phi - 0.0);
}

// With kp (one step of a Kp schedule) VT is formed here from Kp, where it
// is read, and stored for later launches: the block stages (LZ*RE)^2 and
// sin(PHI), which do not depend on Kp, in LDS (coef_ij_lds_bytes), and VT =
// (AVS*(LZ*RE)^2)*sin(PHI) is the product vs_vt forms.  bs, when given, is
// the step's boundary-flux scales, copied to the descriptor's.
// (LZ*RE)^2 [NR+1] and sin(PHI) [NT] into svt, by the block
__device__ inline void stage_vt_factors(const Tab &t, double *svt) {
  const int NR1 = t.nr + 1;
  for (int x = threadIdx.x; x < NR1 + t.nt; x += blockDim.x) {
    if (x < NR1) {
      const double lre = t.lz[x] * kRE;
      svt[x] = lre * lre;
    } else {
      svt[x] = sin(t.phi[x - NR1] - 0.0);
    }
  }
  __syncthreads();
}

// VT(I,J): the array, or with kp (avs) the factors staged in svt
#define VTV(II, JJ) (kp ? (avs * svt[(II)-1]) * svt[t.nr + 1 + (JJ)-1] : GEO2(t.vt, II, JJ))

// DRIFTR's terms of face I (0-based i) on line J: rij = {CR, BNES(I,J) +
// BNES(I+1,J), its reciprocal, outsideMGNP(I,J)} (src/ModRamDrift.f90:
// 118-127) and rsi = {RLZ+MDR/2, its reciprocal}; one definition for the
// table and for the ghost carry, which forms the face-NR row itself
__device__ __forceinline__ double4 radial_rij(const Tab &t, int i, int j, const double *kp, double avs,
                                              const double *svt) {
  const int I = i + 1, J = j + 1, NT = t.nt;
  const double dts = t.dts, mdr = t.mdr, dphi = t.dphi;
  const double rlz = t.rlz[i];
  const int J0 = (J == 1) ? NT - 1 : J - 1;
  const int J1 = (J == NT) ? 2 : J + 1;
  const double vr = dts / mdr / (rlz + 0.5 * mdr) / 2.0 / dphi;  // :67
  const double bsum = GEO2(t.bnes, I, J) + GEO2(t.bnes, I + 1, J);
  const double dv = VTV(I, J0) + VTV(I + 1, J0) - VTV(I, J1) - VTV(I + 1, J1);
  const double cr = vr * dv / bsum + (GEO2(t.eip, I, J) + GEO2(t.eip, I + 1, J)) / bsum * dts / mdr;
  return make_double4(cr, bsum, 1.0 / bsum, (double)MGNP(I, J));
}
__device__ __forceinline__ double2 radial_rsi(const Tab &t, int i) {
  const double rs = t.rlz[i] + 0.5 * t.mdr;
  return make_double2(rs, 1.0 / rs);
}

__device__ inline void coef_ij_unit(const Launch &L, int vb, const double *kp, const double *bs, double avs,
                                    double *svt) {
  {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int nbc = (t.nr * t.nt + blockDim.x - 1) / blockDim.x;  // units of the (I,J) terms
    if (u.b >= nbc) {
      // WAVELO's factors at the step's Kp, one (S,I,J,K) per thread, K >= 2,
      // I >= 2, for the species that take it (src/ModRamWPI.f90:596-630)
      const int m = (u.b - nbc) * blockDim.x + threadIdx.x;
      const int NE = t.ne, NT = t.nt;
      if (m < t.ns * t.nr * NT * NE) {
        const int k = m % NE, r = m / NE, j = r % NT, r2 = r / NT, i = r2 % t.nr, s = r2 / t.nr;
        if (t.wpi[s] && i >= 1 && k >= 1) t.wavef[m] = wave_factor(t, s, i, j, k, kp ? *kp : *t.kpnow);
      }
      return;
    }
    if (bs && u.b == 0)
      for (int s = threadIdx.x; s < t.ns; s += blockDim.x) t.bscale[s] = bs[s];
    if (kp && u.b == 0 && threadIdx.x == 0) *t.kpnow = *kp;  // WAVELO's KP of the step
    if (kp) stage_vt_factors(t, svt);
    const int n = u.b * blockDim.x + threadIdx.x;
    if (n < t.nr * t.nt) {
      const int i = n / t.nt, j = n % t.nt;
      const int I = i + 1, J = j + 1;
      const int NT = t.nt;
      const double dts = t.dts, mdr = t.mdr, dphi = t.dphi;
      const double rlz = t.rlz[i];
      const size_t ij = (size_t)i * NT + j;
      const double mg = (double)MGNP(I, J);
      if (j == 0) {
        t.rrlz[i] = 1.0 / rlz;
        t.rsi[i] = radial_rsi(t, i);
      }
      t.rij[ij] = radial_rij(t, i, j, kp, avs, svt);
      double pa = 0.0, pce = 0.0, pbs = 1.0;
      double drd1 = 0.0, dpd1 = 0.0, drm1 = 0.0, dpm1 = 0.0, db2 = 0.0;
      const double bij = GEO2(t.bnes, I, J);
      if (I >= 2) {
        if (J >= 2) {
          const int J1 = (J == NT) ? 2 : J + 1;
          const double p1 = dts / dphi / 2.0 / mdr / rlz;  // :68
          pa = (VTV(I + 1, J) + VTV(I + 1, J1) - VTV(I - 1, J) - VTV(I - 1, J1)) * p1;
          pce = (GEO2(t.eir, I, J1) + GEO2(t.eir, I, J)) / rlz * dts / dphi;
          pbs = GEO2(t.bnes, I, J) + GEO2(t.bnes, I, J1);
        }
        const int J0 = (J == 1) ? NT - 1 : J - 1;
        const int J2 = (J == NT) ? 2 : J + 1;
        drd1 = (GEO2(t.eip, I, J) * rlz - (VTV(I, J2) - VTV(I, J0)) / 2.0 / dphi) / bij;
        dpd1 = kOME * rlz + ((VTV(I + 1, J) - VTV(I - 1, J)) / 2.0 / mdr - GEO2(t.eir, I, J)) / bij;
        // DRIFTMU's DRM1/DPM1 (:415-416) are the same expressions as DRD1/DPD1
        drm1 = (GEO2(t.eip, I, J) * rlz - (VTV(I, J2) - VTV(I, J0)) / 2.0 / dphi) / bij;
        dpm1 = kOME * rlz + ((VTV(I + 1, J) - VTV(I - 1, J)) / 2.0 / mdr - GEO2(t.eir, I, J)) / bij;
        db2 = GEO2(t.dbdt, I, J) / 2.0 / bij * rlz;  // :431
      }
      t.pij[ij] = make_double4(pa, pce, pbs, 1.0 / pbs);
      double *k8 = t.kij + ij * 8;
      k8[KIJ_DRD1] = drd1;
      k8[KIJ_DPD1] = dpd1;
      k8[KIJ_BIJ] = bij;
      k8[KIJ_RBIJ] = 1.0 / bij;
      k8[KIJ_DRM1] = drm1;
      k8[KIJ_DPM1] = dpm1;
      k8[KIJ_DB2] = db2;
      k8[KIJ_MG] = mg;
      if (kp) {  // VT of this (I,J), and of I=NR+1 from the last row
        double *vt = const_cast<double *>(t.vt);
        GEO2(vt, I, J) = VTV(I, J);
        if (I == t.nr) GEO2(vt, I + 1, J) = VTV(I + 1, J);
      }
    }
    if (kp) __syncthreads();  // svt is rewritten for the block's next unit
  }
}

__global__ void k_coef_ij(Launch L, const double *kp, const double *bs) {
  extern __shared__ double svt[];  // with kp: (LZ*RE)^2 [NR+1], then sin(PHI) [NT]
  const double avs = kp ? vs_avs(*kp) : 0.0;
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) coef_ij_unit(L, vb, kp, bs, avs, svt);
}

__global__ void k_coef(Launch L) {
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int b = u.b;
    const int i = b / t.nt, j = b % t.nt;
    const int I = i + 1, J = j + 1;
    const int NT = t.nt, NPA = t.npa, NR = t.nr;
    const double dts = t.dts, mdr = t.mdr, dphi = t.dphi;
    const double rlz = t.rlz[i];
    const double mg = (double)MGNP(I, J);
    for (int l = threadIdx.x; l < NPA; l += blockDim.x) {
      const int L1 = l + 1;  // 1-based L
      // ---- DRIFTR, all I (:140-144)
      {
        const int J0 = (J == 1) ? NT - 1 : J - 1;
        const int J1 = (J == NT) ? 2 : J + 1;
        const double cgr1 = GEO3(t.fnis, I + 1, J1, L1) + GEO3(t.fnis, I, J1, L1) -
                            GEO3(t.fnis, I + 1, J0, L1) - GEO3(t.fnis, I, J0, L1);
        const double cgr2 = GEO2(t.bnes, I + 1, J1) + GEO2(t.bnes, I, J1) - GEO2(t.bnes, I + 1, J0) -
                            GEO2(t.bnes, I, J0);
        const double cgr3 = cgr1 + (GEO3(t.fnis, I + 1, J, L1) + GEO3(t.fnis, I, J, L1) -
                                    2.0 * GEO3(t.fnhs, I + 1, J, L1) - 2.0 * GEO3(t.fnhs, I, J, L1)) *
                                       cgr2 / 2.0 / (GEO2(t.bnes, I + 1, J) + GEO2(t.bnes, I, J));
        t.q1[((size_t)l * NR + i) * NT + j] = cgr3 / (GEO3(t.fnhs, I, J, L1) + GEO3(t.fnhs, I + 1, J, L1));
      }
      // ---- DRIFTP, I>=2, J>=2 (:232-235)
      {
        double gpa = 0.0, fhs = 1.0;
        if (I >= 2 && J >= 2) {
          const int J1 = (J == NT) ? 2 : J + 1;
          const double gpa1 = GEO3(t.fnis, I, J, L1) + GEO3(t.fnis, I, J1, L1) +
                              (GEO3(t.fnis, I + 1, J1, L1) + GEO3(t.fnis, I + 1, J, L1) -
                               GEO3(t.fnis, I - 1, J, L1) - GEO3(t.fnis, I - 1, J1, L1)) *
                                  rlz / 2.0 / mdr;
          const double gpa2 = rlz / 4.0 / mdr *
                              (GEO3(t.fnis, I, J, L1) + GEO3(t.fnis, I, J1, L1) - 2.0 * GEO3(t.fnhs, I, J, L1) -
                               2.0 * GEO3(t.fnhs, I, J1, L1)) *
                              (GEO2(t.bnes, I + 1, J1) + GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J) -
                               GEO2(t.bnes, I - 1, J1)) /
                              (GEO2(t.bnes, I, J) + GEO2(t.bnes, I, J1));
          gpa = gpa1 + gpa2;
          fhs = GEO3(t.fnhs, I, J, L1) + GEO3(t.fnhs, I, J1, L1);
        }
        t.pl[((size_t)l * NR + i) * NT + j] = make_double4(gpa, fhs, 1.0 / fhs, mg);
      }
      const size_t q = ((size_t)i * NT + j) * NPA + l;
      // ---- DRIFTE, I>=2 (:323-341)
      {
        double gr = 0, gp = 0, drd2 = 0, dpd2 = 0, db1 = 0, di1 = 0;
        const double fnhs = GEO3(t.fnhs, I, J, L1);
        if (I >= 2) {
          const int J0 = (J == 1) ? NT - 1 : J - 1;
          const int J2 = (J == NT) ? 2 : J + 1;
          const double bij = GEO2(t.bnes, I, J);
          const double fnis = GEO3(t.fnis, I, J, L1);
          const double gpa = (1.0 - fnis / 2.0 / fnhs) / bij;
          const double gpr1 = gpa * (GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J)) / 2.0 / mdr;
          const double gpr2 = -fnis / fnhs / rlz;
          const double gpr3 = -(GEO3(t.fnis, I + 1, J, L1) - GEO3(t.fnis, I - 1, J, L1)) / 2.0 / mdr / fnhs;
          const double gpp1 = gpa * (GEO2(t.bnes, I, J2) - GEO2(t.bnes, I, J0)) / 2.0 / dphi;
          const double gpp2 = -(GEO3(t.fnis, I, J2, L1) - GEO3(t.fnis, I, J0, L1)) / 2.0 / dphi / fnhs;
          drd2 = (GEO3(t.fnis, I, J2, L1) - GEO3(t.fnis, I, J0, L1)) / 2.0 / dphi +
                 (fnis - 2.0 * fnhs) * (GEO2(t.bnes, I, J2) - GEO2(t.bnes, I, J0)) / 4.0 / bij / dphi;
          dpd2 = fnis + (GEO3(t.fnis, I + 1, J, L1) - GEO3(t.fnis, I - 1, J, L1)) * rlz / 2.0 / mdr +
                 rlz * (fnis - 2.0 * fnhs) / 4.0 / mdr * (GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J)) / bij;
          db1 = GEO2(t.dbdt, I, J) * (1.0 - fnis / 2.0 / fnhs) * rlz / bij;
          di1 = -GEO3(t.didt, I, J, L1) * rlz / fnhs;
          gr = gpr1 + gpr2 + gpr3;
          gp = gpp1 + gpp2;
        }
        double *e = t.kle + q * 8;
        e[KLE_HS] = fnhs;
        e[KLE_RHS] = 1.0 / fnhs;
        e[KLE_DRD2] = drd2;
        e[KLE_DPD2] = dpd2;
        e[KLE_GR] = gr;
        e[KLE_GP] = gp;
        e[KLE_DB1] = db1;
        e[KLE_DI1] = di1;
      }
      // ---- DRIFTMU, I>=2, L>=2 (:418-432; MUDOT from DRIFTPARA :77-81)
      {
        double cmudot = 0, gmr = 0, gmp = 0, drm2 = 0, dpm2 = 0, di2 = 0;
        const double bhs = GEO3(t.bounhs, I, J, L1);
        if (I >= 2 && L1 >= 2) {
          const int J0 = (J == 1) ? NT - 1 : J - 1;
          const int J1 = (J == NT) ? 2 : J + 1;
          const double bij = GEO2(t.bnes, I, J);
          double mudot = 0.0;
          if (L1 < NPA) {
            const double muboun = t.mu[l] + 0.5 * t.wmu[l];
            mudot = (1.0 - muboun * muboun) * dts / 2.0 / muboun / rlz;
          }
          const double bis = GEO3(t.bounis, I, J, L1);
          cmudot = mudot * bis / bhs;
          const double gmr1 = (GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J)) / 4.0 / mdr / bij;
          const double gmr2 = 1.0 / rlz;
          const double gmr3 = (GEO3(t.bounis, I + 1, J, L1) - GEO3(t.bounis, I - 1, J, L1)) / 2.0 / mdr / bis;
          const double gmp1 = (GEO2(t.bnes, I, J1) - GEO2(t.bnes, I, J0)) / 4.0 / dphi / bij;
          const double gmp2 = (GEO3(t.bounis, I, J1, L1) - GEO3(t.bounis, I, J0, L1)) / 2.0 / dphi / bis;
          drm2 = (GEO3(t.bounis, I, J1, L1) - GEO3(t.bounis, I, J0, L1)) / 2.0 / dphi +
                 (bis - 2.0 * bhs) * (GEO2(t.bnes, I, J1) - GEO2(t.bnes, I, J0)) / 4.0 / bij / dphi;
          dpm2 = bis + (GEO3(t.bounis, I + 1, J, L1) - GEO3(t.bounis, I - 1, J, L1)) * rlz / 2.0 / mdr +
                 (bis - 2.0 * bhs) * rlz / 4.0 / mdr * (GEO2(t.bnes, I + 1, J) - GEO2(t.bnes, I - 1, J)) / bij;
          di2 = GEO3(t.dibndt, I, J, L1) * rlz / bis;
          gmr = gmr1 + gmr2 + gmr3;
          gmp = gmp1 + gmp2;
        }
        double *m = t.klm + q * 8;
        m[KLM_BHS] = bhs;
        m[KLM_RBHS] = 1.0 / bhs;
        m[KLM_DRM2] = drm2;
        m[KLM_DPM2] = dpm2;
        m[KLM_CMUDOT] = cmudot;
        m[KLM_GMR] = gmr;
        m[KLM_GMP] = gmp;
        m[KLM_DI2] = di2;
      }
      t.kll[q] = make_double2(GEO3(t.hdns_f, I, J, L1), 1.0 / GEO3(t.fnhs, I, J, L1));
    }
  }
}

// ---------------------------------------------------------------------------
// DRIFTR's outer ghost F(NR+1) is a per-call buffer that is not reset between
// lines (src/ModRamDrift.f90:112-113,159-167): an outflow line (C(NR)>=0)
// whose face NR-1 is upwinded from outside (C(NR-1)<0, :174-175) reads the
// ghost left by the most recent inflow line in the reference's K,L,J loop
// order, or 0 when no inflow line came before.  Whether a line is inflow
// depends on geometry, fields and grid only, so the carry is precomputed:
// carry[s][p] is the index p'*NT+(J'-1) of the last inflow line before plane
// p=(K-1)*NPA+(L-1), or -1.  One block per (problem, species): each thread
// finds the top inflow J of its planes (four J per memory round trip) into
// LDS, then a block-wide exclusive max-scan over the planes in order (wave
// shuffles, then the waves' totals).
// One (problem, species) of the ghost carry.  The face-NR terms it tests
// (rij(NR-1,.), rsi(NR-1)) are formed here with the table's own code, so the
// carry does not wait for k_coef_ij; rrow holds them (4*NT doubles), svt the
// VT factors when kp is given.  q1s (NPA*NT doubles, or null): LDS for the
// face-NR row of q1, loaded in one pass so that the search for each line's
// last inflow J reads no global memory.
__device__ inline void ghost_unit(const Launch &L, int vb, const double *kp, double avs, double *svt, double *rrow,
                                  int *s_last, int *s_wave, double *q1s = nullptr) {
  const Unit u = unit_of(L, vb);
  const Tab &t = *u.t;
  const int s = u.s, NR = t.nr, NT = t.nt, NPA = t.npa;
  const int P = t.ne * NPA, T = blockDim.x, tid = threadIdx.x;
  if (kp) stage_vt_factors(t, svt);
  if (q1s)
    for (int x = tid; x < NPA * NT; x += T) {
      const int l = x / NT, j = x - l * NT;
      q1s[x] = t.q1[((size_t)l * NR + NR - 1) * NT + j];
    }
  for (int j = tid; j < NT; j += T) {
    const double4 r = radial_rij(t, NR - 1, j, kp, avs, svt);
    rrow[4 * j] = r.x;
    rrow[4 * j + 1] = r.y;
    rrow[4 * j + 2] = r.z;
    rrow[4 * j + 3] = r.w;
  }
  __syncthreads();
  const double2 rs = radial_rsi(t, NR - 1);
  for (int p = tid; p < P; p += T) {
    const int k = p / NPA, l = p - k * NPA;
    const double p4k = t.p4[s * t.ne + k];
    // the highest inflow J of the line, searched down from NT in groups of 4
    auto search = [&](auto q1) {
      int last = -1;
      for (int j0 = NT - 1; j0 >= 0 && last < 0; j0 -= 4) {
        double c[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int j = max(j0 - m, 0);
          c[m] = c_radial_of(make_double4(rrow[4 * j], rrow[4 * j + 1], rrow[4 * j + 2], rrow[4 * j + 3]), rs, q1[j],
                             p4k);
        }
#pragma unroll
        for (int m = 3; m >= 0; --m)
          if (j0 - m >= 0 && c[m] < 0) last = p * NT + j0 - m;
      }
      return last;
    };
    s_last[p] = q1s ? search(q1s + (size_t)l * NT) : search(as_global(t.q1 + ((size_t)l * NR + NR - 1) * NT));
  }
  __syncthreads();
  // planes [p0, p1) of this thread, in order
  const int chunk = (P + T - 1) / T;
  const int p0 = min(P, tid * chunk), p1 = min(P, p0 + chunk);
  int v = -1;
  for (int p = p0; p < p1; ++p) v = max(v, s_last[p]);
  const int lane = tid & 63, w = tid >> 6;
  for (int off = 1; off < 64; off <<= 1) {  // inclusive max-scan in the wave
    const int o = __shfl_up(v, off, 64);
    if (lane >= off) v = max(v, o);
  }
  if (lane == 63) s_wave[w] = v;
  __syncthreads();
  int run = __shfl_up(v, 1, 64);
  if (lane == 0) run = -1;
  for (int ww = 0; ww < w; ++ww) run = max(run, s_wave[ww]);
  // the index of an inflow line grows with p, so the last one is the maximum
  for (int p = p0; p < p1; ++p) {
    t.carry[(size_t)s * P + p] = run;
    run = max(run, s_last[p]);
  }
  __syncthreads();
}
#undef VTV

__global__ void __launch_bounds__(1024) k_ghost(Launch L) {
  extern __shared__ double rrow[];  // 4*NT: ghost_rows_lds_bytes
  __shared__ int s_last[kMaxPlaneCells];
  __shared__ int s_wave[16];
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) ghost_unit(L, vb, nullptr, 0.0, nullptr, rrow, s_last, s_wave);
}

// ---------------------------------------------------------------------------
// DRIFTR, src/ModRamDrift.f90:139-191: one thread per radial line (J,K,L),
// consecutive lanes on consecutive (L,K) positions.  The thread walks I with
// a register window F(I-1..I+2) of old values, forms face I, and updates
// cell I from faces I and I-1.  F(I+3) comes from a ring of registers loaded
// four to eight steps ahead, and the face terms (q1, rij, rsi) of the
// block's lines are staged in LDS first, so a step waits on no memory
// latency.
// MODE 0: sweep; 1: sweep and record the CFL maximum; 2: the last sweep of
// ram_run: also the SUMRC partial sum of src/ModRamRun.f90:246-255 (slot 5),
// then F2(:,:,NT)=F2(:,:,1) and the magnetopause mask (:186,197-201).  In
// mode 2 the J=NT line is not swept: the seam overwrites it with line J=1,
// whose thread stores both.

template <int MODE>
__global__ void __launch_bounds__(kLineThreads, 4) k_r(Launch L) {
  __shared__ double red[kLineThreads / 64];
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
    const bool act = line < NT * P && !(MODE == 2 && j == NT - 1);
    const RadialTerms rt = stage_radial_terms(t, line0, sterm);
    double cmax = 0.0, acc = 0.0;
    if (act) {
      const int q = line - j * P;
      const int l = fast_div(q, NE, t.rne), k = q - l * NE;
      const double beta = t.betalim;
      const size_t di = (size_t)NT * P;  // stride between I and I+1
      const size_t off = (size_t)s * f2_species_stride(t) + (size_t)j * P + q;
      const gptr<const double> fl = as_global(launch_src(L, t) + off);
      const gptr<double> fo = as_global(launch_dst(L, t) + off);
      const gptr<double> fs = fo + (size_t)(NT - 1 - j) * P;  // line J=NT (seam target in mode 2)
      const double *const qrow = rt.q_row(j, l, t.npa), *const rrow = rt.r_row(j), *const rst = rt.rs, *const mgt = rt.mg;
#define RTERM(II) make_double4(rrow[4 * (II)], rrow[4 * (II) + 1], rrow[4 * (II) + 2], rrow[4 * (II) + 3])
#define RSTERM(II) make_double2(rst[2 * (II)], rst[2 * (II) + 1])
      // F(1..4) and the ring's first F(I+3), I = 2..5, in one batch
      double fw[4], pa[4], pb[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) fw[m] = fl[(size_t)m * di];  // NR >= 4
#pragma unroll
      for (int m = 0; m < 4; ++m) pa[m] = (5 + m <= NR) ? fl[(size_t)(4 + m) * di] : 0.0;
      const double p4k = t.p4[s * NE + k];
      const bool in = c_radial_of(RTERM(NR - 1), RSTERM(NR - 1), qrow[NR - 1], p4k) < 0;
      double gh1 = 0.0, gh2 = 0.0;
      if (in) {
        gh1 = ghost_value(t, s, j, k, l, t.conf1);
        gh2 = ghost_value(t, s, j, k, l, t.conf2);
      }
      double wek = 0, wml = 0, ekk = 0;
      if (MODE == 2) {
        wek = t.we[k];
        wml = t.wmu[l];
        ekk = t.ekev[k];
      }
      const bool sum_line = MODE == 2 && k >= 1 && l >= 1 && j <= NT - 2;
      // face 1: FBND(1) = F(2) on inflow lines, 0 otherwise (:156,164)
      const double f1 = fw[0], f2v = fw[1], f3 = fw[2];
      double fnext = fw[3];  // F(4)
      const double4 r1 = RTERM(0);
      const double c1 = c_radial_of(r1, RSTERM(0), qrow[0], p4k);
      if (MODE == 1 && r1.w == 0.0) cmax = fmax(cmax, fmax(fabs(c1), 1E-10));
      double gprev = c1 * (in ? f2v : 0.0);
      if (MODE == 2) {
        const double mg1 = r1.w, mgs = mgt[0];
        if (mg1 == 1.0) fo[0] = 1e-31;
        else if (fo != fl) fo[0] = f1;
        if (j == 0) fs[0] = (mgs == 1.0) ? 1e-31 : f1;
      } else if (fo != fl) {
        fo[0] = f1;  // cell 1 is not updated
      }
      double wm1 = f1, w0 = f2v, w1 = f3, w2 = fnext;  // F(I-1), F(I), F(I+1), F(I+2) at I=2
      // F(NR+1) of face NR-1 when it is upwinded from outside: the line's own
      // ghost on inflow lines, else the one a previous inflow line left
      // (needed only when C(NR-1) < 0; formed here, once, so the steps have
      // no branches)
      double fgh = gh1;
      if (!in && c_radial_of(RTERM(NR - 2), RSTERM(NR - 2), qrow[NR - 2], p4k) < 0) fgh = stale_ghost(t, s, j, k, l);
      // step I with F(I+3) = FN (0 beyond NR); selects only, so the compiler
      // can overlap a step's Courant number with the previous step's limiter
#define RSTEP(IV, FN)                                                              \
  {                                                                                \
    const int I = (IV), i = I - 1;                                                 \
    fnext = (FN);                                                                  \
    const double4 r = RTERM(i);                                                    \
    const double ci = c_radial_of(r, RSTERM(i), qrow[i], p4k);                     \
    if (MODE == 1) cmax = (r.w == 0.0) ? fmax(cmax, fmax(fabs(ci), 1E-10)) : cmax; \
    const bool last = I == NR, neg = ci < 0;                                       \
    const double fp1 = last ? gh1 : w1;                                            \
    /* upwind side for C<0: ghost F(NR+2) at face NR, F(NR+1) at face NR-1 */      \
    const double fneg = last ? gh2 : ((I + 2 <= NR) ? w2 : fgh);                   \
    const double fv = face_value_sel(w0, fp1, neg ? fneg : w0, neg ? fp1 : wm1, ci, ci, beta); \
    const double fb = (last && !in) ? w0 : fv; /* FBND(NR) = F(NR) on outflow lines */ \
    const double g = ci * fb;                                                      \
    double v = w0 - g + gprev;                                                     \
    if (v < 0) v = 1E-15;                                                          \
    gprev = g;                                                                     \
    if (MODE == 2) {                                                               \
      acc += sum_line ? ekk * (v * wek * wml) : 0.0;                               \
      fo[(size_t)i * di] = (r.w == 1.0) ? 1e-31 : v;                               \
      if (j == 0) fs[(size_t)i * di] = (mgt[i] == 1.0) ? 1e-31 : v;                \
    } else {                                                                       \
      fo[(size_t)i * di] = v;                                                      \
    }                                                                              \
    wm1 = w0;                                                                      \
    w0 = w1;                                                                       \
    w1 = w2;                                                                       \
    w2 = fnext;                                                                    \
  }
      // steps I0..I0+3 read ring pa, I0+4..I0+7 ring pb; each ring is
      // reloaded four steps before its next use
      for (int I0 = 2; I0 <= NR; I0 += 8) {
#pragma unroll
        for (int m = 0; m < 4; ++m) pb[m] = (I0 + 7 + m <= NR) ? fl[(size_t)(I0 + 6 + m) * di] : 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (I0 + m <= NR) RSTEP(I0 + m, pa[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m) pa[m] = (I0 + 11 + m <= NR) ? fl[(size_t)(I0 + 10 + m) * di] : 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (I0 + 4 + m <= NR) RSTEP(I0 + 4 + m, pb[m]);
      }
#undef RSTEP
#undef RTERM
#undef RSTERM
    }
    if (MODE == 1) {
      const double m = block_max(cmax, red);
      if (threadIdx.x == 0 && m > 0) atomic_max_pos(&t.cflmax[(size_t)s * (2 + NE + t.npa) + 0], m);
    }
    if (MODE == 2) {
      double a[1] = {acc}, tot[1];
      block_sum_n<1>(a, red, tot);
      if (threadIdx.x == 0) t.partials[((size_t)kSumRadial * t.ns + s) * t.maxblk + u.b] = tot[0];
    }
    __syncthreads();  // the face terms of this block are read
  }
}

// ---------------------------------------------------------------------------
// DRIFTP, src/ModRamDrift.f90:229-272: one thread per ring (I>=2,K,L),
// periodic in J.  Face NT is formed first (FBND(1)=FBND(NT),
// CDriftP(1)=CDriftP(NT), :261-262), then faces and cells J=2..NT-1 in order
// with a register window, then cell NT and F2(I,1)=F2(I,NT) (:272).
// MODE 0: sweep; 1: sweep and record the CFL maximum.
template <int MODE>
__global__ void __launch_bounds__(kLineThreads) k_p(Launch L) {
  __shared__ double red[kLineThreads / 64];
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
    AzimTerms at = {};
    if (line0 < nring) at = stage_azim_terms(t, line0, min(line0 + kLineThreads, nring) - 1, sterm);
    double cmax = 0.0;
    if (line >= nring && out_of_place(L)) {
      // out of place: the I=1 rows are not swept; extra lines copy them
      const int x = line - nring;
      if (x < NT * P) {
        const size_t off = (size_t)s * f2_species_stride(t) + x;
        launch_dst(L, t)[off] = launch_src(L, t)[off];
      }
    } else if (line < nring) {
      const int i1 = fast_div(line, P, t.rp);
      const int q = line - i1 * P;
      const int i = i1 + 1;
      const int l = fast_div(q, NE, t.rne), k = q - l * NE;
      const double p2k = t.p2[((size_t)s * NR + i) * NE + k];
      const double beta = t.betalim, omet = t.omet;
      const size_t off = (size_t)s * f2_species_stride(t) + (size_t)i * NT * P + q;  // F(J) at off+(J-1)*P
      const gptr<const double> fl = as_global(launch_src(L, t) + off);
      const gptr<double> fo = as_global(launch_dst(L, t) + off);
      const double *const ap = at.pij_row(i1), *const gp = at.pl_row(i1, l, t.npa);
      // Courant number of face J (src/ModRamDrift.f90:236-239)
#define CAZIM(JJ, OUT, MG)                                                 \
  {                                                                        \
    const double *a_ = ap + 4 * ((JJ)-1), *g_ = gp + 4 * ((JJ)-1);         \
    const double u_ = div_rn(p2k * g_[0], g_[1], g_[2]);                   \
    const double num_ = a_[0] - u_ - a_[1];                                \
    OUT = div_rn(num_, a_[2], a_[3]) + omet;                               \
    MG = g_[3];                                                            \
  }
      // F(1..4), F(NT-1), F(NT) and the ring's first F(J+3), J = 2..5
      const double F1 = fl[0], F2o = fl[P], F3 = fl[2 * (size_t)P];
      const double FNm1 = fl[(size_t)(NT - 2) * P], FN = fl[(size_t)(NT - 1) * P];
      double wm1 = F1, w0 = F2o, w1 = F3, w2 = fl[3 * (size_t)P];  // F(J-1..J+2) at J=2; NT >= 4
      double pa[4], pb[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) pa[m] = (5 + m <= NT) ? fl[(size_t)(4 + m) * P] : 0.0;
      double cN, mgN;
      CAZIM(NT, cN, mgN);
      if (MODE == 1 && mgN == 0.0) cmax = fmax(cmax, fmax(fabs(cN), 1E-10));
      // face NT: J1 = 2; upwind N = NT+2 -> 3 (:250-251)
      const double gN = cN * face_value(FN, F2o, (cN < 0) ? F3 : FN, (cN < 0) ? F2o : FNm1, cN, cN, beta);
      double gprev = gN;
      // step J with F(J+3) = FN_ (0 beyond NT); selects only
#define PSTEP(JV, FN_)                                                          \
  {                                                                             \
    const int J = (JV);                                                         \
    const double fnext = (FN_);                                                 \
    double cp, mg;                                                              \
    CAZIM(J, cp, mg);                                                           \
    if (MODE == 1) cmax = (mg == 0.0) ? fmax(cmax, fmax(fabs(cp), 1E-10)) : cmax; \
    /* C<0: N = J+2, or at J = NT-1 N = NT+1 -> 2, N-1 = 1 */                   \
    const bool neg = cp < 0, inr = J + 2 <= NT;                                 \
    const double fn = neg ? (inr ? w2 : F2o) : w0;                              \
    const double fnm1 = neg ? (inr ? w1 : F1) : wm1;                            \
    const double g = cp * face_value_sel(w0, w1, fn, fnm1, cp, cp, beta);      \
    double v = w0 - g + gprev;                                                  \
    if (v < 0) v = 1E-15;                                                       \
    fo[(size_t)(J - 1) * P] = v;                                                \
    gprev = g;                                                                  \
    wm1 = w0;                                                                   \
    w0 = w1;                                                                    \
    w1 = w2;                                                                    \
    w2 = fnext;                                                                 \
  }
      // steps J0..J0+3 read ring pa, J0+4..J0+7 ring pb; each ring is
      // reloaded four steps before its next use
      for (int J0 = 2; J0 <= NT - 1; J0 += 8) {
#pragma unroll
        for (int m = 0; m < 4; ++m) pb[m] = (J0 + 7 + m <= NT) ? fl[(size_t)(J0 + 6 + m) * P] : 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (J0 + m <= NT - 1) PSTEP(J0 + m, pa[m]);
#pragma unroll
        for (int m = 0; m < 4; ++m) pa[m] = (J0 + 11 + m <= NT) ? fl[(size_t)(J0 + 10 + m) * P] : 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (J0 + 4 + m <= NT - 1) PSTEP(J0 + 4 + m, pb[m]);
      }
#undef PSTEP
#undef CAZIM
      double v = FN - gN + gprev;
      if (v < 0) v = 1E-15;
      fo[(size_t)(NT - 1) * P] = v;
      fo[0] = v;
    }
    if (MODE == 1) {
      const double m = block_max(cmax, red);
      if (threadIdx.x == 0 && m > 0) atomic_max_pos(&t.cflmax[(size_t)s * (2 + NE + t.npa) + 1], m);
    }
    __syncthreads();  // the block's face terms are read
  }
}

// ---------------------------------------------------------------------------
// Plane kernel: one block per (species, I>=2, J), holding the (K,L) plane in
// LDS.  Thread n < Q*NE (Q = plane_chunks) owns column K = n%NE of the
// consecutive rows of chunk q = n/NE (NPA/Q rows, the first NPA%Q chunks
// one more); only the owner of a cell ever writes it.  The reference
// divides both face fluxes of cell K by that cell's width,
// CDriftE(K)/WE(K)*FBND(K) and CDriftE(K-1)/WE(K)*FBND(K-1)
// (src/ModRamDrift.f90:364; :460 for Mu): the owner of face K
// forms F(K) - C(K)/WE(K)*FBND(K) and the term C(K)/WE(K+1)*FBND(K) its
// upper neighbour adds.  Along L the neighbour is the same thread (the
// chunk's last row hands its term to the next chunk through LDS); along K
// it is the next lane, reached by a wavefront shift (the last lane of a
// wave hands its terms to the next wave through LDS).
//
// LDS rows hold the energy stencil's positions 0..NE+2 of the row:
// [F(0), F(1), F(2) .. F(NE), F(NE+1)=0, F(NE+2)=0] with DRIFTE's
// relativistic ghosts in positions 0 and 1 (refreshed before each DRIFTE);
// the cells K=1 have a column of their own.  So every lane reads its
// energy stencil at fixed offsets.
struct PlaneFwd {
  double *F;   // NPA rows of NE+3 positions
  double *F1;  // NPA: the cells K=1
  double *X;   // T: terms handed between chunks (DRIFTMU) and waves (DRIFTE)
  double *cf, *cm, *cl, *kt, *lt, *red, *ex, *pc;
  unsigned long long *cmx;
  double *cx;  // Coulomb kernels only: 7*NPA per-L terms, then COULMU's RL (NE*NPA)
};

__device__ __forceinline__ PlaneFwd plane_fwd_lds(double *sm, int ne, int npa, int threads) {
  PlaneFwd d;
  d.F = sm;
  d.F1 = d.F + (size_t)npa * (ne + 3);
  d.X = d.F1 + npa;
  d.cf = d.X + threads;       // 8*NPA: per-L DRIFTE terms (KleField)
  d.cm = d.cf + 8 * npa;      // 8*NPA: per-L DRIFTMU terms (KlmField)
  d.cl = d.cm + 8 * npa;      // 2*NPA: HDNS, 1/FNHS
  d.kt = d.cl + 2 * npa;      // KT_N*NE
  d.lt = d.kt + KT_N * ne;    // LT_N*NPA
  d.red = d.lt + LT_N * npa;  // 128: block reductions (up to 16 waves x 7 sums: 112)
  d.ex = d.red + 128;         // 64: 2^(j/64) of exp_tab
  d.pc = d.ex + 64;           // PC_N: the plane's scalars (PcField)
  d.cmx = (unsigned long long *)(d.pc + PC_N);  // NE + kMuSpread*NPA
  d.cx = (double *)(d.cmx + ne + kMuSpread * npa);
  return d;
}
static_assert(KT_N == 11 && LT_N == 4 && PC_N == 24, "plane_fwd_lds_bytes counts these tables");

struct FwdCtx {
  int s, i, j, NE, NPA, NT, S;  // S = NE+3 positions per LDS row
  int cex, upa;                 // the species' charge-exchange kind, UPA(I)
  size_t ij;
  double rlz, rrlz, qs, beta;
  int ke, q, l0, R, nr;  // column, chunk, its first row, the longest chunk's rows, this thread's rows (0: idle)
  double *col;           // cell (ke, l) is col[l*cst]
  int cst;
};

// x, hidden from the optimiser: row indices derived from it are formed
// afresh in each phase instead of being kept in registers across phases.
__device__ __forceinline__ int fresh(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// The value of x in lane `lane` of the wave (a wave-uniform value).
__device__ __forceinline__ double lane_value(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// A block-uniform x (read from LDS) held in scalar registers.
__device__ __forceinline__ double uniform(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// x, formed here: keeps the compiler from sinking the arithmetic that
// makes a row's result down to the barrier, where it would hold every
// operand of every row instead of one result per row.
__device__ __forceinline__ double settle(double x) {
  asm("" : "+v"(x));
  return x;
}

// The value of v in lane-1 (lane 0: -0.0, which adds as an exact zero).
__device__ __forceinline__ double from_lane_below(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)0x80000000, (int)(b >> 32), 0x138, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Block set-up: the context of plane b of species s, this thread's cells
// (from src, device layout of one species) and the plane's tables.
template <int MAXC>
__device__ inline void fwd_setup(const Tab &t, const Unit &u, double *sm, const double *src, bool need_loss,
                                 bool need_kl, FwdCtx &c, PlaneFwd &d) {
  const int NE = t.ne, NPA = t.npa, NT = t.nt;
  const int tid = threadIdx.x;
  c.s = u.s;
  c.i = 1 + u.b / NT;  // I = i+1 >= 2
  c.j = u.b - (c.i - 1) * NT;
  c.NE = NE;
  c.NPA = NPA;
  c.NT = NT;
  c.S = NE + 3;
  c.ij = (size_t)c.i * NT + c.j;
  c.rlz = t.rlz[c.i];
  c.rrlz = t.rrlz[c.i];
  c.qs = (double)t.charge[c.s];  // x/QS == x*QS for QS = +-1
  c.cex = t.cex[c.s];
  c.upa = t.upa[c.i];
  c.beta = t.betalim;
  const int Q = t.plane_chunks, base = NPA / Q, extra = NPA - base * Q;
  c.R = base + (extra > 0 ? 1 : 0);
  c.q = fast_div(tid, NE, t.rne);
  c.ke = tid - c.q * NE;
  c.l0 = c.q * base + min(c.q, extra);
  c.nr = (c.q < Q) ? base + (c.q < extra ? 1 : 0) : 0;
  d = plane_fwd_lds(sm, NE, NPA, blockDim.x);
  c.col = (c.ke == 0) ? d.F1 : d.F + c.ke + 1;
  c.cst = (c.ke == 0) ? 1 : c.S;
  // the plane's base is uniform; cells are 32-bit offsets from it
  const gptr<const double> gF = as_global(src + (size_t)c.s * f2_species_stride(t) + c.ij * ((size_t)NE * NPA));
  double vf[MAXC];
#pragma unroll
  for (int n = 0; n < MAXC; ++n) vf[n] = (n < c.nr) ? gF[(unsigned)((c.l0 + n) * NE + c.ke)] : 0.0;
  plane_tables(t, c.s, c.i, c.j, d, need_loss, need_kl);
  for (int l = tid; l < NPA; l += blockDim.x) {  // F(NE+1) = F(NE+2) = 0
    d.F[l * c.S + NE + 1] = 0.0;
    d.F[l * c.S + NE + 2] = 0.0;
  }
#pragma unroll
  for (int n = 0; n < MAXC; ++n)
    if (n < c.nr) c.col[(c.l0 + n) * c.cst] = vf[n];
  __syncthreads();
}

// The plane to dst (device layout of one species' plane), read from LDS
// after a barrier: consecutive threads store consecutive cells, so every
// cache line of the plane is written whole by one instruction.
__device__ __forceinline__ void fwd_store(const FwdCtx &c, const PlaneFwd &d, float rne, double *dst) {
  const gptr<double> g = as_global(dst);
  const int NE = c.NE, P = NE * c.NPA, S = c.S;
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    const int l = fast_div(i, NE, rne), k = i - l * NE;
    g[(unsigned)i] = (k == 0) ? d.F1[l] : d.F[l * S + k + 1];
  }
}

// Rows of one pass of the sweeps: a thread's rows go in passes of at most
// kPassRows, each ending in a barrier before its rows are rewritten, so the
// registers that hold a pass's results do not grow with R.
constexpr int kPassRows = 8;

// DRIFTE on the plane (src/ModRamDrift.f90:333-369).  F(1),F(0) relativistic
// ghosts from F(2), once per row; F(NE+1)=F(NE+2)=0; K=1 is not updated.
// Each thread owns one K, so the per-K factors sit in registers and the CFL
// maximum of its cells is one value.  The per-L terms are in d.cf.  Rows
// are independent, so a pass rewrites its rows while the next pass reads
// others.
//
// COUL: COULEN instead (src/ModRamCoul.f90:154-217), the same limited
// upwind sweep with CccolE(K) = (COULE+COULI)(S,K,L)*XNE(L), its own ghost
// factors (the square roots of the inverse ratios of DRIFTE's) and rows
// L=2..NPA only.
//
// TAPE: also write each updated cell's clamp decision (1: the forward
// clamped it) to cb (this plane's bytes, [L][K]), for the adjoint.
template <int MAXC, bool COUL = false, bool TAPE = false>
__device__ void plane_e(const Tab &t, const FwdCtx &c, const PlaneFwd &d, bool rec, unsigned char *cb = nullptr) {
  const int NE = c.NE, NPA = c.NPA, S = c.S;
  if (COUL) rec = false;
  {
    const double *rel = COUL ? t.crel + 5 * c.s : d.pc + PC_REL;
    for (int l = threadIdx.x; l < NPA; l += blockDim.x) {
      double *row = d.F + l * S;
      const double g1 = row[2] * rel[0] / rel[1] * rel[2];
      row[1] = g1;
      row[0] = g1 * rel[3] / rel[0] * rel[4];
    }
    if (rec)
      for (int k = threadIdx.x; k < NE; k += blockDim.x) d.cmx[k] = 0ull;
  }
  __syncthreads();
  const double *k8 = d.pc + PC_KIJ;
  const double drd1 = uniform(k8[KIJ_DRD1]), dpd1 = uniform(k8[KIJ_DPD1]);
  const double bij = uniform(k8[KIJ_BIJ]), rbij = uniform(k8[KIJ_RBIJ]);
  const double rlz = uniform(c.rlz), rrlz = uniform(c.rrlz), beta = uniform(c.beta);
  const int k = c.ke;
  // QS = +-1 folded into the prefix: the chain of correctly rounded
  // quotients is odd, so the sign comes out exactly as the reference's /QS
  const double ekp = d.kt[KT_EKP * NE + k] * c.qs, edot = d.kt[KT_EDOT * NE + k];
  const double de = d.kt[KT_DE * NE + k], rde = d.kt[KT_RDE * NE + k];
  const double we = d.kt[KT_WE * NE + k], rwe = d.kt[KT_RWE * NE + k];
  const bool top = k + 1 >= NE;  // its upper term goes to a lane that does not use it
  const double we1 = top ? 1.0 : d.kt[KT_WE * NE + k + 1], rwe1 = top ? 1.0 : d.kt[KT_RWE * NE + k + 1];
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  const int l0 = fresh(c.l0);
  double cmax = 0.0;
#pragma unroll 1
  for (int p0 = 0; p0 < MAXC; p0 += kPassRows) {
    if (p0 >= c.R) break;
    // the terms lane 63 hands to the next wave's lane 0, per pass and wave
    double *xw = d.X + ((p0 / kPassRows) * nw + (threadIdx.x >> 6)) * kPassRows;
    double v[kPassRows];
#pragma unroll
    for (int n = 0; n < kPassRows; ++n) {
      if (p0 + n >= c.nr) break;
      const int l = l0 + p0 + n;
      const double *p = d.F + l * S + k;  // positions K-1, K, K+1, K+2
      const double ce = COUL ? t.cc[((size_t)c.s * NPA + l) * NE + k] * d.cx[l]
                             : c_energy(d.cf + l * 8, ekp, edot, drd1, dpd1, bij, rbij, rlz, rrlz);
      if (!COUL) cmax = fmax(cmax, fabs(ce));
      const bool neg = ce < 0;
      const double f0 = p[1], fp1 = p[2];
      const double a = p[neg ? 3 : 0];
      const double fn = neg ? a : f0, fnm1 = neg ? fp1 : a;
      const double fb = face_value_sel(f0, fp1, fn, fnm1, ce, div_rn(ce, de, rde), beta);
      const double h = div_rn(ce, we1, rwe1) * fb;
      // the lower neighbour is lane-1 of the same chunk (same rows), or for
      // lane 0 the last lane of the wave before, whose term comes after the
      // barrier
      v[n] = settle((f0 - div_rn(ce, we, rwe) * fb) + from_lane_below(h));
      if (lane == 63) xw[n] = h;
    }
    __syncthreads();
    if (k >= 1) {
      const bool first = lane == 0;  // its lower neighbour is the last lane of the wave before
#pragma unroll
      for (int n = 0; n < kPassRows; ++n) {
        const int j = p0 + n;
        if (j >= c.nr) break;
        double w = v[n];
        if (first) w = w + xw[n - kPassRows];
        if (TAPE) cb[(l0 + j) * NE + k] = (w < 0) ? 1 : 0;
        if (w < 0) w = 1E-15;
        if (!COUL || l0 + j >= 1) c.col[(l0 + j) * S] = w;
      }
    }
  }
  if (rec) {
    if (c.nr > 0) atomic_max_pos(&d.cmx[k], fmax(cmax, 1E-10));
    __syncthreads();
    unsigned long long *g = t.cflmax + (size_t)c.s * (2 + NE + NPA);
    for (int kk = threadIdx.x; kk < NE; kk += blockDim.x) atomicMax(&g[2 + kk], d.cmx[kk]);
  }
  __syncthreads();
}

// DRIFTMU on the plane (src/ModRamDrift.f90:413-466).  F(1)=F(2) for the
// stencil; limiter faces L=2..NPA-2; CDriftMu(1)=0, FBND(1)=0,
// FBND(NPA-1)=F(NPA); top cell F2(NPA)=F2(NPA-1)*FNHS(NPA)*MU(NPA)/FNHS(NPA-1)/MU(NPA-1).
// A thread walks its rows upwards with the stencil in registers; the
// per-L terms are read from d.cm and d.lt.  The rows around the chunk are
// read before any row is rewritten; the chunk's first row takes the term
// of the chunk below after the last pass, and the top row is set last
// from the finished row below it (the same thread's, by plane_chunks_for).
// TAPE: as plane_e's.
template <int MAXC, bool TAPE = false>
__device__ void plane_mu(const Tab &t, const FwdCtx &c, const PlaneFwd &d, bool rec, unsigned char *cb = nullptr) {
  const int NE = c.NE, NPA = c.NPA;
  if (rec) {
    for (int l = threadIdx.x; l < kMuSpread * NPA; l += blockDim.x) d.cmx[NE + l] = 0ull;
    __syncthreads();
  }
  const double *k8 = d.pc + PC_KIJ;
  const double drm1 = uniform(k8[KIJ_DRM1]), dpm1 = uniform(k8[KIJ_DPM1]), db2 = uniform(k8[KIJ_DB2]);
  const double bij = uniform(k8[KIJ_BIJ]), rbij = uniform(k8[KIJ_RBIJ]);
  const double rlz = uniform(c.rlz), rrlz = uniform(c.rrlz), beta = uniform(c.beta);
  const int k = c.ke;
  const double mkp = d.kt[KT_MKP * NE + k] * c.qs;  // QS folded in, as in plane_e
  double *col = c.col;
  const int cst = c.cst, l0 = fresh(c.l0), R = c.R;
  // reads past the plane repeat the top row; they feed no kept result
  auto ld = [&](int l) { return col[min(l, NPA - 1) * cst]; };
  // stencil window F(L-1), F(L), F(L+1) of row L = l0+j+1, and the two rows
  // above the chunk
  const int nr = c.nr;
  double fm = ld(max(l0 - 1, 0)), f0 = ld(l0), f1 = ld(l0 + 1);
  const double up0 = ld(l0 + nr), up1 = ld(l0 + nr + 1);
  double hprev = -0.0, v0 = 0.0;
#pragma unroll 1
  for (int p0 = 0; p0 < MAXC; p0 += kPassRows) {
    if (p0 >= R) break;
    double v[kPassRows];
#pragma unroll
    for (int n = 0; n < kPassRows; ++n) {
      const int j = p0 + n;
      if (j >= nr) break;
      const int l = l0 + j, L1 = l + 1;
      // row j+2 of the chunk is not yet rewritten; above the chunk, the
      // rows read at the start
      const double f2 = (j + 2 < nr) ? ld(l + 2) : (j + 2 == nr ? up0 : up1);
      const double cm = c_mu(d.cm + l * 8, mkp, drm1, dpm1, db2, bij, rbij, rlz, rrlz);
      // the row's CFL maximum in kMuSpread slots (by K), so that fewer lanes
      // of a wave update one LDS address
      if (rec && l >= 1)
        atomic_max_pos(&d.cmx[NE + kMuSpread * l + (k & (kMuSpread - 1))], fmax(fabs(cm), 1E-32));
      const bool neg = cm < 0;
      const double a = neg ? f2 : (L1 <= 2 ? f0 : fm);  // F(1) = F(2)
      const double fn = neg ? a : f0, fnm1 = neg ? f1 : a;
      const int lw = (L1 <= NPA - 1) ? l + 1 : l;  // WMU(L+1) of the face's upper cell (in range)
      const double fl = face_value_sel(f0, f1, fn, fnm1, cm, div_rn(cm, d.lt[LT_DMU * NPA + l], d.lt[LT_RDMU * NPA + l]), beta);
      const double fb = (L1 == NPA - 1) ? f1 : fl;  // FBND(NPA-1) = F(NPA)
      // faces 1 and NPA carry nothing: face 1 has CDriftMu(1)=0, FBND(1)=0
      // (its term 0/WMU(2)*0 = 0); row NPA is the top cell
      const bool face = L1 >= 2 && L1 <= NPA - 1;
      const double on = div_rn(cm, d.lt[LT_WMU * NPA + l], d.lt[LT_RWMU * NPA + l]) * fb;
      const double up = div_rn(cm, d.lt[LT_WMU * NPA + lw], d.lt[LT_RWMU * NPA + lw]) * fb;
      v[n] = settle((f0 - (face ? on : 0.0)) + hprev);
      hprev = face ? up : 0.0;
      fm = f0;
      f0 = f1;
      f1 = f2;
    }
    if (p0 == 0) v0 = v[0];
    if (p0 + kPassRows >= R) d.X[threadIdx.x] = hprev;  // the chunk's last row's term, for the next chunk
    __syncthreads();
    // the pass's rows but the chunk's first and the top row
#pragma unroll
    for (int n = 0; n < kPassRows; ++n) {
      const int j = p0 + n, l = l0 + j;
      if (j >= nr) break;
      if (j >= 1 && l <= NPA - 2) {
        double w = v[n];
        if (TAPE) cb[l * NE + k] = (w < 0) ? 1 : 0;
        if (w < 0) w = 1E-15;
        col[l * cst] = w;
      }
    }
  }
  // the chunk's first row with the term of the chunk below (row 1 of the
  // plane, the first chunk's, is not updated)
  if (c.nr > 0 && c.q >= 1) {
    double w = v0 + d.X[threadIdx.x - NE];
    if (TAPE) cb[l0 * NE + k] = (w < 0) ? 1 : 0;
    if (w < 0) w = 1E-15;
    col[l0 * cst] = w;
  }
  if (c.nr > 0 && NPA - 1 < l0 + c.nr) {  // this thread holds the top row and the row below it
    const double hN = d.pc[PC_HN], hN1 = d.pc[PC_HN1];
    const double muN = d.pc[PC_MUN], muN1 = d.pc[PC_MUN1];
    col[(NPA - 1) * cst] = col[(NPA - 2) * cst] * hN * muN / hN1 / muN1;
  }
  if (rec) {
    __syncthreads();
    unsigned long long *g = t.cflmax + (size_t)c.s * (2 + NE + NPA);
    for (int ll = 1 + threadIdx.x; ll < NPA; ll += blockDim.x) {
      unsigned long long m = d.cmx[NE + kMuSpread * ll];
      for (int q = 1; q < kMuSpread; ++q) m = max(m, d.cmx[NE + kMuSpread * ll + q]);
      atomicMax(&g[2 + NE + ll], m);
    }
  }
  __syncthreads();
}

// The loss chain of ram_run between the two drift halves (src/ModRamRun.f90:
// 77-159): SUMRC, [WAVELO, SUMRC], [CEX, SUMRC], ATMOL, SUMRC, ATMOL, SUMRC,
// [CEX, SUMRC], [WAVELO, SUMRC], with the SUMRC partials of this plane block b
// in slots 0..6.  An operator a species does not take multiplies by 1, which
// leaves its cells and so its slot's sum bit for bit those of the slot before.
// WAVELO's factor is one per thread (it depends on I, J, K only), so a
// thread's partial sums after it are its sums before it times the factor
// (to rounding: exactly when the factor is 1).  Each thread rewrites only
// its own cells.
template <int MAXC>
__device__ void plane_loss(const Tab &t, const FwdCtx &c, const PlaneFwd &d, int b) {
  const int NE = c.NE, NPA = c.NPA, NT = c.NT;
  const bool cex = c.cex != 0;
  const int upa = c.upa;
  const bool sum_row = c.j <= NT - 2;
  const int k = c.ke;
  const double sigv = d.kt[KT_SIGV * NE + k], latl = d.kt[KT_LATL * NE + k];
  const double wk = d.kt[KT_WE * NE + k], ek = d.kt[KT_EKEV * NE + k];
  const double wave = d.kt[KT_WAVE * NE + k];  // 1 for K=1 and without WAVELO
  const double dts = d.pc[PC_DTS];
  const double ekw = ek * wk;
  double a[5] = {0, 0, 0, 0, 0}, tot[7];
  const int l0 = fresh(c.l0);
#pragma unroll
  for (int n = 0; n < MAXC; ++n) {
    if (n >= c.nr) break;
    const int l = l0 + n;
    const bool w_on = sum_row && k >= 1 && l >= 1;
    const bool cx = cex && k >= 1 && l >= 1;
    const bool atm = k >= 1 && l + 1 >= upa;
    double *pf = &c.col[l * c.cst];
    double f = *pf;
    // CHARGE=EXP(-ALPHA), ALPHA=10**Y*V*HDNS*DTs (src/ModRamLoss.f90:47-49);
    // ATLOS**(1/FNHS) (:501) as exp(log(ATLOS)*(1/FNHS)); WAVELO's
    // exp(-DTs/TAU) on L>=2 (src/ModRamWPI.f90:630); a factor of 1 leaves a
    // cell outside the operator's range unchanged, and a weight of 0 leaves
    // it out of the sums
    const double cxf = cx ? exp_tab(-(sigv * d.cl[2 * l] * dts), d.ex) : 1.0;
    const double af = atm ? exp_tab(latl * d.cl[2 * l + 1], d.ex) : 1.0;
    const double wf = (l == 0) ? 1.0 : wave;
    const double q = w_on ? ekw * d.lt[LT_WMU * NPA + l] : 0.0;  // EKEV*WE*WMU
    a[0] = fma(f, q, a[0]);
    f = f * wf;
    f = f * cxf;
    a[1] = fma(f, q, a[1]);
    f = f * af;
    a[2] = fma(f, q, a[2]);
    f = f * af;
    a[3] = fma(f, q, a[3]);
    f = f * cxf;
    a[4] = fma(f, q, a[4]);
    *pf = f * wf;
  }
  // slots: before WAVELO, after it, after CEX, ATMOL, ATMOL, CEX, after WAVELO
  double s7[7] = {a[0], a[0] * wave, a[1], a[2], a[3], a[4], a[4] * wave};
  block_sum_n<7>(s7, d.red, tot);
  if (threadIdx.x < 7) {
    double v = tot[0];
    for (int m = 1; m < 7; ++m)
      if (threadIdx.x == m) v = tot[m];
    t.partials[((size_t)threadIdx.x * t.ns + c.s) * t.maxblk + b] = v;
  }
}

// One of the loss operators alone on this thread's cells: CHAREXCHANGE
// (src/ModRamLoss.f90:467-475: K>=2, L>=2), ATMOL (:495-503: K>=2,
// L>=int(UPA(I))) or WAVELO (src/ModRamWPI.f90:612-633: K>=2, L>=2; the
// factor is 1 for K=1).
template <int SEQ, int MAXC>
__device__ void plane_one_loss(const FwdCtx &c, const PlaneFwd &d) {
  const int NE = c.NE, k = c.ke;
  if (k < 1) return;
  if (SEQ == SEQ_CEX && c.cex == 0) return;
#pragma unroll
  for (int n = 0; n < MAXC; ++n) {
    if (n >= c.nr) break;
    const int l = c.l0 + n;
    {
      double &f = c.col[l * c.cst];
      if constexpr (SEQ == SEQ_CEX) {
        if (l >= 1) f = f * cex_factor(d, NE, k, l, d.pc[PC_DTS]);
      } else if constexpr (SEQ == SEQ_ATM) {
        if (l + 1 >= c.upa) f = f * atm_factor(d, NE, k, l);
      } else {
        if (l >= 1) f = f * d.kt[KT_WAVE * NE + k];
      }
    }
  }
}

// ---- Coulomb collisions (src/ModRamCoul.f90) -----------------------------
// The plane's per-L terms of COULEN and COULMU in d.cx, in the reference's
// roundings: XNE(L) = NECR(I,J)*BANE(L) with the bounce factor
// BANE(L) = (1-FNIS/2/FNHS)/(1-MU^2) (:167-172; rows L >= NPA-10 take
// BANE(NPA-11), the sequential copy of :170-172), BASCNE(L) =
// NECR(I,J)*BOUNIS/2/BOUNHS (:263), FNHS(L), BOUNHS(L), MU(L) and the
// correctly rounded reciprocals of FNHS and MU.  The caller puts a barrier
// before they are read.
__device__ void plane_coul_tables(const Tab &t, const FwdCtx &c, const PlaneFwd &d) {
  const int NPA = c.NPA, I = c.i + 1, J = c.j + 1;
  const double necr = t.necr[c.i + (size_t)t.nr * c.j];
  for (int l = threadIdx.x; l < NPA; l += blockDim.x) {
    const int L = l + 1, lb = min(L, NPA - 11);
    const double mub = t.mu[lb - 1];
    const double bane = (1. - GEO3(t.fnis, I, J, lb) / 2. / GEO3(t.fnhs, I, J, lb)) / (1. - mub * mub);
    const double bh = GEO3(t.bounhs, I, J, L);
    d.cx[l] = necr * bane;
    d.cx[NPA + l] = necr * GEO3(t.bounis, I, J, L) / 2. / bh;
    const double fh = GEO3(t.fnhs, I, J, L), mu = t.mu[l];
    d.cx[2 * NPA + l] = fh;
    d.cx[3 * NPA + l] = bh;
    d.cx[4 * NPA + l] = mu;
    d.cx[5 * NPA + l] = 1.0 / fh;  // RN(1/FNHS) and RN(1/MU) for div_rn (1/MU(1) is not used)
    d.cx[6 * NPA + l] = 1.0 / mu;
  }
}

// COULMU on the plane (src/ModRamCoul.f90:255-292): for every K = 2..NE the
// implicit pitch-angle step of the line L = 1..NPA, solved by the
// reference's recurrences in its order (RK(1)=0, RL(1)=-1; RK, RL forward
// over L = 2..NPA-1; back substitution; F2(NPA) = F2(NPA-1); the solution
// is F2/FNHS/MU, scaled back).  A line's recurrences are sequential, so one
// thread owns a whole column: threads 1..NE-1 (the first chunk's), with RK
// in the column's own cells and RL in d.cx.  Negative cells become 1e-15
// when clamp (TimeRamElapsed > 0).  The quotients are correctly rounded
// (div_rn): by FNHS and MU from their tabulated reciprocals, and RK and RL
// from one reciprocal of DENOM per row, so that a row holds one full
// division.  ATA and GTA of the rows ahead are read four rows at a time, the
// next four in flight while these are used.
__device__ void plane_coulmu(const Tab &t, const FwdCtx &c, const PlaneFwd &d, bool clamp) {
  const int NE = c.NE, NPA = c.NPA, S = c.S;
  const double *xb = d.cx + NPA, *fh = d.cx + 2 * NPA, *bh = d.cx + 3 * NPA, *mu = d.cx + 4 * NPA;
  const double *rfh = d.cx + 5 * NPA, *rmu = d.cx + 6 * NPA;
  const int last = NPA - 2;  // row of L = NPA-1
  {
    // the block first forms each cell's GN (into its RL slot, which the
    // recurrence reads before it writes RL there) and RP (in place); they do
    // not depend on the recurrence, which is left AN, DENOM, RK and RL
    const double2 *cg = t.cag + (size_t)c.s * NPA * NE;
    const int w = NE - 1, ncell = w * last;
    for (int e = threadIdx.x; e < ncell; e += blockDim.x) {
      const int q = e / w, l = 1 + q, kk = 1 + (e - q * w);
      const double f = fh[l], rf = rfh[l];
      d.cx[7 * NPA + l * NE + kk] = div_rn(cg[l * NE + kk].y * xb[l - 1], f, rf) * bh[l - 1];
      double *cl = d.F + kk + 1 + l * S;
      *cl = div_rn(div_rn(*cl, f, rf), mu[l], rmu[l]);
    }
  }
  __syncthreads();
  const int k = threadIdx.x;
  if (k >= 1 && k < NE) {
    double *cell = d.F + k + 1;  // cell (K, row l) is cell[l*S]; RP there on entry
    double *rlv = d.cx + 7 * NPA + k;  // GN, then RL, of row l at rlv[l*NE]
    const gptr<const double> ag = as_global((const double *)(t.cag + (size_t)c.s * NPA * NE + k));
    double rk = 0.0, rl = -1.0;
    cell[0] = 0.0;
    rlv[0] = -1.0;
    constexpr int B = 4;
    double2 cur[B], nxt[B];
#pragma unroll
    for (int m = 0; m < B; ++m) cur[m] = (1 + m <= last) ? ld_d2(ag, (unsigned)((1 + m) * NE)) : make_double2(0.0, 0.0);
#pragma unroll 1
    for (int l0 = 1; l0 <= last; l0 += B) {
#pragma unroll
      for (int m = 0; m < B; ++m)
        nxt[m] = (l0 + B + m <= last) ? ld_d2(ag, (unsigned)((l0 + B + m) * NE)) : make_double2(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < B; ++m) {
        const int l = l0 + m;
        if (l > last) break;
        const double an = div_rn(cur[m].x * xb[l], fh[l], rfh[l]) * bh[l];
        const double gn = rlv[l * NE];
        const double bn = an + gn;
        const double rp = cell[l * S];
        const double den = bn + gn * rl + 1;
        const double rden = 1.0 / den;
        rk = div_rn(rp + gn * rk, den, rden);
        rl = div_rn(-an, den, rden);
        cell[l * S] = rk;
        rlv[l * NE] = rl;
      }
#pragma unroll
      for (int m = 0; m < B; ++m) cur[m] = nxt[m];
    }
    double f = rk / (1 + rl);
    cell[last * S] = f;
#pragma unroll 4
    for (int l = last - 1; l >= 0; --l) {
      f = cell[l * S] - rlv[l * NE] * f;
      cell[l * S] = f;
    }
    cell[(NPA - 1) * S] = cell[last * S];
#pragma unroll 4
    for (int l = 0; l < NPA; ++l) {
      double v = cell[l * S] * fh[l] * mu[l];
      if (clamp && v < 0.) v = 1E-15;
      cell[l * S] = v;
    }
  }
  __syncthreads();
}

// This block's SUMRC partial of the plane (I>=2, K>=2, L>=2, J<=NT-1,
// src/ModRamRun.f90:246-255) into slot `slot`, with the weights of the loss
// chain's sums.
template <int MAXC>
__device__ void plane_sum_slot(const Tab &t, const FwdCtx &c, const PlaneFwd &d, int b, int slot) {
  const int NE = c.NE, NPA = c.NPA, k = c.ke;
  double a[1] = {0.0}, tot[1];
  if (c.j <= c.NT - 2 && k >= 1) {
    const double ekw = d.kt[KT_EKEV * NE + k] * d.kt[KT_WE * NE + k];
#pragma unroll
    for (int n = 0; n < MAXC; ++n) {
      if (n >= c.nr) break;
      const int l = c.l0 + n;
      if (l >= 1) a[0] = fma(c.col[l * c.cst], ekw * d.lt[LT_WMU * NPA + l], a[0]);
    }
  }
  block_sum_n<1>(a, d.red, tot);
  if (threadIdx.x == 0) t.partials[((size_t)slot * t.ns + c.s) * t.maxblk + b] = tot[0];
}

// MAXC (8 or 16) bounds the rows a thread owns; the sweeps take them in
// passes of kPassRows, so both bounds fit the register budget of 4 waves
// per SIMD with blocks of up to 1024 threads.
template <int SEQ, int MAXC>
__global__ void __launch_bounds__(kPlaneMaxThreads, 4) k_plane(Launch L, int rec) {
  extern __shared__ double sm[];
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int NE = t.ne, NPA = t.npa, NT = t.nt;
    if (u.b >= blocks_kl(t)) {  // out of place: copy plane (I=1, J) of the unswept row
      const size_t off = (size_t)u.s * f2_species_stride(t) + (size_t)(u.b - blocks_kl(t)) * NE * NPA;
      const double *src = launch_src(L, t) + off;
      double *dst = launch_dst(L, t) + off;
      block_copy<8>(dst, src, NE * NPA);
      continue;
    }
    constexpr bool need_loss = SEQ == SEQ_CEX || SEQ == SEQ_ATM || SEQ == SEQ_STEP || SEQ == SEQ_LOSS ||
                               SEQ == SEQ_TAPE || SEQ == SEQ_WAVE || SEQ == SEQ_STEPC;
    constexpr bool need_kl = SEQ == SEQ_E || SEQ == SEQ_MU || SEQ == SEQ_STEP || SEQ == SEQ_TAPE ||
                             SEQ == SEQ_STEPC;
    FwdCtx c;
    PlaneFwd d;
    fwd_setup<MAXC>(t, u, sm, launch_src(L, t), need_loss, need_kl, c, d);
    const size_t po = (size_t)c.s * f2_species_stride(t) + c.ij * ((size_t)NE * NPA);
    double *gF = launch_dst(L, t) + po;
    const bool inside = d.pc[PC_KIJ + KIJ_MG] == 0.0;
    const bool rec_here = (rec & 1) && inside;  // bit 0: record the CFL maxima
    const bool clamp = (rec & 2) != 0;          // bit 1: COULMU's clamp (TimeRamElapsed > 0)
    if constexpr (seq_coul(SEQ)) {
      plane_coul_tables(t, c, d);
      __syncthreads();
    }

    if constexpr (SEQ == SEQ_E) {
      plane_e<MAXC>(t, c, d, rec_here);
    } else if constexpr (SEQ == SEQ_MU) {
      plane_mu<MAXC>(t, c, d, rec_here);
    } else if constexpr (SEQ == SEQ_CEX || SEQ == SEQ_ATM || SEQ == SEQ_WAVE) {
      plane_one_loss<SEQ, MAXC>(c, d);
    } else if constexpr (SEQ == SEQ_SUM) {
      // SUMRC over I>=2, K>=2, L>=2, J<=NT-1 (src/ModRamRun.f90:246-255), slot 0
      double a[1] = {0.0}, tot[1];
      const int k = c.ke;
      if (c.j <= NT - 2 && k >= 1) {
#pragma unroll
        for (int n = 0; n < MAXC; ++n) {
          if (n >= c.nr) break;
          const int l = c.l0 + n;
          if (l >= 1)
            a[0] += d.kt[KT_EKEV * NE + k] * (c.col[l * c.cst] * d.kt[KT_WE * NE + k] * d.lt[LT_WMU * NPA + l]);
        }
      }
      block_sum_n<1>(a, d.red, tot);
      if (threadIdx.x == 0) t.partials[((size_t)0 * t.ns + c.s) * t.maxblk + u.b] = tot[0];
    } else if constexpr (SEQ == SEQ_COULEN) {
      plane_e<MAXC, true>(t, c, d, false);
    } else if constexpr (SEQ == SEQ_COULMU) {
      plane_coulmu(t, c, d, clamp);
    } else if constexpr (SEQ == SEQ_STEPC) {
      // src/ModRamRun.f90:72-89 and 160-172: the Coulomb operators sit
      // between the drift halves and the loss chain, with a SUMRC after each
      plane_e<MAXC>(t, c, d, rec_here);
      plane_mu<MAXC>(t, c, d, rec_here);
      plane_sum_slot<MAXC>(t, c, d, u.b, kSumCoul);
      plane_e<MAXC, true>(t, c, d, false);
      plane_sum_slot<MAXC>(t, c, d, u.b, kSumCoul + 1);
      plane_coulmu(t, c, d, clamp);
      plane_loss<MAXC>(t, c, d, u.b);
      plane_coulmu(t, c, d, clamp);
      plane_sum_slot<MAXC>(t, c, d, u.b, kSumCoul + 2);
      plane_e<MAXC, true>(t, c, d, false);
      plane_sum_slot<MAXC>(t, c, d, u.b, kSumCoul + 3);
      plane_mu<MAXC>(t, c, d, false);
      plane_e<MAXC>(t, c, d, false);
    } else if constexpr (SEQ == SEQ_LOSS) {
      plane_loss<MAXC>(t, c, d, u.b);
    } else if constexpr (SEQ == SEQ_TAPE) {
      // src is tape slot 0 (this phase's input); phase m's output, the input
      // of phase m+1, goes to slot m+1, the last to dst.  A thread stores
      // its own cells, which only it writes.
      // The clamp decisions of the four drift phases go to the bytes of the
      // tape slot 7 after slot 0, one byte plane per phase.
      const size_t ns = f2_elems(t);
      double *slot = const_cast<double *>(launch_src(L, t)) + po;
      unsigned char *cb = reinterpret_cast<unsigned char *>(const_cast<double *>(launch_src(L, t)) + 7 * ns) + po;
      // each phase ends with a barrier, so its output is whole in LDS; the
      // loss chain rewrites cells at once, so a barrier follows the store
      // before it (the later phases rewrite cells only after barriers)
      plane_e<MAXC, false, true>(t, c, d, rec_here, cb);
      fwd_store(c, d, t.rne, slot + ns);
      plane_mu<MAXC, true>(t, c, d, rec_here, cb + ns);
      fwd_store(c, d, t.rne, slot + 2 * ns);
      __syncthreads();
      plane_loss<MAXC>(t, c, d, u.b);
      fwd_store(c, d, t.rne, slot + 3 * ns);
      plane_mu<MAXC, true>(t, c, d, false, cb + 2 * ns);
      fwd_store(c, d, t.rne, slot + 4 * ns);
      plane_e<MAXC, false, true>(t, c, d, false, cb + 3 * ns);
    } else {  // SEQ_STEP
      plane_e<MAXC>(t, c, d, rec_here);
      plane_mu<MAXC>(t, c, d, rec_here);
      plane_loss<MAXC>(t, c, d, u.b);
      plane_mu<MAXC>(t, c, d, false);
      plane_e<MAXC>(t, c, d, false);
    }

    if constexpr (SEQ != SEQ_SUM) {
      if constexpr (SEQ == SEQ_CEX || SEQ == SEQ_ATM || SEQ == SEQ_WAVE) __syncthreads();  // cells rewritten at once
      fwd_store(c, d, t.rne, gF);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// DtDrift*(S) from the CFL maxima (cflmask bits 1 R, 2 P, 4 E, 8 Mu): min
// over cells of FracCFL*DTs*w/ctemp equals FracCFL*DTs*w / max(ctemp) because
// correctly rounded division is monotone; initial values 1e5 (R,P) and 1e4
// (E,Mu) as the reference's (:115,223,308,404).  The maxima are cleared after
// use for the next recording.
// SUMRC bookkeeping (src/ModRamRun.f90:77-175,239-257), summode 1: one SUMRC
// from slot 0, or with kSumFromRadial from the radial slot (SETRC:=S,
// ELORC=old-S, and with an accumulator in the bits at kSumAccShift that
// LS* += ELORC); summode 2: the full ram_run chain from
// the loss chain's slots 0..6 and the radial slot 7, with Coulomb collisions
// also the slots 8..11 of their operators.  One block per (problem, species).
constexpr int kFinNV = 2 + kSumSlots;  // the DtDriftE and DtDriftMu minima, then the SUMRC slot sums
__device__ inline void finalize_unit(const Launch &L, int vb, int cflmask, int summode, double *red) {
  constexpr int NV = kFinNV;
  {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int s = u.s, tid = threadIdx.x, T = blockDim.x;
    const int NE = t.ne, NPA = t.npa;
    const double a = t.fraccfl * t.dts;
    unsigned long long *m = t.cflmax + (size_t)s * (2 + NE + NPA);
    // every load of the block first: thread 0's scalars, the CFL maxima, the
    // plane and radial blocks' partial sums
    double e0[4] = {0.0, 0.0, 0.0, 0.0}, lswae = 0.0, lscoe = 0.0, lscsc = 0.0;
    unsigned long long m01[2] = {0ull, 0ull};
    int cex = 0, wpi = 0;
    const bool coul = t.coul != 0;
    if (tid == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) e0[q] = t.energy[s * 4 + q];
      lswae = t.lsx[s * 3];
      lscoe = t.lsx[s * 3 + 1];
      lscsc = t.lsx[s * 3 + 2];
      m01[0] = m[0];
      m01[1] = m[1];
      cex = t.cex[s];
      wpi = t.wpi[s];
    }
    double v[NV];
    v[0] = 10000.0;
    v[1] = 10000.0;
#pragma unroll
    for (int q = 2; q < NV; ++q) v[q] = 0.0;
    if (cflmask & 4)
      for (int k = tid; k < NE; k += T) {
        const double c = __longlong_as_double((long long)m[2 + k]);
        if (c > 0) v[0] = fmin(v[0], a * t.de[k] / c);
      }
    if (cflmask & 8)
      for (int l = 1 + tid; l < NPA; l += T) {
        const double c = __longlong_as_double((long long)m[2 + NE + l]);
        if (c > 0) v[1] = fmin(v[1], a * t.dmu[l] / c);
      }
    const int mode = summode & kSumModeMask;
    if (mode != 0) {
      // slots 0..6 (and with Coulomb collisions 8..11) from the plane
      // blocks, slot 7 from the R blocks; a single SUMRC reads slot 0, or
      // the radial slot
      const int nbk = blocks_kl(t), nbr = blocks_r(t);
      const double *pp = t.partials + (size_t)s * t.maxblk;
      const size_t slot = (size_t)t.ns * t.maxblk;
      if (mode == 1 && (summode & kSumFromRadial)) {
        for (int b = tid; b < nbr; b += T) v[2] += pp[kSumRadial * slot + b];
      } else {
        const int nslot = (mode == 1) ? 1 : kSumRadial;
        for (int b = tid; b < nbk; b += T)
          for (int sl = 0; sl < nslot; ++sl) v[2 + sl] += pp[sl * slot + b];
      }
      if (mode == 2) {
        for (int b = tid; b < nbr; b += T) v[2 + kSumRadial] += pp[kSumRadial * slot + b];
        if (coul)
          for (int b = tid; b < nbk; b += T)
#pragma unroll
            for (int sl = kSumCoul; sl < kSumCoul + 4; ++sl) v[2 + sl] += pp[sl * slot + b];
      }
    }
    // one block reduction (fixed order): minima of v[0..1], sums of the rest
    for (int off = 32; off > 0; off >>= 1) {
      v[0] = fmin(v[0], __shfl_down(v[0], off, 64));
      v[1] = fmin(v[1], __shfl_down(v[1], off, 64));
#pragma unroll
      for (int q = 2; q < NV; ++q) v[q] += __shfl_down(v[q], off, 64);
    }
    const int lane = tid & 63, w = tid >> 6, nw = (T + 63) >> 6;
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < NV; ++q) red[w * NV + q] = v[q];
    __syncthreads();
    if (tid == 0) {
      double r[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) r[q] = red[q];
      for (int ww = 1; ww < nw; ++ww) {
        r[0] = fmin(r[0], red[ww * NV]);
        r[1] = fmin(r[1], red[ww * NV + 1]);
#pragma unroll
        for (int q = 2; q < NV; ++q) r[q] += red[ww * NV + q];
      }
      if (cflmask & 1) {
        const double c = __longlong_as_double((long long)m01[0]);
        t.dtdrift[s * 4 + 0] = (c > 0) ? fmin(100000.0, a / c) : 100000.0;
      }
      if (cflmask & 2) {
        const double c = __longlong_as_double((long long)m01[1]);
        t.dtdrift[s * 4 + 1] = (c > 0) ? fmin(100000.0, a / c) : 100000.0;
      }
      if (cflmask & 4) t.dtdrift[s * 4 + 2] = r[0];
      if (cflmask & 8) t.dtdrift[s * 4 + 3] = r[1];
      if (mode != 0) {
        const double *S = r + 2;
        double *e = t.energy + s * 4;
        const double old = e0[0];
        if (mode == 1) {
          const double elorc = old - S[0];
          e[0] = S[0];
          t.partials[((size_t)kSumSlots * t.ns + s) * t.maxblk] = elorc;  // ELORC
          // LS* += ELORC as ram_run books it after the operator
          const int acc = (summode >> kSumAccShift) & 15;
          if (acc >= SUM_LSDR && acc <= SUM_LSATM) e[acc] = e0[acc] + elorc;
          else if (acc == SUM_LSWAE) t.lsx[s * 3] = lswae + elorc;
          else if (acc == SUM_LSCOE) t.lsx[s * 3 + 1] = lscoe + elorc;
          else if (acc == SUM_LSCSC) t.lsx[s * 3 + 2] = lscsc + elorc;
        } else {
          // the accumulators of src/ModRamRun.f90:77-175 from consecutive
          // slot sums, each operator's only where the reference runs it
          double lsdr = e0[1], lscha = e0[2], lsatm = e0[3];
          double prev = old;
          if (coul) {  // :77-89
            lsdr += prev - S[kSumCoul];
            prev = S[kSumCoul];
            lscoe += prev - S[kSumCoul + 1];
            prev = S[kSumCoul + 1];
            lscsc += prev - S[0];
          } else {
            lsdr += prev - S[0];
          }
          prev = S[0];
          if (wpi) {
            lswae += prev - S[1];
            prev = S[1];
          }
          if (cex) {
            lscha += prev - S[2];
            prev = S[2];
          }
          lsatm += prev - S[3];
          prev = S[3];
          lsatm += prev - S[4];
          prev = S[4];
          if (cex) {
            lscha += prev - S[5];
            prev = S[5];
          }
          if (wpi) {
            lswae += prev - S[6];
            prev = S[6];
          }
          if (coul) {  // :160-167
            lscsc += prev - S[kSumCoul + 2];
            prev = S[kSumCoul + 2];
            lscoe += prev - S[kSumCoul + 3];
            prev = S[kSumCoul + 3];
          }
          lsdr += prev - S[kSumRadial];
          e[0] = S[kSumRadial];
          e[1] = lsdr;
          e[2] = lscha;
          e[3] = lsatm;
          t.lsx[s * 3] = lswae;
          if (coul) {
            t.lsx[s * 3 + 1] = lscoe;
            t.lsx[s * 3 + 2] = lscsc;
          }
        }
      }
    }
    __syncthreads();  // every read of the maxima is done
    if (cflmask)
      for (int e = tid; e < 2 + NE + NPA; e += T) m[e] = 0ull;
    __syncthreads();  // red is reused by the block's next unit
  }
}

__global__ void __launch_bounds__(kFinalizeThreads) k_finalize(Launch L, int cflmask, int summode) {
  __shared__ double red[(kFinalizeThreads / 64) * kFinNV];
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) finalize_unit(L, vb, cflmask, summode, red);
}

// The head of a step in one launch (each part on its own is a latency-bound
// launch of a few microseconds): blocks below Lc.nv form the step's
// Courant-number terms (k_coef_ij), the next Lf.nv finish the previous step
// of a schedule (k_finalize, one block per species; none when Lf.nv = 0),
// the last Lg.nv form the step's ghost carry (k_ghost).  The parts touch
// disjoint data.  Blocks of kFinalizeThreads, as k_finalize's and
// k_ghost's, so the sums and scans are theirs.  Dynamic LDS: the VT factors
// (NR+1+NT), the ghost's face row (4*NT), then (stage_q1) the face-NR row of
// q1 (NPA*NT).
__global__ void __launch_bounds__(kFinalizeThreads) k_step_head(Launch Lc, const double *kp, const double *bs,
                                                                Launch Lf, int cflmask, int summode, Launch Lg,
                                                                int nvt, int stage_q1) {
  extern __shared__ double dyn[];
  __shared__ double red[(kFinalizeThreads / 64) * kFinNV];
  __shared__ int s_last[kMaxPlaneCells];
  __shared__ int s_wave[kFinalizeThreads / 64];
  const double avs = kp ? vs_avs(*kp) : 0.0;
  int b = blockIdx.x;
  if (b < Lc.nv) {
    coef_ij_unit(Lc, b, kp, bs, avs, dyn);
  } else if ((b -= Lc.nv) < Lf.nv) {
    finalize_unit(Lf, b, cflmask, summode, red);
  } else {
    const int nt = Lg.tabs[0].nt;
    ghost_unit(Lg, b - Lf.nv, kp, avs, dyn, dyn + nvt, s_last, s_wave, stage_q1 ? dyn + nvt + 4 * nt : nullptr);
  }
}

// VT(I,J)=AVS*(LZ(I)*RE)**2*SIN(PHI(J)-PHIOFS), src/ModRamRun.f90:45-54
// Volland-Stern VT(I,J) from Kp (*kp_ptr, or kp_val); with bs_src, also the
// step's boundary-flux scales bs_src[0..ns) into bs_dst (one launch per step
// of a schedule instead of a kernel and a copy).
__global__ void k_vs(double *vt, const double *lz, const double *phi, const double *kp_ptr,
                     double kp_val, int nr1, int nt, const double *bs_src, double *bs_dst, int ns) {
  if (bs_src && blockIdx.x == 0)
    for (int s = threadIdx.x; s < ns; s += blockDim.x) bs_dst[s] = bs_src[s];
  const double avs = vs_avs(kp_ptr ? *kp_ptr : kp_val);
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nr1 * nt; n += gridDim.x * blockDim.x) {
    const int i = n % nr1, j = n / nr1;
    vt[n] = vs_vt(avs, lz[i], phi[j]);
  }
}

// One step of an ensemble's schedule: member m gets VT for kp[step*M+m] (the
// formula of k_vs) and boundary scales bscale[step*stride + bs_off[m] + s].
// One unit per member.
__global__ void k_sched(Launch L, const double *kp, const double *bscale, const int *bs_off, const int *step,
                        int nmembers, int stride_bs) {
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int m = (int)(u.t - L.tabs);
    const int n = *step;
    const double avs = vs_avs(kp[(size_t)n * nmembers + m]);
    if (threadIdx.x == 0) *t.kpnow = kp[(size_t)n * nmembers + m];  // WAVELO's KP of the step
    double *vt = const_cast<double *>(t.vt);
    const int nr1 = t.nr + 1;
    for (int x = threadIdx.x; x < nr1 * t.nt; x += blockDim.x) {
      const int i = x % nr1, j = x / nr1;
      vt[x] = vs_vt(avs, t.lz[i], t.phi[j]);
    }
    if (bscale)
      for (int s = threadIdx.x; s < t.ns; s += blockDim.x)
        t.bscale[s] = bscale[(size_t)n * stride_bs + bs_off[m] + s];
  }
}

__global__ void k_tick(int *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) step[0] = step[0] + 1;
}

__global__ void k_setstep(int *step, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) step[0] = v;
}

// Whole-F2 copy of every problem of an ensemble launch, source to
// destination as the launch names them (tape slots or F2); each unit is one
// problem, its nb blocks striding over it.
// ---------------------------------------------------------------------------
// Implicit pitch-angle diffusion of the species of job (blockIdx.y picks
// one): WPADIF (src/ModRamWPI.f90:643-714) and FLCscatter
// (src/ModRamLoss.f90:513-575) are the same solve with different
// coefficients D = coef (NR,NT,NE,NPA in F2's device layout [I][J][L][K]).
// One thread per (I>=2, J, K>=2) line, blocks of one wave,
// consecutive threads on consecutive K, so each row of the lines of a wave
// is one contiguous run.  In the reference's order: F = F2/FACMU with
// FACMU(L) = FNHS(I,J,L)*MU(L); RK(1)=0, RL(1)=-1; for L = 2..NPA-1
// AN = D(L)/DMU(L)*DTs/FACMU(L)/WMU(L), GN the same of D(L-1)/DMU(L-1),
// DENOM = AN+GN+GN*RL(L-1)+1, RK = (F+GN*RK(L-1))/DENOM, RL = -AN/DENOM;
// back substitution, F2(NPA) = F2(NPA-1), F2 = F*FACMU.  Every quotient is
// correctly rounded: by DMU and WMU from their tabulated reciprocals, by
// FACMU and DENOM from one reciprocal each per row (div_rn).  RK and RL of
// the line stay in LDS (2*NPA doubles per thread) for the substitution; F,
// D and FNHS of the next eight rows are in flight while eight are solved:
// the recurrence is serial, so its speed is the latency of each row.
constexpr int kPitchAhead = 8;
__global__ void __launch_bounds__(kPitchThreads) k_pitch_diffusion(const Tab *tp, PitchJob job) {
  extern __shared__ double sm[];
  const Tab &t = *tp;
  const int s = job.s[blockIdx.y];
  const double *__restrict__ coef = job.coef[blockIdx.y];
  const int NE = t.ne, NPA = t.npa, NT = t.nt, NR = t.nr, nk = NE - 1;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (NR - 1) * NT * nk) return;
  const int ij = n / nk, k = 1 + (n - ij * nk);
  const int i = 1 + ij / NT, j = ij - (i - 1) * NT;
  const size_t po = ((size_t)i * NT + j) * ((size_t)NE * NPA) + k;
  const gptr<double> f = as_global(t.f2 + (size_t)s * f2_species_stride(t) + po);
  const gptr<const double> d = as_global(coef + po);
  const gptr<const double> fh = as_global(&GEO3(t.fnhs, i + 1, j + 1, 1));
  const size_t fst = (size_t)(NR + 1) * NT;  // FNHS's stride along L
  double *rkv = sm + threadIdx.x, *rlv = sm + (size_t)NPA * kPitchThreads + threadIdx.x;  // row l at [l*64]
  const double dts = t.dts;
  const int last = NPA - 2;                  // row of L = NPA-1
  constexpr int B = kPitchAhead;
  double cf[B], cd[B], ch[B], nf[B], nd[B], nh[B];
#pragma unroll
  for (int m = 0; m < B; ++m) {
    const int l = min(1 + m, last);
    cf[m] = f[(unsigned)(l * NE)];
    cd[m] = d[(unsigned)(l * NE)];
    ch[m] = fh[l * fst];
  }
  double pprev = div_rn(d[0], t.dmu[0], t.rdmu[0]);  // D(1)/DMU(1)
  double rk = 0.0, rl = -1.0;
  rkv[0] = 0.0;
  rlv[0] = -1.0;
#pragma unroll 1
  for (int l0 = 1; l0 <= last; l0 += B) {
#pragma unroll
    for (int m = 0; m < B; ++m) {
      const int l = min(l0 + B + m, last);
      nf[m] = f[(unsigned)(l * NE)];
      nd[m] = d[(unsigned)(l * NE)];
      nh[m] = fh[l * fst];
    }
#pragma unroll
    for (int m = 0; m < B; ++m) {
      const int l = l0 + m;
      if (l > last) break;
      const double fac = ch[m] * t.mu[l], rfac = 1.0 / fac;
      const double wmu = t.wmu[l], rwmu = t.rwmu[l];
      const double p = div_rn(cd[m], t.dmu[l], t.rdmu[l]);
      const double an = div_rn(div_rn(p * dts, fac, rfac), wmu, rwmu);
      const double gn = div_rn(div_rn(pprev * dts, fac, rfac), wmu, rwmu);
      const double bn = an + gn;
      const double rp = div_rn(cf[m], fac, rfac);
      const double den = bn + gn * rl + 1;
      const double rden = 1.0 / den;
      rk = div_rn(rp + gn * rk, den, rden);
      rl = div_rn(-an, den, rden);
      rkv[l * kPitchThreads] = rk;
      rlv[l * kPitchThreads] = rl;
      pprev = p;
    }
#pragma unroll
    for (int m = 0; m < B; ++m) {
      cf[m] = nf[m];
      cd[m] = nd[m];
      ch[m] = nh[m];
    }
  }
  // back substitution on the unscaled values, each row stored scaled
  double v = rk / (1 + rl);
  f[(unsigned)(last * NE)] = v * (fh[last * fst] * t.mu[last]);
  f[(unsigned)((NPA - 1) * NE)] = v * (fh[(NPA - 1) * fst] * t.mu[NPA - 1]);
#pragma unroll 8
  for (int l = last - 1; l >= 0; --l) {
    v = rkv[l * kPitchThreads] - rlv[l * kPitchThreads] * v;
    f[(unsigned)(l * NE)] = v * (fh[l * fst] * t.mu[l]);
  }
}

__global__ void k_copy_ens(Launch L) {
  for (int vb = blockIdx.x; vb < L.nv; vb += gridDim.x) {
    const Unit u = unit_of(L, vb);
    const Tab &t = *u.t;
    const int nb = L.vmap ? L.jobs[L.vmap[vb] >> kVmapShift].nb : L.nb;
    const size_t n = (size_t)t.ns * f2_species_stride(t);
    const double *src = launch_src(L, t);
    double *dst = launch_dst(L, t);
    for (size_t m = (size_t)u.b * blockDim.x + threadIdx.x; m < n; m += (size_t)nb * blockDim.x) dst[m] = src[m];
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

// explicit instantiations used by the host code
template __global__ void k_r<0>(Launch);
template __global__ void k_r<1>(Launch);
template __global__ void k_r<2>(Launch);
template __global__ void k_p<0>(Launch);
template __global__ void k_p<1>(Launch);
template __global__ void k_plane<SEQ_WAVE, 8>(Launch, int);
template __global__ void k_plane<SEQ_WAVE, 16>(Launch, int);
template __global__ void k_plane<SEQ_E, 8>(Launch, int);
template __global__ void k_plane<SEQ_MU, 8>(Launch, int);
template __global__ void k_plane<SEQ_CEX, 8>(Launch, int);
template __global__ void k_plane<SEQ_ATM, 8>(Launch, int);
template __global__ void k_plane<SEQ_SUM, 8>(Launch, int);
template __global__ void k_plane<SEQ_STEP, 8>(Launch, int);
template __global__ void k_plane<SEQ_LOSS, 8>(Launch, int);
template __global__ void k_plane<SEQ_TAPE, 8>(Launch, int);
template __global__ void k_plane<SEQ_E, 16>(Launch, int);
template __global__ void k_plane<SEQ_MU, 16>(Launch, int);
template __global__ void k_plane<SEQ_CEX, 16>(Launch, int);
template __global__ void k_plane<SEQ_ATM, 16>(Launch, int);
template __global__ void k_plane<SEQ_SUM, 16>(Launch, int);
template __global__ void k_plane<SEQ_STEP, 16>(Launch, int);
template __global__ void k_plane<SEQ_LOSS, 16>(Launch, int);
template __global__ void k_plane<SEQ_TAPE, 16>(Launch, int);
template __global__ void k_plane<SEQ_COULEN, 8>(Launch, int);
template __global__ void k_plane<SEQ_COULEN, 16>(Launch, int);
template __global__ void k_plane<SEQ_COULMU, 8>(Launch, int);
template __global__ void k_plane<SEQ_COULMU, 16>(Launch, int);
template __global__ void k_plane<SEQ_STEPC, 8>(Launch, int);
template __global__ void k_plane<SEQ_STEPC, 16>(Launch, int);

}  // namespace ramgpu
