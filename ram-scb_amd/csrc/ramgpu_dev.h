// ramgpu_dev.h -- device-side data of the RAM split step (gfx950).
//
// F2 is kept on the device per species in a [I][J][L][K] layout (K fastest):
// the (K,L) plane of one (I,J) cell is one contiguous run of NE*NPA doubles,
// so the energy and pitch-angle sweeps (plus the pointwise loss operators
// that sit between them) work on whole contiguous planes, and the radial and
// local-time sweeps work on tiles of KW consecutive (L,K) positions across
// the full (I,J) plane.
//
// Coefficient tables hold the reference's intermediate quantities (its
// GPA, GPR1.., DRD1.., CMUDOT.. terms) computed in the reference's expression
// order, plus correctly rounded reciprocals of every divisor that is reused
// across cells, so that each Courant number is rebuilt per cell with exactly
// the reference's sequence of roundings.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace ramgpu {

constexpr int kMaxSpecies = 8;
constexpr int kMaxOps = 16;

// sweep-program op codes
enum Op : int {
  OP_NONE = 0,
  OP_R = 1,     // DRIFTR   (IJ kernel); arg=1 records the CFL maxima
  OP_P = 2,     // DRIFTP   (IJ kernel); arg=1 records the CFL maxima
  OP_SEAM = 3,  // F2(:,:,NT)=F2(:,:,1) and magnetopause mask (IJ kernel)
  OP_E = 4,     // DRIFTE   (KL kernel); arg=1 records the CFL maxima
  OP_MU = 5,    // DRIFTMU  (KL kernel); arg=1 records the CFL maxima
  OP_CEX = 6,   // CHAREXCHANGE (KL kernel)
  OP_ATM = 7,   // ATMOL    (KL kernel)
  OP_SUM = 8,   // SUMRC partial sum into slot `arg` (either kernel)
  OP_LOSS = 9,  // fused SUMRC,CEX,SUMRC,ATMOL,SUMRC,ATMOL,SUMRC,CEX,SUMRC chain
};

struct Program {
  int n;
  int op[kMaxOps];
  int arg[kMaxOps];
};

// All device pointers and scalars one launch needs (passed by value).
struct Tab {
  int ns, nr, nt, ne, npa;
  double dts, betalim, fraccfl, mdr, dphi, conf1, conf2;
  double omet;  // OME*DTs/DPHI of DRIFTP
  // species
  int charge[kMaxSpecies];
  int cex[kMaxSpecies];
  double rmas[kMaxSpecies];
  // 1-D grid
  const double *rlz, *ekev, *ebnd, *we, *de, *grel, *grbnd, *v, *mu, *wmu, *dmu;
  const int *upa;  // int(UPA(I)), 1-based values
  // geometry in the reference's Fortran layout
  const double *fnhs, *fnis, *bounhs, *bounis, *hdns_f, *didt, *dibndt;
  const double *bnes, *dbdt, *vt, *eir, *eip;
  const int *mgnp;
  const double *fgeos;   // (NS,NT,NE,NPA)
  const double *bscale;  // (NS)
  // reciprocals of grid widths: 1/DE(K), 1/WE(K), 1/DMU(L), 1/WMU(L)
  double *rde, *rwe, *rdmu, *rwmu;
  // per I
  double *rrlz;        // 1/RLZ(I)
  double *rsum, *rrsum;  // RLZ(I)+0.5*MDR and reciprocal (DRIFTR)
  // per (I,J), index i*NT+j
  double *cr, *bsr, *rbsr;          // DRIFTR: CR, BNES(I,J)+BNES(I+1,J), 1/..
  double *pa, *pce, *pbs, *rpbs;    // DRIFTP: VT-term*P1, EIR term, BNES sum, 1/..
  double *drd1, *dpd1, *bij, *rbij; // DRIFTE DRD1/DPD1, BNES(I,J), 1/BNES(I,J)
  double *drm1, *dpm1, *db2;        // DRIFTMU DRM1/DPM1/dBdt2
  // per (L,I,J), index (l*NR+i)*NT+j  -- read by the (I,J)-plane sweeps
  double *q1;                       // DRIFTR CGR3/(FNHS(I)+FNHS(I+1))
  double *gpa, *fhs, *rfhs;         // DRIFTP GPA1+GPA2, FNHS(J)+FNHS(J1), 1/..
  // per (I,J,L), index (i*NT+j)*NPA+l  -- read by the (K,L)-plane sweeps
  double *gr, *gp, *drd2, *dpd2, *db1, *di1, *hs, *rhs;        // DRIFTE
  double *cmudot, *gmr, *gmp, *drm2, *dpm2, *di2, *bhs, *rbhs;  // DRIFTMU
  double *hdns, *invfnhs;                                       // loss
  // per species
  double *p4;     // [s][k] DRIFTR P4
  double *ekp;    // [s][k] EBND*1e3*(GRBND+1)/2/GRBND (DRIFTE EDT1 prefix)
  double *mkp;    // [s][k] EKEV*1e3*(GREL+1)/2/GREL   (DRIFTMU EDT prefix)
  double *p2;     // [s][i][k]
  double *edot;   // [s][i][k]
  double *sigv;   // [s][k] 10**Y*V(S,K) of CEPARA
  double *atlos;  // [s][i][k]
  double *rel;    // [s][8]: GREL1, GREL2, sq1, GRZERO, sq0
  int *carry;     // [s][NE*NPA] DRIFTR ghost carry
  double *latl;   // [s][i][k] log(ATLOS): ATLOS**(1/FNHS) = exp(log(ATLOS)/FNHS)
  // CFL maxima of max(|C|,floor) over cells inside the magnetopause, as the
  // bit patterns of non-negative doubles (atomicMax order = double order):
  // [s][0] R, [s][1] P, [s][2+K] E per K, [s][2+NE+L] Mu per L
  unsigned long long *cflmax;
  double *dtdrift;      // [s][4]
  // float reciprocals for exact small-integer division (fast_div)
  float rnt, rne;
  // state
  double *f2;
  double *partials;  // [slot][s][maxblk]
  int maxblk;
  double *energy;    // [s][4]: SETRC, LSDR, LSCHA, LSATM
};

inline __host__ __device__ size_t f2_species_stride(const Tab &t) {
  return (size_t)t.nr * t.nt * t.npa * t.ne;
}

}  // namespace ramgpu
