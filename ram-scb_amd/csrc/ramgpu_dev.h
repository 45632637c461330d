// ramgpu_dev.h -- device-side data of the RAM split step (gfx950).
//
// F2 is kept on the device per species in a [I][J][L][K] layout (K fastest):
// the (K,L) plane of one (I,J) cell is one contiguous run of NE*NPA doubles.
// The radial and local-time sweeps run one thread per line with consecutive
// lanes on consecutive (L,K) positions, so every load and store of a wave is
// one contiguous run; the energy and pitch-angle sweeps (plus the pointwise
// loss operators between them) stage whole (K,L) planes in LDS.
//
// Coefficient tables hold the reference's intermediate quantities (its
// GPA, GPR1.., DRD1.., CMUDOT.. terms) computed in the reference's expression
// order, plus correctly rounded reciprocals of every divisor that is reused
// across cells, so that each Courant number is rebuilt per cell with exactly
// the reference's sequence of roundings.  Tables read together are packed
// together (one base address, consecutive fields).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace ramgpu {

constexpr int kMaxSpecies = 8;
constexpr int kLineThreads = 256;      // R/P sweep kernels: one line per thread
constexpr int kPlaneThreads = 512;     // plane adjoint kernels: one (K,L) plane per block
constexpr int kPlaneMaxThreads = 1024; // forward plane kernels: block size chosen per problem
constexpr int kMaxCellsPerThread = 16; // cells of one plane per thread, in every mapping
constexpr int kMaxPlaneCells = kPlaneThreads * kMaxCellsPerThread;  // NE*NPA bound, checked at creation
// SUMRC partial-sum slots of one ram_run: 0..6 the loss chain's (plane
// blocks), 7 the last SUMRC (radial blocks), 8..11 the Coulomb operators'
// (plane blocks: after the first drift half, after COULEN; after the second
// COULMU, after the second COULEN)
constexpr int kSumSlots = 12;
constexpr int kSumRadial = 7;
constexpr int kSumCoul = 8;
// k_finalize's summode: bits 0-1 the mode (0 none, 1 one SUMRC, 2 the
// fused step's chain); with mode 1, kSumFromRadial reads the radial slot
// and bits kSumAccShift.. name the accumulator the ELORC goes to
constexpr int kSumModeMask = 3;
constexpr int kSumFromRadial = 1 << 2;
constexpr int kSumAccShift = 4;
enum SumAcc { SUM_NONE = 0, SUM_LSDR = 1, SUM_LSCHA = 2, SUM_LSATM = 3, SUM_LSWAE = 4, SUM_LSCOE = 5,
              SUM_LSCSC = 6 };
constexpr int kMuSpread = 4;           // LDS slots per pitch-angle row of DRIFTMU's CFL maximum (by K mod 4)
constexpr int kFinalizeThreads = 1024;  // one block per (problem, species) of k_finalize (and of k_ghost)

// packed per-(I,J) terms of the plane sweeps, kij[(i*NT+j)*8 + ..]
enum KijField { KIJ_DRD1, KIJ_DPD1, KIJ_BIJ, KIJ_RBIJ, KIJ_DRM1, KIJ_DPM1, KIJ_DB2, KIJ_MG };
// packed per-(I,J,L) DRIFTE terms, kle[((i*NT+j)*NPA+l)*8 + ..]
enum KleField { KLE_HS, KLE_RHS, KLE_DRD2, KLE_DPD2, KLE_GR, KLE_GP, KLE_DB1, KLE_DI1 };
// packed per-(I,J,L) DRIFTMU terms, klm[((i*NT+j)*NPA+l)*8 + ..]
enum KlmField { KLM_BHS, KLM_RBHS, KLM_DRM2, KLM_DPM2, KLM_CMUDOT, KLM_GMR, KLM_GMP, KLM_DI2 };

// All device pointers and scalars of one problem (one ensemble member).
// Lives in device memory; kernels read it through a pointer.
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
  double *bscale;        // (NS)
  const double *lz, *phi;  // LZ(NR+1), PHI(NT) for the device-side VT of a Kp schedule
  // reciprocals of grid widths: 1/DE(K), 1/WE(K), 1/DMU(L), 1/WMU(L)
  double *rde, *rwe, *rdmu, *rwmu;
  double *rrlz;    // [i] 1/RLZ(I)
  double2 *rsi;    // [i] {RLZ(I)+0.5*MDR, 1/..}                       (DRIFTR)
  double4 *rij;    // [i*NT+j] {CR, BNES(I,J)+BNES(I+1,J), 1/.., outsideMGNP(I,J)}
  double *q1;      // [(l*NR+i)*NT+j] CGR3/(FNHS(I)+FNHS(I+1))         (DRIFTR)
  double4 *pij;    // [i*NT+j] {VT term*P1, EIR term, BNES(I,J)+BNES(I,J1), 1/..}
  double4 *pl;     // [(l*NR+i)*NT+j] {GPA1+GPA2, FNHS(J)+FNHS(J1), 1/.., outsideMGNP(I,J)}
  double *kij;     // [(i*NT+j)*8] KijField
  double *kle;     // [((i*NT+j)*NPA+l)*8] KleField
  double *klm;     // [((i*NT+j)*NPA+l)*8] KlmField
  double2 *kll;    // [(i*NT+j)*NPA+l] {HDNS, 1/FNHS}                   (loss)
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
  // WAVELO (electron lifetimes) of the species with wpi[s] set: WALOS1/2/3
  // as wl[(q*NS + s)*NR*NE + i*NE + k], q = 0,1,2; the plasmapause L of each
  // local time; KP of the current step (written by ramgpu_set_kp, the Kp
  // schedules and ramgpu_set_wave_params)
  int wpi[kMaxSpecies];
  double *wl;
  double *wavef;  // [((s*NR + i)*NT + j)*NE + k] exp(-DTs/TAU) of the step (k_coef_ij)
  double *rlpp;   // [j]
  double *kpnow;  // 1 value
  double *lsx;    // [s][3]: LSWAE, LSCOE, LSCSC
  // Coulomb collisions (DoUseCoulomb): the COULPARA tables of every species,
  // (COULE+COULI)(S,K,L) as cc[(s*NPA + l)*NE + k] and {ATA, GTA}(S,K,L) as
  // cag[(s*NPA + l)*NE + k]; COULEN's ghost factors per species, crel[s*5 +
  // ..] = GREL(S,1), GREL(S,2), SQRT((GREL1^2-1)/(GREL2^2-1)), GRZERO(S),
  // SQRT((GRZERO^2-1)/(GREL1^2-1)); NECR (NR,NT) in the reference's layout
  int coul;
  double *cc;
  double2 *cag;
  double *crel;
  const double *necr;
  double *ex64;   // [j] 2^(j/64), the table of the loss factors' exponential
  // CFL maxima of max(|C|,floor) over cells inside the magnetopause, as the
  // bit patterns of non-negative doubles (atomicMax order = double order):
  // [s][0] R, [s][1] P, [s][2+K] E per K, [s][2+NE+L] Mu per L
  unsigned long long *cflmax;
  double *dtdrift;      // [s][4]
  // float reciprocals for exact small-integer division (fast_div)
  float rnt, rne, rp;   // 1/NT, 1/NE, 1/(NE*NPA)
  // row mapping of the forward plane kernels: thread n < plane_chunks*NE
  // owns K = n%NE of the rows of chunk n/NE (plane_chunks_for); the
  // threads above, in a block of any size, are idle
  int plane_chunks;
  // state
  double *f2;
  double *partials;  // [slot][s][maxblk]
  int maxblk;
  double *energy;    // [s][4]: SETRC, LSDR, LSCHA, LSATM
  // reverse mode of an ensemble (null otherwise): the problem's tape of
  // F2-sized slots, its adjoint lam, the adjoint of each step's FGEOS
  // ([step][NS,NT,NE,NPA]), and its flux-misfit points, weights
  // ([p][NPA]), targets and per-step losses ([step][p])
  double *tape, *lam, *bbar;
  const int *mpts;
  const double *mw, *mtarget;
  double *mloss;
  int mnpts;
};

// One unit of a launch: blocks [b0, b0+nb) of species s of problem tabs[tab].
struct Job {
  int tab, s, b0, nb;
};
// vmap entries: a virtual block's job and its block within the job
constexpr int kVmapShift = 16;
inline __host__ __device__ int vmap_entry(int job, int b) { return (job << kVmapShift) | b; }

// A launch over several units.  With vmap == nullptr the launch covers
// species s0, s0+1, .. of tabs[0], nb blocks each, species after species,
// or with ilv the species of each run of 8 blocks in turn (so the species'
// blocks of one plane land on one XCD, whose L2 then serves the plane's
// species-independent tables to all of them: workgroups go to the 8 XCDs
// in turn); otherwise vmap[vb] is the vmap_entry of virtual block vb.
// Blocks loop over virtual blocks with stride gridDim.x, so the grid may be
// smaller than nv.
// src/dst (sweep kernels): read F2 from src and write to dst (device layout
// of the whole problem) instead of updating tabs[..].f2 in place; null for in
// place.  ssl/dsl name them per problem instead (ensembles): n > 0 is slot
// n-1 of the problem's tape, dsl < 0 its own F2 written out of place, 0 not
// used.
struct Launch {
  const Tab *tabs;
  const Job *jobs;
  const int *vmap;
  int nv, nb, s0, ilv;
  const double *src;
  double *dst;
  int ssl, dsl;
  int bstep;  // reverse mode of an ensemble: the step whose FGEOS adjoint (Tab::bbar) DRIFTR's adjoint feeds
};

inline __host__ __device__ size_t f2_species_stride(const Tab &t) {
  return (size_t)t.nr * t.nt * t.npa * t.ne;
}

// the input and output F2 of a launch for problem t, and whether they differ
inline __host__ __device__ const double *launch_src(const Launch &L, const Tab &t) {
  if (L.ssl > 0) return t.tape + (size_t)(L.ssl - 1) * t.ns * f2_species_stride(t);
  return L.src ? L.src : t.f2;
}
inline __host__ __device__ double *launch_dst(const Launch &L, const Tab &t) {
  if (L.dsl > 0) return t.tape + (size_t)(L.dsl - 1) * t.ns * f2_species_stride(t);
  if (L.dsl < 0) return t.f2;
  return L.dst ? L.dst : t.f2;
}
inline __host__ __device__ bool out_of_place(const Launch &L) { return L.dsl != 0 || (L.dst && L.dst != L.src); }

// block counts of the kernels, per (problem, species)
inline __host__ __device__ int blocks_r(const Tab &t) {
  return (t.nt * t.ne * t.npa + kLineThreads - 1) / kLineThreads;
}
inline __host__ __device__ int blocks_p(const Tab &t) {
  return ((t.nr - 1) * t.ne * t.npa + kLineThreads - 1) / kLineThreads;
}
inline __host__ __device__ int blocks_kl(const Tab &t) { return (t.nr - 1) * t.nt; }

// LDS bytes of the azimuthal sweep: pl of a block's (I,L) pairs and pij of
// its I values
inline __host__ __device__ size_t azim_lds_bytes(int nt, int ne, int npa) {
  const int npair = (kLineThreads - 1) / ne + 2, ni = (kLineThreads - 1) / (ne * npa) + 2;
  return (size_t)(npair + ni) * nt * 4 * sizeof(double);
}
// LDS bytes of the radial sweep: the face terms of a block's lines, q1 of
// its (J,L) pairs, rij of its J values, rsi and the flags rij(I,NT).w
inline __host__ __device__ size_t radial_lds_bytes(int nr, int ne, int npa) {
  const int npair = (kLineThreads - 1) / ne + 2, nj = (kLineThreads - 1) / (ne * npa) + 2;
  return (size_t)(npair + 4 * nj + 3) * nr * sizeof(double);
}

// LDS bytes of the plane kernel
// LDS of the forward plane kernels; mu_slots: LDS slots per row of DRIFTMU's
// CFL maximum
inline __host__ __device__ size_t plane_lds_bytes(int ne, int npa, int mu_slots = kMuSpread) {
  const size_t P = (size_t)ne * npa;
  return (2 * P + 16 * (size_t)npa + 2 * (size_t)npa + 11 * (size_t)ne + 4 * (size_t)npa + 128 + 64 + 24 +
          2 * (size_t)npa + (size_t)ne + (size_t)mu_slots * npa) * sizeof(double);
}
// Chunks of rows of the forward plane kernels in blocks of T threads: one
// chunk of NE threads per floor(T/NE), at most NPA/2 so that every chunk
// holds two rows or more (the top row L=NPA then never opens a chunk:
// DRIFTMU sets it from the row below, which the same thread holds).  The
// NPA rows are shared out as evenly as they go, the first NPA % chunks
// chunks one row longer.  0 when T < NE.
inline __host__ __device__ int plane_chunks_for(int ne, int npa, int threads) {
  const int q = threads / ne;
  return q < npa / 2 ? q : npa / 2;
}
// the longest chunk's rows
inline __host__ __device__ int plane_max_rows(int npa, int chunks) { return (npa + chunks - 1) / chunks; }
// LDS of the forward plane kernels (PlaneFwd): NPA rows of NE+3 energy
// stencil slots, the K=1 column, T exchange slots, the per-L DRIFTE and
// DRIFTMU terms (16 NPA), HDNS and 1/FNHS (2 NPA), KT_N=11 per-K and LT_N=4
// per-L tables, 128 reduction slots, the 64-entry exp table, PC_N=24 plane
// scalars and the CFL slots (NE + kMuSpread*NPA).
inline __host__ __device__ size_t plane_fwd_lds_bytes(int ne, int npa, int threads) {
  return ((size_t)npa * (ne + 3) + (size_t)npa + (size_t)threads + 16 * (size_t)npa + 2 * (size_t)npa +
          11 * (size_t)ne + 4 * (size_t)npa + 128 + 64 + 24 + (size_t)ne + (size_t)kMuSpread * npa) *
         sizeof(double);
}
// The Coulomb operators' LDS after the forward plane kernels': per L the
// bounce-averaged density XNE, BASCNE, FNHS, BOUNHS, MU, 1/FNHS and 1/MU
// (7 NPA), and COULMU's RL recurrence of every column (NE*NPA).
inline __host__ __device__ size_t plane_coul_lds_bytes(int ne, int npa, int threads) {
  return plane_fwd_lds_bytes(ne, npa, threads) + (7 * (size_t)npa + (size_t)ne * npa) * sizeof(double);
}
// LDS of the pitch-angle line solve (k_pitch_diffusion): RK and RL of each
// of its 64 lines
constexpr int kPitchThreads = 64;
// the species of one k_pitch_diffusion launch and their coefficient tables
struct PitchJob {
  int n;
  int s[kMaxSpecies];
  const double *coef[kMaxSpecies];
};
inline __host__ __device__ size_t pitch_lds_bytes(int npa) { return (size_t)2 * npa * kPitchThreads * sizeof(double); }
// the adjoint kernels record no CFL maxima: their adjoint plane takes the
// place of the CFL slots (NE + mu_slots*NPA doubles)
inline __host__ __device__ size_t plane_bwd_lds_bytes(int ne, int npa) {
  return plane_lds_bytes(ne, npa, 0) - (size_t)ne * sizeof(double) + (size_t)ne * npa * sizeof(double);
}

}  // namespace ramgpu
