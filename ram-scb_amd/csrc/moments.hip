// moments.hip -- the pressure moments (ANISCH) and the flux conversion (FLUX)
// on the device-resident F2: the block of ram_run after its species loop
// (src/ModRamRun.f90:207-222) and ANISCH (:265-415) without its
// wave-particle branch (:422-, taken only with DoUseWPI) and file output.
// A host that needs PPERT/PPART every step downloads NS*NR*NT values of each
// instead of F2.
#include <algorithm>
#include <vector>

#include "ctx.h"

namespace ramgpu {

constexpr int kMomThreads = 256;  // one (S,I,J) plane per block
constexpr int kMomChunk = 512;    // cells of a plane staged per pass (whole rows of NE)
constexpr int kMomMaxNE = 4 * kMomThreads;
constexpr int kXcds = 8;          // MI355X: blocks go round-robin to 8 XCDs, each with its own L2

// rows of a staged pass and the dynamic LDS of k_moments (doubles): the
// pass's terms, the per-K terms, then FNHS(I,J,L) and its reciprocal
__host__ __device__ inline int mom_rows(int ne) { return ne < kMomChunk ? kMomChunk / ne : 1; }
inline size_t mom_lds_doubles(int ne, int npa) {
  return 3 * (size_t)mom_rows(ne) * ne + 4 * (size_t)ne + 2 * (size_t)npa;
}

// The quotient tables beside FFACTOR (each n doubles, [s][i][l][k]):
// RN(1/FFACTOR) and RN(WMU(L)/FFACTOR), so that the kernel's divisions are
// div_rn's exact reciprocal form: ERNM = RN(RN(WMU/FFACTOR)/FNHS) and
// FLUX = RN(RN(F2/FFACTOR)/FNHS) as the reference's two divisions.
__global__ void k_mom_tables(double *ff, const double *__restrict__ wmu, int ne, int npa, size_t n) {
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
    const int l = (int)((x / ne) % npa);
    ff[n + x] = 1.0 / ff[x];
    ff[2 * n + x] = wmu[l] / ff[x];
  }
}

// One block per (S,I,J) plane of F2 ([S][I][J][L][K], NE*NPA contiguous),
// in passes of mom_rows(NE) pitch-angle rows: every thread forms the terms
// F2*EPME, F2*EPMA, F2*ERNM of its cells (and FLUX = F2/FFACTOR/FNHS) into
// LDS; then the thread of each energy K adds its column's terms row by row,
// so SUME, SUMA and SUMN are the reference's sequential sums over
// L=1..int(UPA(I)-1).  Thread 0 finally adds the per-K terms in K order into
// the energy windows K=2..6, 7..10, 11..25, 26..30, 31..NE (:297-298; each
// window ends at min(khi, NE): the reference reads past NE when NE < 30).
// FLUX is 0 outside I=2..NR, K=2..NE, L=2..NPA, J=1..NT-1, where the
// reference's FLUX keeps the zero of its allocation (src/ModRamInit.f90:78).
// ff: FFACTOR as [s][i][l][k], then its k_mom_tables quotients; mom: PPERT, PPART, EDENT, RNHTT, each
// (NS,NR,NT) in the reference's layout.  MAXK: energies per thread (NE <=
// MAXK*kMomThreads).
template <int MAXK>
__global__ void __launch_bounds__(kMomThreads, MAXK == 1 ? 8 : 4)
    k_moments(double *__restrict__ f2, int ns, int nr, int nt, int ne, int npa, int s0, int nsp, int anisch,
              const double *__restrict__ ff, const double *__restrict__ ernh, const double *__restrict__ epp,
              const double *__restrict__ ekev, const double *__restrict__ wmu, const double *__restrict__ mu,
              const double *__restrict__ fnhs, const int *__restrict__ upa, double *__restrict__ flux,
              double *__restrict__ mom) {
  extern __shared__ double sm[];
  // XCD-aware: the blocks an XCD receives (blockIdx mod 8) take one
  // contiguous range of planes, so the FFACTOR tables of an (S,I), shared by
  // its NT planes, are fetched into about one L2; a block past the last
  // plane exits whole
  const int nplane = nsp * (nr - 1) * nt, per = (nplane + kXcds - 1) / kXcds;
  const int p = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (p >= nplane) return;
  const int j = p % nt;
  const int i = 1 + (p / nt) % (nr - 1);  // I = 2..NR
  const int s = s0 + p / (nt * (nr - 1));
  const int R = mom_rows(ne), C = R * ne;
  double *pe = sm, *pa = sm + C, *pn = sm + 2 * C, *term = sm + 3 * C;
  double *fnl = term + 4 * ne, *rfl = fnl + npa;  // FNHS(I,J,L), RN(1/FNHS(I,J,L))
  const size_t P = (size_t)ne * npa, nff = (size_t)ns * nr * P;
  const size_t plane = (((size_t)s * nr + i) * nt + j) * P;
  double *F = f2 + plane;
  double *FX = flux ? flux + plane : nullptr;
  const double *ffp = ff + ((size_t)s * nr + i) * P, *rffp = ffp + nff, *wfp = ffp + 2 * nff;
  const double *fn = fnhs + i + (size_t)(nr + 1) * j;  // FNHS(I,J,L) = fn[fst*(L-1)]
  const size_t fst = (size_t)(nr + 1) * nt;
  for (int l = threadIdx.x; l < npa; l += kMomThreads) {
    const double v = fn[fst * l];
    fnl[l] = v;
    rfl[l] = 1.0 / v;
  }
  __syncthreads();
  const int u = upa[i] - 1;  // int(UPA(I)-1): UPA holds whole numbers
  const bool jflux = j <= nt - 2;
  const int lend = FX ? npa : (anisch ? std::min(u, npa) : 0);
  double se[MAXK], sa[MAXK], sn[MAXK];
#pragma unroll
  for (int q = 0; q < MAXK; ++q) se[q] = sa[q] = sn[q] = 0.0;
  // a thread's cells of a pass (R*NE <= CPT*kMomThreads)
  constexpr int CPT = MAXK == 1 ? 2 : 4;
  int cr[CPT], ck[CPT];  // row in the pass and energy of each cell (the same in every pass)
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int c = threadIdx.x + q * kMomThreads;
    cr[q] = c / ne;
    ck[q] = c - cr[q] * ne;
  }
  for (int l0 = 0; l0 < lend; l0 += R) {
    const int rows = std::min(R, lend - l0);
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int c = threadIdx.x + q * kMomThreads;
      if (c >= rows * ne) continue;
      const int k = ck[q], l = l0 + cr[q];
      const bool kin = k >= 1;  // K = 2..NE
      double f;
      if (l == 0 && anisch && kin) {
        f = F[ne + k];  // F2(S,I,J,K,1) = F2(S,I,J,K,2)
        F[k] = f;
      } else {
        f = F[(size_t)l * ne + k];
      }
      const double fnv = fnl[l], rfnv = rfl[l];
      double te = 0.0, ta = 0.0, tn = 0.0;
      if (anisch && kin && l < u) {
        const double ernm = div_rn(wfp[(size_t)l * ne + k], fnv, rfnv);  // WMU/FFACTOR/FNHS
        const double epma = ernm * mu[l] * mu[l];
        const double epme = ernm - epma;
        te = f * epme;
        ta = f * epma;
        tn = f * ernm;
      }
      pe[c] = te;
      pa[c] = ta;
      pn[c] = tn;
      if (FX) {
        const size_t x = (size_t)l * ne + k;
        FX[x] = (kin && l >= 1 && jflux) ? div_rn(div_rn(f, ffp[x], rffp[x]), fnv, rfnv) : 0.0;
      }
    }
    __syncthreads();
    if (anisch) {
#pragma unroll
      for (int q = 0; q < MAXK; ++q) {
        const int k = threadIdx.x + q * kMomThreads;
        // rows at L > int(UPA(I)-1) hold +0 terms: adding +0 leaves a sum
        // that starts at +0 unchanged, so the whole pass is added
        if (k >= 1 && k < ne)
#pragma unroll 4
          for (int r = 0; r < rows; ++r) {
            se[q] = se[q] + pe[r * ne + k];
            sa[q] = sa[q] + pa[r * ne + k];
            sn[q] = sn[q] + pn[r * ne + k];
          }
      }
    }
    __syncthreads();
  }
  if (!anisch) return;
#pragma unroll
  for (int q = 0; q < MAXK; ++q) {
    const int k = threadIdx.x + q * kMomThreads;
    if (k >= 1 && k < ne) {
      const double e = epp[s + (size_t)ns * k], h = ernh[s + (size_t)ns * k];
      term[k] = e * se[q];                    // EPP(S,K)*SUME
      term[ne + k] = e * sa[q];               // EPP(S,K)*SUMA
      term[2 * ne + k] = h * sn[q];           // ERNH(S,K)*SUMN
      term[3 * ne + k] = h * ekev[k] * sn[q]; // ERNH(S,K)*EKEV(K)*SUMN
    }
  }
  __syncthreads();
  // thread 4w+q adds quantity q over energy window w in K order (a window's
  // own sum, from 0, as the reference's PPER, PPAR, RNHT, EDEN); thread 0
  // then scales the windows and adds them in order (:375-397)
  double *wsum = pe;  // the passes' terms are no longer needed
  if (threadIdx.x < 20) {
    const int q = threadIdx.x & 3, w = threadIdx.x >> 2;
    auto khi = [ne](int m) { return std::min(m == 0 ? 6 : m == 1 ? 10 : m == 2 ? 25 : m == 3 ? 30 : ne, ne); };
    const int klo = w == 0 ? 2 : khi(w - 1) + 1;  // 1-based
    double v = 0.0;
    for (int K = klo; K <= khi(w); ++K) v = v + term[q * ne + K - 1];
    wsum[threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  constexpr double kRfac = 4 * 3.1415926535897932384626433832795 / (2.998E8 * 100);  // 4*PI/(CS*100)
  double tot[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int w = 0; w < 5; ++w) {
    // RNHT*RFAC, EDEN*RFAC, 2*RFAC*PPAR, RFAC*PPER, then the totals
    tot[0] = tot[0] + kRfac * wsum[4 * w];
    tot[1] = tot[1] + (2 * kRfac) * wsum[4 * w + 1];
    tot[2] = tot[2] + wsum[4 * w + 3] * kRfac;
    tot[3] = tot[3] + wsum[4 * w + 2] * kRfac;
  }
  const size_t n3 = (size_t)ns * nr * nt, x = s + (size_t)ns * (i + (size_t)nr * j);
#pragma unroll
  for (int q = 0; q < 4; ++q) mom[q * n3 + x] = tot[q];
}

}  // namespace ramgpu

using namespace ramgpu_host;
using ramgpu::k_moments;
using ramgpu::kMomMaxNE;
using ramgpu::kMomThreads;
using ramgpu::kXcds;
using ramgpu::mom_lds_doubles;

namespace {

int alloc_zero(ramgpu_ctx *c, double **p, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess)
    return fail(RAMGPU_ENOMEM, "hipMalloc of %zu bytes failed", n * sizeof(double));
  c->allocs.push_back(q);
  *p = (double *)q;
  HIP_TRY(hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(double), c->stream));
  return RAMGPU_OK;
}

// ANISCH of species s0..s0+nsp-1 (anisch != 0) and/or FLUX of them into
// d_flux (flux), one launch.
int run_moments(ramgpu_ctx *c, int s0, int nsp, int anisch, bool flux) {
  if (!c->have_grid) return fail(RAMGPU_ESTATE, "ramgpu_set_grid has not been called");
  if (!c->have_geom) return fail(RAMGPU_ESTATE, "ramgpu_set_geometry has not been called");
  if (!c->have_f2) return fail(RAMGPU_ESTATE, "ramgpu_put_f2 has not been called");
  if (!c->d_ff || !c->d_ernh || !c->d_epp || !c->d_mom)
    return fail(RAMGPU_ESTATE, "ramgpu_set_moment_factors has not been called (or its allocation failed)");
  if (c->ne > kMomMaxNE) return fail(RAMGPU_EINVAL, "moments: NE=%d above %d", c->ne, kMomMaxNE);
  int rc;
  if (flux && !c->d_flux && (rc = alloc_zero(c, &c->d_flux, (size_t)c->ns * c->nr * c->nt * c->ne * c->npa)))
    return rc;
  if (c->mom_dirty) {  // the quotient tables, after new factors or a new grid (WMU)
    const size_t n = (size_t)c->ns * c->nr * c->ne * c->npa;
    hipLaunchKernelGGL(ramgpu::k_mom_tables, dim3(256), dim3(256), 0, c->stream, c->d_ff,
                       (const double *)c->d_wmu, c->ne, c->npa, n);
    if ((rc = launch_check("k_mom_tables"))) return rc;
    c->mom_dirty = false;
  }
  if (c->nr < 2) return RAMGPU_OK;
  const int nplane = nsp * (c->nr - 1) * c->nt;
  const int nblk = (nplane + kXcds - 1) / kXcds * kXcds;  // every XCD gets ceil(nplane/8) planes' blocks
  const size_t lds = mom_lds_doubles(c->ne, c->npa) * sizeof(double);
  tbeg(c, RAMGPU_KERNEL_MOMENTS);
  if (c->ne <= kMomThreads)
    hipLaunchKernelGGL(k_moments<1>, dim3(nblk), dim3(kMomThreads), lds, c->stream, c->t.f2, c->ns, c->nr,
                       c->nt, c->ne, c->npa, s0, nsp, anisch, (const double *)c->d_ff, (const double *)c->d_ernh,
                       (const double *)c->d_epp, (const double *)c->d_ekev, (const double *)c->d_wmu,
                       (const double *)c->d_mu, (const double *)c->d_fnhs, (const int *)c->d_upa,
                       flux ? c->d_flux : nullptr, c->d_mom);
  else
    hipLaunchKernelGGL(k_moments<4>, dim3(nblk), dim3(kMomThreads), lds, c->stream, c->t.f2, c->ns, c->nr,
                       c->nt, c->ne, c->npa, s0, nsp, anisch, (const double *)c->d_ff, (const double *)c->d_ernh,
                       (const double *)c->d_epp, (const double *)c->d_ekev, (const double *)c->d_wmu,
                       (const double *)c->d_mu, (const double *)c->d_fnhs, (const int *)c->d_upa,
                       flux ? c->d_flux : nullptr, c->d_mom);
  tend(c);
  return launch_check("k_moments");
}

// the (S, 2..NR, 1..NT) entries of species s0..s0+nsp-1 into the caller's
// (NS,NR,NT) arrays (any may be NULL)
int download_moments(ramgpu_ctx *c, int s0, int nsp, double *ppert, double *ppart, double *edent,
                     double *rnhtt) {
  const size_t n3 = (size_t)c->ns * c->nr * c->nt;
  std::vector<double> m(4 * n3);
  HIP_TRY(hipMemcpyAsync(m.data(), c->d_mom, m.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double *out[4] = {ppert, ppart, edent, rnhtt};
  for (int q = 0; q < 4; ++q) {
    if (!out[q]) continue;
    for (int j = 0; j < c->nt; ++j)
      for (int i = 1; i < c->nr; ++i)
        for (int s = s0; s < s0 + nsp; ++s) {
          const size_t x = s + (size_t)c->ns * (i + (size_t)c->nr * j);
          out[q][x] = m[q * n3 + x];
        }
  }
  return RAMGPU_OK;
}

}  // namespace

int ramgpu_set_moment_factors(ramgpu_ctx *c, const double *ffactor, const double *ernh, const double *epp) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!ffactor || !ernh || !epp) return fail(RAMGPU_EINVAL, "ffactor, ernh and epp must not be null");
  const size_t ns = c->ns, nr = c->nr, ne = c->ne, npa = c->npa;
  // each buffer on its own test: a retry after a failed allocation completes the set
  if ((!c->d_ff && (rc = alloc_zero(c, &c->d_ff, 3 * ns * nr * ne * npa))) ||
      (!c->d_ernh && (rc = alloc_zero(c, &c->d_ernh, ns * ne))) ||
      (!c->d_epp && (rc = alloc_zero(c, &c->d_epp, ns * ne))) ||
      (!c->d_mom && (rc = alloc_zero(c, &c->d_mom, 4 * ns * nr * c->nt))))
    return rc;
  // FFACTOR(NS,NR,NE,NPA) -> [s][i][l][k], so that a plane's K run is contiguous
  std::vector<double> t(ns * nr * ne * npa);
  for (size_t l = 0; l < npa; ++l)
    for (size_t k = 0; k < ne; ++k)
      for (size_t i = 0; i < nr; ++i)
        for (size_t s = 0; s < ns; ++s)
          t[((s * nr + i) * npa + l) * ne + k] = ffactor[s + ns * (i + nr * (k + ne * l))];
  HIP_TRY(hipMemcpyAsync(c->d_ff, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_ernh, ernh, ns * ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_epp, epp, ns * ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->mom_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_anisch(ramgpu_ctx *c, int s, double *ppert, double *ppart, double *edent, double *rnhtt) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = run_moments(c, s - 1, 1, 1, false))) return rc;
  return download_moments(c, s - 1, 1, ppert, ppart, edent, rnhtt);
}

int ramgpu_flux(ramgpu_ctx *c, double *flux) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!flux) return fail(RAMGPU_EINVAL, "flux is null");
  if ((rc = run_moments(c, 0, c->ns, 0, true))) return rc;
  return download_f2(c, flux, c->d_flux);
}

int ramgpu_moments(ramgpu_ctx *c, double *ppert, double *ppart, double *edent, double *rnhtt, double *flux) {
  int rc = check_ctx(c);
  if (rc || (rc = run_moments(c, 0, c->ns, 1, flux != nullptr))) return rc;
  if ((rc = download_moments(c, 0, c->ns, ppert, ppart, edent, rnhtt))) return rc;
  return flux ? download_f2(c, flux, c->d_flux) : RAMGPU_OK;
}
