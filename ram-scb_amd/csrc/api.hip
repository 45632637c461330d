// api.hip -- C ABI of the gfx950 RAM split step (declared in include/ramgpu.h).
//
// One context = one (NS,NR,NT,NE,NPA) problem resident on one device, one
// HIP stream.  Inputs arrive in the reference's Fortran layouts and are kept
// on the device as uploaded; F2 is transposed once on upload/download into
// the device layout [S][I][J][L][K].  Derived coefficient tables are rebuilt
// lazily when an input they depend on changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.h"

using namespace ramgpu;

namespace {
// 2^(j/64), j = 0..63, correctly rounded (the table of exp_tab)
const double kExp2Tab[64] = {
    1.0, 1.0108892860517005, 1.0218971486541166, 1.0330248790212284, 1.0442737824274138, 1.0556451783605572,
    1.0671404006768237, 1.0787607977571199, 1.0905077326652577, 1.102382583307841, 1.1143867425958924,
    1.1265216186082418, 1.1387886347566916, 1.1511892299529827, 1.1637248587775775, 1.1763969916502812,
    1.189207115002721, 1.202156731452703, 1.215247359980469, 1.22848053610687, 1.241857812073484,
    1.255380757024691, 1.2690509571917332, 1.2828700160787783, 1.2968395546510096, 1.3109612115247644,
    1.3252366431597413, 1.339667524053303, 1.3542555469368927, 1.3690024229745905, 1.383909881963832,
    1.3989796725383112, 1.4142135623730951, 1.42961333839197, 1.4451808069770467, 1.460917794180647,
    1.4768261459394993, 1.4929077282912648, 1.5091644275934228, 1.5255981507445384, 1.5422108254079407,
    1.559004400237837, 1.5759808451078865, 1.593142151342267, 1.6104903319492543, 1.6280274218573478,
    1.645755478153965, 1.6636765803267364, 1.681792830507429, 1.7001063537185235, 1.718619298122478,
    1.7373338352737062, 1.7562521603732995, 1.7753764925265212, 1.7947090750031072, 1.8142521755003989,
    1.8340080864093424, 1.8539791250833855, 1.8741676341103, 1.8945759815869656, 1.9152065613971474,
    1.9360617934922943, 1.9571441241754002, 1.978456026387951};
}  // namespace

namespace ramgpu_host {

thread_local std::string g_err;

const std::string &last_error() { return g_err; }

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace ramgpu_host

using namespace ramgpu_host;

// Staging of ramgpu_ram_run_schedule_noalloc: the Kp schedule and boundary
// scales of one call (nsteps * (1 + NS) doubles at most), in a fixed array of
// every device's memory instead of a per-call allocation.  Host code never
// names the array itself (its name is a host-side stand-in, not a device
// address): it uses the address hipGetSymbolAddress returns on each device.
namespace {
constexpr size_t kNoAllocDoubles = size_t(1) << 16;  // 512 KiB per device
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
std::mutex g_scratch_mu;                  // guards g_scratch_ptr, serialises the calls
double *g_scratch_ptr[kMaxDevices] = {};  // the array's address on each device

// The array's address on `device`, which must be current; looked up once per
// device (in ramgpu_create, or the entry point's first call there).
int scratch_address(int device, double **out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(RAMGPU_EINVAL, "device %d outside 0..%d of the staging table", device, kMaxDevices - 1);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(RAMGPU_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    }
    g_scratch_ptr[device] = (double *)p;
  }
  *out = g_scratch_ptr[device];
  return RAMGPU_OK;
}
}  // namespace


namespace ramgpu_host {

template <class T>
int dev_alloc(ramgpu_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return fail(RAMGPU_ENOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
  (void)hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T));
  c->allocs.push_back(q);
  *p = (T *)q;
  return RAMGPU_OK;
}

int check_ctx(ramgpu_ctx *c) {
  if (!c) return fail(RAMGPU_EINVAL, "null context");
  if (hipSetDevice(c->device) != hipSuccess) return fail(RAMGPU_EHIP, "hipSetDevice(%d) failed", c->device);
  return RAMGPU_OK;
}

int check_species(ramgpu_ctx *c, int s) {
  if (s < 1 || s > c->ns) return fail(RAMGPU_EINVAL, "species index %d outside 1..%d", s, c->ns);
  return RAMGPU_OK;
}

int ready(ramgpu_ctx *c) {
  if (!c->have_grid) return fail(RAMGPU_ESTATE, "ramgpu_set_grid has not been called");
  if (!c->have_geom) return fail(RAMGPU_ESTATE, "ramgpu_set_geometry has not been called");
  if (!c->have_vt) return fail(RAMGPU_ESTATE, "no electric field (ramgpu_set_efield/ramgpu_set_kp)");
  if (!c->have_bnd) return fail(RAMGPU_ESTATE, "ramgpu_set_boundary has not been called");
  if (!c->have_f2) return fail(RAMGPU_ESTATE, "ramgpu_put_f2 has not been called");
  for (int s = 0; s < c->ns; ++s) {
    if (!c->have_species[s]) return fail(RAMGPU_ESTATE, "species %d not set", s + 1);
    if (c->t.wpi[s] && !c->have_wave)
      return fail(RAMGPU_ESTATE, "species %d takes WAVELO: ramgpu_set_wave_params has not been called", s + 1);
  }
  if (c->t.coul && !c->have_necr)
    return fail(RAMGPU_ESTATE, "Coulomb collisions are on: ramgpu_set_necr has not been called");
  return RAMGPU_OK;
}

int launch_check(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(RAMGPU_EHIP, "launch of %s failed: %s", what, hipGetErrorString(e));
  return RAMGPU_OK;
}

void tbeg(ramgpu_ctx *c, int kind) {
  if (!c->timing) return;
  if (2 * c->tused + 2 > c->tev.size()) {
    for (int n = 0; n < 2; ++n) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        c->timing = false;
        return;
      }
      c->tev.push_back(e);
    }
  }
  if (c->tkind.size() < c->tused + 1) c->tkind.resize(c->tused + 1);
  c->tkind[c->tused] = kind;
  (void)hipEventRecord(c->tev[2 * c->tused], c->stream);
}

void tend(ramgpu_ctx *c) {
  if (!c->timing) return;
  (void)hipEventRecord(c->tev[2 * c->tused + 1], c->stream);
  ++c->tused;
}

// The device copy of the descriptor follows every change of the host image
// (species, parameters, grid scalars); uploaded synchronously, before any
// launch that reads it.
int sync_tab(ramgpu_ctx *c) {
  if (!c->tab_dirty) return RAMGPU_OK;
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->d_tab, &c->t, sizeof(Tab), hipMemcpyHostToDevice));
  c->tab_dirty = false;
  return RAMGPU_OK;
}

// species s0..s0+nsp-1 (0-based) of this context, nb blocks each
// Whether a species of the problem takes WAVELO.
bool takes_wave(const ramgpu_ctx *c) {
  for (int s = 0; s < c->ns; ++s)
    if (c->t.wpi[s]) return true;
  return false;
}

// Units of the per-step coefficient pass in blocks of `threads`: the (I,J)
// terms, then WAVELO's factors (one (S,I,J,K) per thread) when a species
// takes it.
int coef_units(const ramgpu_ctx *c, int threads) {
  int n = (c->nr * c->nt + threads - 1) / threads;
  if (takes_wave(c)) n += (int)(((long long)c->ns * c->nr * c->nt * c->ne + threads - 1) / threads);
  return n;
}

Launch launch_of(const ramgpu_ctx *c, int s0, int nsp, int nb) {
  Launch L{};
  L.tabs = c->d_tab;
  L.jobs = nullptr;
  L.vmap = nullptr;
  L.nb = nb;
  L.s0 = s0;
  L.nv = nsp * nb;
  return L;
}

// Rebuild derived tables (species tables, Courant-number terms) and the
// DRIFTR ghost carry when their inputs changed.
int prepare(ramgpu_ctx *c) {
  int rc = ready(c);
  if (rc || (c->t.coul && (rc = coul_tables(c)))) return rc;
  Tab &t = c->t;
  // a pending finalize goes with the step's k_coef_ij, or alone before any
  // table changes
  if (c->fin_mask >= 0 && (c->tables_dirty || !c->efield_dirty)) {
    const int m = c->fin_mask;
    c->fin_mask = -1;
    if ((rc = run_finalize(c, m, c->fin_mode, 0, c->ns))) return rc;
  }
  if (c->tables_dirty) {
    t.omet = 7.3E-5 * t.dts / t.dphi;  // OME*DTs/DPHI, src/ModRamDrift.f90:239
    c->tab_dirty = true;
  }
  if ((rc = sync_tab(c))) return rc;
  if (c->tables_dirty) {
    tbeg(c, RAMGPU_KERNEL_PREPARE);
    const Launch ls = launch_of(c, 0, c->ns, 1);
    hipLaunchKernelGGL(k_species, dim3(ls.nv), dim3(256), 0, c->stream, ls);
    if ((rc = launch_check("k_species"))) return rc;
    const Launch lc = launch_of(c, 0, 1, c->nr * c->nt);
    hipLaunchKernelGGL(k_coef, dim3(lc.nv), dim3(((c->npa + 63) / 64) * 64 > 256 ? 256 : ((c->npa + 63) / 64) * 64),
                       0, c->stream, lc);
    if ((rc = launch_check("k_coef"))) return rc;
    tend(c);
    c->tables_dirty = false;
    c->efield_dirty = true;
  }
  if (c->efield_dirty) {
    // the step's Courant-number terms, the previous schedule step's finalize
    // if pending, and the ghost carry, in one launch
    tbeg(c, RAMGPU_KERNEL_PREPARE);
    const int nvt = c->step_kp ? c->nr + 1 + c->nt : 0;
    size_t lds = (size_t)(nvt + 4 * c->nt) * sizeof(double);
    // the ghost's q1 row in LDS where it fits the default 64 KB with the
    // kernel's static LDS
    const size_t q1b = (size_t)c->npa * c->nt * sizeof(double);
    const int stage_q1 = c->head_static && c->head_static + lds + q1b <= 65536;
    if (stage_q1) lds += q1b;
    const Launch lc = launch_of(c, 0, 1, coef_units(c, kFinalizeThreads));
    const Launch lf = launch_of(c, 0, c->fin_mask >= 0 ? c->ns : 0, 1);
    const Launch lg = launch_of(c, 0, c->ns, 1);
    hipLaunchKernelGGL(k_step_head, dim3(lc.nv + lf.nv + lg.nv), dim3(kFinalizeThreads), lds, c->stream, lc,
                       c->step_kp, c->step_bs, lf, c->fin_mask, c->fin_mode, lg, nvt, stage_q1);
    c->fin_mask = -1;
    c->step_kp = c->step_bs = nullptr;
    tend(c);
    if ((rc = launch_check("k_step_head"))) return rc;
    c->efield_dirty = false;
    c->ghost_dirty = false;
    c->cfl_valid = false;
  }
  if (c->ghost_dirty) {
    tbeg(c, RAMGPU_KERNEL_PREPARE);
    const Launch l = launch_of(c, 0, c->ns, 1);
    hipLaunchKernelGGL(k_ghost, dim3(l.nv), dim3(1024), (size_t)4 * c->nt * sizeof(double), c->stream, l);
    tend(c);
    if ((rc = launch_check("k_ghost"))) return rc;
    c->ghost_dirty = false;
  }
  return RAMGPU_OK;
}

// The pending finalize of a schedule's last step, alone.
int flush_finalize(ramgpu_ctx *c) {
  if (c->fin_mask < 0) return RAMGPU_OK;
  const int m = c->fin_mask;
  c->fin_mask = -1;
  return run_finalize(c, m, c->fin_mode, 0, c->ns);
}

// One step of a Kp schedule (device pointers to its Kp and, if given, its
// boundary scales): VT and the boundary scales are formed with the step's
// Courant-number terms in k_coef_ij, at once.
int apply_step_inputs(ramgpu_ctx *c, const double *d_kp_n, const double *d_bs_n) {
  c->step_kp = d_kp_n;
  c->step_bs = d_bs_n;
  c->efield_dirty = true;
  if (d_bs_n) c->ghost_dirty = true;
  const int rc = prepare(c);
  c->step_kp = c->step_bs = nullptr;  // never left pointing into the caller's buffers
  return rc;
}

// mode: 0 sweep, 1 sweep + CFL record, 2 last sweep of ram_run (SUMRC slot 5,
// seam, mask)
int run_r(ramgpu_ctx *c, int mode, int s0, int nsp, const double *src, double *dst) {
  Launch L = launch_of(c, s0, nsp, blocks_r(c->t));
  L.src = src;
  L.dst = dst;
  const size_t lds = radial_lds_bytes(c->nr, c->ne, c->npa);
  tbeg(c, RAMGPU_KERNEL_R);
  if (mode == 0)
    hipLaunchKernelGGL(k_r<0>, dim3(L.nv), dim3(kLineThreads), lds, c->stream, L);
  else if (mode == 1)
    hipLaunchKernelGGL(k_r<1>, dim3(L.nv), dim3(kLineThreads), lds, c->stream, L);
  else
    hipLaunchKernelGGL(k_r<2>, dim3(L.nv), dim3(kLineThreads), lds, c->stream, L);
  tend(c);
  return launch_check("k_r");
}

int run_p(ramgpu_ctx *c, int rec, int s0, int nsp, const double *src, double *dst) {
  // out of place, extra lines copy the unswept I=1 rows
  const bool copy = dst && dst != src;
  const int P = c->ne * c->npa;
  const int nb = copy ? ((c->nr - 1) * P + c->nt * P + kLineThreads - 1) / kLineThreads : blocks_p(c->t);
  Launch L = launch_of(c, s0, nsp, nb);
  L.src = src;
  L.dst = dst;
  tbeg(c, RAMGPU_KERNEL_P);
  if (rec)
    hipLaunchKernelGGL(k_p<1>, dim3(L.nv), dim3(kLineThreads), azim_lds_bytes(c->nt, c->ne, c->npa), c->stream, L);
  else
    hipLaunchKernelGGL(k_p<0>, dim3(L.nv), dim3(kLineThreads), azim_lds_bytes(c->nt, c->ne, c->npa), c->stream, L);
  tend(c);
  return launch_check("k_p");
}

// Shape of the forward plane kernels for an NE x NPA plane: blocks of T =
// 256, 512, 768 or 1024 threads (whole waves on each of the 4 SIMDs), each
// thread holding one K of the rows of its chunk (plane_chunks_for), the
// row bound MAXC 8 or 16.  A plane's phases are VALU-bound chains between
// barriers, so a block's time follows the rows of its busiest SIMD (the
// longest chunk a wave holds, summed over the waves on that SIMD).  Two
// blocks resident on a CU sweep two planes at once and keep it busy; a
// third added nothing, and one block alone, whose barriers stall the whole
// CU, ran at 0.65 of one of two (MI355X, 35x72, 48x90 and 64x90 planes).
// Residency is limited by the 16 wave slots of the register budget (4
// waves per SIMD) and by the LDS (lds_cu per CU).  Ties go to the larger
// block.  A function of (NE, NPA, LDS) only, so a problem keeps its
// mapping in any launch.
PlaneShape plane_shape(int ne, int npa, size_t lds_cu) {
  PlaneShape best{0, 0, 0};
  double bcost = 0;
  for (int T = 256; T <= kPlaneMaxThreads; T += 256) {
    const PlaneShape sh = plane_shape_for(ne, npa, T);
    if (sh.threads == 0) continue;
    const size_t lds = plane_fwd_lds_bytes(ne, npa, T);
    const int res = std::min(16 / (T / 64), (int)(lds_cu / lds));
    if (res < 1) continue;
    const int q = plane_chunks_for(ne, npa, T), base = npa / q, extra = npa % q;
    int simd[4] = {0, 0, 0, 0};
    for (int w = 0; w < T / 64; ++w) {  // rows of wave w: its first (longest) chunk's
      const int c0 = (64 * w) / ne;
      if (c0 < q) simd[w & 3] += base + (c0 < extra ? 1 : 0);
    }
    const int busiest = std::max(std::max(simd[0], simd[1]), std::max(simd[2], simd[3]));
    // planes a CU sweeps at a time (two are enough to keep it busy); one
    // block alone stalls the CU at each barrier
    const double cost = busiest / (res >= 2 ? 2.0 : 0.65);
    if (best.threads == 0 || cost <= bcost + 1e-9) {
      best = sh;
      bcost = cost;
    }
  }
  return best;
}

// The forward plane shape with blocks of T threads ({0,0,0} if none fits
// the built kernels): rows is the longest chunk's.
PlaneShape plane_shape_for(int ne, int npa, int T) {
  const int q = plane_chunks_for(ne, npa, T);
  if (q < 1 || T % 64 || T > kPlaneMaxThreads) return {0, 0, 0};
  const int rows = plane_max_rows(npa, q);
  if (rows > kMaxCellsPerThread) return {0, 0, 0};
  return {T, rows <= 8 ? 8 : 16, rows};
}

template <int MAXC>
void launch_plane_t(int seq, int threads, dim3 grid, size_t lds, hipStream_t st, const Launch &L, int rec) {
  const dim3 blk(threads);
  switch (seq) {
    case SEQ_E: hipLaunchKernelGGL((k_plane<SEQ_E, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_MU: hipLaunchKernelGGL((k_plane<SEQ_MU, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_CEX: hipLaunchKernelGGL((k_plane<SEQ_CEX, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_ATM: hipLaunchKernelGGL((k_plane<SEQ_ATM, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_SUM: hipLaunchKernelGGL((k_plane<SEQ_SUM, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_LOSS: hipLaunchKernelGGL((k_plane<SEQ_LOSS, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_TAPE: hipLaunchKernelGGL((k_plane<SEQ_TAPE, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_WAVE: hipLaunchKernelGGL((k_plane<SEQ_WAVE, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_COULEN: hipLaunchKernelGGL((k_plane<SEQ_COULEN, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_COULMU: hipLaunchKernelGGL((k_plane<SEQ_COULMU, MAXC>), grid, blk, lds, st, L, rec); break;
    case SEQ_STEPC: hipLaunchKernelGGL((k_plane<SEQ_STEPC, MAXC>), grid, blk, lds, st, L, rec); break;
    default: hipLaunchKernelGGL((k_plane<SEQ_STEP, MAXC>), grid, blk, lds, st, L, rec); break;
  }
}

void launch_plane(int seq, PlaneShape sh, dim3 grid, size_t lds, hipStream_t st, const Launch &L, int rec) {
  if (sh.maxc <= 8)
    launch_plane_t<8>(seq, sh.threads, grid, lds, st, L, rec);
  else
    launch_plane_t<16>(seq, sh.threads, grid, lds, st, L, rec);
}

int run_plane(ramgpu_ctx *c, int seq, int rec, int s0, int nsp, const double *src, double *dst) {
  // out of place (not SUMRC), extra blocks copy the unswept I=1 planes
  const bool copy = dst && dst != src && seq != SEQ_SUM;
  Launch L = launch_of(c, s0, nsp, blocks_kl(c->t) + (copy ? c->nt : 0));
  L.src = src;
  L.dst = dst;
  L.ilv = 1;  // a plane's species on one XCD
  const size_t lds = seq_coul(seq) ? plane_coul_lds_bytes(c->ne, c->npa, c->plane.threads)
                                    : plane_fwd_lds_bytes(c->ne, c->npa, c->plane.threads);
  tbeg(c, RAMGPU_KERNEL_PLANE);
  launch_plane(seq, c->plane, dim3(L.nv), lds, c->stream, L, rec);
  tend(c);
  return launch_check("k_plane");
}

int run_finalize(ramgpu_ctx *c, int cflmask, int summode, int s0, int nsp) {
  const Launch L = launch_of(c, s0, nsp, 1);
  tbeg(c, RAMGPU_KERNEL_FINALIZE);
  hipLaunchKernelGGL(k_finalize, dim3(L.nv), dim3(kFinalizeThreads), 0, c->stream, L, cflmask, summode);
  tend(c);
  return launch_check("k_finalize");
}

int copy_dtdrift(ramgpu_ctx *c, int s, int which, double *out) {
  if (!out) return RAMGPU_OK;
  HIP_TRY(hipMemcpyAsync(out, c->t.dtdrift + (s - 1) * 4 + which, sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

// which: 0 R, 1 P, 2 E, 3 Mu.  Each call records its CFL limit, as the
// reference's operators do.
int drift_op(ramgpu_ctx *c, int s, int which, double *dtdrift) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  switch (which) {
    case 0: rc = run_r(c, 1, s - 1, 1); break;
    case 1: rc = run_p(c, 1, s - 1, 1); break;
    case 2: rc = run_plane(c, SEQ_E, 1, s - 1, 1); break;
    default: rc = run_plane(c, SEQ_MU, 1, s - 1, 1); break;
  }
  if (rc || (rc = run_finalize(c, 1 << which, 0, s - 1, 1))) return rc;
  return copy_dtdrift(c, s, which, dtdrift);
}

int set_f64(ramgpu_ctx *c, double *dst, const double *src, size_t n, const char *name) {
  if (!src) return fail(RAMGPU_EINVAL, "%s is null", name);
  HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return RAMGPU_OK;
}

int check_dims(int nr, int nt, int ne, int npa) {
  const long long P = (long long)ne * npa;
  if (P > (long long)kPlaneThreads * kMaxCellsPerThread)
    return fail(RAMGPU_EINVAL, "NE*NPA=%lld above %d", P, kPlaneThreads * kMaxCellsPerThread);
  // plane kernels: a thread owns one K of ceil(NPA/(512/NE)) rows
  if (ne > kPlaneThreads || npa > kPlaneThreads ||
      (npa + kPlaneThreads / ne - 1) / (kPlaneThreads / ne) > kMaxCellsPerThread)
    return fail(RAMGPU_EINVAL, "plane %dx%d too large for one block of %d threads", ne, npa, kPlaneThreads);
  if ((long long)nt * P >= (1LL << 22) || (long long)nr * nt >= (1LL << 22))
    return fail(RAMGPU_EINVAL, "grid %dx%dx%dx%d: NT*NE*NPA must stay below 2^22", nr, nt, ne, npa);
  return RAMGPU_OK;
}

// Let the plane kernels and the radial sweep use the device's whole LDS
// (less each kernel's static LDS); refuse planes that do not fit it.
int set_plane_lds_limit(size_t limit) {
#define PLANE_K(S) (const void *)k_plane<S, 8>, (const void *)k_plane<S, 16>
  const void *fns[] = {
      PLANE_K(SEQ_E), PLANE_K(SEQ_MU), PLANE_K(SEQ_CEX), PLANE_K(SEQ_ATM), PLANE_K(SEQ_SUM),
      PLANE_K(SEQ_STEP), PLANE_K(SEQ_LOSS), PLANE_K(SEQ_TAPE), PLANE_K(SEQ_WAVE),
      PLANE_K(SEQ_COULEN), PLANE_K(SEQ_COULMU), PLANE_K(SEQ_STEPC),
      (const void *)k_r<0>, (const void *)k_r<1>, (const void *)k_r<2>,
      (const void *)k_r_bwd<0>, (const void *)k_r_bwd<1>, (const void *)k_p<0>, (const void *)k_p<1>,
      (const void *)k_p_bwd, (const void *)k_pitch_diffusion};
#undef PLANE_K
  for (const void *f : fns) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, f) != hipSuccess ||
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(limit - fa.sharedSizeBytes)) !=
            hipSuccess) {
      (void)hipGetLastError();
      return fail(RAMGPU_EHIP, "cannot raise the LDS limit of a kernel to %zu bytes", limit);
    }
  }
  return RAMGPU_OK;
}

}  // namespace ramgpu_host

int ramgpu_host::fill_out(ramgpu_ctx *c, ramgpu_species_out *out) {
  if (!out) return RAMGPU_OK;
  double e[4 * kMaxSpecies], d[4 * kMaxSpecies], x[3 * kMaxSpecies];
  HIP_TRY(hipMemcpyAsync(e, c->t.energy, (size_t)c->ns * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(d, c->t.dtdrift, (size_t)c->ns * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(x, c->t.lsx, (size_t)c->ns * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int s = 0; s < c->ns; ++s) {
    out[s].setrc = e[s * 4 + 0];
    out[s].lsdr = e[s * 4 + 1];
    out[s].lscha = e[s * 4 + 2];
    out[s].lsatm = e[s * 4 + 3];
    out[s].dtdrift_r = d[s * 4 + 0];
    out[s].dtdrift_p = d[s * 4 + 1];
    out[s].dtdrift_e = d[s * 4 + 2];
    out[s].dtdrift_mu = d[s * 4 + 3];
    out[s].lswae = x[s * 3 + 0];
    out[s].lscoe = x[s * 3 + 1];
    out[s].lscsc = x[s * 3 + 2];
  }
  return RAMGPU_OK;
}

// One ram_run call for all species: R, P, plane(E, Mu, loss chain, Mu, E),
// P, R(+SUMRC, seam, mask), finalize.  The CFL limits depend on the tables
// only, so they are recorded in the first sweep of each kind after a table
// change and reused until the next one.
int ramgpu_host::step_once(ramgpu_ctx *c) {
  int cflmask, rc;
  if ((rc = step_body(c, &cflmask))) return rc;
  if (cflmask >= 0 && (rc = run_finalize(c, cflmask, 2, 0, c->ns))) return rc;
  return RAMGPU_OK;
}

// The sweeps of one step; its finalize (cflmask, SUMRC mode 2) is the
// caller's, except for the composed step (WPADIF or FLC on), which
// finalizes species by species and returns cflmask -1.
int ramgpu_host::step_body(ramgpu_ctx *c, int *cflmask) {
  int rc;
  if ((rc = prepare(c))) return rc;
  const int rec = c->cfl_valid ? 0 : 1;
  const int ns = c->ns;
  if (takes_pitch(c)) {
    if ((rc = step_composed(c, rec))) return rc;
    *cflmask = -1;
    c->cfl_valid = true;
    return RAMGPU_OK;
  }
  // with Coulomb collisions the plane kernel also runs COULEN and COULMU,
  // flag bit 1 being COULMU's clamp (TimeRamElapsed > 0)
  const int seq = c->t.coul ? SEQ_STEPC : SEQ_STEP;
  const int flags = rec | ((c->t.coul && c->time_elapsed > 0.) ? 2 : 0);
  if ((rc = run_r(c, rec, 0, ns)) || (rc = run_p(c, rec, 0, ns)) ||
      (rc = run_plane(c, seq, flags, 0, ns)) || (rc = run_p(c, 0, 0, ns)) || (rc = run_r(c, 2, 0, ns)))
    return rc;
  *cflmask = rec ? 15 : 0;
  c->cfl_valid = true;
  return RAMGPU_OK;
}

extern "C" {

const char *ramgpu_last_error(void) { return ramgpu_host::last_error().c_str(); }

int ramgpu_create(ramgpu_ctx **out, int ns, int nr, int nt, int ne, int npa, int device) {
  if (!out) return fail(RAMGPU_EINVAL, "null output pointer");
  *out = nullptr;
  int rc0;
  if (ns < 1 || ns > kMaxSpecies) return fail(RAMGPU_EINVAL, "NS=%d outside 1..%d", ns, kMaxSpecies);
  if (nr < 4 || nt < 4 || ne < 4 || npa < 4)
    return fail(RAMGPU_EINVAL, "grid %dx%dx%dx%d: every axis needs at least 4 points", nr, nt, ne, npa);
  if (npa > 1024) return fail(RAMGPU_EINVAL, "NPA=%d above 1024", npa);
  if ((rc0 = check_dims(nr, nt, ne, npa))) return rc0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(RAMGPU_EHIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(RAMGPU_EINVAL, "device %d outside 0..%d", device, ndev - 1);
  HIP_TRY(hipSetDevice(device));
  ramgpu_ctx *c = new ramgpu_ctx();
  c->device = device;
  c->ns = ns;
  c->nr = nr;
  c->nt = nt;
  c->ne = ne;
  c->npa = npa;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->lds_limit = prop.sharedMemPerBlock;
  int rc = RAMGPU_OK;
  auto cleanup = [&](int code) {
    ramgpu_destroy(c);
    return code;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(RAMGPU_EHIP, "hipStreamCreate failed"));
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return cleanup(fail(RAMGPU_EHIP, "hipEventCreate failed"));
  const size_t nr1 = nr + 1;
  const size_t g3 = nr1 * nt * npa, g2 = nr1 * nt;
  const size_t n3 = (size_t)nr * nt * npa, n2 = (size_t)nr * nt;
  const size_t nf2 = (size_t)ns * nr * nt * ne * npa;
  Tab &t = c->t;
  t.ns = ns; t.nr = nr; t.nt = nt; t.ne = ne; t.npa = npa;
  t.maxblk = std::max(blocks_kl(t), blocks_r(t));  // partial sums per (slot, species)
  if (plane_bwd_lds_bytes(ne, npa) > c->lds_limit)
    return cleanup(fail(RAMGPU_EINVAL, "NE*NPA=%d plane does not fit the device LDS (%zu bytes)", ne * npa,
                        c->lds_limit));
  if (radial_lds_bytes(nr, ne, npa) + 1024 > c->lds_limit)
    return cleanup(fail(RAMGPU_EINVAL, "NR=%d: the radial sweep's face terms do not fit the device LDS", nr));
  if (azim_lds_bytes(nt, ne, npa) + 1024 > c->lds_limit)
    return cleanup(fail(RAMGPU_EINVAL, "NT=%d: the azimuthal sweep's face terms do not fit the device LDS", nt));
  // gfx950: the LDS of a CU is the largest a block may use
  c->plane = plane_shape(ne, npa, c->lds_limit);
  if (c->plane.threads == 0) return cleanup(fail(RAMGPU_EINVAL, "no plane kernel shape for %dx%d", ne, npa));
  t.plane_chunks = plane_chunks_for(ne, npa, c->plane.threads);
  c->d_pdc.assign(ns, nullptr);
  c->d_flc.assign(ns, nullptr);
  c->d_flc_mem.assign(ns, nullptr);
  c->pd_wpi.assign(ns, 0);
  c->pd_emic.assign(ns, 0);
#define A(p, n) \
  if ((rc = dev_alloc(c, &p, n))) return cleanup(rc)
  A(c->d_rlz, nr1); A(c->d_ekev, ne); A(c->d_ebnd, ne); A(c->d_we, ne); A(c->d_de, ne);
  A(c->d_grel, (size_t)ns * ne); A(c->d_grbnd, (size_t)ns * ne); A(c->d_v, (size_t)ns * ne);
  A(c->d_mu, npa); A(c->d_wmu, npa); A(c->d_dmu, npa); A(c->d_upa, nr);
  A(c->d_fnhs, g3); A(c->d_fnis, g3); A(c->d_bounhs, g3); A(c->d_bounis, g3); A(c->d_hdnsf, g3);
  A(c->d_didt, g3); A(c->d_dibndt, g3);
  A(c->d_bnes, g2); A(c->d_dbdt, g2); A(c->d_vt, g2); A(c->d_eir, g2); A(c->d_eip, g2);
  A(c->d_mgnp, n2); A(c->d_fgeos, (size_t)ns * nt * ne * npa); A(c->d_bscale, ns);
  A(c->d_lz, nr1); A(c->d_phi, nt); A(c->d_kp, 1);
  A(t.rde, ne); A(t.rwe, ne); A(t.rdmu, npa); A(t.rwmu, npa);
  A(t.rrlz, nr); A(t.rsi, nr); A(t.rij, n2); A(t.pij, n2); A(t.kij, 8 * n2);
  A(t.q1, n3); A(t.pl, n3); A(t.kle, 8 * n3); A(t.klm, 8 * n3); A(t.kll, n3);
  A(t.p4, (size_t)ns * ne); A(t.ekp, (size_t)ns * ne); A(t.mkp, (size_t)ns * ne);
  A(t.p2, (size_t)ns * nr * ne); A(t.edot, (size_t)ns * nr * ne); A(t.sigv, (size_t)ns * ne);
  A(t.atlos, (size_t)ns * nr * ne); A(t.rel, (size_t)ns * 8);
  A(t.carry, (size_t)ns * ne * npa);
  A(t.latl, (size_t)ns * nr * ne);
  A(t.ex64, 64);
  A(c->d_tab, 1);
  A(t.cflmax, (size_t)ns * (2 + ne + npa));
  A(t.dtdrift, (size_t)ns * 4);
  A(t.f2, nf2);
  A(c->d_stage, nf2);
  A(t.partials, (size_t)(kSumSlots + 1) * ns * t.maxblk);  // + ELORC slot of ramgpu_sumrc
  A(t.energy, (size_t)ns * 4);
  A(t.wl, (size_t)3 * ns * nr * ne); A(t.rlpp, nt); A(t.lsx, (size_t)ns * 3);
  A(t.wavef, (size_t)ns * nr * nt * ne);
  if (hipMemcpy(t.ex64, kExp2Tab, sizeof(kExp2Tab), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(RAMGPU_EHIP, "hipMemcpy failed"));
#undef A
  t.rlz = c->d_rlz; t.ekev = c->d_ekev; t.ebnd = c->d_ebnd; t.we = c->d_we; t.de = c->d_de;
  t.grel = c->d_grel; t.grbnd = c->d_grbnd; t.v = c->d_v; t.mu = c->d_mu; t.wmu = c->d_wmu;
  t.dmu = c->d_dmu; t.upa = c->d_upa;
  t.fnhs = c->d_fnhs; t.fnis = c->d_fnis; t.bounhs = c->d_bounhs; t.bounis = c->d_bounis;
  t.hdns_f = c->d_hdnsf; t.didt = c->d_didt; t.dibndt = c->d_dibndt; t.bnes = c->d_bnes;
  t.dbdt = c->d_dbdt; t.vt = c->d_vt; t.eir = c->d_eir; t.eip = c->d_eip; t.mgnp = c->d_mgnp;
  t.fgeos = c->d_fgeos; t.bscale = c->d_bscale; t.lz = c->d_lz; t.phi = c->d_phi; t.kpnow = c->d_kp;
  t.dts = 5.0; t.betalim = 1.5; t.fraccfl = 0.8;
  t.rnt = 1.0f / (float)nt;
  t.rne = 1.0f / (float)ne;
  t.rp = 1.0f / (float)(ne * npa);
  std::vector<double> ones(ns, 1.0);
  if (hipMemcpy(c->d_bscale, ones.data(), ns * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(RAMGPU_EHIP, "hipMemcpy failed"));
  if ((rc = set_plane_lds_limit(c->lds_limit))) return cleanup(rc);
  {
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, (const void *)k_step_head) == hipSuccess) c->head_static = fa.sharedSizeBytes;
    (void)hipGetLastError();
  }
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    double *sp = nullptr;
    if ((rc = scratch_address(device, &sp))) return cleanup(rc);
  }
  *out = c;
  return RAMGPU_OK;
}

int ramgpu_destroy(ramgpu_ctx *c) {
  if (!c) return RAMGPU_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  for (void *p : c->allocs) (void)hipFree(p);
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  delete c;
  return RAMGPU_OK;
}

int ramgpu_set_species(ramgpu_ctx *c, int s, int charge, int cex_kind, double rmas) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  if (charge == 0) return fail(RAMGPU_EINVAL, "species %d: zero charge", s);
  if (cex_kind < RAMGPU_CEX_NONE || cex_kind > RAMGPU_CEX_O)
    return fail(RAMGPU_EINVAL, "species %d: unknown CEX kind %d", s, cex_kind);
  if (!(rmas > 0)) return fail(RAMGPU_EINVAL, "species %d: mass must be positive", s);
  c->t.charge[s - 1] = charge;
  c->t.cex[s - 1] = cex_kind;
  c->t.rmas[s - 1] = rmas;
  c->have_species[s - 1] = true;
  c->coul_dirty = true;
  c->tables_dirty = true;
  c->tab_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_params(ramgpu_ctx *c, double dts, double betalim, double fraccfl) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!(dts > 0)) return fail(RAMGPU_EINVAL, "DTs must be positive");
  c->t.dts = dts;
  c->t.betalim = betalim;
  c->t.fraccfl = fraccfl;
  c->coul_dirty = true;  // COULPARA's tables scale with DTs
  c->tables_dirty = true;
  c->tab_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_plane_shape(ramgpu_ctx *c, int threads, int maxc) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (threads == 0 && maxc == 0) {  // back to the model's choice
    const PlaneShape sh = plane_shape(c->ne, c->npa, c->lds_limit);
    if (c->t.coul && (rc = coul_check_lds(c, sh.threads))) return rc;
    c->plane = sh;
  } else {
    if (threads < 64 || threads % 64 || threads > kPlaneMaxThreads || (maxc != 8 && maxc != 16))
      return fail(RAMGPU_EINVAL, "plane shape %d threads, %d rows: not a built kernel shape", threads, maxc);
    const PlaneShape sh = plane_shape_for(c->ne, c->npa, threads);
    if (sh.threads == 0 || sh.rows > maxc)
      return fail(RAMGPU_EINVAL, "plane shape %d threads holds %d rows per thread, above %d", threads, sh.rows, maxc);
    const int rows = sh.rows;
    if (plane_fwd_lds_bytes(c->ne, c->npa, threads) > c->lds_limit)
      return fail(RAMGPU_EINVAL, "plane shape %d threads: the plane's LDS exceeds the device's", threads);
    if (c->t.coul && (rc = coul_check_lds(c, threads))) return rc;
    c->plane = {threads, maxc, rows};
  }
  c->t.plane_chunks = plane_chunks_for(c->ne, c->npa, c->plane.threads);
  c->tab_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_get_plane_shape(ramgpu_ctx *c, int *threads, int *maxc) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (threads) *threads = c->plane.threads;
  if (maxc) *maxc = c->plane.maxc;
  return RAMGPU_OK;
}

int ramgpu_set_grid(ramgpu_ctx *c, double mdr, double dphi, double conf1, double conf2,
                    const double *rlz, const double *ekev, const double *ebnd, const double *we,
                    const double *de, const double *grel, const double *grbnd, const double *v,
                    const double *mu, const double *wmu, const double *dmu, const double *upa) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const size_t ns = c->ns, ne = c->ne, npa = c->npa;
  if ((rc = set_f64(c, c->d_rlz, rlz, c->nr + 1, "rlz")) || (rc = set_f64(c, c->d_ekev, ekev, ne, "ekev")) ||
      (rc = set_f64(c, c->d_ebnd, ebnd, ne, "ebnd")) || (rc = set_f64(c, c->d_we, we, ne, "we")) ||
      (rc = set_f64(c, c->d_de, de, ne, "de")) || (rc = set_f64(c, c->d_grel, grel, ns * ne, "grel")) ||
      (rc = set_f64(c, c->d_grbnd, grbnd, ns * ne, "grbnd")) || (rc = set_f64(c, c->d_v, v, ns * ne, "v")) ||
      (rc = set_f64(c, c->d_mu, mu, npa, "mu")) || (rc = set_f64(c, c->d_wmu, wmu, npa, "wmu")) ||
      (rc = set_f64(c, c->d_dmu, dmu, npa, "dmu")))
    return rc;
  if (!upa) return fail(RAMGPU_EINVAL, "upa is null");
  std::vector<int> u(c->nr);
  for (int i = 0; i < c->nr; ++i) u[i] = (int)upa[i];  // int(UPA(I),kind=4), src/ModRamLoss.f90:499
  HIP_TRY(hipMemcpyAsync(c->d_upa, u.data(), u.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->t.mdr = mdr;
  c->t.dphi = dphi;
  c->t.conf1 = conf1;
  c->t.conf2 = conf2;
  c->mom_dirty = true;  // the moments' WMU/FFACTOR table
  c->h_ekev.assign(ekev, ekev + ne);
  c->h_we.assign(we, we + ne);
  c->h_grel.assign(grel, grel + ns * ne);
  c->h_grbnd.assign(grbnd, grbnd + ns * ne);
  c->h_v.assign(v, v + ns * ne);
  c->h_mu.assign(mu, mu + npa);
  c->h_wmu.assign(wmu, wmu + npa);
  c->h_dmu.assign(dmu, dmu + npa);
  c->coul_dirty = true;
  c->have_grid = true;
  c->tables_dirty = true;
  c->tab_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_geometry(ramgpu_ctx *c, const double *fnhs, const double *fnis, const double *bounhs,
                        const double *bounis, const double *hdns, const double *didt,
                        const double *dibndt, const double *bnes, const double *dbdt,
                        const double *eir, const double *eip, const int *outside_mgnp) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const size_t g3 = (size_t)(c->nr + 1) * c->nt * c->npa, g2 = (size_t)(c->nr + 1) * c->nt;
  if ((rc = set_f64(c, c->d_fnhs, fnhs, g3, "fnhs")) || (rc = set_f64(c, c->d_fnis, fnis, g3, "fnis")) ||
      (rc = set_f64(c, c->d_bounhs, bounhs, g3, "bounhs")) ||
      (rc = set_f64(c, c->d_bounis, bounis, g3, "bounis")) ||
      (rc = set_f64(c, c->d_hdnsf, hdns, g3, "hdns")) || (rc = set_f64(c, c->d_didt, didt, g3, "didt")) ||
      (rc = set_f64(c, c->d_dibndt, dibndt, g3, "dibndt")) ||
      (rc = set_f64(c, c->d_bnes, bnes, g2, "bnes")) || (rc = set_f64(c, c->d_dbdt, dbdt, g2, "dbdt")) ||
      (rc = set_f64(c, c->d_eir, eir, g2, "eir")) || (rc = set_f64(c, c->d_eip, eip, g2, "eip")))
    return rc;
  if (!outside_mgnp) return fail(RAMGPU_EINVAL, "outside_mgnp is null");
  HIP_TRY(hipMemcpyAsync(c->d_mgnp, outside_mgnp, (size_t)c->nr * c->nt * sizeof(int),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->have_geom = true;
  c->tables_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_efield(ramgpu_ctx *c, const double *vt) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = set_f64(c, c->d_vt, vt, (size_t)(c->nr + 1) * c->nt, "vt"))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->have_vt = true;
  c->efield_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_kp(ramgpu_ctx *c, double kp, const double *lz, const double *phi) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (lz || phi) {
    if ((rc = set_f64(c, c->d_lz, lz, c->nr + 1, "lz")) || (rc = set_f64(c, c->d_phi, phi, c->nt, "phi")))
      return rc;
    c->have_lzphi = true;
  }
  if (!c->have_lzphi) return fail(RAMGPU_EINVAL, "ramgpu_set_kp needs lz and phi on its first call");
  hipLaunchKernelGGL(k_vs, dim3(1), dim3(256), 0, c->stream, c->d_vt, (const double *)c->d_lz,
                     (const double *)c->d_phi, (const double *)nullptr, kp, c->nr + 1, c->nt,
                     (const double *)nullptr, (double *)nullptr, 0);
  if ((rc = launch_check("k_vs"))) return rc;
  c->kp_host = kp;  // WAVELO's KP follows the convection's
  HIP_TRY(hipMemcpyAsync(c->d_kp, &c->kp_host, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->have_vt = true;
  c->efield_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_boundary(ramgpu_ctx *c, const double *fgeos) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = set_f64(c, c->d_fgeos, fgeos, (size_t)c->ns * c->nt * c->ne * c->npa, "fgeos"))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->have_bnd = true;
  c->ghost_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_boundary_scale(ramgpu_ctx *c, const double *scale) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = set_f64(c, c->d_bscale, scale, c->ns, "scale"))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->ghost_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_put_f2(ramgpu_ctx *c, const double *f2) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const size_t n = (size_t)c->ns * c->nr * c->nt * c->ne * c->npa;
  if ((rc = set_f64(c, c->d_stage, f2, n, "f2"))) return rc;
  hipLaunchKernelGGL(k_f2_in, dim3(1024), dim3(256), 0, c->stream, c->t.f2, (const double *)c->d_stage,
                     c->ns, c->nr, c->nt, c->ne, c->npa);
  if ((rc = launch_check("k_f2_in"))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->have_f2 = true;
  return RAMGPU_OK;
}

int ramgpu_get_f2(ramgpu_ctx *c, double *f2) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!f2) return fail(RAMGPU_EINVAL, "f2 is null");
  const size_t n = (size_t)c->ns * c->nr * c->nt * c->ne * c->npa;
  hipLaunchKernelGGL(k_f2_out, dim3(1024), dim3(256), 0, c->stream, c->d_stage, (const double *)c->t.f2,
                     c->ns, c->nr, c->nt, c->ne, c->npa);
  if ((rc = launch_check("k_f2_out"))) return rc;
  HIP_TRY(hipMemcpyAsync(f2, c->d_stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

int ramgpu_set_energy_state(ramgpu_ctx *c, const double *setrc, const double *lsdr,
                            const double *lscha, const double *lsatm) {
  int rc = check_ctx(c);
  if (rc) return rc;
  std::vector<double> e(c->ns * 4);
  HIP_TRY(hipMemcpy(e.data(), c->t.energy, e.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < c->ns; ++s) {
    if (setrc) e[s * 4 + 0] = setrc[s];
    if (lsdr) e[s * 4 + 1] = lsdr[s];
    if (lscha) e[s * 4 + 2] = lscha[s];
    if (lsatm) e[s * 4 + 3] = lsatm[s];
  }
  HIP_TRY(hipMemcpy(c->t.energy, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice));
  return RAMGPU_OK;
}

int ramgpu_get_energy_state(ramgpu_ctx *c, double *setrc, double *lsdr, double *lscha, double *lsatm) {
  int rc = check_ctx(c);
  if (rc) return rc;
  std::vector<double> e(c->ns * 4);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(e.data(), c->t.energy, e.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < c->ns; ++s) {
    if (setrc) setrc[s] = e[s * 4 + 0];
    if (lsdr) lsdr[s] = e[s * 4 + 1];
    if (lscha) lscha[s] = e[s * 4 + 2];
    if (lsatm) lsatm[s] = e[s * 4 + 3];
  }
  return RAMGPU_OK;
}

int ramgpu_driftpara(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

int ramgpu_driftr(ramgpu_ctx *c, int s, double *dt) { return drift_op(c, s, 0, dt); }
int ramgpu_driftp(ramgpu_ctx *c, int s, double *dt) { return drift_op(c, s, 1, dt); }
int ramgpu_drifte(ramgpu_ctx *c, int s, double *dt) { return drift_op(c, s, 2, dt); }
int ramgpu_driftmu(ramgpu_ctx *c, int s, double *dt) { return drift_op(c, s, 3, dt); }

int ramgpu_driftend(ramgpu_ctx *c) { return check_ctx(c); }

int ramgpu_cepara(ramgpu_ctx *c, int s) { return ramgpu_driftpara(c, s); }

int ramgpu_charexchange(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  if ((rc = run_plane(c, SEQ_CEX, 0, s - 1, 1))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

int ramgpu_atmol(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  if ((rc = run_plane(c, SEQ_ATM, 0, s - 1, 1))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

int ramgpu_set_wavelo(ramgpu_ctx *c, int s, const double *walos1, const double *walos2, const double *walos3) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  if (!walos1 && !walos2 && !walos3) {
    c->t.wpi[s - 1] = 0;
    c->tab_dirty = true;
    return RAMGPU_OK;
  }
  if (!walos1 || !walos2 || !walos3) return fail(RAMGPU_EINVAL, "WALOS1, WALOS2 and WALOS3 must all be given");
  const size_t nr = c->nr, ne = c->ne, nik = nr * ne;
  std::vector<double> w(3 * nik);
  const double *src[3] = {walos1, walos2, walos3};
  for (int q = 0; q < 3; ++q)
    for (size_t k = 0; k < ne; ++k)
      for (size_t i = 0; i < nr; ++i) w[q * nik + i * ne + k] = src[q][i + nr * k];  // (NR,NE) -> [i][k]
  for (int q = 0; q < 3; ++q)
    HIP_TRY(hipMemcpy(c->t.wl + ((size_t)q * c->ns + (s - 1)) * nik, w.data() + q * nik, nik * sizeof(double),
                      hipMemcpyHostToDevice));
  c->t.wpi[s - 1] = 1;
  c->tab_dirty = true;
  c->efield_dirty = true;  // the step's factors are formed with the Courant-number terms
  return RAMGPU_OK;
}

int ramgpu_set_wave_params(ramgpu_ctx *c, double kp, double kpmax12, const double *lz, const double *necr,
                           int use_plasmasphere) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!lz) return fail(RAMGPU_EINVAL, "lz is null");
  if (use_plasmasphere && !necr) return fail(RAMGPU_EINVAL, "a plasmasphere model needs NECR");
  // the plasmapause of WAVELO, src/ModRamWPI.f90:598-603
  std::vector<double> rl(c->nt);
  for (int j = 0; j < c->nt; ++j) {
    rl[j] = 5.39 - 0.382 * kpmax12;
    if (use_plasmasphere)
      for (int i = 1; i < c->nr; ++i)
        if (necr[i + (size_t)c->nr * j] > 50.) rl[j] = lz[i];
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(c->t.rlpp, rl.data(), rl.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_lz, lz, (size_t)(c->nr + 1) * sizeof(double), hipMemcpyHostToDevice));  // WAVELO's LZ(I)
  c->kp_host = kp;
  HIP_TRY(hipMemcpy(c->d_kp, &c->kp_host, sizeof(double), hipMemcpyHostToDevice));
  c->have_wave = true;
  c->efield_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_wavelo(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  if (!c->t.wpi[s - 1]) return fail(RAMGPU_ESTATE, "species %d has no WAVELO tables (ramgpu_set_wavelo)", s);
  if ((rc = run_plane(c, SEQ_WAVE, 0, s - 1, 1))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

int ramgpu_set_loss_state(ramgpu_ctx *c, const double *lswae, const double *lscoe, const double *lscsc) {
  int rc = check_ctx(c);
  if (rc) return rc;
  std::vector<double> e(c->ns * 3);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(e.data(), c->t.lsx, e.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < c->ns; ++s) {
    if (lswae) e[s * 3 + 0] = lswae[s];
    if (lscoe) e[s * 3 + 1] = lscoe[s];
    if (lscsc) e[s * 3 + 2] = lscsc[s];
  }
  HIP_TRY(hipMemcpy(c->t.lsx, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice));
  return RAMGPU_OK;
}

int ramgpu_get_loss_state(ramgpu_ctx *c, double *lswae, double *lscoe, double *lscsc) {
  int rc = check_ctx(c);
  if (rc) return rc;
  std::vector<double> e(c->ns * 3);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(e.data(), c->t.lsx, e.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < c->ns; ++s) {
    if (lswae) lswae[s] = e[s * 3 + 0];
    if (lscoe) lscoe[s] = e[s * 3 + 1];
    if (lscsc) lscsc[s] = e[s * 3 + 2];
  }
  return RAMGPU_OK;
}

int ramgpu_sumrc(ramgpu_ctx *c, int s, double *setrc, double *elorc) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = ready(c)) || (rc = sync_tab(c))) return rc;
  // every plane block (rows I>=2) writes its slot-0 partial
  if ((rc = run_plane(c, SEQ_SUM, 0, s - 1, 1))) return rc;
  if ((rc = run_finalize(c, 0, 1, s - 1, 1))) return rc;
  double e[4], el;
  HIP_TRY(hipMemcpyAsync(e, c->t.energy + (s - 1) * 4, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&el, c->t.partials + ((size_t)kSumSlots * c->ns + (s - 1)) * c->t.maxblk,
                         sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (setrc) *setrc = e[0];
  if (elorc) *elorc = el;
  return RAMGPU_OK;
}



int ramgpu_ram_run(ramgpu_ctx *c, int nsteps, ramgpu_species_out *out) {
  int rc = check_ctx(c);
  if (rc || (rc = ready(c))) return rc;
  if (nsteps < 0) return fail(RAMGPU_EINVAL, "nsteps < 0");
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (int n = 0; n < nsteps; ++n) {
    if ((rc = step_once(c))) return rc;
    c->time_elapsed += 2.0 * c->t.dts;  // src/ModRamScbRun.f90:88, after each ram_run
  }
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  return fill_out(c, out);
}

}  // extern "C"

namespace {
// Checks shared by the two schedule entry points.
int schedule_args(ramgpu_ctx *c, int nsteps, const double *kp) {
  int rc = check_ctx(c);
  if (rc || (rc = ready(c))) return rc;
  if (nsteps < 0) return fail(RAMGPU_EINVAL, "nsteps < 0");
  if (!kp) return fail(RAMGPU_EINVAL, "kp schedule is null");
  if (!c->have_lzphi) return fail(RAMGPU_ESTATE, "ramgpu_set_kp(lz, phi) must precede a Kp schedule");
  return RAMGPU_OK;
}

// The steps of a Kp schedule already staged on the device (d_kp[nsteps],
// d_bs[nsteps*NS] or null).  Each step's finalize runs with the next step's
// first launch (prepare), the last one before the end of the timed run.
// Returns after the stream has finished, so the staging may be reused.
int run_schedule_staged(ramgpu_ctx *c, int nsteps, const double *d_kp, const double *d_bs) {
  int rc = RAMGPU_OK;
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (int n = 0; n < nsteps && rc == RAMGPU_OK; ++n) {
    if ((rc = apply_step_inputs(c, d_kp + n, d_bs ? d_bs + (size_t)n * c->ns : nullptr))) break;
    int cflmask = 0;
    if ((rc = step_body(c, &cflmask))) break;
    if (cflmask >= 0) {
      c->fin_mask = cflmask;
      c->fin_mode = 2;
    }
    c->time_elapsed += 2.0 * c->t.dts;  // src/ModRamScbRun.f90:88, after each ram_run
  }
  const int rf = flush_finalize(c);
  if (rc == RAMGPU_OK) rc = rf;
  (void)hipEventRecord(c->ev1, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == RAMGPU_OK)
    rc = fail(RAMGPU_EHIP, "hipStreamSynchronize failed after a Kp schedule");
  return rc;
}
}  // namespace

extern "C" {

int ramgpu_ram_run_schedule(ramgpu_ctx *c, int nsteps, const double *kp, const double *bscale,
                            ramgpu_species_out *out) {
  int rc = schedule_args(c, nsteps, kp);
  if (rc) return rc;
  // the whole schedule is staged on the device before the timed steps
  double *d_kp = nullptr, *d_bs = nullptr;
  HIP_TRY(hipMalloc(&d_kp, std::max(nsteps, 1) * sizeof(double)));
  if (bscale && hipMalloc(&d_bs, std::max(nsteps, 1) * c->ns * sizeof(double)) != hipSuccess) {
    (void)hipFree(d_kp);
    return fail(RAMGPU_ENOMEM, "hipMalloc of the boundary-scale schedule failed");
  }
  if (hipMemcpyAsync(d_kp, kp, nsteps * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      (bscale && hipMemcpyAsync(d_bs, bscale, (size_t)nsteps * c->ns * sizeof(double), hipMemcpyHostToDevice,
                                c->stream) != hipSuccess) ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    rc = fail(RAMGPU_EHIP, "staging the Kp schedule failed");
  if (rc == RAMGPU_OK) rc = run_schedule_staged(c, nsteps, d_kp, d_bs);
  (void)hipFree(d_kp);
  if (d_bs) (void)hipFree(d_bs);
  if (rc) return rc;
  return fill_out(c, out);
}

int ramgpu_ram_run_schedule_noalloc(ramgpu_ctx *c, int nsteps, const double *kp, const double *bscale,
                                    ramgpu_species_out *out) {
  int rc = schedule_args(c, nsteps, kp);
  if (rc) return rc;
  const size_t need = (size_t)nsteps * (1 + (bscale ? (size_t)c->ns : 0));
  if (need > kNoAllocDoubles)
    return fail(RAMGPU_EINVAL, "schedule of %d steps needs %zu staged doubles, above the %zu of the fixed array",
                nsteps, need, kNoAllocDoubles);
  // one call at a time per process: the calls share every device's array
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  double *sp = nullptr;
  if ((rc = scratch_address(c->device, &sp))) return rc;
  double *d_kp = sp, *d_bs = bscale ? sp + nsteps : nullptr;
  if (nsteps > 0 &&
      (hipMemcpyAsync(d_kp, kp, nsteps * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
       (d_bs && hipMemcpyAsync(d_bs, bscale, (size_t)nsteps * c->ns * sizeof(double), hipMemcpyHostToDevice,
                               c->stream) != hipSuccess)))
    return fail(RAMGPU_EHIP, "staging the Kp schedule failed");
  if ((rc = run_schedule_staged(c, nsteps, d_kp, d_bs))) return rc;
  return fill_out(c, out);
}

int ramgpu_last_run_ms(ramgpu_ctx *c, double *ms) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!ms) return fail(RAMGPU_EINVAL, "ms is null");
  HIP_TRY(hipEventSynchronize(c->ev1));
  float f = 0.f;
  HIP_TRY(hipEventElapsedTime(&f, c->ev0, c->ev1));
  *ms = f;
  return RAMGPU_OK;
}

int ramgpu_copy_probe(int device, size_t bytes, int reps, double *gbps) {
  if (!gbps || reps < 1 || bytes < 4096) return fail(RAMGPU_EINVAL, "bad copy probe arguments");
  HIP_TRY(hipSetDevice(device));
  const size_t n2 = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  HIP_TRY(hipMalloc(&a, n2 * 16));
  if (hipMalloc(&b, n2 * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(RAMGPU_ENOMEM, "copy probe allocation failed");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemset(a, 0, n2 * 16);
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, b, (const double2 *)a, n2);  // warm
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, b, (const double2 *)a, n2);
  (void)hipEventRecord(e1, 0);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) return fail(RAMGPU_EHIP, "copy probe: %s", hipGetErrorString(err));
  *gbps = 2.0 * (double)(n2 * 16) * reps / (ms * 1e-3) / 1e9;
  return RAMGPU_OK;
}

int ramgpu_kernel_timing(ramgpu_ctx *c, int enable) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->timing = enable != 0;
  c->tused = 0;
  return RAMGPU_OK;
}

int ramgpu_kernel_times(ramgpu_ctx *c, double *total_ms, int *launches) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!total_ms || !launches) return fail(RAMGPU_EINVAL, "null output array");
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < RAMGPU_KERNEL_KINDS; ++k) {
    total_ms[k] = 0.0;
    launches[k] = 0;
  }
  for (size_t n = 0; n < c->tused; ++n) {
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, c->tev[2 * n], c->tev[2 * n + 1]));
    total_ms[c->tkind[n]] += f;
    launches[c->tkind[n]] += 1;
  }
  c->tused = 0;
  return RAMGPU_OK;
}

}  // extern "C"
