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
#include <string>
#include <vector>

#include "ramgpu.h"
#include "ramgpu_dev.h"

namespace ramgpu {
__global__ void k_species(Tab t);
__global__ void k_coef_ij(Tab t);
__global__ void k_coef(Tab t);
__global__ void k_ghost(Tab t);
__global__ void k_ij(Tab t, Program prog, int s0, int ntiles, int kws);
__global__ void k_kl(Tab t, Program prog, int s0);
__global__ void k_finalize(Tab t, int cflmask, int summode, int s0, int nblk_kl, int nblk_ij,
                           double *elorc_out);
__global__ void k_vs(double *vt, const double *lz, const double *phi, const double *kp_ptr,
                     double kp_val, int nr1, int nt);
__global__ void k_copy(double2 *dst, const double2 *src, size_t n2);
__global__ void k_f2_in(double *dst, const double *src, int ns, int nr, int nt, int ne, int npa);
__global__ void k_f2_out(double *dst, const double *src, int ns, int nr, int nt, int ne, int npa);
}  // namespace ramgpu

using namespace ramgpu;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(x)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) return fail(RAMGPU_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

}  // namespace

struct ramgpu_ctx {
  int device = 0;
  int ns = 0, nr = 0, nt = 0, ne = 0, npa = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  size_t lds_limit = 65536;
  Tab t{};
  std::vector<void *> allocs;
  // device buffers with host-visible handles
  double *d_rlz, *d_ekev, *d_ebnd, *d_we, *d_de, *d_grel, *d_grbnd, *d_v, *d_mu, *d_wmu, *d_dmu;
  int *d_upa;
  double *d_fnhs, *d_fnis, *d_bounhs, *d_bounis, *d_hdnsf, *d_didt, *d_dibndt;
  double *d_bnes, *d_dbdt, *d_vt, *d_eir, *d_eip, *d_fgeos, *d_bscale;
  int *d_mgnp;
  double *d_lz, *d_phi, *d_kp;
  double *d_stage;  // Fortran-layout staging buffer for F2 transfers
  double *d_elorc;
  bool have_grid = false, have_geom = false, have_vt = false, have_bnd = false, have_f2 = false;
  bool have_species[kMaxSpecies] = {};
  bool have_lzphi = false;
  bool tables_dirty = true, ghost_dirty = true;
  bool efield_dirty = true;  // only the VT-dependent per-(I,J) terms are stale
  bool cfl_valid = false;  // DtDrift* current for the present tables
  // opt-in per-kernel timing: event pairs around every launch, by kind
  bool timing = false;
  std::vector<hipEvent_t> tev;  // 2 per recorded launch
  std::vector<int> tkind;
  size_t tused = 0;
  int kws = 3;            // IJ tile width 2^kws (consecutive (L,K) positions)
  int ij_threads = 256, kl_threads = 512;
};

namespace {

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
  for (int s = 0; s < c->ns; ++s)
    if (!c->have_species[s]) return fail(RAMGPU_ESTATE, "species %d not set", s + 1);
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

size_t ij_lds(const ramgpu_ctx *c) {
  const size_t kw = (size_t)1 << c->kws;
  return (2 * (size_t)c->nr * c->nt * kw + 16) * sizeof(double) + (size_t)c->nt * kw * sizeof(int);
}
size_t kl_lds(const ramgpu_ctx *c) {
  return (3 * (size_t)c->ne * c->npa + 16 + c->ne + c->npa) * sizeof(double);
}

// Rebuild derived tables (species tables, Courant-number terms) and the
// DRIFTR ghost carry when their inputs changed.
int prepare(ramgpu_ctx *c) {
  int rc = ready(c);
  if (rc) return rc;
  Tab &t = c->t;
  if (c->tables_dirty) {
    t.omet = 7.3E-5 * t.dts / t.dphi;  // OME*DTs/DPHI, src/ModRamDrift.f90:239
    tbeg(c, RAMGPU_KERNEL_PREPARE);
    hipLaunchKernelGGL(k_species, dim3(c->ns), dim3(256), 0, c->stream, t);
    if ((rc = launch_check("k_species"))) return rc;
    const int bt = ((c->npa + 63) / 64) * 64;
    hipLaunchKernelGGL(k_coef, dim3(c->nr * c->nt), dim3(bt), 0, c->stream, t);
    if ((rc = launch_check("k_coef"))) return rc;
    tend(c);
    c->tables_dirty = false;
    c->efield_dirty = true;
  }
  if (c->efield_dirty) {
    tbeg(c, RAMGPU_KERNEL_PREPARE);
    hipLaunchKernelGGL(k_coef_ij, dim3((c->nr * c->nt + 255) / 256), dim3(256), 0, c->stream, t);
    tend(c);
    if ((rc = launch_check("k_coef_ij"))) return rc;
    c->efield_dirty = false;
    c->ghost_dirty = true;
    c->cfl_valid = false;
  }
  if (c->ghost_dirty) {
    tbeg(c, RAMGPU_KERNEL_PREPARE);
    hipLaunchKernelGGL(k_ghost, dim3(c->ns), dim3(1024), 0, c->stream, t);
    tend(c);
    if ((rc = launch_check("k_ghost"))) return rc;
    c->ghost_dirty = false;
  }
  return RAMGPU_OK;
}

int ntiles(const ramgpu_ctx *c) {
  const int kw = 1 << c->kws;
  return (c->ne * c->npa + kw - 1) / kw;
}

int run_ij(ramgpu_ctx *c, const Program &p, int s0, int nsp) {
  const int nt = ntiles(c);
  tbeg(c, RAMGPU_KERNEL_IJ);
  hipLaunchKernelGGL(k_ij, dim3(nsp * nt), dim3(c->ij_threads), ij_lds(c), c->stream, c->t, p, s0,
                     nt, c->kws);
  tend(c);
  return launch_check("k_ij");
}

int run_kl(ramgpu_ctx *c, const Program &p, int s0, int nsp) {
  const int nb = nsp * (c->nr - 1) * c->nt;
  tbeg(c, RAMGPU_KERNEL_KL);
  hipLaunchKernelGGL(k_kl, dim3(nb), dim3(c->kl_threads), kl_lds(c), c->stream, c->t, p, s0);
  tend(c);
  return launch_check("k_kl");
}

int run_finalize(ramgpu_ctx *c, int cflmask, int summode, int s0, int nsp, double *elorc) {
  tbeg(c, RAMGPU_KERNEL_FINALIZE);
  hipLaunchKernelGGL(k_finalize, dim3(nsp), dim3(256), 0, c->stream, c->t, cflmask, summode, s0,
                     (c->nr - 1) * c->nt, ntiles(c), elorc);
  tend(c);
  return launch_check("k_finalize");
}

Program prog1(int op, int arg = 0) {
  Program p{};
  p.n = 1;
  p.op[0] = op;
  p.arg[0] = arg;
  return p;
}

int copy_dtdrift(ramgpu_ctx *c, int s, int which, double *out) {
  if (!out) return RAMGPU_OK;
  HIP_TRY(hipMemcpyAsync(out, c->t.dtdrift + (s - 1) * 4 + which, sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

int drift_op(ramgpu_ctx *c, int s, int op, int which, double *dtdrift) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  const Program p = prog1(op, 1);  // each call records its CFL limit, as the reference's
  rc = (op == OP_R || op == OP_P) ? run_ij(c, p, s - 1, 1) : run_kl(c, p, s - 1, 1);
  if (rc || (rc = run_finalize(c, 1 << which, 0, s - 1, 1, nullptr))) return rc;
  return copy_dtdrift(c, s, which, dtdrift);
}

int set_f64(ramgpu_ctx *c, double *dst, const double *src, size_t n, const char *name) {
  if (!src) return fail(RAMGPU_EINVAL, "%s is null", name);
  HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return RAMGPU_OK;
}

}  // namespace

extern "C" {

const char *ramgpu_last_error(void) { return g_err.c_str(); }

int ramgpu_create(ramgpu_ctx **out, int ns, int nr, int nt, int ne, int npa, int device) {
  if (!out) return fail(RAMGPU_EINVAL, "null output pointer");
  *out = nullptr;
  if (ns < 1 || ns > kMaxSpecies) return fail(RAMGPU_EINVAL, "NS=%d outside 1..%d", ns, kMaxSpecies);
  if (nr < 4 || nt < 4 || ne < 4 || npa < 4)
    return fail(RAMGPU_EINVAL, "grid %dx%dx%dx%d: every axis needs at least 4 points", nr, nt, ne, npa);
  if (npa > 1024) return fail(RAMGPU_EINVAL, "NPA=%d above 1024", npa);
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
  A(t.rrlz, nr); A(t.rsum, nr); A(t.rrsum, nr);
  A(t.cr, n2); A(t.bsr, n2); A(t.rbsr, n2); A(t.pa, n2); A(t.pce, n2); A(t.pbs, n2); A(t.rpbs, n2);
  A(t.drd1, n2); A(t.dpd1, n2); A(t.bij, n2); A(t.rbij, n2); A(t.drm1, n2); A(t.dpm1, n2); A(t.db2, n2);
  A(t.q1, n3); A(t.gpa, n3); A(t.fhs, n3); A(t.rfhs, n3);
  A(t.gr, n3); A(t.gp, n3); A(t.drd2, n3); A(t.dpd2, n3); A(t.db1, n3); A(t.di1, n3); A(t.hs, n3);
  A(t.rhs, n3); A(t.cmudot, n3); A(t.gmr, n3); A(t.gmp, n3); A(t.drm2, n3); A(t.dpm2, n3);
  A(t.di2, n3); A(t.bhs, n3); A(t.rbhs, n3); A(t.hdns, n3); A(t.invfnhs, n3);
  A(t.p4, (size_t)ns * ne); A(t.ekp, (size_t)ns * ne); A(t.mkp, (size_t)ns * ne);
  A(t.p2, (size_t)ns * nr * ne); A(t.edot, (size_t)ns * nr * ne); A(t.sigv, (size_t)ns * ne);
  A(t.atlos, (size_t)ns * nr * ne); A(t.rel, (size_t)ns * 8);
  A(t.carry, (size_t)ns * ne * npa);
  A(t.latl, (size_t)ns * nr * ne);
  t.maxblk = std::max((nr - 1) * nt, ne * npa);  // covers KL planes and IJ tiles (kw>=1)
  A(t.cflmax, (size_t)ns * (2 + ne + npa));
  A(t.dtdrift, (size_t)ns * 4);
  A(t.f2, nf2);
  A(c->d_stage, nf2);
  A(t.partials, (size_t)8 * ns * t.maxblk);
  A(t.energy, (size_t)ns * 4);
  A(c->d_elorc, ns);
#undef A
  t.ns = ns; t.nr = nr; t.nt = nt; t.ne = ne; t.npa = npa;
  t.rlz = c->d_rlz; t.ekev = c->d_ekev; t.ebnd = c->d_ebnd; t.we = c->d_we; t.de = c->d_de;
  t.grel = c->d_grel; t.grbnd = c->d_grbnd; t.v = c->d_v; t.mu = c->d_mu; t.wmu = c->d_wmu;
  t.dmu = c->d_dmu; t.upa = c->d_upa;
  t.fnhs = c->d_fnhs; t.fnis = c->d_fnis; t.bounhs = c->d_bounhs; t.bounis = c->d_bounis;
  t.hdns_f = c->d_hdnsf; t.didt = c->d_didt; t.dibndt = c->d_dibndt; t.bnes = c->d_bnes;
  t.dbdt = c->d_dbdt; t.vt = c->d_vt; t.eir = c->d_eir; t.eip = c->d_eip; t.mgnp = c->d_mgnp;
  t.fgeos = c->d_fgeos; t.bscale = c->d_bscale;
  t.dts = 5.0; t.betalim = 1.5; t.fraccfl = 0.8;
  t.rnt = 1.0f / (float)nt;
  t.rne = 1.0f / (float)ne;
  std::vector<double> ones(ns, 1.0);
  if (hipMemcpy(c->d_bscale, ones.data(), ns * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(RAMGPU_EHIP, "hipMemcpy failed"));
  // launch shapes: IJ tile width 2^kws chosen so the tile fits the LDS budget
  while (c->kws > 0 && ij_lds(c) > std::min<size_t>(c->lds_limit, 64 * 1024)) --c->kws;
  if (ij_lds(c) > c->lds_limit)
    return cleanup(fail(RAMGPU_EINVAL, "NR*NT=%d plane does not fit the device LDS", nr * nt));
  if (kl_lds(c) > c->lds_limit)
    return cleanup(fail(RAMGPU_EINVAL, "NE*NPA=%d plane does not fit the device LDS", ne * npa));
  (void)hipFuncSetAttribute((const void *)k_ij, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_limit);
  (void)hipFuncSetAttribute((const void *)k_kl, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_limit);
  *out = c;
  return RAMGPU_OK;
}

int ramgpu_destroy(ramgpu_ctx *c) {
  if (!c) return RAMGPU_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void *p : c->allocs) (void)hipFree(p);
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
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
  c->tables_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_params(ramgpu_ctx *c, double dts, double betalim, double fraccfl) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!(dts > 0)) return fail(RAMGPU_EINVAL, "DTs must be positive");
  c->t.dts = dts;
  c->t.betalim = betalim;
  c->t.fraccfl = fraccfl;
  c->tables_dirty = true;
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
  c->have_grid = true;
  c->tables_dirty = true;
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
                     (const double *)c->d_phi, (const double *)nullptr, kp, c->nr + 1, c->nt);
  if ((rc = launch_check("k_vs"))) return rc;
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

int ramgpu_driftr(ramgpu_ctx *c, int s, double *dt) { return drift_op(c, s, OP_R, 0, dt); }
int ramgpu_driftp(ramgpu_ctx *c, int s, double *dt) { return drift_op(c, s, OP_P, 1, dt); }
int ramgpu_drifte(ramgpu_ctx *c, int s, double *dt) { return drift_op(c, s, OP_E, 2, dt); }
int ramgpu_driftmu(ramgpu_ctx *c, int s, double *dt) { return drift_op(c, s, OP_MU, 3, dt); }

int ramgpu_driftend(ramgpu_ctx *c) { return check_ctx(c); }

int ramgpu_cepara(ramgpu_ctx *c, int s) { return ramgpu_driftpara(c, s); }

int ramgpu_charexchange(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  if ((rc = run_kl(c, prog1(OP_CEX), s - 1, 1))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

int ramgpu_atmol(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  if ((rc = run_kl(c, prog1(OP_ATM), s - 1, 1))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

int ramgpu_sumrc(ramgpu_ctx *c, int s, double *setrc, double *elorc) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = ready(c))) return rc;
  // slot 0 partials: every KL block writes one (rows I>=2); zero the rest
  HIP_TRY(hipMemsetAsync(c->t.partials + (size_t)(s - 1) * c->t.maxblk, 0,
                         c->t.maxblk * sizeof(double), c->stream));
  if ((rc = run_kl(c, prog1(OP_SUM, 0), s - 1, 1))) return rc;
  if ((rc = run_finalize(c, 0, 1, s - 1, 1, c->d_elorc))) return rc;
  double e[4], el;
  HIP_TRY(hipMemcpyAsync(e, c->t.energy + (s - 1) * 4, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&el, c->d_elorc + (s - 1), sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (setrc) *setrc = e[0];
  if (elorc) *elorc = el;
  return RAMGPU_OK;
}

// One ram_run call for all species: IJ(R,P) -> KL(E,Mu,loss chain,Mu,E) ->
// IJ(P,R,SUMRC,seam/mask) -> finalize.  The CFL limits depend on the tables
// only, so they are recorded in the first sweep of each kind after a table
// change and reused until the next one.
static int step_once(ramgpu_ctx *c) {
  int rc;
  if ((rc = prepare(c))) return rc;
  const int rec = c->cfl_valid ? 0 : 1;
  Program a{};
  a.n = 2;
  a.op[0] = OP_R;
  a.arg[0] = rec;
  a.op[1] = OP_P;
  a.arg[1] = rec;
  if ((rc = run_ij(c, a, 0, c->ns))) return rc;
  Program k{};
  k.n = 5;
  k.op[0] = OP_E;
  k.arg[0] = rec;
  k.op[1] = OP_MU;
  k.arg[1] = rec;
  k.op[2] = OP_LOSS;
  k.arg[2] = 0;
  k.op[3] = OP_MU;
  k.op[4] = OP_E;
  if ((rc = run_kl(c, k, 0, c->ns))) return rc;
  Program b{};
  b.n = 4;
  b.op[0] = OP_P;
  b.op[1] = OP_R;
  b.op[2] = OP_SUM;
  b.arg[2] = 5;
  b.op[3] = OP_SEAM;
  if ((rc = run_ij(c, b, 0, c->ns))) return rc;
  if ((rc = run_finalize(c, rec ? 15 : 0, 2, 0, c->ns, nullptr))) return rc;
  c->cfl_valid = true;
  return RAMGPU_OK;
}

static int fill_out(ramgpu_ctx *c, ramgpu_species_out *out) {
  if (!out) return RAMGPU_OK;
  std::vector<double> e(c->ns * 4), d(c->ns * 4);
  HIP_TRY(hipMemcpyAsync(e.data(), c->t.energy, e.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(d.data(), c->t.dtdrift, d.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
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
  }
  return RAMGPU_OK;
}

int ramgpu_ram_run(ramgpu_ctx *c, int nsteps, ramgpu_species_out *out) {
  int rc = check_ctx(c);
  if (rc || (rc = ready(c))) return rc;
  if (nsteps < 0) return fail(RAMGPU_EINVAL, "nsteps < 0");
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (int n = 0; n < nsteps; ++n)
    if ((rc = step_once(c))) return rc;
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  return fill_out(c, out);
}

int ramgpu_ram_run_schedule(ramgpu_ctx *c, int nsteps, const double *kp, const double *bscale,
                            ramgpu_species_out *out) {
  int rc = check_ctx(c);
  if (rc || (rc = ready(c))) return rc;
  if (nsteps < 0) return fail(RAMGPU_EINVAL, "nsteps < 0");
  if (!kp) return fail(RAMGPU_EINVAL, "kp schedule is null");
  if (!c->have_lzphi) return fail(RAMGPU_ESTATE, "ramgpu_set_kp(lz, phi) must precede a Kp schedule");
  // the whole schedule is staged on the device before the timed steps
  double *d_kp = nullptr, *d_bs = nullptr;
  HIP_TRY(hipMalloc(&d_kp, std::max(nsteps, 1) * sizeof(double)));
  HIP_TRY(hipMemcpyAsync(d_kp, kp, nsteps * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (bscale) {
    HIP_TRY(hipMalloc(&d_bs, std::max(nsteps, 1) * c->ns * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(d_bs, bscale, (size_t)nsteps * c->ns * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (int n = 0; n < nsteps && rc == RAMGPU_OK; ++n) {
    tbeg(c, RAMGPU_KERNEL_VS);
    hipLaunchKernelGGL(k_vs, dim3(1), dim3(256), 0, c->stream, c->d_vt, (const double *)c->d_lz,
                       (const double *)c->d_phi, (const double *)(d_kp + n), 0.0, c->nr + 1, c->nt);
    tend(c);
    if ((rc = launch_check("k_vs"))) break;
    c->efield_dirty = true;
    if (d_bs) {
      if (hipMemcpyAsync(c->d_bscale, d_bs + (size_t)n * c->ns, c->ns * sizeof(double),
                         hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
        rc = fail(RAMGPU_EHIP, "boundary scale copy failed");
        break;
      }
      c->ghost_dirty = true;
    }
    rc = step_once(c);
  }
  (void)hipEventRecord(c->ev1, c->stream);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d_kp);
  if (d_bs) (void)hipFree(d_bs);
  if (rc) return rc;
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
