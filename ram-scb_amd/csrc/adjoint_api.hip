// adjoint_api.hip -- C ABI of reverse-mode differentiation (include/ramgpu.h).
//
// Operator level: ramgpu_<op>_bwd is the backward of the autograd-style pair
// (ramgpu_<op>, ramgpu_<op>_bwd): evaluated at the context's current F2 (the
// operator's input), it maps dL/d(output) to dL/d(input) in place and
// accumulates dL/dFGEOS (DRIFTR).  Step level: ramgpu_ram_run_bwd, the same
// for one whole ram_run step.  Run level: ramgpu_adjoint_run, the gradient of
// a virtual-satellite flux misfit over a multi-step run with respect to the
// initial F2 and to the boundary FGEOS of every step, with checkpoint-every-k
// recomputation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ctx.h"

using namespace ramgpu;
using namespace ramgpu_host;

namespace ramgpu_host {

template <int BT>
void set_bwd_lds_limit_t(size_t limit) {
  const void *fns[] = {(const void *)k_plane_bwd<SEQ_E, BT>,   (const void *)k_plane_bwd<SEQ_MU, BT>,
                       (const void *)k_plane_bwd<SEQ_CEX, BT>, (const void *)k_plane_bwd<SEQ_ATM, BT>,
                       (const void *)k_plane_bwd<SEQ_LOSS, BT>, (const void *)k_plane_bwd<SEQ_STEP, BT>};
  for (const void *f : fns) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)limit);
}

// cells per thread of the plane adjoints (cell n of a thread: tid + n*512)
// threads of the plane adjoints: cell n of a thread is tid + n*threads, at
// most 8 cells per thread (NE*NPA <= 8192)
int plane_bwd_threads(int ne, int npa) { return (ne * npa <= 8 * kPlaneThreads) ? kPlaneThreads : kPlaneMaxThreads; }

template <int BT>
void launch_plane_bwd_t(int seq, dim3 grid, size_t lds, hipStream_t st, const Launch &L, const double *x,
                        double *adj) {
  const dim3 blk(BT);
  switch (seq) {
    case SEQ_E: hipLaunchKernelGGL((k_plane_bwd<SEQ_E, BT>), grid, blk, lds, st, L, x, adj); break;
    case SEQ_MU: hipLaunchKernelGGL((k_plane_bwd<SEQ_MU, BT>), grid, blk, lds, st, L, x, adj); break;
    case SEQ_CEX: hipLaunchKernelGGL((k_plane_bwd<SEQ_CEX, BT>), grid, blk, lds, st, L, x, adj); break;
    case SEQ_ATM: hipLaunchKernelGGL((k_plane_bwd<SEQ_ATM, BT>), grid, blk, lds, st, L, x, adj); break;
    case SEQ_LOSS: hipLaunchKernelGGL((k_plane_bwd<SEQ_LOSS, BT>), grid, blk, lds, st, L, x, adj); break;
    default: hipLaunchKernelGGL((k_plane_bwd<SEQ_STEP, BT>), grid, blk, lds, st, L, x, adj); break;
  }
}

void set_plane_bwd_lds_limit(size_t limit) {
  set_bwd_lds_limit_t<kPlaneThreads>(limit);
  set_bwd_lds_limit_t<kPlaneMaxThreads>(limit);
}

void launch_plane_bwd(int seq, int threads, dim3 grid, size_t lds, hipStream_t st, const Launch &L, const double *x,
                      double *adj) {
  if (threads <= kPlaneThreads)
    launch_plane_bwd_t<kPlaneThreads>(seq, grid, lds, st, L, x, adj);
  else
    launch_plane_bwd_t<kPlaneMaxThreads>(seq, grid, lds, st, L, x, adj);
}

// host Fortran-layout F2 <-> device-layout buffer (through the context's stage)
int upload_f2(ramgpu_ctx *c, double *dst, const double *host) {
  const size_t n = (size_t)c->ns * c->nr * c->nt * c->ne * c->npa;
  HIP_TRY(hipMemcpyAsync(c->d_stage, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_f2_in, dim3(1024), dim3(256), 0, c->stream, dst, (const double *)c->d_stage, c->ns, c->nr,
                     c->nt, c->ne, c->npa);
  return launch_check("k_f2_in");
}

int download_f2(ramgpu_ctx *c, double *host, const double *src) {
  const size_t n = (size_t)c->ns * c->nr * c->nt * c->ne * c->npa;
  hipLaunchKernelGGL(k_f2_out, dim3(1024), dim3(256), 0, c->stream, c->d_stage, src, c->ns, c->nr, c->nt, c->ne,
                     c->npa);
  int rc = launch_check("k_f2_out");
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(host, c->d_stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RAMGPU_OK;
}

}  // namespace ramgpu_host

namespace {

size_t nf2(const ramgpu_ctx *c) { return (size_t)c->ns * c->nr * c->nt * c->ne * c->npa; }
size_t nfg(const ramgpu_ctx *c) { return (size_t)c->ns * c->nt * c->ne * c->npa; }

struct DevBuf {
  std::vector<void *> p;
  ~DevBuf() {
    for (void *q : p) (void)hipFree(q);
  }
  template <class T>
  int alloc(T **out, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
      return fail(RAMGPU_ENOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
    p.push_back(q);
    *out = (T *)q;
    return RAMGPU_OK;
  }
};

// an allocation the caller can do without (no error is recorded when it fails)
template <class T>
bool try_alloc(DevBuf &b, T **out, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  b.p.push_back(q);
  *out = (T *)q;
  return true;
}

// The second stream of ramgpu_adjoint_run and the events that order the two
// streams: fwd (the forward pass done), rec[g%2] (segment g recomputed),
// bwd[g%2] (segment g's adjoint done).  Declared after the run's buffers, so
// that on every exit the second stream has drained before they are freed.
struct Overlap {
  ramgpu_ctx *c;
  hipEvent_t fwd = nullptr, rec[2] = {nullptr, nullptr}, bwd[2] = {nullptr, nullptr};
  explicit Overlap(ramgpu_ctx *cc) : c(cc) {}
  bool init() {
    if (!c->stream2 && hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
      c->stream2 = nullptr;
      (void)hipGetLastError();
      return false;
    }
    for (hipEvent_t *e : {&fwd, &rec[0], &rec[1], &bwd[0], &bwd[1]})
      if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
        *e = nullptr;
        (void)hipGetLastError();
        return false;
      }
    return true;
  }
  ~Overlap() {
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    for (hipEvent_t e : {fwd, rec[0], rec[1], bwd[0], bwd[1]})
      if (e) (void)hipEventDestroy(e);
  }
};

int fill(ramgpu_ctx *c, double *y, double v, size_t n) {
  hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, c->stream, y, v, n);
  return launch_check("k_fill");
}

int copy_d2d(ramgpu_ctx *c, double *dst, const double *src, size_t n) {
  HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return RAMGPU_OK;
}

size_t bwd_lds(const ramgpu_ctx *c) {
  return plane_bwd_lds_bytes(c->ne, c->npa);
}

// backward kernels over species s0..s0+nsp-1
int bwd_r(ramgpu_ctx *c, int final_sweep, const double *x, double *adj, double *bbar, int s0, int nsp) {
  const Launch L = launch_of(c, s0, nsp, blocks_r(c->t));
  tbeg(c, RAMGPU_KERNEL_R_BWD);
  if (final_sweep)
    hipLaunchKernelGGL(k_r_bwd<1>, dim3(L.nv), dim3(kLineThreads), radial_lds_bytes(c->nr, c->ne, c->npa), c->stream,
                       L, x, adj, bbar);
  else
    hipLaunchKernelGGL(k_r_bwd<0>, dim3(L.nv), dim3(kLineThreads), radial_lds_bytes(c->nr, c->ne, c->npa), c->stream,
                       L, x, adj, bbar);
  tend(c);
  return launch_check("k_r_bwd");
}

int bwd_p(ramgpu_ctx *c, const double *x, double *adj, int s0, int nsp) {
  const Launch L = launch_of(c, s0, nsp, blocks_p(c->t));
  tbeg(c, RAMGPU_KERNEL_P_BWD);
  hipLaunchKernelGGL(k_p_bwd, dim3(L.nv), dim3(kLineThreads), azim_lds_bytes(c->nt, c->ne, c->npa), c->stream, L, x,
                     adj);
  tend(c);
  return launch_check("k_p_bwd");
}

int bwd_plane(ramgpu_ctx *c, int seq, const double *x, double *adj, int s0, int nsp) {
  if (bwd_lds(c) > c->lds_limit) return fail(RAMGPU_EINVAL, "plane adjoint does not fit the device LDS");
  if (!c->bwd_lds_set) {
    set_plane_bwd_lds_limit(c->lds_limit);
    c->bwd_lds_set = true;
  }
  const Launch L = launch_of(c, s0, nsp, blocks_kl(c->t));
  const size_t lds = bwd_lds(c);
  tbeg(c, RAMGPU_KERNEL_PLANE_BWD);
  launch_plane_bwd(seq, plane_bwd_threads(c->ne, c->npa), dim3(L.nv), lds, c->stream, L, x, adj);
  tend(c);
  return launch_check("k_plane_bwd");
}

// One ram_run step (all species) from tape slot T_R1 (the step's input,
// filled by the caller) to y, every operator writing its output into the
// next slot, so that the tape holds the input of each operator.  The
// kernels are those of the fused step with the same arithmetic (the plane
// kernel also storing its phase inputs): the state, the SUMRC bookkeeping
// and the CFL limits are those of step_once.  With bookkeeping false (a
// segment recomputed for its tape, whose SUMRC state and CFL limits are
// discarded) only F2 and the tape are formed.
int taped_step(ramgpu_ctx *c, double *tape, double *y, bool bookkeeping = true) {
  int rc;
  if ((rc = prepare(c))) return rc;
  const int rec = (c->cfl_valid || !bookkeeping) ? 0 : 1, ns = c->ns;
  const size_t n = nf2(c);
  auto x = [&](int slot) { return tape + (size_t)slot * n; };
  if ((rc = run_r(c, rec, 0, ns, x(T_R1), x(T_P1))) || (rc = run_p(c, rec, 0, ns, x(T_P1), x(T_E1))) ||
      (rc = run_plane(c, SEQ_TAPE, rec, 0, ns, x(T_E1), x(T_P2))) ||
      (rc = run_p(c, 0, 0, ns, x(T_P2), x(T_R2))) || (rc = run_r(c, 2, 0, ns, x(T_R2), y)) ||
      (bookkeeping && (rc = run_finalize(c, rec ? 15 : 0, 2, 0, ns))))
    return rc;
  if (bookkeeping) c->cfl_valid = true;
  return RAMGPU_OK;
}

// Adjoint of one taped step: adj (dL/d output) -> dL/d input, in place;
// bbar (may be NULL) accumulates dL/dFGEOS of the step.  The tables must be
// those of the step (its VT and boundary scale).
int step_bwd(ramgpu_ctx *c, const double *tape, double *adj, double *bbar) {
  const size_t n = nf2(c);
  const int ns = c->ns;
  auto x = [&](int slot) { return tape + (size_t)slot * n; };
  int rc;
  if ((rc = bwd_r(c, 1, x(T_R2), adj, bbar, 0, ns)) || (rc = bwd_p(c, x(T_P2), adj, 0, ns)) ||
      (rc = bwd_plane(c, SEQ_STEP, x(T_E1), adj, 0, ns)) || (rc = bwd_p(c, x(T_P1), adj, 0, ns)) ||
      (rc = bwd_r(c, 0, x(T_R1), adj, bbar, 0, ns)))
    return rc;
  return RAMGPU_OK;
}

// Operator-level backward at the current F2: which 0 R, 1 P, 2 E, 3 Mu,
// 4 CEX, 5 ATMOL.
int op_bwd(ramgpu_ctx *c, int s, int which, double *adj, double *fgeos_adj) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s)) || (rc = prepare(c))) return rc;
  if (!adj) return fail(RAMGPU_EINVAL, "adj is null");
  DevBuf b;
  double *da = nullptr, *db = nullptr;
  if ((rc = b.alloc(&da, nf2(c)))) return rc;
  if ((rc = upload_f2(c, da, adj))) return rc;
  if (fgeos_adj && which == 0) {
    if ((rc = b.alloc(&db, nfg(c)))) return rc;
    HIP_TRY(hipMemcpyAsync(db, fgeos_adj, nfg(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  const double *x = c->t.f2;
  switch (which) {
    case 0: rc = bwd_r(c, 0, x, da, db, s - 1, 1); break;
    case 1: rc = bwd_p(c, x, da, s - 1, 1); break;
    case 2: rc = bwd_plane(c, SEQ_E, x, da, s - 1, 1); break;
    case 3: rc = bwd_plane(c, SEQ_MU, x, da, s - 1, 1); break;
    case 4: rc = bwd_plane(c, SEQ_CEX, x, da, s - 1, 1); break;
    default: rc = bwd_plane(c, SEQ_ATM, x, da, s - 1, 1); break;
  }
  if (rc || (rc = download_f2(c, adj, da))) return rc;
  if (db) {
    HIP_TRY(hipMemcpyAsync(fgeos_adj, db, nfg(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return RAMGPU_OK;
}

// The context's per-step inputs of a schedule (as ramgpu_ram_run_schedule).
int set_step_inputs(ramgpu_ctx *c, const double *d_kp, const double *d_bs, int n) {
  if (d_kp) {
    return apply_step_inputs(c, d_kp + n, d_bs ? d_bs + (size_t)n * c->ns : nullptr);
  } else if (d_bs) {
    HIP_TRY(hipMemcpyAsync(c->d_bscale, d_bs + (size_t)n * c->ns, c->ns * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream));
    c->ghost_dirty = true;
  }
  return RAMGPU_OK;
}

}  // namespace

extern "C" {

int ramgpu_driftr_bwd(ramgpu_ctx *c, int s, double *adj, double *fgeos_adj) { return op_bwd(c, s, 0, adj, fgeos_adj); }
int ramgpu_driftp_bwd(ramgpu_ctx *c, int s, double *adj) { return op_bwd(c, s, 1, adj, nullptr); }
int ramgpu_drifte_bwd(ramgpu_ctx *c, int s, double *adj) { return op_bwd(c, s, 2, adj, nullptr); }
int ramgpu_driftmu_bwd(ramgpu_ctx *c, int s, double *adj) { return op_bwd(c, s, 3, adj, nullptr); }
int ramgpu_charexchange_bwd(ramgpu_ctx *c, int s, double *adj) { return op_bwd(c, s, 4, adj, nullptr); }
int ramgpu_atmol_bwd(ramgpu_ctx *c, int s, double *adj) { return op_bwd(c, s, 5, adj, nullptr); }

int ramgpu_ram_run_bwd(ramgpu_ctx *c, double *adj, double *fgeos_adj) {
  int rc = check_ctx(c);
  if (rc || (rc = ready(c)) || (rc = forward_only(c)) || (rc = prepare(c))) return rc;
  if (!adj) return fail(RAMGPU_EINVAL, "adj is null");
  const size_t n = nf2(c);
  DevBuf b;
  double *tape = nullptr, *da = nullptr, *db = nullptr, *save_f2 = nullptr, *save_e = nullptr;
  if ((rc = b.alloc(&tape, T_N * n)) || (rc = b.alloc(&da, n)) || (rc = b.alloc(&save_f2, n)) ||
      (rc = b.alloc(&save_e, (size_t)c->ns * 7)))
    return rc;
  if (fgeos_adj) {
    if ((rc = b.alloc(&db, nfg(c)))) return rc;
    HIP_TRY(hipMemcpyAsync(db, fgeos_adj, nfg(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  // the step runs from a copy of the state; the SUMRC state is restored
  const size_t ne4 = (size_t)c->ns * 4, nx3 = (size_t)c->ns * 3;
  if ((rc = copy_d2d(c, tape, c->t.f2, n)) || (rc = copy_d2d(c, save_e, c->t.energy, ne4)) ||
      (rc = copy_d2d(c, save_e + ne4, c->t.lsx, nx3)) || (rc = taped_step(c, tape, save_f2)) ||
      (rc = copy_d2d(c, c->t.energy, save_e, ne4)) || (rc = copy_d2d(c, c->t.lsx, save_e + ne4, nx3)) ||
      (rc = upload_f2(c, da, adj)) || (rc = step_bwd(c, tape, da, db)) || (rc = download_f2(c, adj, da)))
    return rc;
  if (db) {
    HIP_TRY(hipMemcpyAsync(fgeos_adj, db, nfg(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return RAMGPU_OK;
}

int ramgpu_adjoint_run(ramgpu_ctx *c, int nsteps, int ckpt_every, const double *kp, const double *bscale,
                       int npts, const int *pts, const double *w, const double *target, double *loss,
                       double *grad_f2, double *grad_fgeos) {
  int rc = check_ctx(c);
  if (rc || (rc = ready(c)) || (rc = forward_only(c))) return rc;
  if (nsteps < 1 || ckpt_every < 1) return fail(RAMGPU_EINVAL, "nsteps and ckpt_every must be >= 1");
  if (npts < 1 || !pts || !w || !target) return fail(RAMGPU_EINVAL, "misfit points, weights and targets required");
  if (kp && !c->have_lzphi) return fail(RAMGPU_ESTATE, "ramgpu_set_kp(lz, phi) must precede a Kp schedule");
  for (int p = 0; p < npts; ++p) {
    const int *q = pts + 4 * p;
    if (q[0] < 1 || q[0] > c->ns || q[1] < 1 || q[1] > c->nr || q[2] < 1 || q[2] > c->nt || q[3] < 1 ||
        q[3] > c->ne)
      return fail(RAMGPU_EINVAL, "misfit point %d outside the grid", p);
  }
  const size_t n = nf2(c), ng = nfg(c);
  const int nseg = (nsteps + ckpt_every - 1) / ckpt_every;
  const int seglen = std::min(ckpt_every, nsteps);
  const size_t tape_n = (size_t)seglen * T_N * n;
  DevBuf b;
  double *ckpt, *tape, *lam, *bbar = nullptr, *dw, *dtarget, *lossv, *scratch_loss, *final_f2, *final_e;
  double *d_kp = nullptr, *d_bs = nullptr;
  int *dpts;
  if ((rc = b.alloc(&ckpt, (size_t)nseg * n)) || (rc = b.alloc(&tape, tape_n)) ||
      (rc = b.alloc(&lam, n)) || (rc = b.alloc(&dw, (size_t)npts * c->npa)) ||
      (rc = b.alloc(&dtarget, (size_t)nsteps * npts)) || (rc = b.alloc(&lossv, (size_t)nsteps * npts)) ||
      (rc = b.alloc(&scratch_loss, npts)) || (rc = b.alloc(&final_f2, n)) ||
      (rc = b.alloc(&final_e, (size_t)c->ns * 4)) || (rc = b.alloc(&dpts, (size_t)4 * npts)))
    return rc;
  if (grad_fgeos && (rc = b.alloc(&bbar, (size_t)nsteps * ng))) return rc;
  if (kp) {
    if ((rc = b.alloc(&d_kp, nsteps))) return rc;
    HIP_TRY(hipMemcpyAsync(d_kp, kp, nsteps * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  if (bscale) {
    if ((rc = b.alloc(&d_bs, (size_t)nsteps * c->ns))) return rc;
    HIP_TRY(hipMemcpyAsync(d_bs, bscale, (size_t)nsteps * c->ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(dpts, pts, (size_t)4 * npts * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dw, w, (size_t)npts * c->npa * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dtarget, target, (size_t)nsteps * npts * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const int mb = std::max(npts, 1);  // one block per point
  auto misfit = [&](const double *state, int step, double *lout, double *lm) {
    hipLaunchKernelGGL(k_misfit, dim3(mb), dim3(64), 0, c->stream, state, c->nr, c->nt, c->ne, c->npa, npts,
                       (const int *)dpts, (const double *)dw, (const double *)(dtarget + (size_t)step * npts), lout,
                       lm);
    return launch_check("k_misfit");
  };
  // Two tapes when the step tables stay fixed over the run (no Kp or
  // boundary-scale schedule): segment g-1 is then recomputed into one of
  // them on the second stream while segment g's adjoint runs from the
  // other.  The two streams share only the tables, which neither writes,
  // and run the kernels of the serial order on the same data, so the
  // results are the same bits.  Without the memory for a second tape, or
  // with a schedule, the segments go one after the other on one stream.
  Overlap ov(c);
  double *tapes[2] = {tape, tape}, *yend[2] = {nullptr, nullptr};
  const bool overlap = !kp && !bscale && nseg >= 2 && ov.init() && try_alloc(b, &tapes[1], tape_n) &&
                       try_alloc(b, &yend[0], n) && try_alloc(b, &yend[1], n);
  auto buf = [&](int g) { return overlap ? g % 2 : 0; };
  // step s of segment g (steps n0..n1-1): its tape, the input it holds, its output
  auto tape_at = [&](int s, int g) { return tapes[buf(g)] + (size_t)(s - g * ckpt_every) * T_N * n; };
  auto tape_in = [&](int s, int g) { return tape_at(s, g) + (size_t)T_R1 * n; };
  auto step_out = [&](int s, int g, int n1) -> double * {
    if (s < n1 - 1) return tape_in(s + 1, g);
    return (overlap && g < nseg - 1) ? yend[buf(g)] : c->t.f2;  // the next segment's checkpoint
  };
  const int nlast0 = (nseg - 1) * ckpt_every;
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  // forward, keeping a checkpoint at the start of every segment; the last
  // segment is run taped at once, so the backward need not recompute it.
  // A plain step's finalize runs with the next step's first launch; none is
  // left pending when the pass ends, on error too.
  auto forward = [&]() -> int {
    int r;
    for (int s = 0; s < nsteps; ++s) {
      if (s % ckpt_every == 0 && s < nlast0 && (r = copy_d2d(c, ckpt + (size_t)(s / ckpt_every) * n, c->t.f2, n)))
        return r;
      if ((r = set_step_inputs(c, d_kp, d_bs, s))) return r;
      if (s < nlast0) {
        int cflmask = 0;
        if ((r = step_body(c, &cflmask))) return r;
        c->fin_mask = cflmask;
        c->fin_mode = 2;
        if ((r = misfit(c->t.f2, s, lossv + (size_t)s * npts, nullptr))) return r;
      } else {
        const int g = nseg - 1;
        if (s == nlast0 && (r = copy_d2d(c, tape_in(s, g), c->t.f2, n))) return r;
        double *y = step_out(s, g, nsteps);
        if ((r = taped_step(c, tape_at(s, g), y)) || (r = misfit(y, s, lossv + (size_t)s * npts, nullptr)))
          return r;
      }
    }
    return RAMGPU_OK;
  };
  rc = forward();
  const int rf = flush_finalize(c);
  if (rc == RAMGPU_OK) rc = rf;
  if (rc || (rc = copy_d2d(c, final_f2, c->t.f2, n)) || (rc = copy_d2d(c, final_e, c->t.energy, (size_t)c->ns * 4)))
    return rc;
  // backward, segment by segment: recompute the segment with its tape (on
  // the context's current stream), then walk its steps in reverse
  if ((rc = fill(c, lam, 0.0, n))) return rc;
  if (bbar && (rc = fill(c, bbar, 0.0, (size_t)nsteps * ng))) return rc;
  auto recompute = [&](int g) -> int {
    const int n0 = g * ckpt_every, n1 = std::min(nsteps, n0 + ckpt_every);
    int r;
    if ((r = copy_d2d(c, tape_in(n0, g), ckpt + (size_t)g * n, n))) return r;
    for (int s = n0; s < n1; ++s)
      if ((r = set_step_inputs(c, d_kp, d_bs, s)) || (r = taped_step(c, tape_at(s, g), step_out(s, g, n1), false)))
        return r;
    return RAMGPU_OK;
  };
  auto backward = [&](int g) -> int {
    const int n0 = g * ckpt_every, n1 = std::min(nsteps, n0 + ckpt_every);
    int r;
    for (int s = n1 - 1; s >= n0; --s)
      if ((r = misfit(step_out(s, g, n1), s, scratch_loss, lam)) || (r = set_step_inputs(c, d_kp, d_bs, s)) ||
          (r = prepare(c)) || (r = step_bwd(c, tape_at(s, g), lam, bbar ? bbar + (size_t)s * ng : nullptr)))
        return r;
    return RAMGPU_OK;
  };
  if (!overlap) {
    for (int g = nseg - 1; g >= 0; --g)
      if ((g < nseg - 1 && (rc = recompute(g))) || (rc = backward(g))) return rc;
  } else {
    // recompute(g) on the second stream once `after` has passed; its tape is
    // that of segment g+2, whose adjoint must be done
    auto on_side = [&](int g, hipEvent_t after) -> int {
      HIP_TRY(hipStreamWaitEvent(c->stream2, after, 0));
      hipStream_t main = c->stream;
      c->stream = c->stream2;
      const int r = recompute(g);
      c->stream = main;
      if (r) return r;
      HIP_TRY(hipEventRecord(ov.rec[g % 2], c->stream2));
      return RAMGPU_OK;
    };
    HIP_TRY(hipEventRecord(ov.fwd, c->stream));
    if ((rc = on_side(nseg - 2, ov.fwd))) return rc;
    for (int g = nseg - 1; g >= 0; --g) {
      if (g < nseg - 1) HIP_TRY(hipStreamWaitEvent(c->stream, ov.rec[g % 2], 0));
      if ((rc = backward(g))) return rc;
      if (g >= 2) {
        HIP_TRY(hipEventRecord(ov.bwd[g % 2], c->stream));
        if ((rc = on_side(g - 2, ov.bwd[g % 2]))) return rc;
      }
    }
  }
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  // results; the context is left in the state after the run
  if ((rc = copy_d2d(c, c->t.f2, final_f2, n)) || (rc = copy_d2d(c, c->t.energy, final_e, (size_t)c->ns * 4)))
    return rc;
  if (nsteps > 0 && (rc = set_step_inputs(c, d_kp, d_bs, nsteps - 1))) return rc;
  std::vector<double> lv((size_t)nsteps * npts);
  HIP_TRY(hipMemcpyAsync(lv.data(), lossv, lv.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (loss) {
    double tot = 0.0;
    for (double v : lv) tot += v;
    *loss = tot;
  }
  if (grad_f2 && (rc = download_f2(c, grad_f2, lam))) return rc;
  if (grad_fgeos) {
    HIP_TRY(hipMemcpyAsync(grad_fgeos, bbar, (size_t)nsteps * ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return RAMGPU_OK;
}

}  // extern "C"
