// batch.hip -- ensembles: many independent problems stepped together.
//
// The members are ordinary contexts on one device, each with its own grid
// sizes, species, fields and F2.  Every kernel of the split step is ONE
// launch over all members: a job table lists (member, species, first block)
// units and a per-block map sends each virtual block to its unit, so members
// of different sizes share a grid.  The launches of one step are captured
// once into a HIP graph and replayed per step; the Kp / boundary-scale
// schedule is read on the device through a step counter, so every replay is
// the same graph.  An autotuner picks, per kernel, how many resident blocks
// loop over the virtual blocks (persistent grid) or one block per unit.
//
// A member stepped in an ensemble runs exactly the kernels and arithmetic of
// its own ramgpu_ram_run / ramgpu_ram_run_schedule, so its results are
// bit-identical to stepping it alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ctx.h"

using namespace ramgpu;
using namespace ramgpu_host;

namespace {

enum BatchKind { BK_SCHED, BK_COEFIJ, BK_GHOST, BK_R, BK_P, BK_PLANE, BK_FIN, BK_N };

struct Plan {
  Job *jobs = nullptr;
  int *vmap = nullptr;
  int nv = 0;
  int cap = 0;  // 0: one block per virtual block; else the persistent grid size
};

}  // namespace

struct ramgpu_batch {
  int device = 0;
  std::vector<ramgpu_ctx *> m;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Tab *d_tabs = nullptr;
  Plan plan[BK_N];
  size_t plane_lds = 0;
  int plane_maxc = 8;  // the largest any member needs
  int sum_ns = 0;
  std::vector<int> bs_off;  // first species of member m in a schedule row
  int *d_bsoff = nullptr;
  int *d_step = nullptr;
  double *d_kp = nullptr, *d_bs = nullptr;
  int sched_cap = 0;  // steps the schedule buffers hold
  bool use_graph = true;
  hipGraphExec_t exec[2] = {nullptr, nullptr};  // [0] plain step, [1] scheduled step
  std::vector<void *> allocs;
  int ncu = 256;
};

namespace {

int batch_alloc(ramgpu_batch *b, void **p, size_t bytes) {
  if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess)
    return fail(RAMGPU_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
  b->allocs.push_back(*p);
  return RAMGPU_OK;
}

int blocks_of(int kind, const ramgpu_ctx *c) {
  switch (kind) {
    case BK_COEFIJ: return (c->nr * c->nt + 255) / 256;
    case BK_R: return blocks_r(c->t);
    case BK_P: return blocks_p(c->t);
    case BK_PLANE: return blocks_kl(c->t);
    default: return 1;
  }
}

bool per_member(int kind) { return kind == BK_SCHED || kind == BK_COEFIJ; }

// Job table and block map of one kernel kind, members in order.
int build_plan(ramgpu_batch *b, int kind) {
  std::vector<Job> jobs;
  std::vector<int> vmap;
  for (size_t mi = 0; mi < b->m.size(); ++mi) {
    const ramgpu_ctx *c = b->m[mi];
    const int nsp = per_member(kind) ? 1 : c->ns;
    const int nb = blocks_of(kind, c);
    for (int s = 0; s < nsp; ++s) {
      Job j{(int)mi, s, (int)vmap.size(), nb};
      for (int x = 0; x < nb; ++x) vmap.push_back((int)jobs.size());
      jobs.push_back(j);
    }
  }
  Plan &p = b->plan[kind];
  int rc;
  if ((rc = batch_alloc(b, (void **)&p.jobs, jobs.size() * sizeof(Job))) ||
      (rc = batch_alloc(b, (void **)&p.vmap, vmap.size() * sizeof(int))))
    return rc;
  HIP_TRY(hipMemcpy(p.jobs, jobs.data(), jobs.size() * sizeof(Job), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(p.vmap, vmap.data(), vmap.size() * sizeof(int), hipMemcpyHostToDevice));
  p.nv = (int)vmap.size();
  return RAMGPU_OK;
}

Launch launch_for(const ramgpu_batch *b, int kind) {
  const Plan &p = b->plan[kind];
  Launch L{};
  L.tabs = b->d_tabs;
  L.jobs = p.jobs;
  L.vmap = p.vmap;
  L.nv = p.nv;
  return L;
}

unsigned grid_for(const ramgpu_batch *b, int kind) {
  const Plan &p = b->plan[kind];
  return (unsigned)((p.cap > 0 && p.cap < p.nv) ? p.cap : p.nv);
}

void destroy_graphs(ramgpu_batch *b) {
  for (auto &e : b->exec)
    if (e) {
      (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
}

// Bring every member's tables up to date on its own stream and refresh the
// device copies of the member descriptors.
int sync_members(ramgpu_batch *b) {
  int rc;
  std::vector<Tab> tabs;
  for (ramgpu_ctx *c : b->m) {
    if ((rc = check_ctx(c)) || (rc = prepare(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    tabs.push_back(c->t);
  }
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipMemcpy(b->d_tabs, tabs.data(), tabs.size() * sizeof(Tab), hipMemcpyHostToDevice));
  return RAMGPU_OK;
}

// The launches of one ensemble step.  sched: VT / boundary scale of the step
// from the device schedule, then the VT-dependent tables and the ghost carry.
// CFL limits are recorded every step (the same values a member's own
// ram_run reports).
int enqueue_step(ramgpu_batch *b, bool sched) {
  hipStream_t st = b->stream;
  const int M = (int)b->m.size();
  if (sched) {
    const Launch ls = launch_for(b, BK_SCHED);
    hipLaunchKernelGGL(k_sched, dim3(grid_for(b, BK_SCHED)), dim3(256), 0, st, ls, (const double *)b->d_kp,
                       (const double *)b->d_bs, (const int *)b->d_bsoff, (const int *)b->d_step, M, b->sum_ns);
    const Launch lc = launch_for(b, BK_COEFIJ);
    hipLaunchKernelGGL(k_coef_ij, dim3(grid_for(b, BK_COEFIJ)), dim3(256), 0, st, lc);
    const Launch lg = launch_for(b, BK_GHOST);
    hipLaunchKernelGGL(k_ghost, dim3(grid_for(b, BK_GHOST)), dim3(1024), 0, st, lg);
  }
  const Launch lr = launch_for(b, BK_R), lp = launch_for(b, BK_P), lk = launch_for(b, BK_PLANE),
               lf = launch_for(b, BK_FIN);
  hipLaunchKernelGGL(k_r<1>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), 0, st, lr);
  hipLaunchKernelGGL(k_p<1>, dim3(grid_for(b, BK_P)), dim3(kLineThreads), 0, st, lp);
  launch_plane(SEQ_STEP, b->plane_maxc, dim3(grid_for(b, BK_PLANE)), b->plane_lds, st, lk, 1);
  hipLaunchKernelGGL(k_p<0>, dim3(grid_for(b, BK_P)), dim3(kLineThreads), 0, st, lp);
  hipLaunchKernelGGL(k_r<2>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), 0, st, lr);
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(b, BK_FIN)), dim3(256), 0, st, lf, 15, 2);
  if (sched) hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, b->d_step);
  return launch_check("ensemble step");
}

int ensure_graph(ramgpu_batch *b, bool sched) {
  hipGraphExec_t &e = b->exec[sched ? 1 : 0];
  if (e) return RAMGPU_OK;
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_step(b, sched);
  const hipError_t ec = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(RAMGPU_EHIP, "graph capture failed: %s", hipGetErrorString(ec));
  const hipError_t ei = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    e = nullptr;
    return fail(RAMGPU_EHIP, "graph instantiation failed: %s", hipGetErrorString(ei));
  }
  return RAMGPU_OK;
}

int run_steps(ramgpu_batch *b, int nsteps, bool sched) {
  int rc;
  if (b->use_graph && (rc = ensure_graph(b, sched))) return rc;
  for (int n = 0; n < nsteps; ++n) {
    if (b->use_graph) {
      HIP_TRY(hipGraphLaunch(b->exec[sched ? 1 : 0], b->stream));
    } else if ((rc = enqueue_step(b, sched))) {
      return rc;
    }
  }
  return RAMGPU_OK;
}

int finish(ramgpu_batch *b, bool sched, ramgpu_species_out *out) {
  HIP_TRY(hipStreamSynchronize(b->stream));
  size_t o = 0;
  for (ramgpu_ctx *c : b->m) {
    if (sched) c->efield_dirty = true;  // VT and scales now those of the last step
    int rc = fill_out(c, out ? out + o : nullptr);
    if (rc) return rc;
    o += c->ns;
  }
  return RAMGPU_OK;
}

int check_batch(ramgpu_batch *b) {
  if (!b) return fail(RAMGPU_EINVAL, "null ensemble");
  HIP_TRY(hipSetDevice(b->device));
  return RAMGPU_OK;
}

}  // namespace

extern "C" {

int ramgpu_batch_create(ramgpu_batch **out, ramgpu_ctx *const *members, int nmembers) {
  if (!out) return fail(RAMGPU_EINVAL, "null output pointer");
  *out = nullptr;
  if (!members || nmembers < 1) return fail(RAMGPU_EINVAL, "an ensemble needs at least one member");
  const int dev = members[0] ? members[0]->device : -1;
  for (int i = 0; i < nmembers; ++i) {
    if (!members[i]) return fail(RAMGPU_EINVAL, "member %d is null", i);
    if (members[i]->device != dev) return fail(RAMGPU_EINVAL, "member %d is on another device", i);
    for (int j = 0; j < i; ++j)
      if (members[j] == members[i]) return fail(RAMGPU_EINVAL, "member %d appears twice", i);
  }
  HIP_TRY(hipSetDevice(dev));
  ramgpu_batch *b = new ramgpu_batch();
  b->device = dev;
  b->m.assign(members, members + nmembers);
  auto cleanup = [&](int code) {
    ramgpu_batch_destroy(b);
    return code;
  };
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) b->ncu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return cleanup(fail(RAMGPU_EHIP, "stream/event creation failed"));
  int rc;
  if ((rc = batch_alloc(b, (void **)&b->d_tabs, nmembers * sizeof(Tab)))) return cleanup(rc);
  for (ramgpu_ctx *c : b->m) {
    b->bs_off.push_back(b->sum_ns);
    b->sum_ns += c->ns;
    b->plane_lds = std::max(b->plane_lds, plane_lds_bytes(c->ne, c->npa));
    b->plane_maxc = std::max(b->plane_maxc, plane_maxc(c->ne, c->npa));
  }
  if ((rc = batch_alloc(b, (void **)&b->d_bsoff, nmembers * sizeof(int))) ||
      (rc = batch_alloc(b, (void **)&b->d_step, sizeof(int))))
    return cleanup(rc);
  if (hipMemcpy(b->d_bsoff, b->bs_off.data(), nmembers * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(RAMGPU_EHIP, "hipMemcpy failed"));
  for (int k = 0; k < BK_N; ++k)
    if ((rc = build_plan(b, k))) return cleanup(rc);
  *out = b;
  return RAMGPU_OK;
}

int ramgpu_batch_destroy(ramgpu_batch *b) {
  if (!b) return RAMGPU_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  destroy_graphs(b);
  for (void *p : b->allocs) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return RAMGPU_OK;
}

int ramgpu_batch_use_graph(ramgpu_batch *b, int enable) {
  int rc = check_batch(b);
  if (rc) return rc;
  b->use_graph = enable != 0;
  return RAMGPU_OK;
}

int ramgpu_batch_ram_run(ramgpu_batch *b, int nsteps, ramgpu_species_out *out) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (nsteps < 0) return fail(RAMGPU_EINVAL, "nsteps < 0");
  if ((rc = sync_members(b))) return rc;
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  if ((rc = run_steps(b, nsteps, false))) return rc;
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  return finish(b, false, out);
}

int ramgpu_batch_ram_run_schedule(ramgpu_batch *b, int nsteps, const double *kp, const double *bscale,
                                  ramgpu_species_out *out) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (nsteps < 0) return fail(RAMGPU_EINVAL, "nsteps < 0");
  if (!kp) return fail(RAMGPU_EINVAL, "kp schedule is null");
  for (size_t mi = 0; mi < b->m.size(); ++mi)
    if (!b->m[mi]->have_lzphi)
      return fail(RAMGPU_ESTATE, "member %zu: ramgpu_set_kp(lz, phi) must precede a Kp schedule", mi);
  const int M = (int)b->m.size();
  if (nsteps > b->sched_cap) {
    // the graph reads the schedule buffers by address: rebuild it with them
    destroy_graphs(b);
    int cap = std::max(nsteps, 1);
    if ((rc = batch_alloc(b, (void **)&b->d_kp, (size_t)cap * M * sizeof(double))) ||
        (rc = batch_alloc(b, (void **)&b->d_bs, (size_t)cap * b->sum_ns * sizeof(double))))
      return rc;
    b->sched_cap = cap;
  }
  HIP_TRY(hipMemcpyAsync(b->d_kp, kp, (size_t)nsteps * M * sizeof(double), hipMemcpyHostToDevice, b->stream));
  if (bscale) {
    HIP_TRY(hipMemcpyAsync(b->d_bs, bscale, (size_t)nsteps * b->sum_ns * sizeof(double), hipMemcpyHostToDevice,
                           b->stream));
  } else {
    std::vector<double> cur((size_t)nsteps * b->sum_ns);
    for (int mi = 0; mi < M; ++mi) {
      std::vector<double> sc(b->m[mi]->ns);
      HIP_TRY(hipMemcpy(sc.data(), b->m[mi]->d_bscale, sc.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int n = 0; n < nsteps; ++n)
        for (int s = 0; s < b->m[mi]->ns; ++s) cur[(size_t)n * b->sum_ns + b->bs_off[mi] + s] = sc[s];
    }
    HIP_TRY(hipMemcpy(b->d_bs, cur.data(), cur.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  const int zero = 0;
  HIP_TRY(hipMemcpyAsync(b->d_step, &zero, sizeof(int), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if ((rc = sync_members(b))) return rc;
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  if ((rc = run_steps(b, nsteps, true))) return rc;
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  return finish(b, true, out);
}

int ramgpu_batch_last_run_ms(ramgpu_batch *b, double *ms) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!ms) return fail(RAMGPU_EINVAL, "ms is null");
  HIP_TRY(hipEventSynchronize(b->ev1));
  float f = 0.f;
  HIP_TRY(hipEventElapsedTime(&f, b->ev0, b->ev1));
  *ms = f;
  return RAMGPU_OK;
}

// Autotuner: for each sweep kernel (R, P, plane) time one launch of every
// candidate grid -- one block per unit, or 1, 2, 4, 8 resident blocks per CU
// looping over the units -- and keep the fastest.  The members' F2 and
// energy records are saved before and restored after, so tuning leaves the
// ensemble's state unchanged.  ms_best (may be NULL) receives the best
// per-launch times of R, P and plane.
int ramgpu_batch_autotune(ramgpu_batch *b, int reps, double *ms_best) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (reps < 1) reps = 1;
  if ((rc = sync_members(b))) return rc;
  // save state
  std::vector<void *> saved;
  auto release = [&]() {
    for (void *p : saved) (void)hipFree(p);
  };
  for (ramgpu_ctx *c : b->m) {
    const size_t nf = f2_species_stride(c->t) * c->ns * sizeof(double), ne = (size_t)c->ns * 4 * sizeof(double);
    void *p = nullptr, *q = nullptr;
    if (hipMalloc(&p, nf) != hipSuccess || hipMalloc(&q, ne) != hipSuccess) {
      if (p) (void)hipFree(p);
      release();
      return fail(RAMGPU_ENOMEM, "autotune: no room to save the state");
    }
    saved.push_back(p);
    saved.push_back(q);
    HIP_TRY(hipMemcpyAsync(p, c->t.f2, nf, hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(q, c->t.energy, ne, hipMemcpyDeviceToDevice, b->stream));
  }
  const int kinds[3] = {BK_R, BK_P, BK_PLANE};
  const int mult[5] = {0, 1, 2, 4, 8};
  hipEvent_t e0 = b->ev0, e1 = b->ev1;
  for (int ki = 0; ki < 3 && rc == RAMGPU_OK; ++ki) {
    const int kind = kinds[ki];
    Plan &p = b->plan[kind];
    float best = 1e30f;
    int best_cap = 0;
    for (int mi = 0; mi < 5; ++mi) {
      p.cap = mult[mi] * b->ncu;
      const Launch L = launch_for(b, kind);
      float tot = 0.f;
      for (int r = 0; r <= reps; ++r) {  // first launch warms up
        (void)hipEventRecord(e0, b->stream);
        if (kind == BK_R)
          hipLaunchKernelGGL(k_r<0>, dim3(grid_for(b, kind)), dim3(kLineThreads), 0, b->stream, L);
        else if (kind == BK_P)
          hipLaunchKernelGGL(k_p<0>, dim3(grid_for(b, kind)), dim3(kLineThreads), 0, b->stream, L);
        else
          launch_plane(SEQ_STEP, b->plane_maxc, dim3(grid_for(b, kind)), b->plane_lds, b->stream, L, 0);
        (void)hipEventRecord(e1, b->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
          rc = fail(RAMGPU_EHIP, "autotune launch failed");
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) tot += ms;
      }
      if (rc) break;
      const float avg = tot / reps;
      if (avg < best) {
        best = avg;
        best_cap = p.cap;
      }
    }
    p.cap = best_cap;
    if (ms_best) ms_best[ki] = best;
  }
  // restore state; the tuned shapes change the captured graphs
  size_t o = 0;
  for (ramgpu_ctx *c : b->m) {
    const size_t nf = f2_species_stride(c->t) * c->ns * sizeof(double), ne = (size_t)c->ns * 4 * sizeof(double);
    (void)hipMemcpyAsync(c->t.f2, saved[o], nf, hipMemcpyDeviceToDevice, b->stream);
    (void)hipMemcpyAsync(c->t.energy, saved[o + 1], ne, hipMemcpyDeviceToDevice, b->stream);
    o += 2;
  }
  (void)hipStreamSynchronize(b->stream);
  release();
  destroy_graphs(b);
  return rc;
}

// Grid caps in use: caps[0..2] for R, P, plane (0 = one block per unit).
int ramgpu_batch_get_shapes(ramgpu_batch *b, int *caps) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!caps) return fail(RAMGPU_EINVAL, "caps is null");
  caps[0] = b->plan[BK_R].cap;
  caps[1] = b->plan[BK_P].cap;
  caps[2] = b->plan[BK_PLANE].cap;
  return RAMGPU_OK;
}

int ramgpu_batch_set_shapes(ramgpu_batch *b, const int *caps) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!caps) return fail(RAMGPU_EINVAL, "caps is null");
  for (int i = 0; i < 3; ++i)
    if (caps[i] < 0) return fail(RAMGPU_EINVAL, "negative grid cap");
  b->plan[BK_R].cap = caps[0];
  b->plan[BK_P].cap = caps[1];
  b->plan[BK_PLANE].cap = caps[2];
  destroy_graphs(b);
  return RAMGPU_OK;
}

}  // extern "C"
