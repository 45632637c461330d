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
#include <cstdlib>
#include <vector>

#include "ctx.h"

using namespace ramgpu;
using namespace ramgpu_host;

namespace {

// kernel kinds of an ensemble launch; the last four serve reverse mode: P and
// plane out of place (with the copy lines / planes of the unswept I=1 rows),
// one unit per member (misfit) and a whole-F2 copy per member
enum BatchKind { BK_SCHED, BK_COEFIJ, BK_GHOST, BK_R, BK_P, BK_PLANE, BK_FIN, BK_POOP, BK_PLOOP, BK_MEM, BK_COPY, BK_N };

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
  size_t plane_lds = 0, radial_lds = 0, azim_lds = 0, ghost_lds = 0;
  int mpts_max = 0;  // most satellite points of a member (ensemble adjoint)
  // forward plane blocks of every member: the largest block a member
  // chose and the largest row bound among them
  PlaneShape plane{64, 8, 0};
  std::vector<char> wave;  // each member's WAVELO at creation (its coefficient units)
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
  // opt-in per-launch timing (eager steps only): event pairs by kernel kind
  bool timing = false;
  std::vector<hipEvent_t> tev;
  std::vector<int> tkind;
  size_t tused = 0;
};

namespace {

int batch_alloc(ramgpu_batch *b, void **p, size_t bytes) {
  if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess)
    return fail(RAMGPU_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
  b->allocs.push_back(*p);
  return RAMGPU_OK;
}

int blocks_of(int kind, const ramgpu_ctx *c) {
  const int P = c->ne * c->npa;
  switch (kind) {
    case BK_COEFIJ: return coef_units(c, 256);
    case BK_R: return blocks_r(c->t);
    case BK_P: return blocks_p(c->t);
    case BK_PLANE: return blocks_kl(c->t);
    case BK_POOP: return ((c->nr - 1) * P + c->nt * P + kLineThreads - 1) / kLineThreads;
    case BK_PLOOP: return blocks_kl(c->t) + c->nt;
    case BK_COPY: return 256;
    default: return 1;
  }
}

bool per_member(int kind) { return kind == BK_SCHED || kind == BK_COEFIJ || kind == BK_MEM || kind == BK_COPY; }

// Job table and block map of one kernel kind, members in order.
int build_plan(ramgpu_batch *b, int kind) {
  std::vector<Job> jobs;
  std::vector<int> vmap;
  for (size_t mi = 0; mi < b->m.size(); ++mi) {
    const ramgpu_ctx *c = b->m[mi];
    const int nsp = per_member(kind) ? 1 : c->ns;
    const int nb = blocks_of(kind, c);
    if (nb >= (1 << kVmapShift) || jobs.size() + nsp >= (1u << (31 - kVmapShift)))
      return fail(RAMGPU_EINVAL, "ensemble too large for the launch tables");
    const int j0 = (int)jobs.size();
    for (int s = 0; s < nsp; ++s) jobs.push_back(Job{(int)mi, s, (int)vmap.size(), nb});
    if (kind == BK_PLANE) {
      // the species of each run of 8 planes in turn: a plane's species land
      // on one XCD, whose L2 serves them the plane's tables (Launch::ilv)
      for (int g = 0; g < nb; g += 8)
        for (int s = 0; s < nsp; ++s)
          for (int x = g; x < std::min(nb, g + 8); ++x) vmap.push_back(vmap_entry(j0 + s, x));
    } else {
      for (int s = 0; s < nsp; ++s)
        for (int x = 0; x < nb; ++x) vmap.push_back(vmap_entry(j0 + s, x));
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
  for (size_t m = 0; m < b->m.size(); ++m) {
    ramgpu_ctx *c = b->m[m];
    if ((rc = check_ctx(c)) || (rc = forward_only(c)) || (rc = prepare(c))) return rc;
    if (c->plane.threads > b->plane.threads || c->plane.maxc > b->plane.maxc)
      return fail(RAMGPU_ESTATE, "member %zu changed its plane shape after joining the ensemble", m);
    if ((char)takes_wave(c) != b->wave[m])
      return fail(RAMGPU_ESTATE, "member %zu switched WAVELO after joining the ensemble", m);
    HIP_TRY(hipStreamSynchronize(c->stream));
    tabs.push_back(c->t);
  }
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipMemcpy(b->d_tabs, tabs.data(), tabs.size() * sizeof(Tab), hipMemcpyHostToDevice));
  return RAMGPU_OK;
}

// Event pair around one launch of kind `kind` while timing is on.
void btbeg(ramgpu_batch *b, int kind) {
  if (!b->timing) return;
  if (2 * b->tused + 2 > b->tev.size()) {
    for (int n = 0; n < 2; ++n) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        b->timing = false;
        return;
      }
      b->tev.push_back(e);
    }
  }
  if (b->tkind.size() < b->tused + 1) b->tkind.resize(b->tused + 1);
  b->tkind[b->tused] = kind;
  (void)hipEventRecord(b->tev[2 * b->tused], b->stream);
}

void btend(ramgpu_batch *b) {
  if (!b->timing) return;
  (void)hipEventRecord(b->tev[2 * b->tused + 1], b->stream);
  ++b->tused;
}

// The launches of one ensemble step.  sched: VT / boundary scale of the step
// from the device schedule, then the VT-dependent tables and the ghost carry.
// CFL limits are recorded every step (the same values a member's own
// ram_run reports).
int enqueue_step(ramgpu_batch *b, bool sched) {
  hipStream_t st = b->stream;
  const int M = (int)b->m.size();
  if (sched) {
    btbeg(b, RAMGPU_KERNEL_PREPARE);
    const Launch ls = launch_for(b, BK_SCHED);
    hipLaunchKernelGGL(k_sched, dim3(grid_for(b, BK_SCHED)), dim3(256), 0, st, ls, (const double *)b->d_kp,
                       (const double *)b->d_bs, (const int *)b->d_bsoff, (const int *)b->d_step, M, b->sum_ns);
    const Launch lc = launch_for(b, BK_COEFIJ);
    hipLaunchKernelGGL(k_coef_ij, dim3(grid_for(b, BK_COEFIJ)), dim3(256), 0, st, lc, (const double *)nullptr,
                     (const double *)nullptr);
    const Launch lg = launch_for(b, BK_GHOST);
    hipLaunchKernelGGL(k_ghost, dim3(grid_for(b, BK_GHOST)), dim3(1024), b->ghost_lds, st, lg);
    btend(b);
  }
  const Launch lr = launch_for(b, BK_R), lp = launch_for(b, BK_P), lk = launch_for(b, BK_PLANE),
               lf = launch_for(b, BK_FIN);
  btbeg(b, RAMGPU_KERNEL_R);
  hipLaunchKernelGGL(k_r<1>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), b->radial_lds, st, lr);
  btend(b);
  btbeg(b, RAMGPU_KERNEL_P);
  hipLaunchKernelGGL(k_p<1>, dim3(grid_for(b, BK_P)), dim3(kLineThreads), b->azim_lds, st, lp);
  btend(b);
  btbeg(b, RAMGPU_KERNEL_PLANE);
  launch_plane(SEQ_STEP, b->plane, dim3(grid_for(b, BK_PLANE)), b->plane_lds, st, lk, 1);
  btend(b);
  btbeg(b, RAMGPU_KERNEL_P);
  hipLaunchKernelGGL(k_p<0>, dim3(grid_for(b, BK_P)), dim3(kLineThreads), b->azim_lds, st, lp);
  btend(b);
  btbeg(b, RAMGPU_KERNEL_R);
  hipLaunchKernelGGL(k_r<2>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), b->radial_lds, st, lr);
  btend(b);
  btbeg(b, RAMGPU_KERNEL_FINALIZE);
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(b, BK_FIN)), dim3(kFinalizeThreads), 0, st, lf, 15, 2);
  if (sched) hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, b->d_step);
  btend(b);
  return launch_check("ensemble step");
}


// ---- reverse mode ----------------------------------------------------------
// The launches of the ensemble adjoint name each member's buffers by tape
// slot (Launch::ssl/dsl, 1-based; 0 = the member's F2 as source, -1 as an
// out-of-place destination), so one launch serves members of any size.

Launch slot_launch(const ramgpu_batch *b, int kind, int ssl, int dsl, int bstep = 0) {
  Launch L = launch_for(b, kind);
  L.ssl = ssl;
  L.dsl = dsl;
  L.bstep = bstep;
  return L;
}

// VT / boundary scale of step `step` from the device schedule, then the
// VT-dependent tables and the ghost carry
void enq_sched(ramgpu_batch *b, int step) {
  hipStream_t st = b->stream;
  hipLaunchKernelGGL(k_setstep, dim3(1), dim3(64), 0, st, b->d_step, step);
  const Launch ls = launch_for(b, BK_SCHED);
  hipLaunchKernelGGL(k_sched, dim3(grid_for(b, BK_SCHED)), dim3(256), 0, st, ls, (const double *)b->d_kp,
                     (const double *)b->d_bs, (const int *)b->d_bsoff, (const int *)b->d_step, (int)b->m.size(),
                     b->sum_ns);
  const Launch lc = launch_for(b, BK_COEFIJ);
  hipLaunchKernelGGL(k_coef_ij, dim3(grid_for(b, BK_COEFIJ)), dim3(256), 0, st, lc, (const double *)nullptr,
                     (const double *)nullptr);
  const Launch lg = launch_for(b, BK_GHOST);
  hipLaunchKernelGGL(k_ghost, dim3(grid_for(b, BK_GHOST)), dim3(1024), b->ghost_lds, st, lg);
}

void enq_copy(ramgpu_batch *b, int ssl, int dsl) {
  const Launch L = slot_launch(b, BK_COPY, ssl, dsl);
  hipLaunchKernelGGL(k_copy_ens, dim3(L.nv), dim3(256), 0, b->stream, L);
}

void enq_misfit(ramgpu_batch *b, int ssl, int step, int adjoint) {
  const Launch L = slot_launch(b, BK_MEM, ssl, 0);
  hipLaunchKernelGGL(k_misfit_ens, dim3(L.nv, std::max(b->mpts_max, 1)), dim3(64), 0, b->stream, L, step, adjoint);
}

// One step with every operator's input on the tape group starting at slot g
// (0-based), the step's output to dsl_out; the arithmetic of enqueue_step.
// With bookkeeping false (a segment recomputed for its tape) the CFL limits
// and the SUMRC chain are not formed: only the states are needed.
void enq_taped_step(ramgpu_batch *b, int g, int dsl_out, bool bookkeeping = true) {
  hipStream_t st = b->stream;
  Launch L = slot_launch(b, BK_R, g + T_R1 + 1, g + T_P1 + 1);
  if (bookkeeping)
    hipLaunchKernelGGL(k_r<1>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), b->radial_lds, st, L);
  else
    hipLaunchKernelGGL(k_r<0>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), b->radial_lds, st, L);
  L = slot_launch(b, BK_POOP, g + T_P1 + 1, g + T_E1 + 1);
  if (bookkeeping)
    hipLaunchKernelGGL(k_p<1>, dim3(L.nv), dim3(kLineThreads), b->azim_lds, st, L);
  else
    hipLaunchKernelGGL(k_p<0>, dim3(L.nv), dim3(kLineThreads), b->azim_lds, st, L);
  L = slot_launch(b, BK_PLOOP, g + T_E1 + 1, g + T_P2 + 1);
  launch_plane(SEQ_TAPE, b->plane, dim3(L.nv), b->plane_lds, st, L, bookkeeping ? 1 : 0);
  L = slot_launch(b, BK_POOP, g + T_P2 + 1, g + T_R2 + 1);
  hipLaunchKernelGGL(k_p<0>, dim3(L.nv), dim3(kLineThreads), b->azim_lds, st, L);
  L = slot_launch(b, BK_R, g + T_R2 + 1, dsl_out);
  hipLaunchKernelGGL(k_r<2>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), b->radial_lds, st, L);
  if (bookkeeping) {
    const Launch lf = launch_for(b, BK_FIN);
    hipLaunchKernelGGL(k_finalize, dim3(grid_for(b, BK_FIN)), dim3(kFinalizeThreads), 0, st, lf, 15, 2);
  }
}

// adjoint of the step taped at group g: Tab::lam in place, Tab::bbar[step]
void enq_step_bwd(ramgpu_batch *b, int g, int step, size_t bwd_lds, int bwd_threads) {
  hipStream_t st = b->stream;
  Launch L = slot_launch(b, BK_R, g + T_R2 + 1, 0, step);
  hipLaunchKernelGGL(k_r_bwd<1>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), b->radial_lds, st, L, nullptr, nullptr,
                     nullptr);
  L = slot_launch(b, BK_P, g + T_P2 + 1, 0);
  hipLaunchKernelGGL(k_p_bwd, dim3(grid_for(b, BK_P)), dim3(kLineThreads), b->azim_lds, st, L, nullptr, nullptr);
  L = slot_launch(b, BK_PLANE, g + T_E1 + 1, 0);
  launch_plane_bwd(SEQ_STEP, bwd_threads, dim3(grid_for(b, BK_PLANE)), bwd_lds, st, L, nullptr, nullptr);
  L = slot_launch(b, BK_P, g + T_P1 + 1, 0);
  hipLaunchKernelGGL(k_p_bwd, dim3(grid_for(b, BK_P)), dim3(kLineThreads), b->azim_lds, st, L, nullptr, nullptr);
  L = slot_launch(b, BK_R, g + T_R1 + 1, 0, step);
  hipLaunchKernelGGL(k_r_bwd<0>, dim3(grid_for(b, BK_R)), dim3(kLineThreads), b->radial_lds, st, L, nullptr, nullptr,
                     nullptr);
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
  const bool graph = b->use_graph && !b->timing;  // per-launch events need eager launches
  if (graph && (rc = ensure_graph(b, sched))) return rc;
  for (int n = 0; n < nsteps; ++n) {
    if (graph) {
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

// Upload a per-member Kp / boundary-scale schedule of nsteps steps (bscale
// may be NULL: the members' current scales) and reset the step counter.
int load_schedule(ramgpu_batch *b, int nsteps, const double *kp, const double *bscale) {
  const int M = (int)b->m.size();
  int rc;
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
    const int rc = forward_only(members[i]);
    if (rc) return rc;
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
    b->radial_lds = std::max(b->radial_lds, radial_lds_bytes(c->nr, c->ne, c->npa));
    b->azim_lds = std::max(b->azim_lds, azim_lds_bytes(c->nt, c->ne, c->npa));
    b->ghost_lds = std::max(b->ghost_lds, (size_t)4 * c->nt * sizeof(double));
    b->plane.threads = std::max(b->plane.threads, c->plane.threads);
    b->wave.push_back((char)takes_wave(c));
  }
  // one block size for all members, the largest any member chose: each
  // member keeps its own rows per thread (its Tab's plane_chunks), so its
  // planes are swept exactly as when it runs alone; the threads above its
  // chunks are idle
  {
    size_t lds = 0;
    int maxc = 8;
    for (ramgpu_ctx *c : b->m) {
      maxc = std::max(maxc, c->plane.maxc);
      lds = std::max(lds, plane_fwd_lds_bytes(c->ne, c->npa, b->plane.threads));
    }
    if (lds > b->m[0]->lds_limit) return cleanup(fail(RAMGPU_EINVAL, "the members' planes do not fit one block's LDS"));
    b->plane = {b->plane.threads, maxc, 0};
    b->plane_lds = lds;
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
  for (hipEvent_t e : b->tev) (void)hipEventDestroy(e);
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
  if ((rc = load_schedule(b, nsteps, kp, bscale))) return rc;
  if ((rc = sync_members(b))) return rc;
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  if ((rc = run_steps(b, nsteps, true))) return rc;
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  return finish(b, true, out);
}

int ramgpu_batch_kernel_timing(ramgpu_batch *b, int enable) {
  int rc = check_batch(b);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->timing = enable != 0;
  b->tused = 0;
  return RAMGPU_OK;
}

int ramgpu_batch_kernel_times(ramgpu_batch *b, double *total_ms, int *launches) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!total_ms || !launches) return fail(RAMGPU_EINVAL, "null output array");
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (int k = 0; k < RAMGPU_KERNEL_KINDS; ++k) {
    total_ms[k] = 0.0;
    launches[k] = 0;
  }
  for (size_t n = 0; n < b->tused; ++n) {
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, b->tev[2 * n], b->tev[2 * n + 1]));
    total_ms[b->tkind[n]] += f;
    launches[b->tkind[n]] += 1;
  }
  b->tused = 0;
  return RAMGPU_OK;
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

// Reverse mode of an ensemble: the flux-misfit gradient of every member
// (ramgpu_adjoint_run for each member, all members in the same launches).
// Tape slots of member m, each one F2 of that member: nseg checkpoints,
// then either one tape group per step of a segment (the segment is taped in
// one recomputation) or, when that does not fit the device, seglen states
// plus one group (the segment's states are recomputed, then each step is
// re-taped just before its adjoint).  The arithmetic is the same either way.
int ramgpu_batch_adjoint_run(ramgpu_batch *b, int nsteps, int ckpt_every, const double *kp, const double *bscale,
                             const int *npts, const int *const *pts, const double *const *w,
                             const double *const *target, double *loss, double *const *grad_f2,
                             double *const *grad_fgeos) {
  int rc = check_batch(b);
  if (rc) return rc;
  const int M = (int)b->m.size();
  if (nsteps < 1 || ckpt_every < 1) return fail(RAMGPU_EINVAL, "nsteps and ckpt_every must be >= 1");
  if (!kp) return fail(RAMGPU_EINVAL, "kp schedule is null");
  if (!npts || !pts || !w || !target) return fail(RAMGPU_EINVAL, "misfit points, weights and targets required");
  for (int mi = 0; mi < M; ++mi) {
    const ramgpu_ctx *c = b->m[mi];
    if (!c->have_lzphi) return fail(RAMGPU_ESTATE, "member %d: ramgpu_set_kp(lz, phi) must precede a Kp schedule", mi);
    if (npts[mi] < 1 || !pts[mi] || !w[mi] || !target[mi])
      return fail(RAMGPU_EINVAL, "member %d: misfit points, weights and targets required", mi);
    for (int p = 0; p < npts[mi]; ++p) {
      const int *q = pts[mi] + 4 * p;
      if (q[0] < 1 || q[0] > c->ns || q[1] < 1 || q[1] > c->nr || q[2] < 1 || q[2] > c->nt || q[3] < 1 ||
          q[3] > c->ne)
        return fail(RAMGPU_EINVAL, "member %d: misfit point %d outside the grid", mi, p);
    }
  }
  const int nseg = (nsteps + ckpt_every - 1) / ckpt_every, seglen = std::min(ckpt_every, nsteps);
  // plane adjoint LDS
  size_t bwd_lds = 0, lds_limit = b->m[0]->lds_limit;
  int bwd_threads = kPlaneThreads;
  for (ramgpu_ctx *c : b->m) {
    bwd_lds = std::max(bwd_lds, plane_bwd_lds_bytes(c->ne, c->npa));
    bwd_threads = std::max(bwd_threads, plane_bwd_threads(c->ne, c->npa));
  }
  if (bwd_lds > lds_limit) return fail(RAMGPU_EINVAL, "plane adjoint does not fit the device LDS");
  set_plane_bwd_lds_limit(lds_limit);
  // memory plan
  size_t nall = 0, ngall = 0;
  for (ramgpu_ctx *c : b->m) {
    nall += (size_t)c->ns * c->nr * c->nt * c->ne * c->npa;
    ngall += (size_t)c->ns * c->nt * c->ne * c->npa;
  }
  size_t freeb = 0, totb = 0;
  HIP_TRY(hipMemGetInfo(&freeb, &totb));
  const size_t fixed = 8 * (2 * nall + (grad_fgeos ? (size_t)nsteps * ngall : 0)) + ((size_t)64 << 20);
  const size_t one = (size_t)(nseg + seglen * T_N) * nall * 8, two = (size_t)(nseg + seglen + T_N) * nall * 8;
  // two-level: states, then one re-taped step at a time (forced with
  // RAMGPU_ADJOINT_RETAPE=1, which tests the scheme on small ensembles)
  const char *force = getenv("RAMGPU_ADJOINT_RETAPE");
  bool split;
  if (force && force[0] == '1' && fixed + two <= freeb * 0.92) split = true;
  else if (fixed + one <= freeb * 0.92) split = false;
  else if (fixed + two <= freeb * 0.92) split = true;
  else return fail(RAMGPU_ENOMEM, "ensemble adjoint needs %.1f GB, %.1f GB free", (fixed + two) / 1e9, freeb / 1e9);
  const int nslots = split ? nseg + seglen + T_N : nseg + seglen * T_N;
  const int CK = 0, ST = nseg, TG = split ? nseg + seglen : nseg;
  auto group = [&](int s, int n0) { return split ? TG : TG + (s - n0) * T_N; };
  // buffers
  std::vector<void *> bufs;
  auto release = [&]() {
    (void)hipStreamSynchronize(b->stream);
    for (void *q : bufs) (void)hipFree(q);
  };
  auto alloc = [&](void **q, size_t bytes) {
    if (hipMalloc(q, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    bufs.push_back(*q);
    return true;
  };
  std::vector<double *> fin(M), fin_e(M);
  for (int mi = 0; mi < M; ++mi) {
    ramgpu_ctx *c = b->m[mi];
    const size_t n = (size_t)c->ns * c->nr * c->nt * c->ne * c->npa, ng = (size_t)c->ns * c->nt * c->ne * c->npa;
    const size_t np = npts[mi];
    void *tape, *lam, *bb = nullptr, *mp, *mw, *mt, *ml, *f, *fe;
    if (!alloc(&tape, (size_t)nslots * n * 8) || !alloc(&lam, n * 8) ||
        (grad_fgeos && grad_fgeos[mi] && !alloc(&bb, (size_t)nsteps * ng * 8)) || !alloc(&mp, np * 4 * sizeof(int)) ||
        !alloc(&mw, np * c->npa * 8) || !alloc(&mt, (size_t)nsteps * np * 8) || !alloc(&ml, (size_t)nsteps * np * 8) ||
        !alloc(&f, n * 8) || !alloc(&fe, (size_t)c->ns * 4 * 8)) {
      release();
      return fail(RAMGPU_ENOMEM, "ensemble adjoint: hipMalloc failed (member %d)", mi);
    }
    HIP_TRY(hipMemcpy(mp, pts[mi], np * 4 * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(mw, w[mi], np * c->npa * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(mt, target[mi], (size_t)nsteps * np * 8, hipMemcpyHostToDevice));
    c->t.tape = (double *)tape;
    c->t.lam = (double *)lam;
    c->t.bbar = (double *)bb;
    c->t.mpts = (const int *)mp;
    c->t.mw = (const double *)mw;
    c->t.mtarget = (const double *)mt;
    c->t.mloss = (double *)ml;
    c->t.mnpts = (int)np;
    b->mpts_max = std::max(b->mpts_max, (int)np);
    fin[mi] = (double *)f;
    fin_e[mi] = (double *)fe;
  }
  auto detach = [&]() {  // the members' descriptors no longer name the buffers
    for (ramgpu_ctx *c : b->m) {
      c->t.tape = c->t.lam = c->t.bbar = c->t.mloss = nullptr;
      c->t.mpts = nullptr;
      c->t.mw = c->t.mtarget = nullptr;
      c->t.mnpts = 0;
      c->tab_dirty = true;
      c->efield_dirty = true;  // VT and scales are those of the schedule
    }
    release();
    return sync_members(b);
  };
  if ((rc = load_schedule(b, nsteps, kp, bscale)) || (rc = sync_members(b))) {
    detach();
    return rc;
  }
  hipStream_t st = b->stream;
  const int nlast0 = (nseg - 1) * ckpt_every;
  HIP_TRY(hipEventRecord(b->ev0, st));
  // forward, keeping a checkpoint at the start of every segment; the last
  // segment is taped at once when whole segments are taped
  for (int s = 0; s < nsteps; ++s) {
    enq_sched(b, s);
    if (!split && s >= nlast0) {
      if (s == nlast0) enq_copy(b, 0, group(s, nlast0) + T_R1 + 1);
      const int out = (s == nsteps - 1) ? -1 : group(s + 1, nlast0) + T_R1 + 1;
      enq_taped_step(b, group(s, nlast0), out);
      enq_misfit(b, out < 0 ? 0 : out, s, 0);
    } else {
      if (s % ckpt_every == 0) enq_copy(b, 0, CK + s / ckpt_every + 1);
      enqueue_step(b, false);
      enq_misfit(b, 0, s, 0);
    }
  }
  for (int mi = 0; mi < M; ++mi) {
    ramgpu_ctx *c = b->m[mi];
    const size_t n = (size_t)c->ns * c->nr * c->nt * c->ne * c->npa;
    HIP_TRY(hipMemcpyAsync(fin[mi], c->t.f2, n * 8, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(fin_e[mi], c->t.energy, (size_t)c->ns * 32, hipMemcpyDeviceToDevice, st));
    const size_t ng = (size_t)c->ns * c->nt * c->ne * c->npa;
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, st, c->t.lam, 0.0, n);
    if (c->t.bbar) hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, st, c->t.bbar, 0.0, (size_t)nsteps * ng);
  }
  // backward, segment by segment
  for (int g = nseg - 1; g >= 0; --g) {
    const int n0 = g * ckpt_every, n1 = std::min(nsteps, n0 + ckpt_every);
    auto out_of = [&](int s) { return (s == n1 - 1) ? -1 : (split ? ST + (s + 1 - n0) + 1 : group(s + 1, n0) + T_R1 + 1); };
    if (split) {  // the segment's states
      enq_copy(b, CK + g + 1, -1);
      for (int s = n0; s < n1; ++s) {
        enq_copy(b, 0, ST + (s - n0) + 1);
        enq_sched(b, s);
        enqueue_step(b, false);
      }
    } else if (g < nseg - 1) {  // the segment taped
      enq_copy(b, CK + g + 1, group(n0, n0) + T_R1 + 1);
      for (int s = n0; s < n1; ++s) {
        enq_sched(b, s);
        enq_taped_step(b, group(s, n0), out_of(s), false);
      }
    }
    for (int s = n1 - 1; s >= n0; --s) {
      const int out = out_of(s);
      if (split) {
        enq_copy(b, ST + (s - n0) + 1, TG + T_R1 + 1);
        enq_sched(b, s);
        enq_taped_step(b, TG, out, false);  // rewrites the step's output with the same values
      } else {
        enq_sched(b, s);
      }
      enq_misfit(b, out < 0 ? 0 : out, s, 1);
      enq_step_bwd(b, group(s, n0), s, bwd_lds, bwd_threads);
    }
    if ((rc = launch_check("ensemble adjoint"))) {
      detach();
      return rc;
    }
  }
  HIP_TRY(hipEventRecord(b->ev1, st));
  enq_sched(b, nsteps - 1);  // the members keep the inputs of the last step
  // results; the members are left in the state after the run
  for (int mi = 0; mi < M; ++mi) {
    ramgpu_ctx *c = b->m[mi];
    const size_t n = (size_t)c->ns * c->nr * c->nt * c->ne * c->npa, ng = (size_t)c->ns * c->nt * c->ne * c->npa;
    const size_t np = npts[mi];
    HIP_TRY(hipMemcpyAsync(c->t.f2, fin[mi], n * 8, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->t.energy, fin_e[mi], (size_t)c->ns * 32, hipMemcpyDeviceToDevice, st));
    std::vector<double> lv((size_t)nsteps * np);
    HIP_TRY(hipMemcpyAsync(lv.data(), c->t.mloss, lv.size() * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (loss) {
      double tot = 0.0;
      for (double v : lv) tot += v;
      loss[mi] = tot;
    }
    if (grad_f2 && grad_f2[mi]) {
      HIP_TRY(hipStreamSynchronize(c->stream));
      if ((rc = download_f2(c, grad_f2[mi], c->t.lam))) {
        detach();
        return rc;
      }
    }
    if (c->t.bbar) {
      HIP_TRY(hipMemcpy(grad_fgeos[mi], c->t.bbar, (size_t)nsteps * ng * 8, hipMemcpyDeviceToHost));
    }
  }
  return detach();
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
          hipLaunchKernelGGL(k_r<0>, dim3(grid_for(b, kind)), dim3(kLineThreads), b->radial_lds, b->stream, L);
        else if (kind == BK_P)
          hipLaunchKernelGGL(k_p<0>, dim3(grid_for(b, kind)), dim3(kLineThreads), b->azim_lds, b->stream, L);
        else
          launch_plane(SEQ_STEP, b->plane, dim3(grid_for(b, kind)), b->plane_lds, b->stream, L, 0);
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
