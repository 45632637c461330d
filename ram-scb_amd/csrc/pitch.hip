// pitch.hip -- the implicit pitch-angle diffusions of the split step, WPADIF
// (src/ModRamWPI.f90:643-714) and FLCscatter (src/ModRamLoss.f90:513-575),
// and the step that takes them: ram_run's species loop as one launch per
// operator (src/ModRamRun.f90:66-183), for the operators the fused plane
// kernel does not hold.
#include <hip/hip_runtime.h>

#include <vector>

#include "ctx.h"

using namespace ramgpu;

namespace {

using namespace ramgpu_host;

// a species' coefficient table (device layout [I][J][L][K]) from d1 + d2
// (Fortran (NR,NT,NE,NPA), d2 may be null), allocated on first use
int upload_coef(ramgpu_ctx *c, double **slot, const double *d1, const double *d2) {
  const size_t n = (size_t)c->nr * c->nt * c->ne * c->npa;
  if (pitch_lds_bytes(c->npa) > c->lds_limit)
    return fail(RAMGPU_EINVAL, "NPA=%d: the pitch-angle line solve's LDS exceeds the device's", c->npa);
  if (!*slot) {
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess)
      return fail(RAMGPU_ENOMEM, "allocation of a pitch-angle diffusion table failed");
    c->allocs.push_back(p);
    *slot = (double *)p;
  }
  std::vector<double> sum(d1, d1 + n);
  if (d2)
    for (size_t e = 0; e < n; ++e) sum[e] = d1[e] + d2[e];  // ATAW+ATAC, src/ModRamWPI.f90:675
  HIP_TRY(hipMemcpyAsync(c->d_stage, sum.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_f2_in, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, *slot, c->d_stage, 1,
                     c->nr, c->nt, c->ne, c->npa);
  int rc = launch_check("k_f2_in");
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));  // the host sum goes out of scope
  return RAMGPU_OK;
}

// the line solves of the species of job in one launch
int run_pitch(ramgpu_ctx *c, const PitchJob &job) {
  if (job.n == 0) return RAMGPU_OK;
  const int lines = (c->nr - 1) * c->nt * (c->ne - 1);
  tbeg(c, RAMGPU_KERNEL_PITCH);
  hipLaunchKernelGGL(k_pitch_diffusion, dim3((lines + kPitchThreads - 1) / kPitchThreads, job.n),
                     dim3(kPitchThreads), pitch_lds_bytes(c->npa), c->stream, c->d_tab, job);
  tend(c);
  return launch_check("k_pitch_diffusion");
}

int run_pitch(ramgpu_ctx *c, int s, const double *coef) {
  PitchJob job{};
  job.n = 1;
  job.s[0] = s;
  job.coef[0] = coef;
  return run_pitch(c, job);
}

// SUMRC of every species, each ELORC booked to accumulator acc (a species
// whose F2 the operator before left alone books exactly 0, its sum being
// the same as at its previous SUMRC)
int sumrc_into(ramgpu_ctx *c, int acc) {
  int rc = run_plane(c, SEQ_SUM, 0, 0, c->ns);
  return rc ? rc : run_finalize(c, 0, 1 | (acc << kSumAccShift), 0, c->ns);
}

}  // namespace

namespace ramgpu_host {

int takes_pitch(const ramgpu_ctx *c) {
  for (int s = 0; s < c->ns; ++s)
    if (c->pd_wpi[s] || c->pd_emic[s] || c->d_flc[s]) return 1;
  return 0;
}

int forward_only(const ramgpu_ctx *c) {
  if (c->t.coul || takes_pitch(c))
    return fail(RAMGPU_ESTATE,
                "Coulomb collisions, WPADIF or FLC scattering are on: the reverse mode and ensembles do not take "
                "them");
  return RAMGPU_OK;
}

// One ram_run step with one launch per operator (src/ModRamRun.f90:66-183):
// the species' chains are independent, so each operator runs for every
// species that takes it at once, in the reference's order, and each SUMRC
// for every species (booking 0 where the operator did not run).  The chain
// ends with DRIFTR's SUMRC, the seam copy, the magnetopause mask and
// DtDrift* (the CFL maxima recorded by the first sweeps when rec).
int step_composed(ramgpu_ctx *c, int rec) {
  int rc = RAMGPU_OK;
  const int ns = c->ns;
  const bool coul = c->t.coul != 0;
  const int clamp = (coul && c->time_elapsed > 0.) ? 2 : 0;
  PitchJob wpd{}, emic{}, flc{};
  bool wave[kMaxSpecies] = {}, any_wave = false, any_cex = false;
  for (int s = 0; s < ns; ++s) {
    if (c->pd_wpi[s]) {
      wpd.s[wpd.n] = s;
      wpd.coef[wpd.n++] = c->d_pdc[s];
    } else if (c->t.wpi[s]) {
      wave[s] = any_wave = true;
    }
    if (c->pd_emic[s]) {
      emic.s[emic.n] = s;
      emic.coef[emic.n++] = c->d_pdc[s];
    }
    if (c->d_flc[s] && c->time_elapsed >= c->dt_bc) {
      flc.s[flc.n] = s;
      flc.coef[flc.n++] = c->d_flc[s];
    }
    any_cex = any_cex || c->t.cex[s] != 0;
  }
#define TRY(x) \
  if ((rc = (x))) return rc
  // [WAVELO or WPADIF], [EMIC], [CEX], [FLC]: the loss half, forward and
  // then reversed around the two ATMOLs; CHAREXCHANGE leaves species
  // without CEX alone, so it runs for all
  auto half = [&](bool forward) -> int {
    for (int q = 0; q < 4; ++q) {
      const int op = forward ? q : 3 - q;
      if (op == 0 && (any_wave || wpd.n)) {
        for (int s = 0; s < ns; ++s)
          if (wave[s]) TRY(run_plane(c, SEQ_WAVE, 0, s, 1));
        TRY(run_pitch(c, wpd));
        TRY(sumrc_into(c, SUM_LSWAE));
      } else if (op == 1 && emic.n) {
        TRY(run_pitch(c, emic));
        TRY(sumrc_into(c, SUM_LSWAE));
      } else if (op == 2 && any_cex) {
        TRY(run_plane(c, SEQ_CEX, 0, 0, ns));
        TRY(sumrc_into(c, SUM_LSCHA));
      } else if (op == 3 && flc.n) {
        TRY(run_pitch(c, flc));  // no SUMRC after FLCscatter
      }
    }
    return RAMGPU_OK;
  };
  TRY(run_r(c, rec, 0, ns));
  TRY(run_p(c, rec, 0, ns));
  TRY(run_plane(c, SEQ_E, rec, 0, ns));
  TRY(run_plane(c, SEQ_MU, rec, 0, ns));
  TRY(sumrc_into(c, SUM_LSDR));
  if (coul) {
    TRY(run_plane(c, SEQ_COULEN, 0, 0, ns));
    TRY(sumrc_into(c, SUM_LSCOE));
    TRY(run_plane(c, SEQ_COULMU, clamp, 0, ns));
    TRY(sumrc_into(c, SUM_LSCSC));
  }
  TRY(half(true));
  for (int a = 0; a < 2; ++a) {
    TRY(run_plane(c, SEQ_ATM, 0, 0, ns));
    TRY(sumrc_into(c, SUM_LSATM));
  }
  TRY(half(false));
  if (coul) {
    TRY(run_plane(c, SEQ_COULMU, clamp, 0, ns));
    TRY(sumrc_into(c, SUM_LSCSC));
    TRY(run_plane(c, SEQ_COULEN, 0, 0, ns));
    TRY(sumrc_into(c, SUM_LSCOE));
  }
  TRY(run_plane(c, SEQ_MU, 0, 0, ns));
  TRY(run_plane(c, SEQ_E, 0, 0, ns));
  TRY(run_p(c, 0, 0, ns));
  TRY(run_r(c, 2, 0, ns));
  TRY(run_finalize(c, rec ? 15 : 0, 1 | kSumFromRadial | (SUM_LSDR << kSumAccShift), 0, ns));
#undef TRY
  return RAMGPU_OK;
}

}  // namespace ramgpu_host

extern "C" {

int ramgpu_set_wpadif(ramgpu_ctx *c, int s, const double *d1, const double *d2, int at_wpi, int at_emic) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  if (!d1) {
    c->pd_wpi[s - 1] = c->pd_emic[s - 1] = 0;
    return RAMGPU_OK;
  }
  if ((rc = upload_coef(c, &c->d_pdc[s - 1], d1, d2))) return rc;
  c->pd_wpi[s - 1] = at_wpi ? 1 : 0;
  c->pd_emic[s - 1] = at_emic ? 1 : 0;
  return RAMGPU_OK;
}

int ramgpu_wpadif(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  if (!c->d_pdc[s - 1]) return fail(RAMGPU_ESTATE, "species %d has no WPADIF coefficients (ramgpu_set_wpadif)", s);
  if ((rc = prepare(c))) return rc;
  return run_pitch(c, s - 1, c->d_pdc[s - 1]);
}

int ramgpu_set_flc(ramgpu_ctx *c, int s, const double *coef, double dt_bc) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  c->dt_bc = dt_bc;
  if (!coef) {
    c->d_flc[s - 1] = nullptr;  // the table stays allocated for a later call
    return RAMGPU_OK;
  }
  double *slot = c->d_flc_mem[s - 1];
  if ((rc = upload_coef(c, &slot, coef, nullptr))) return rc;
  c->d_flc_mem[s - 1] = slot;
  c->d_flc[s - 1] = slot;
  return RAMGPU_OK;
}

int ramgpu_flcscatter(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  if (!c->d_flc[s - 1]) return fail(RAMGPU_ESTATE, "species %d has no FLC coefficients (ramgpu_set_flc)", s);
  if (c->time_elapsed < c->dt_bc) return RAMGPU_OK;  // src/ModRamLoss.f90:526
  if ((rc = prepare(c))) return rc;
  return run_pitch(c, s - 1, c->d_flc[s - 1]);
}

}  // extern "C"
