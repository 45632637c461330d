// pitch.hip -- the implicit pitch-angle diffusions of the split step, WPADIF
// (src/ModRamWPI.f90:643-714) and FLCscatter (src/ModRamLoss.f90:513-575),
// and the step that takes them: ram_run's species loop as one launch per
// operator (src/ModRamRun.f90:66-183) where the fused plane kernel does not
// hold the operator.
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

// the line solve of species s (0-based) with coefficient table coef
int run_pitch(ramgpu_ctx *c, int s, const double *coef) {
  const int lines = (c->nr - 1) * c->nt * (c->ne - 1);
  tbeg(c, RAMGPU_KERNEL_PITCH);
  hipLaunchKernelGGL(k_pitch_diffusion, dim3((lines + 63) / 64), dim3(64), 0, c->stream, c->d_tab, s, coef,
                     c->d_stage);
  tend(c);
  return launch_check("k_pitch_diffusion");
}

// SUMRC of species s, its ELORC booked to accumulator acc
int sumrc_into(ramgpu_ctx *c, int s, int acc) {
  int rc = run_plane(c, SEQ_SUM, 0, s, 1);
  return rc ? rc : run_finalize(c, 0, 1 | (acc << kSumAccShift), s, 1);
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

// One ram_run step with one launch per operator, species after species,
// in the order of src/ModRamRun.f90:66-183 and with its SUMRC bookkeeping;
// each species ends with DRIFTR's SUMRC, seam copy and magnetopause mask
// and its DtDrift* (the CFL maxima recorded by its first sweeps when rec).
int step_composed(ramgpu_ctx *c, int rec) {
  int rc = RAMGPU_OK;
  const bool coul = c->t.coul != 0;
  const int clamp = (coul && c->time_elapsed > 0.) ? 2 : 0;
#define TRY(x) \
  if ((rc = (x))) return rc
  for (int s = 0; s < c->ns; ++s) {
    const bool wave = c->t.wpi[s] != 0, wpd = c->pd_wpi[s] != 0, emic = c->pd_emic[s] != 0;
    const bool cex = c->t.cex[s] != 0, flc = c->d_flc[s] && c->time_elapsed >= c->dt_bc;
    // [WAVELO or WPADIF], [EMIC], [CEX], [FLC]: the loss half, once forward
    // and once reversed around the two ATMOLs
    auto half = [&](bool forward) -> int {
      for (int q = 0; q < 4; ++q) {
        const int op = forward ? q : 3 - q;
        if (op == 0 && (wave || wpd)) {
          TRY(wpd ? run_pitch(c, s, c->d_pdc[s]) : run_plane(c, SEQ_WAVE, 0, s, 1));
          TRY(sumrc_into(c, s, SUM_LSWAE));
        } else if (op == 1 && emic) {
          TRY(run_pitch(c, s, c->d_pdc[s]));
          TRY(sumrc_into(c, s, SUM_LSWAE));
        } else if (op == 2 && cex) {
          TRY(run_plane(c, SEQ_CEX, 0, s, 1));
          TRY(sumrc_into(c, s, SUM_LSCHA));
        } else if (op == 3 && flc) {
          TRY(run_pitch(c, s, c->d_flc[s]));  // no SUMRC after FLCscatter
        }
      }
      return RAMGPU_OK;
    };
    TRY(run_r(c, rec, s, 1));
    TRY(run_p(c, rec, s, 1));
    TRY(run_plane(c, SEQ_E, rec, s, 1));
    TRY(run_plane(c, SEQ_MU, rec, s, 1));
    TRY(sumrc_into(c, s, SUM_LSDR));
    if (coul) {
      TRY(run_plane(c, SEQ_COULEN, 0, s, 1));
      TRY(sumrc_into(c, s, SUM_LSCOE));
      TRY(run_plane(c, SEQ_COULMU, clamp, s, 1));
      TRY(sumrc_into(c, s, SUM_LSCSC));
    }
    TRY(half(true));
    for (int a = 0; a < 2; ++a) {
      TRY(run_plane(c, SEQ_ATM, 0, s, 1));
      TRY(sumrc_into(c, s, SUM_LSATM));
    }
    TRY(half(false));
    if (coul) {
      TRY(run_plane(c, SEQ_COULMU, clamp, s, 1));
      TRY(sumrc_into(c, s, SUM_LSCSC));
      TRY(run_plane(c, SEQ_COULEN, 0, s, 1));
      TRY(sumrc_into(c, s, SUM_LSCOE));
    }
    TRY(run_plane(c, SEQ_MU, 0, s, 1));
    TRY(run_plane(c, SEQ_E, 0, s, 1));
    TRY(run_p(c, 0, s, 1));
    TRY(run_r(c, 2, s, 1));
    TRY(run_finalize(c, rec ? 15 : 0, 1 | kSumFromRadial | (SUM_LSDR << kSumAccShift), s, 1));
  }
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
