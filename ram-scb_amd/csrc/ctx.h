// ctx.h -- host-side context of the gfx950 RAM split step (internal to the
// library: api.hip, batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ramgpu.h"
#include "ramgpu_dev.h"
#include "sweep_common.h"

namespace ramgpu {
__global__ void k_sched(Launch L, const double *kp, const double *bscale, const int *bs_off, const int *step,
                        int nmembers, int stride_bs);
__global__ void k_tick(int *step);
__global__ void k_setstep(int *step, int v);
__global__ void k_copy_ens(Launch L);
__global__ void k_pitch_diffusion(const Tab *tp, PitchJob job);
template <int FINAL> __global__ void k_r_bwd(Launch L, const double *xin, double *adj, double *bbar);
__global__ void k_p_bwd(Launch L, const double *xin, double *adj);
template <int SEQ, int BT> __global__ void k_plane_bwd(Launch L, const double *xin, double *adj);
__global__ void k_misfit(const double *f2, int nr, int nt, int ne, int npa, int npts, const int *pts,
                         const double *w, const double *target, double *loss_out, double *lam);
__global__ void k_misfit_ens(Launch L, int step, int adjoint);
__global__ void k_axpy(double *y, const double *x, double a, size_t n);
__global__ void k_fill(double *y, double v, size_t n);
__global__ void k_species(Launch L);
__global__ void k_coef_ij(Launch L, const double *kp, const double *bs);
__global__ void k_step_head(Launch Lc, const double *kp, const double *bs, Launch Lf, int cflmask, int summode,
                            Launch Lg, int nvt, int stage_q1);
__global__ void k_coef(Launch L);
__global__ void k_ghost(Launch L);
template <int MODE> __global__ void k_r(Launch L);
template <int MODE> __global__ void k_p(Launch L);
template <int SEQ, int MAXC> __global__ void k_plane(Launch L, int rec);
__global__ void k_finalize(Launch L, int cflmask, int summode);
__global__ void k_vs(double *vt, const double *lz, const double *phi, const double *kp_ptr,
                     double kp_val, int nr1, int nt, const double *bs_src, double *bs_dst, int ns);
__global__ void k_copy(double2 *dst, const double2 *src, size_t n2);
__global__ void k_f2_in(double *dst, const double *src, int ns, int nr, int nt, int ne, int npa);
__global__ void k_f2_out(double *dst, const double *src, int ns, int nr, int nt, int ne, int npa);
}  // namespace ramgpu


using ramgpu::Tab;
using ramgpu::kMaxSpecies;

struct PlaneShape {
  int threads, maxc, rows;  // block size, row bound (template) and rows per thread of the forward plane kernels
};

struct ramgpu_ctx {
  int device = 0;
  int ns = 0, nr = 0, nt = 0, ne = 0, npa = 0;
  hipStream_t stream = nullptr;
  // a second stream (default priority) on which ramgpu_adjoint_run
  // recomputes a segment while the previous one's adjoint runs; created by
  // its first use
  hipStream_t stream2 = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  size_t lds_limit = 65536;
  PlaneShape plane{ramgpu::kPlaneThreads, 16, 0};
  Tab t{};            // host image of the problem descriptor
  Tab *d_tab = nullptr;  // its device copy, read by every kernel
  bool tab_dirty = true;
  std::vector<void *> allocs;
  // device buffers with host-visible handles
  double *d_rlz, *d_ekev, *d_ebnd, *d_we, *d_de, *d_grel, *d_grbnd, *d_v, *d_mu, *d_wmu, *d_dmu;
  int *d_upa;
  double *d_fnhs, *d_fnis, *d_bounhs, *d_bounis, *d_hdnsf, *d_didt, *d_dibndt;
  double *d_bnes, *d_dbdt, *d_vt, *d_eir, *d_eip, *d_fgeos, *d_bscale;
  int *d_mgnp;
  double *d_lz, *d_phi, *d_kp;
  double *d_stage;  // Fortran-layout staging buffer for F2 transfers
  // moments (ANISCH, FLUX): FFACTOR as [s][i][l][k], ERNH and EPP (NS,NE),
  // PPERT/PPART/EDENT/RNHTT (4 x (NS,NR,NT)), FLUX in F2's layout; allocated
  // by their first use
  double *d_ff = nullptr, *d_ernh = nullptr, *d_epp = nullptr, *d_mom = nullptr, *d_flux = nullptr;
  bool mom_dirty = true;  // d_ff's quotient tables (RN(1/FFACTOR), RN(WMU/FFACTOR)) to be formed
  bool have_grid = false, have_geom = false, have_vt = false, have_bnd = false, have_f2 = false;
  bool have_species[kMaxSpecies] = {};
  bool have_lzphi = false;
  bool have_wave = false;  // ramgpu_set_wave_params called (WAVELO's KP and plasmapause)
  double kp_host = 0.0;    // staging of the Kp copied to Tab::kpnow
  bool tables_dirty = true, ghost_dirty = true;
  bool efield_dirty = true;  // only the VT-dependent per-(I,J) terms are stale
  // one step of a Kp schedule (device pointers to its Kp and boundary
  // scales), applied by the next prepare (apply_step_inputs)
  const double *step_kp = nullptr, *step_bs = nullptr;
  // the finalize (CFL mask, SUMRC mode) of the previous step of a Kp
  // schedule, run with the next step's k_coef_ij (prepare); -1 when none.
  // Only the schedule loops of ramgpu_ram_run_schedule and
  // ramgpu_adjoint_run set it, and they flush it before they return.
  int fin_mask = -1, fin_mode = 0;
  bool cfl_valid = false;  // DtDrift* current for the present tables
  // opt-in per-kernel timing: event pairs around every launch, by kind
  bool timing = false;
  std::vector<hipEvent_t> tev;  // 2 per recorded launch
  std::vector<int> tkind;
  size_t tused = 0;
  bool bwd_lds_set = false;  // LDS attribute of the plane adjoint kernels set
  size_t head_static = 0;    // static LDS of k_step_head (0: unknown, q1 row not staged)
  // Coulomb collisions: host copies of the grid that COULPARA reads (kept by
  // ramgpu_set_grid), VBND, the tables of the last COULPARA
  // ([table][s][l*NE + k]), NECR on the device, TimeRamElapsed
  std::vector<double> h_ekev, h_we, h_grel, h_grbnd, h_v, h_mu, h_wmu, h_dmu, h_vbnd, h_coul;
  bool coul_dirty = true, have_necr = false;
  double *d_necr = nullptr;
  double time_elapsed = 0.0;
  double told = 0.0;  // TimeRamElapsed at the last ramgpu_computehI (TOld)
  // WPADIF and FLCscatter: per species the WPADIF table (device layout) and
  // whether ram_run takes it at the WPI positions (in place of WAVELO) and at
  // the EMIC positions; the FLC table (null: no FLC) and its allocation; Dt_bc
  std::vector<double *> d_pdc, d_flc, d_flc_mem;
  std::vector<char> pd_wpi, pd_emic;
  double dt_bc = 0.0;
};

namespace ramgpu_host {

int fail(int code, const char *fmt, ...);
const std::string &last_error();

#define HIP_TRY(x)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) return fail(RAMGPU_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

int check_ctx(ramgpu_ctx *c);
int ready(ramgpu_ctx *c);
int launch_check(const char *what);
void tbeg(ramgpu_ctx *c, int kind);
void tend(ramgpu_ctx *c);
int sync_tab(ramgpu_ctx *c);
ramgpu::Launch launch_of(const ramgpu_ctx *c, int s0, int nsp, int nb);
int prepare(ramgpu_ctx *c);
int apply_step_inputs(ramgpu_ctx *c, const double *d_kp_n, const double *d_bs_n);
int flush_finalize(ramgpu_ctx *c);
int fill_out(ramgpu_ctx *c, ramgpu_species_out *out);
int check_species(ramgpu_ctx *c, int s);
// sweeps of species s0..s0+nsp-1, in place on F2 or from src into dst
int run_r(ramgpu_ctx *c, int mode, int s0, int nsp, const double *src = nullptr, double *dst = nullptr);
int run_p(ramgpu_ctx *c, int rec, int s0, int nsp, const double *src = nullptr, double *dst = nullptr);
int run_plane(ramgpu_ctx *c, int seq, int rec, int s0, int nsp, const double *src = nullptr,
              double *dst = nullptr);
int run_finalize(ramgpu_ctx *c, int cflmask, int summode, int s0, int nsp);
int step_once(ramgpu_ctx *c);
int step_body(ramgpu_ctx *c, int *cflmask);
// cells per thread the plane kernels need for an NE x NPA plane (8 or 16)
PlaneShape plane_shape(int ne, int npa, size_t lds_cu);
int coef_units(const ramgpu_ctx *c, int threads);
bool takes_wave(const ramgpu_ctx *c);
PlaneShape plane_shape_for(int ne, int npa, int threads);
// launch k_plane<seq, maxc>
void launch_plane(int seq, PlaneShape sh, dim3 grid, size_t lds, hipStream_t st, const ramgpu::Launch &L, int rec);
// tape slots: the inputs of the operators of one ram_run step, in order
// (T_E1..T_E2 consecutive: the plane kernels address them from T_E1)
// T_CLAMP: the plane phases' clamp decisions, one byte plane per drift
// phase (E1, MU1, MU2, E2) in the slot's bytes
enum TapeSlot { T_R1, T_P1, T_E1, T_MU1, T_LOSS, T_MU2, T_E2, T_P2, T_R2, T_CLAMP, T_N };
// plane adjoint kernels
int plane_bwd_threads(int ne, int npa);
void set_plane_bwd_lds_limit(size_t limit);
void launch_plane_bwd(int seq, int threads, dim3 grid, size_t lds, hipStream_t st, const ramgpu::Launch &L,
                      const double *x, double *adj);
// Coulomb collisions: the COULPARA tables (formed when stale), and whether
// the plane kernels' LDS with them fits blocks of `threads`
int coul_tables(ramgpu_ctx *c);
int coul_check_lds(const ramgpu_ctx *c, int threads);
// RAMGPU_ESTATE for a context with Coulomb collisions, WPADIF or FLC
// scattering on: the reverse mode and ensembles step without them
int forward_only(const ramgpu_ctx *c);
// whether a species takes WPADIF or FLCscatter in ram_run, and the step that
// then runs: one launch per operator for all species, finalized at its end
int takes_pitch(const ramgpu_ctx *c);
int step_composed(ramgpu_ctx *c, int rec);
// host Fortran-layout F2 <-> device-layout buffer, through the context's stage
int upload_f2(ramgpu_ctx *c, double *dst, const double *host);
int download_f2(ramgpu_ctx *c, double *host, const double *src);

}  // namespace ramgpu_host
