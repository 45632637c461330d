/*
 * ramgpu.h -- C ABI of the MI355X (gfx950) RAM kinetic split step.
 *
 * Drop-in boundary for the reference's operator-split advection step.  Every
 * entry point returns RAMGPU_OK (0) or a nonzero RAMGPU_E* status; the text of
 * the last failure on the calling thread is ramgpu_last_error().  Arrays cross
 * the boundary as plain host pointers in the reference's own Fortran
 * (column-major, 1-based) layouts, so a Fortran host passes its module arrays
 * unchanged through ISO_C_BINDING:
 *   F2(NS,NR,NT,NE,NPA)                          src/ModRamInit.f90:72
 *   FNHS,FNIS,BOUNHS,BOUNIS,HDNS,dIdt,dIbndt (NR+1,NT,NPA)   :75-77
 *   BNES,dBdt,VT,EIR,EIP (NR+1,NT)                           :75-77,129
 *   outsideMGNP(NR,NT) (int)                                  :68
 *   FGEOS(NS,NT,NE,NPA)                                       :133
 *   GREL,GRBND,V(NS,NE); RLZ(NR+1); EKEV,EBND,WE,DE(NE); MU,WMU,DMU(NPA);
 *   UPA(NR)                                                   :100-106
 * The library never frees or keeps host pointers.  It owns the device copies
 * inside an opaque context; F2 stays resident on the device between calls.
 * Species indices are 1-based, as the reference's S argument.
 *
 * Threading: one context is used by one host thread at a time.  The
 * reference calls its operators from an OpenMP loop over species
 * (src/ModRamRun.f90:64); a host that keeps that loop gives each thread its
 * own context, or serialises calls on a shared one.
 */
#ifndef RAMGPU_H
#define RAMGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RAMGPU_OK 0
#define RAMGPU_EINVAL 1   /* bad argument (size, index, null pointer)     */
#define RAMGPU_ENOMEM 2   /* device or host allocation failed             */
#define RAMGPU_EHIP 3     /* a HIP runtime call failed                    */
#define RAMGPU_ESTATE 4   /* call order violated (e.g. no grid uploaded)  */

/* charge-exchange cross-section fits selected by species name in CEPARA,
 * src/ModRamLoss.f90:37-83 */
#define RAMGPU_CEX_NONE 0
#define RAMGPU_CEX_H 1
#define RAMGPU_CEX_HE 2
#define RAMGPU_CEX_O 3

typedef struct ramgpu_ctx ramgpu_ctx;

/* Per-species energy bookkeeping of one ram_run call (SUMRC accumulators,
 * src/ModRamRun.f90:77-175,231-259) and the CFL limits of the drift
 * operators (DtDriftR/P/E/Mu, src/ModRamDrift.f90:115,223,308,404). */
typedef struct {
  double setrc;                       /* SETRC(S) after the step          */
  double lsdr, lscha, lsatm;          /* LSDR/LSCHA/LSATM(S), cumulative  */
  double dtdrift_r, dtdrift_p, dtdrift_e, dtdrift_mu;
} ramgpu_species_out;

const char *ramgpu_last_error(void);

/* Context holding one (NS,NR,NT,NE,NPA) problem on device `device`.
 * Replaces the allocation in ram_allocate, src/ModRamInit.f90:67-138. */
int ramgpu_create(ramgpu_ctx **ctx, int ns, int nr, int nt, int ne, int npa, int device);
int ramgpu_destroy(ramgpu_ctx *ctx);

/* species(S): charge state s_charge, CEX kind (0 when species%CEX is false),
 * rest mass RMAS(S) in kg (src/ModRamSpecies.f90:11-30, ModRamInit.f90:418-420) */
int ramgpu_set_species(ramgpu_ctx *ctx, int s, int charge, int cex_kind, double rmas);

/* ModRamTiming::DTs, ModRamParams::BetaLim, ModRamVariables::FracCFL */
int ramgpu_set_params(ramgpu_ctx *ctx, double dts, double betalim, double fraccfl);

/* Grid of ARRAYS, src/ModRamInit.f90:364-587 */
int ramgpu_set_grid(ramgpu_ctx *ctx, double mdr, double dphi, double conf1, double conf2,
                    const double *rlz, const double *ekev, const double *ebnd,
                    const double *we, const double *de, const double *grel,
                    const double *grbnd, const double *v, const double *mu,
                    const double *wmu, const double *dmu, const double *upa);

/* Bounce-averaged geometry and induced fields supplied by the SCB side
 * (computehI, src/ModRamScb.f90:567-619). */
int ramgpu_set_geometry(ramgpu_ctx *ctx, const double *fnhs, const double *fnis,
                        const double *bounhs, const double *bounis, const double *hdns,
                        const double *didt, const double *dibndt, const double *bnes,
                        const double *dbdt, const double *eir, const double *eip,
                        const int *outside_mgnp);

/* Convection potential VT(NR+1,NT) [V] (src/ModRamRun.f90:45-54). */
int ramgpu_set_efield(ramgpu_ctx *ctx, const double *vt);

/* Volland-Stern VT computed on the device from Kp and the grid's L shells and
 * local times, exactly the formula of src/ModRamRun.f90:45-54 with PHIOFS=0:
 * lz(NR+1) and phi(NT) as built by ARRAYS. */
int ramgpu_set_kp(ramgpu_ctx *ctx, double kp, const double *lz, const double *phi);

/* Outer-boundary distribution FGEOS(NS,NT,NE,NPA) (src/ModRamBoundary.f90:288-295). */
int ramgpu_set_boundary(ramgpu_ctx *ctx, const double *fgeos);

/* Per-species multiplier of FGEOS (1 by default): the boundary seen by DRIFTR
 * is FGEOS(S,...)*scale(S), formed before the CONF1/CONF2 ghost factors. */
int ramgpu_set_boundary_scale(ramgpu_ctx *ctx, const double *scale /* (NS) */);

/* Whole F2(NS,NR,NT,NE,NPA) host <-> device. */
int ramgpu_put_f2(ramgpu_ctx *ctx, const double *f2);
int ramgpu_get_f2(ramgpu_ctx *ctx, double *f2);

/* SETRC/LSDR/LSCHA/LSATM(S) state (zero after create, as ram_init sets them,
 * src/ModRamInit.f90:147). */
int ramgpu_set_energy_state(ramgpu_ctx *ctx, const double *setrc, const double *lsdr,
                            const double *lscha, const double *lsatm);
int ramgpu_get_energy_state(ramgpu_ctx *ctx, double *setrc, double *lsdr, double *lscha,
                            double *lsatm);

/* ---- operators, one per reference subroutine (S is 1-based) ------------- */

/* DRIFTPARA(S), src/ModRamDrift.f90:36-88: builds the drift coefficient
 * tables (for every species; they are cached until an input changes). */
int ramgpu_driftpara(ramgpu_ctx *ctx, int s);
/* DRIFTR/DRIFTP/DRIFTE/DRIFTMU(S), src/ModRamDrift.f90:95,204,285,382.
 * dtdrift (may be NULL) receives DtDriftR/P/E/Mu(S). */
int ramgpu_driftr(ramgpu_ctx *ctx, int s, double *dtdrift);
int ramgpu_driftp(ramgpu_ctx *ctx, int s, double *dtdrift);
int ramgpu_drifte(ramgpu_ctx *ctx, int s, double *dtdrift);
int ramgpu_driftmu(ramgpu_ctx *ctx, int s, double *dtdrift);
/* DRIFTEND, src/ModRamDrift.f90:23-30 (tables stay cached; no-op kept for
 * interface parity). */
int ramgpu_driftend(ramgpu_ctx *ctx);
/* CEPARA/CHAREXCHANGE/ATMOL(S), src/ModRamLoss.f90:19,457,485 */
int ramgpu_cepara(ramgpu_ctx *ctx, int s);
int ramgpu_charexchange(ramgpu_ctx *ctx, int s);
int ramgpu_atmol(ramgpu_ctx *ctx, int s);
/* SUMRC(S), src/ModRamRun.f90:231-259: setrc receives the new SETRC(S),
 * elorc = old - new.  Either may be NULL. */
int ramgpu_sumrc(ramgpu_ctx *ctx, int s, double *setrc, double *elorc);

/* ---- the fused split step ----------------------------------------------- */

/* `nsteps` consecutive ram_run calls (src/ModRamRun.f90:64-205, optional
 * processes off): per species R,P,E,Mu drifts, charge exchange, atmospheric
 * loss twice, charge exchange, Mu,E,P,R drifts with the SUMRC bookkeeping,
 * then the periodic local-time copy and the magnetopause mask.  out (may be
 * NULL) receives NS records for the last step. */
int ramgpu_ram_run(ramgpu_ctx *ctx, int nsteps, ramgpu_species_out *out);

/* The same with a per-step Kp and boundary scale schedule held on the host:
 * before step n the device recomputes VT for kp[n] (as ramgpu_set_kp, with
 * the lz/phi given there) and applies bscale[n*NS + s-1] (may be NULL). */
int ramgpu_ram_run_schedule(ramgpu_ctx *ctx, int nsteps, const double *kp,
                            const double *bscale, ramgpu_species_out *out);

/* Synchronise and return the device time (ms) of the last ram_run call,
 * measured with HIP events on the library's stream. */
int ramgpu_last_run_ms(ramgpu_ctx *ctx, double *ms);

/* Opt-in per-kernel timing (no counterpart in the reference: measurement
 * support).  While enabled, every launch is bracketed by HIP events on the
 * library's stream; ramgpu_kernel_times returns, per kernel kind, the summed
 * device time and the launch count since the last query, then resets. */
#define RAMGPU_KERNEL_PREPARE 0  /* species tables, Courant terms, ghost carry */
#define RAMGPU_KERNEL_IJ 1       /* radial / azimuthal sweeps                 */
#define RAMGPU_KERNEL_KL 2       /* energy / pitch-angle sweeps, losses       */
#define RAMGPU_KERNEL_FINALIZE 3 /* DtDrift* and SUMRC bookkeeping            */
#define RAMGPU_KERNEL_VS 4       /* Volland-Stern VT from Kp                  */
#define RAMGPU_KERNEL_KINDS 5
int ramgpu_kernel_timing(ramgpu_ctx *ctx, int enable);
/* Measurement support: a plain device copy of `bytes` (read + write) run
 * `reps` times on device `device`; gbps receives the achieved read+write
 * bandwidth.  Used as the measured HBM peak and to calibrate counters. */
int ramgpu_copy_probe(int device, size_t bytes, int reps, double *gbps);
int ramgpu_kernel_times(ramgpu_ctx *ctx, double *total_ms, int *launches);

#ifdef __cplusplus
}
#endif
#endif
