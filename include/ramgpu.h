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
  double lswae, lscoe, lscsc;         /* LSWAE (WAVELO), LSCOE/LSCSC (Coulomb
                                         energy loss / scattering), cumulative */
} ramgpu_species_out;

const char *ramgpu_last_error(void);

/* Context holding one (NS,NR,NT,NE,NPA) problem on device `device`.
 * Replaces the allocation in ram_allocate, src/ModRamInit.f90:67-138.
 * Sizes: every axis >= 4; NE*NPA <= 8192, one (K,L) plane in one block's
 * LDS; NT*NE*NPA and NR*NT below 2^22; and NR, NT small enough that the
 * face terms a sweep block stages in LDS fit the device (160 KB on MI355X:
 * NR <= 149 and NT <= 62 at NE = NPA = 4, NR <= 1017 and NT <= 462 at
 * 35 x 72).  Returns RAMGPU_EINVAL otherwise. */
int ramgpu_create(ramgpu_ctx **ctx, int ns, int nr, int nt, int ne, int npa, int device);
int ramgpu_destroy(ramgpu_ctx *ctx);

/* species(S): charge state s_charge, CEX kind (0 when species%CEX is false),
 * rest mass RMAS(S) in kg (src/ModRamSpecies.f90:11-30, ModRamInit.f90:418-420) */
int ramgpu_set_species(ramgpu_ctx *ctx, int s, int charge, int cex_kind, double rmas);

/* ModRamTiming::DTs, ModRamParams::BetaLim, ModRamVariables::FracCFL */
int ramgpu_set_params(ramgpu_ctx *ctx, double dts, double betalim, double fraccfl);
/* Launch shape of the energy / pitch-angle plane kernels (no counterpart in
 * the reference; a tuning control): blocks of `threads` (a multiple of 64,
 * at most 1024) in which each thread owns one K of ceil(NPA/floor(threads/NE))
 * rows, at most `maxc` (8 or 16; 16 with at most 512 threads).  0, 0
 * restores the shape the library picks from NE, NPA and the device's LDS.
 * The SUMRC sums add the plane's cells in an order that follows the shape, so
 * ensembles keep each member's shape. */
int ramgpu_set_plane_shape(ramgpu_ctx *ctx, int threads, int maxc);
int ramgpu_get_plane_shape(ramgpu_ctx *ctx, int *threads, int *maxc);

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

/* The geometry of the context back to the host, in the shapes of
 * ramgpu_set_geometry (ModRamVariables' FNHS ... EIP, outsideMGNP); any
 * pointer may be null. */
int ramgpu_get_geometry(ramgpu_ctx *ctx, double *fnhs, double *fnis, double *bounhs,
                        double *bounis, double *hdns, double *didt, double *dibndt,
                        double *bnes, double *dbdt, double *eir, double *eip,
                        int *outside_mgnp);

/* computehI (src/ModRamScb.f90:144-643) after its field-line step: from the
 * field lines on the RAM grid, x, y, z [RE] and B [units of bnormal], each
 * (nthe, NR, NT) as xRAM(k,i,j) with i = 1..NR the rows I = 2..NR+1, the line
 * coordinate chi(nthe) (chiVal, strictly increasing) and the equator sample
 * ieq (nThetaEquator, 1-based): the h and I integrals and the Rairden density
 * bounce average of every line with outside_mgnp(NR,NT) = 0, the rows from
 * scale_at(NT) (ScaleAt, 0 none; may be null) scaled as the reference does
 * beyond the SCB domain, its corrections, Steffen interpolation to pabn(NPA)
 * over pa(NPA) [deg], the Gaussian smoothing when smooth != 0 (needs NR, NT
 * >= 9), and FNHS, FNIS, BOUNHS, BOUNIS, HDNS, BNES, dIdt, dIbndt, dBdt with
 * DthI = TimeRamElapsed - TOld, written into the context's geometry;
 * EIR(1,:) = EIP(1,:) = 0; outsideMGNP taken from outside_mgnp; TOld set to
 * TimeRamElapsed.  lz(NR+1) is LZ.  The integrals are exact for the
 * piecewise-linear B the reference's quadrature samples (GSL cquad, 1e-3). */
int ramgpu_computehI(ramgpu_ctx *ctx, int nthe, int ieq, const double *x, const double *y,
                     const double *z, const double *b, const double *chi, const double *pa,
                     const double *pabn, const double *lz, const int *outside_mgnp,
                     const int *scale_at, double bnormal, int smooth);

/* TOld (src/ModRamTiming.f90:20), the TimeRamElapsed of the last
 * computehI: set by ramgpu_computehI, restored from a restart file
 * (src/ModRamRestart.f90 writes and reads it). */
int ramgpu_set_told(ramgpu_ctx *ctx, double told);
int ramgpu_get_told(ramgpu_ctx *ctx, double *told);

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

/* LSWAE/LSCOE/LSCSC(S) (src/ModRamRun.f90:84-98,151-168; zero after
 * create).  Any pointer may be NULL. */
int ramgpu_set_loss_state(ramgpu_ctx *ctx, const double *lswae, const double *lscoe, const double *lscsc);
int ramgpu_get_loss_state(ramgpu_ctx *ctx, double *lswae, double *lscoe, double *lscsc);

/* Electron lifetimes (WAVELO, src/ModRamWPI.f90:580-636), taken in ram_run by
 * the species with species%WPI set while DoUseWPI is off (the reference's
 * default for electrons, src/ModRamSpecies.f90:49; src/ModRamRun.f90:91-99,
 * 151-159).  ramgpu_set_wavelo gives species s the lifetime tables WALOS1,
 * WALOS2, WALOS3 (NR,NE) of WAVEPARA1/2 (src/ModRamWPI.f90:18-182) and
 * switches WAVELO on for it (all three NULL: off).  ramgpu_set_wave_params
 * sets KP and the plasmapause, 5.39-0.382*Kpmax12 or, with a plasmasphere
 * model (use_plasmasphere, DoUsePlasmasphere), the largest LZ(I), I=2..NR,
 * with NECR(I,J) (NR,NT) above 50; lz(NR+1) is LZ.  KP follows every later
 * ramgpu_set_kp and each step of a Kp schedule. */
int ramgpu_set_wavelo(ramgpu_ctx *ctx, int s, const double *walos1, const double *walos2,
                      const double *walos3);
int ramgpu_set_wave_params(ramgpu_ctx *ctx, double kp, double kpmax12, const double *lz,
                           const double *necr, int use_plasmasphere);

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
/* WAVELO(S), src/ModRamWPI.f90:580-636 (the species must have its tables) */
int ramgpu_wavelo(ramgpu_ctx *ctx, int s);
/* SUMRC(S), src/ModRamRun.f90:231-259: setrc receives the new SETRC(S),
 * elorc = old - new.  Either may be NULL. */
int ramgpu_sumrc(ramgpu_ctx *ctx, int s, double *setrc, double *elorc);

/* ---- Coulomb collisions (src/ModRamCoul.f90) ---------------------------- */

/* DoUseCoulomb (src/ModRamParams.f90:110, read by IM_set_parameters.f90:95):
 * use = 1 switches COULEN and COULMU into ramgpu_ram_run's species loop as
 * ram_run calls them (src/ModRamRun.f90:81-89,160-167, with LSCOE and LSCSC
 * in ramgpu_species_out), with vbnd = VBND(NS,NE), the speed at the energy
 * bin boundaries (src/ModRamInit.f90:448,457); use = 0 switches them off
 * (vbnd may be NULL).  Needs NPA >= 13 (COULEN copies BANE(NPA-11) into
 * rows NPA-10..NPA).  The reverse-mode entry points and ensembles refuse a
 * context with Coulomb collisions on (RAMGPU_ESTATE). */
int ramgpu_set_coulomb(ramgpu_ctx *ctx, int use, const double *vbnd);
/* The plasmaspheric electron density NECR(NR,NT) [cm^-3] the Coulomb
 * operators read (src/ModRamVariables.f90, filled by the reference's
 * plasmasphere models, src/ModRamPlasmasphere.f90:90,294-299); required
 * before a step with Coulomb collisions. */
int ramgpu_set_necr(ramgpu_ctx *ctx, const double *necr);
/* TimeRamElapsed [s] (src/ModRamTiming.f90:19; 0 after create): COULMU sets
 * negative cells to 1e-15 only when it is positive (src/ModRamCoul.f90:286).
 * Each step of ramgpu_ram_run and of the schedules advances it by 2*DTs, as
 * src/ModRamScbRun.f90:88 does after every ram_run call. */
int ramgpu_set_time(ramgpu_ctx *ctx, double time_elapsed);
int ramgpu_get_time(ramgpu_ctx *ctx, double *time_elapsed);
/* COULPARA(S), src/ModRamCoul.f90:17-125: builds the Coulomb tables (for
 * every species; cached until the grid, a species, VBND or DTs changes). */
int ramgpu_coulpara(ramgpu_ctx *ctx, int s);
/* The tables of species s as COULPARA leaves them: COULE(S,:,:),
 * COULI, ATA, GTA, CEDR, CIDR, each (NE,NPA); any pointer may be NULL. */
int ramgpu_get_coulomb_tables(ramgpu_ctx *ctx, int s, double *coule, double *couli, double *ata,
                              double *gta, double *cedr, double *cidr);
/* COULEN(S), src/ModRamCoul.f90:133-221, and COULMU(S), :229-296, on the
 * device-resident F2 (Coulomb collisions must be on). */
int ramgpu_coulen(ramgpu_ctx *ctx, int s);
int ramgpu_coulmu(ramgpu_ctx *ctx, int s);

/* ---- implicit pitch-angle diffusion: WPADIF, FLCscatter ------------------ */

/* The WPADIF coefficients of species s (src/ModRamWPI.f90:643-714): the
 * bounce-averaged diffusion coefficient d1 + d2, both (NR,NT,NE,NPA) --
 * ATAW + ATAC for electrons (hiss and chorus), ATAW_emic_h + ATAW_emic_he
 * for ions (EMIC) -- d2 may be NULL.  In ramgpu_ram_run the species then
 * takes WPADIF at the WPI positions in place of WAVELO when at_wpi
 * (species%WPI and DoUseWPI, src/ModRamRun.f90:91-99,151-159) and at the
 * EMIC positions when at_emic (species%EMIC and DoUseEMIC, :101-106,
 * 144-149).  d1 NULL: the species takes no WPADIF. */
int ramgpu_set_wpadif(ramgpu_ctx *ctx, int s, const double *d1, const double *d2, int at_wpi, int at_emic);
/* WPADIF(S) on the device-resident F2 with the species' coefficients. */
int ramgpu_wpadif(ramgpu_ctx *ctx, int s);
/* FLC_coef(S,:,:,:,:) (NR,NT,NE,NPA) of species s, the field-line-curvature
 * pitch-angle diffusion coefficient of PARA_FLC (src/ModRamLoss.f90:342),
 * and Dt_bc (src/ModRamTiming.f90): ramgpu_ram_run then takes FLCscatter
 * for the species around its atmospheric losses (species%FLC and DoUseFLC,
 * src/ModRamRun.f90:114-119,130-135); coef NULL: no FLC for the species. */
int ramgpu_set_flc(ramgpu_ctx *ctx, int s, const double *coef, double dt_bc);
/* FLCscatter(S), src/ModRamLoss.f90:513-575: nothing while TimeRamElapsed
 * < Dt_bc. */
int ramgpu_flcscatter(ramgpu_ctx *ctx, int s);

/* ---- moments of the device-resident F2 ---------------------------------- */

/* The moment factors of ARRAYS (src/ModRamInit.f90:560-576): FFACTOR
 * (NS,NR,NE,NPA), ERNH and EPP (NS,NE).  Needed before the calls below. */
int ramgpu_set_moment_factors(ramgpu_ctx *ctx, const double *ffactor, const double *ernh,
                              const double *epp);
/* ANISCH(S), src/ModRamRun.f90:265-415, without its wave-particle branch
 * (:422-, DoUseWPI) and its file output.  Sets F2(S,2:NR,:,2:NE,1) =
 * F2(S,2:NR,:,2:NE,2) on the device as the reference does, and writes the
 * (S, 2..NR, 1..NT) entries of PPERT and PPART, and of EDENT and RNHTT (the
 * energy and number densities of its wans file), all (NS,NR,NT) arrays;
 * other entries are left alone, any pointer may be NULL.  The energy
 * windows end at min(khi, NE) (the reference reads past NE when NE < 30). */
int ramgpu_anisch(ramgpu_ctx *ctx, int s, double *ppert, double *ppart, double *edent, double *rnhtt);
/* FLUX(NS,NR,NT,NE,NPA) = F2/FFACTOR/FNHS, src/ModRamRun.f90:208-222, of
 * the current F2; 0 outside I=2..NR, K=2..NE, L=2..NPA, J=1..NT-1, as the
 * reference's FLUX (zeroed at allocation, src/ModRamInit.f90:78). */
int ramgpu_flux(ramgpu_ctx *ctx, double *flux);
/* The block of ram_run after its species loop (src/ModRamRun.f90:207-222):
 * ANISCH of every species, then FLUX, in one launch.  flux may be NULL: then
 * only the NS*NR*NT moments leave the device. */
int ramgpu_moments(ramgpu_ctx *ctx, double *ppert, double *ppart, double *edent, double *rnhtt,
                   double *flux);

/* ---- the fused split step ----------------------------------------------- */

/* `nsteps` consecutive ram_run calls (src/ModRamRun.f90:64-205): per species
 * R,P,E,Mu drifts, [COULEN, COULMU], [WAVELO or WPADIF], [EMIC WPADIF],
 * [charge exchange], [FLCscatter], atmospheric loss twice, and the same
 * back in reverse, Mu,E,P,R drifts with the SUMRC bookkeeping, then the
 * periodic local-time copy and the magnetopause mask.  Each optional
 * process as its set-up call switches it on: WAVELO (ramgpu_set_wavelo),
 * Coulomb collisions (ramgpu_set_coulomb), WPADIF and EMIC
 * (ramgpu_set_wpadif), FLC scattering (ramgpu_set_flc).  Without WPADIF and
 * FLC the middle of the step is one fused plane kernel; with them the step
 * runs one launch per operator for all species at once (the same results).
 * out (may be NULL) receives NS records for the last step. */
int ramgpu_ram_run(ramgpu_ctx *ctx, int nsteps, ramgpu_species_out *out);

/* The same with a per-step Kp and boundary scale schedule held on the host:
 * before step n the device recomputes VT for kp[n] (as ramgpu_set_kp, with
 * the lz/phi given there) and applies bscale[n*NS + s-1] (may be NULL). */
int ramgpu_ram_run_schedule(ramgpu_ctx *ctx, int nsteps, const double *kp,
                            const double *bscale, ramgpu_species_out *out);

/* ramgpu_ram_run_schedule without any allocation (no counterpart in the
 * reference: it is for callers that cannot allocate device memory while they
 * step, such as tight forecast loops and harnesses that forbid hipMalloc).
 * The schedule is staged in a fixed array of 65536 doubles that the library
 * keeps on every device it runs on until the process ends, instead of a
 * per-call allocation: nsteps * (1 + NS) doubles with bscale, nsteps without,
 * must fit it (RAMGPU_EINVAL otherwise; e.g. 13107 steps of 4 species).  The
 * calls of all contexts on a device share that array, which keeps the last
 * call's schedule: the library serialises these calls within a process and
 * returns only after its stream has finished reading the array, so calls from
 * several host threads wait for each other; concurrent calls from several
 * processes on one device are not coordinated.  Results are bit-identical to
 * ramgpu_ram_run_schedule's. */
int ramgpu_ram_run_schedule_noalloc(ramgpu_ctx *ctx, int nsteps, const double *kp,
                                    const double *bscale, ramgpu_species_out *out);

/* Synchronise and return the device time (ms) of the last ram_run call,
 * measured with HIP events on the library's stream. */
int ramgpu_last_run_ms(ramgpu_ctx *ctx, double *ms);

/* Opt-in per-kernel timing (no counterpart in the reference: measurement
 * support).  While enabled, every launch is bracketed by HIP events on the
 * library's stream; ramgpu_kernel_times returns, per kernel kind, the summed
 * device time and the launch count since the last query, then resets. */
#define RAMGPU_KERNEL_PREPARE 0  /* species tables, Courant terms, ghost carry */
#define RAMGPU_KERNEL_R 1        /* radial sweeps (DRIFTR)                    */
#define RAMGPU_KERNEL_PLANE 2    /* energy / pitch-angle sweeps, losses, SUMRC */
#define RAMGPU_KERNEL_FINALIZE 3 /* DtDrift* and SUMRC bookkeeping            */
#define RAMGPU_KERNEL_VS 4       /* Volland-Stern VT from Kp; a schedule forms
                                     it with the Courant-number terms, timed
                                     under RAMGPU_KERNEL_PREPARE            */
#define RAMGPU_KERNEL_P 5        /* azimuthal sweeps (DRIFTP)                 */
#define RAMGPU_KERNEL_R_BWD 6    /* adjoint of DRIFTR                         */
#define RAMGPU_KERNEL_P_BWD 7    /* adjoint of DRIFTP                         */
#define RAMGPU_KERNEL_PLANE_BWD 8 /* adjoints of DRIFTE, DRIFTMU, losses      */
#define RAMGPU_KERNEL_MOMENTS 9  /* ANISCH pressure moments and FLUX          */
#define RAMGPU_KERNEL_PITCH 10   /* WPADIF and FLCscatter line solves          */
#define RAMGPU_KERNEL_KINDS 11
int ramgpu_kernel_timing(ramgpu_ctx *ctx, int enable);
/* Measurement support: a plain device copy of `bytes` (read + write) run
 * `reps` times on device `device`; gbps receives the achieved read+write
 * bandwidth.  Used as the measured HBM peak and to calibrate counters. */
int ramgpu_copy_probe(int device, size_t bytes, int reps, double *gbps);
int ramgpu_kernel_times(ramgpu_ctx *ctx, double *total_ms, int *launches);

/* ---- reverse-mode differentiation ------------------------------------ */

/* The split step is piecewise linear in (F2, FGEOS): every Courant number,
 * CHARGE and ATLOS factor depends on fields, geometry and grid only, so the
 * limiter branch, upwind side and clamp F2<0 -> 1e-15 of each cell are fixed
 * by the operator's input.  The backward entry points are exact
 * vector-Jacobian products at that input.  Subgradient conventions: a
 * clamped cell passes no derivative (a cell exactly at 0 is not clamped);
 * at a kink of the limiter max(min(BetaLim*R,1), min(R,BetaLim)) a tie of
 * min or max takes the derivative of its first argument; R <= 0 and
 * |F(I+1)-F(I)| <= 1e-27 take the upwind value.  (No counterpart in the
 * reference, which has no derivatives: these serve data assimilation.)
 *
 * Operator pairs: ramgpu_<op> (forward, above) and ramgpu_<op>_bwd
 * (backward).  The backward is evaluated at the context's current F2, which
 * must be the operator's input; F2 is not changed.  adj (NS,NR,NT,NE,NPA)
 * holds dL/d(output) on entry and dL/d(input) on return; fgeos_adj
 * (NS,NT,NE,NPA, may be NULL) accumulates dL/dFGEOS (DRIFTR only). */
int ramgpu_driftr_bwd(ramgpu_ctx *ctx, int s, double *adj, double *fgeos_adj);
int ramgpu_driftp_bwd(ramgpu_ctx *ctx, int s, double *adj);
int ramgpu_drifte_bwd(ramgpu_ctx *ctx, int s, double *adj);
int ramgpu_driftmu_bwd(ramgpu_ctx *ctx, int s, double *adj);
int ramgpu_charexchange_bwd(ramgpu_ctx *ctx, int s, double *adj);
int ramgpu_atmol_bwd(ramgpu_ctx *ctx, int s, double *adj);
/* Backward of one ram_run step at the current F2 (the step's input; F2 and
 * the SUMRC state are left unchanged). */
int ramgpu_ram_run_bwd(ramgpu_ctx *ctx, double *adj, double *fgeos_adj);

/* Gradient of a virtual-satellite flux misfit over an nsteps ram_run run
 * starting from the current F2, with the schedule of ramgpu_ram_run_schedule
 * (kp, bscale may be NULL):
 *   L = sum_n sum_p ( log(omega_p(F2 after step n)) - target[n*npts+p] )^2,
 *   omega_p = sum_L w[p*NPA + L-1] * F2(S_p,I_p,J_p,K_p,L),
 * with pts[4p..4p+3] = S_p,I_p,J_p,K_p (1-based).  (omega is the
 * omnidirectional flux of src/ModRamRun.f90:214 and src/ModRamSats.f90:661
 * when w = 4*pi*WMU/(FFACTOR*FNHS).)  loss receives L, grad_f2
 * (NS,NR,NT,NE,NPA) dL/dF2 at the start, grad_fgeos (may be NULL)
 * [nsteps](NS,NT,NE,NPA) dL/dFGEOS of each step.  Checkpoints every
 * ckpt_every steps; each segment is recomputed once with every operator's
 * input on a tape.  The context is left in the state after the run. */
int ramgpu_adjoint_run(ramgpu_ctx *ctx, int nsteps, int ckpt_every, const double *kp,
                       const double *bscale, int npts, const int *pts, const double *w,
                       const double *target, double *loss, double *grad_f2, double *grad_fgeos);

/* ---- ensembles ------------------------------------------------------- */

/* An ensemble of independent problems stepped together (no counterpart in the
 * reference, which runs one problem per process: it serves ensemble
 * forecasts and ensemble data assimilation).  Members are contexts created
 * on one device, each with its own grid sizes, species, fields and F2; the
 * ensemble does not own them and must be destroyed before them.  Every kernel
 * of a step is one launch over all members, and a step is one captured HIP
 * graph, replayed per step.  A member's results are bit-identical to those
 * of stepping it alone with ramgpu_ram_run / ramgpu_ram_run_schedule. */
typedef struct ramgpu_batch ramgpu_batch;
int ramgpu_batch_create(ramgpu_batch **batch, ramgpu_ctx *const *members, int nmembers);
int ramgpu_batch_destroy(ramgpu_batch *batch);
/* nsteps ram_run steps of every member; out (may be NULL) receives the
 * members' ramgpu_species_out records, member after member (NS_m each). */
int ramgpu_batch_ram_run(ramgpu_batch *batch, int nsteps, ramgpu_species_out *out);
/* With per-member schedules: before step n member m gets VT for
 * kp[n*M + m] (as ramgpu_set_kp) and boundary scales
 * bscale[n*sum(NS) + off_m + s-1], off_m = NS_0+..+NS_(m-1) (may be NULL:
 * the members' current scales are kept). */
int ramgpu_batch_ram_run_schedule(ramgpu_batch *batch, int nsteps, const double *kp,
                                  const double *bscale, ramgpu_species_out *out);
/* Device time (ms) of the last ensemble run, HIP events on its stream. */
int ramgpu_batch_last_run_ms(ramgpu_batch *batch, double *ms);
/* Opt-in per-launch timing of the ensemble's steps, as ramgpu_kernel_timing
 * (measurement support; steps are launched eagerly, not from the graph, while
 * it is on) and the summed times and launch counts per kernel kind since the
 * last query (RAMGPU_KERNEL_* indices). */
int ramgpu_batch_kernel_timing(ramgpu_batch *batch, int enable);
int ramgpu_batch_kernel_times(ramgpu_batch *batch, double *total_ms, int *launches);
/* 1 (default): replay a captured graph per step; 0: launch eagerly. */
int ramgpu_batch_use_graph(ramgpu_batch *batch, int enable);
/* Autotuner: times every candidate launch grid of the R, P and plane
 * kernels (one block per work unit, or 1/2/4/8 resident blocks per CU
 * looping over the units) with `reps` timed launches each and keeps the
 * fastest.  The members' state is restored afterwards.  ms_best (may be
 * NULL) receives the best per-launch times of R, P and plane. */
int ramgpu_batch_autotune(ramgpu_batch *batch, int reps, double *ms_best /* [3] */);
/* Launch grid caps of R, P, plane (0 = one block per work unit). */
int ramgpu_batch_get_shapes(ramgpu_batch *batch, int *caps /* [3] */);
int ramgpu_batch_set_shapes(ramgpu_batch *batch, const int *caps /* [3] */);
/* Reverse mode of an ensemble: ramgpu_adjoint_run for every member at once
 * (each kernel one launch over all members), from the members' current F2,
 * with the per-member schedule of ramgpu_batch_ram_run_schedule (kp
 * required, bscale may be NULL).  Member m has npts[m] misfit points
 * pts[m][4p..4p+3], weights w[m][p*NPA_m + L-1] and targets
 * target[m][n*npts[m]+p], as ramgpu_adjoint_run; loss[m] receives its loss,
 * grad_f2[m] (may be NULL, or hold NULLs) its dL/dF2 at the start, and
 * grad_fgeos[m] (likewise) its [nsteps](NS,NT,NE,NPA) dL/dFGEOS.  A member's
 * results equal those of ramgpu_adjoint_run on that member alone.  When the
 * taped segments of all members do not fit the device, each segment's states
 * are kept and every step re-taped just before its adjoint (one more forward
 * step per step, the same results). */
int ramgpu_batch_adjoint_run(ramgpu_batch *batch, int nsteps, int ckpt_every, const double *kp,
                             const double *bscale, const int *npts, const int *const *pts,
                             const double *const *w, const double *const *target, double *loss,
                             double *const *grad_f2, double *const *grad_fgeos);

#ifdef __cplusplus
}
#endif
#endif
