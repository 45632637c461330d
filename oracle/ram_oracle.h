/*
 * ram_oracle.h -- CPU restatement of the RAM kinetic split step (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This is the oracle the HIP path is checked against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the timed CPU baseline -- never as the product.
 *
 * Every array uses the reference's Fortran (column-major, 1-based) layout so
 * that the same host buffers can be handed to this oracle, to the reference's
 * own Fortran (oracle/_ref) and to the C-ABI of the GPU library:
 *   F2(NS,NR,NT,NE,NPA)                    src/ModRamInit.f90:72
 *   FNHS,FNIS,BOUNHS,BOUNIS,HDNS,dIdt,dIbndt (NR+1,NT,NPA)   :75-77
 *   BNES,dBdt,VT,EIR,EIP (NR+1,NT)                           :129
 *   FGEOS(NS,NT,NE,NPA)                                       :133
 *   outsideMGNP(NR,NT) (int)
 *   GREL,GRBND,V (NS,NE)
 */
#ifndef RAM_ORACLE_H
#define RAM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Charge-exchange cross-section family of a species
 * (src/ModRamLoss.f90:37-83, selected by species%s_name). */
enum { ORACLE_CEX_NONE = 0, ORACLE_CEX_H = 1, ORACLE_CEX_HE = 2, ORACLE_CEX_O = 3 };

typedef struct {
  int ns, nr, nt, ne, npa;
  /* scalar parameters: ModRamTiming::DTs, ModRamParams::BetaLim,
   * ModRamVariables::FracCFL, MDR, DPHI, CONF1, CONF2 */
  double dts, betalim, fraccfl, mdr, dphi, conf1, conf2;
  /* species */
  const int *charge;     /* (NS)  species%s_charge                  */
  const int *cex_kind;   /* (NS)  ORACLE_CEX_* ; 0 => species%CEX false */
  const double *rmas;    /* (NS)  RMAS                              */
  /* 1-D grids */
  const double *rlz;     /* (NR+1) */
  const double *ekev, *ebnd, *we, *de;   /* (NE) */
  const double *grel, *grbnd, *v;        /* (NS,NE) */
  const double *mu, *wmu, *dmu;          /* (NPA) */
  const double *upa;                     /* (NR) real-valued index */
  /* geometry and fields */
  const double *fnhs, *fnis, *bounhs, *bounis, *hdns, *didt, *dibndt;
  const double *bnes, *dbdt, *vt, *eir, *eip;
  const int *outside_mgnp;
  const double *fgeos;
  /* state */
  double *f2;
  double *dtdrift_r, *dtdrift_p, *dtdrift_e, *dtdrift_mu; /* (NS) */
  double *setrc, *elorc, *lsdr, *lscha, *lsatm;          /* (NS) */
  /* electron lifetimes, WAVELO (src/ModRamWPI.f90:580-636), in ram_run for
   * species with species%WPI and DoUseWPI off; wpi NULL: off for all */
  const int *wpi;                          /* (NS) species%WPI              */
  const double *walos1, *walos2, *walos3;  /* (NS,NR,NE) WAVEPARA1/2 tables */
  const double *lz;                        /* (NR+1) LZ                     */
  const double *rlpp;                      /* (NT) plasmapause L of WAVELO  */
  double kp;                               /* KP                            */
  double *lswae;                           /* (NS) LSWAE                    */
  /* Coulomb collisions (src/ModRamCoul.f90), in ram_run when use_coulomb
   * (DoUseCoulomb) is set: VBND (NS,NE), the plasmasphere density NECR
   * (NR,NT), TimeRamElapsed, the tables COULPARA writes, COULE, COULI, ATA,
   * GTA, CEDR, CIDR (NS,NE,NPA), and the accumulators LSCOE, LSCSC (NS) */
  const double *vbnd, *necr;
  double time_elapsed;
  double *coule, *couli, *ata, *gta, *cedr, *cidr;
  int use_coulomb;
  double *lscoe, *lscsc;
  /* WPADIF and FLCscatter in ram_run (src/ModRamRun.f90:91-111,122-150):
   * per species the WPADIF coefficients pd1[s-1] + pd2[s-1] (NR,NT,NE,NPA;
   * pd2[s-1] may be NULL), taken at the WPI positions in place of WAVELO
   * when pd_wpi[s-1] (species%WPI and DoUseWPI) and at the EMIC positions
   * when pd_emic[s-1] (species%EMIC and DoUseEMIC); FLC_coef(S,...) as
   * flc[s-1] (NULL: no FLC) with Dt_bc.  NULL arrays: none of them. */
  const double *const *pd1, *const *pd2;
  const int *pd_wpi, *pd_emic;
  const double *const *flc;
  double dt_bc;
} oracle_state;

/* Per-species drift tables built by DRIFTPARA (src/ModRamDrift.f90:15-19). */
typedef struct {
  int qs;
  double *vr, *p1;           /* (NR)      */
  double *p2, *edot;         /* (NR,NE)   */
  double *mudot;             /* (NR,NPA)  */
  double *cdr, *cdp, *cde, *cdmu; /* (NR,NT,NE,NPA) */
} oracle_drift_scratch;

/* Per-species loss tables built by CEPARA (src/ModRamLoss.f90:19-170). */
typedef struct {
  double *charge;  /* (NR,NT,NE,NPA) for this species */
  double *atlos;   /* (NR,NE) */
} oracle_loss_scratch;

int  oracle_drift_alloc(const oracle_state *st, oracle_drift_scratch *d);
void oracle_drift_free(oracle_drift_scratch *d);
int  oracle_loss_alloc(const oracle_state *st, oracle_loss_scratch *l);
void oracle_loss_free(oracle_loss_scratch *l);

/* Species index s is 1-based, as in the reference. */
void oracle_driftpara(const oracle_state *st, int s, oracle_drift_scratch *d);
void oracle_driftr(oracle_state *st, int s, oracle_drift_scratch *d);
void oracle_driftp(oracle_state *st, int s, oracle_drift_scratch *d);
void oracle_drifte(oracle_state *st, int s, oracle_drift_scratch *d);
void oracle_driftmu(oracle_state *st, int s, oracle_drift_scratch *d);
void oracle_cepara(const oracle_state *st, int s, oracle_loss_scratch *l);
void oracle_charexchange(oracle_state *st, int s, const oracle_loss_scratch *l);
void oracle_atmol(oracle_state *st, int s, const oracle_loss_scratch *l);
void oracle_sumrc(oracle_state *st, int s);
/* WAVELO(S), src/ModRamWPI.f90:580-636, with the plasmapause RLpp(J) of
 * :596-603 given in st->rlpp. */
void oracle_wavelo(oracle_state *st, int s);
/* COULPARA(S), COULEN(S), COULMU(S), src/ModRamCoul.f90:17-125,133-221,
 * 229-296. */
void oracle_coulpara(oracle_state *st, int s);
void oracle_coulen(oracle_state *st, int s);
void oracle_coulmu(oracle_state *st, int s);
/* WPADIF(S), src/ModRamWPI.f90:643-714, with D = d1 + d2 (NR,NT,NE,NPA):
 * ATAW + ATAC for electrons, ATAW_emic_h + ATAW_emic_he for ions.
 * FLCscatter(S), src/ModRamLoss.f90:513-575, with coef = FLC_coef(S,...)
 * (NR,NT,NE,NPA); no change while st->time_elapsed < dt_bc. */
void oracle_wpadif(oracle_state *st, int s, const double *d1, const double *d2);
void oracle_flcscatter(oracle_state *st, int s, const double *coef, double dt_bc);

/* One ram_run call (src/ModRamRun.f90:64-205): WAVELO for species with wpi
 * set, WPADIF and FLCscatter as pd_wpi, pd_emic and flc give them, Coulomb
 * collisions with use_coulomb (their tables from COULPARA, and the COULMU
 * clamp of TimeRamElapsed > 0 from st->time_elapsed, which this call does
 * not advance).  nthreads>1 runs the species
 * loop with OpenMP as the reference does (src/ModRamRun.f90:64).  Returns
 * DtsNext before the DtsMin floor in *dts_next (may be NULL). */
int oracle_ram_run(oracle_state *st, int nthreads, double *dts_next);

/* ANISCH(S) without its wave-particle branch (src/ModRamRun.f90:265-415):
 * PPERT, PPART (and EDENT, RNHTT of its wans file, may be NULL), all
 * (NS,NR,NT), from FFACTOR (NS,NR,NE,NPA), ERNH and EPP (NS,NE); sets
 * F2(S,2:NR,:,2:NE,1) = F2(S,2:NR,:,2:NE,2) as the reference does. */
void oracle_anisch(oracle_state *st, int s, const double *ffactor, const double *ernh, const double *epp,
                   double *ppert, double *ppart, double *edent, double *rnhtt);
/* FLUX = F2/FFACTOR/FNHS over I=2..NR, K=2..NE, L=2..NPA, J=1..NT-1
 * (src/ModRamRun.f90:208-222); FLUX (NS,NR,NT,NE,NPA). */
void oracle_flux(const oracle_state *st, const double *ffactor, double *flux);

#ifdef __cplusplus
}
#endif
#endif
