/*
 * ram_oracle.c -- CPU restatement of RAM's kinetic split step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  The product never loads it.
 *
 * Each routine follows the reference's loop order and expression order
 * (left-to-right Fortran evaluation, integer powers as repeated products),
 * so that, compiled without FMA contraction, it reproduces the reference's
 * arithmetic.  The reference files are cited per routine.  Indices written
 * in comments are the reference's 1-based ones; the macros below map them to
 * 0-based storage in Fortran (column-major) order.
 */
#include "ram_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/ModRamConst.f90:13-22 (decimal literals rounded to double) */
static const double RAM_Q = 1.602E-19;
static const double RAM_CS = 2.998E8;
/* Earth's angular velocity, src/ModRamDrift.f90:224,309,405 */
static const double RAM_OME = 7.3E-5;

#define NR1 (st->nr + 1)
/* F2(S,I,J,K,L) */
#define F2(S, I, J, K, L)                                                     \
  st->f2[((S)-1) + (size_t)st->ns *                                           \
                       (((I)-1) + (size_t)st->nr *                            \
                                      (((J)-1) + (size_t)st->nt *             \
                                                     (((K)-1) + (size_t)st->ne * ((L)-1))))]
/* geometry (NR+1,NT,NPA) */
#define G3(A, I, J, L) (A)[((I)-1) + (size_t)NR1 * (((J)-1) + (size_t)st->nt * ((L)-1))]
/* fields (NR+1,NT) */
#define G2(A, I, J) (A)[((I)-1) + (size_t)NR1 * ((J)-1)]
#define MGNP(I, J) st->outside_mgnp[((I)-1) + (size_t)st->nr * ((J)-1)]
#define FGEOS(S, J, K, L)                                                     \
  st->fgeos[((S)-1) + (size_t)st->ns *                                        \
                          (((J)-1) + (size_t)st->nt * (((K)-1) + (size_t)st->ne * ((L)-1)))]
#define SK(A, S, K) (A)[((S)-1) + (size_t)st->ns * ((K)-1)]
#define V1(A, I) (A)[(I)-1]
/* drift scratch (NR,NT,NE,NPA), (NR,NE), (NR,NPA) */
#define C4(A, I, J, K, L)                                                     \
  (A)[((I)-1) + (size_t)st->nr * (((J)-1) + (size_t)st->nt * (((K)-1) + (size_t)st->ne * ((L)-1)))]
#define RK(A, I, K) (A)[((I)-1) + (size_t)st->nr * ((K)-1)]

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* The reference's beta limiter, LIMITER=MAX(MIN(BetaLim*R,1.),MIN(R,BetaLim))
 * (src/ModRamDrift.f90:178). */
static double limiter(double r, double beta) {
  return dmax(dmin(beta * r, 1.0), dmin(r, beta));
}

int oracle_drift_alloc(const oracle_state *st, oracle_drift_scratch *d) {
  size_t n4 = (size_t)st->nr * st->nt * st->ne * st->npa;
  memset(d, 0, sizeof(*d));
  d->vr = calloc(st->nr, sizeof(double));
  d->p1 = calloc(st->nr, sizeof(double));
  d->p2 = calloc((size_t)st->nr * st->ne, sizeof(double));
  d->edot = calloc((size_t)st->nr * st->ne, sizeof(double));
  d->mudot = calloc((size_t)st->nr * st->npa, sizeof(double));
  d->cdr = calloc(n4, sizeof(double));
  d->cdp = calloc(n4, sizeof(double));
  d->cde = calloc(n4, sizeof(double));
  d->cdmu = calloc(n4, sizeof(double));
  if (!d->vr || !d->p1 || !d->p2 || !d->edot || !d->mudot || !d->cdr || !d->cdp ||
      !d->cde || !d->cdmu) {
    oracle_drift_free(d);
    return -1;
  }
  return 0;
}

void oracle_drift_free(oracle_drift_scratch *d) {
  free(d->vr); free(d->p1); free(d->p2); free(d->edot); free(d->mudot);
  free(d->cdr); free(d->cdp); free(d->cde); free(d->cdmu);
  memset(d, 0, sizeof(*d));
}

int oracle_loss_alloc(const oracle_state *st, oracle_loss_scratch *l) {
  size_t n4 = (size_t)st->nr * st->nt * st->ne * st->npa;
  l->charge = calloc(n4, sizeof(double));
  l->atlos = calloc((size_t)st->nr * st->ne, sizeof(double));
  if (!l->charge || !l->atlos) {
    oracle_loss_free(l);
    return -1;
  }
  return 0;
}

void oracle_loss_free(oracle_loss_scratch *l) {
  free(l->charge); free(l->atlos);
  l->charge = l->atlos = 0;
}

/* DRIFTPARA, src/ModRamDrift.f90:36-88. */
void oracle_driftpara(const oracle_state *st, int s, oracle_drift_scratch *d) {
  const double dts = st->dts, mdr = st->mdr, dphi = st->dphi;
  d->qs = st->charge[s - 1];
  const double qs = (double)d->qs;
  for (int i = 1; i <= st->nr; ++i) {
    const double rlz = V1(st->rlz, i);
    /* :65-69 -- the inner J loop rewrites the same value NT times */
    d->vr[i - 1] = dts / mdr / (rlz + 0.5 * mdr) / 2.0 / dphi;
    d->p1[i - 1] = dts / dphi / 2.0 / mdr / rlz;
    for (int k = 1; k <= st->ne; ++k) {
      const double g = SK(st->grel, s, k);
      RK(d->p2, i, k) = dts * V1(st->ekev, k) * 1000.0 * (g + 1.0) / g / (rlz * rlz) / dphi / qs;
    }
  }
  for (int i = 1; i <= st->nr; ++i) {
    const double rlz = V1(st->rlz, i);
    for (int l = 1; l <= st->npa - 1; ++l) { /* :77-80 */
      const double muboun = V1(st->mu, l) + 0.5 * V1(st->wmu, l);
      RK(d->mudot, i, l) = (1.0 - muboun * muboun) * dts / 2.0 / muboun / rlz;
    }
    RK(d->mudot, i, st->npa) = 0.0; /* :81 */
    for (int k = 1; k <= st->ne; ++k) { /* :82-84 */
      const double gb = SK(st->grbnd, s, k);
      RK(d->edot, i, k) = V1(st->ebnd, k) * dts / rlz * (gb + 1.0) / gb / 2.0;
    }
  }
}

/* DRIFTR, src/ModRamDrift.f90:95-198.  The ghost cells F(NR+1), F(NR+2) of
 * the line buffer are NOT reset between lines: in the reference they keep
 * the values of the last inflow line (F is allocated and zeroed once per
 * call, :112-113), and an outflow line whose last interior face is upwinded
 * from outside (sgn(NR-1)=-1, :174-175) reads that stale F(NR+1).  The
 * restatement keeps this behaviour. */
void oracle_driftr(oracle_state *st, int s, oracle_drift_scratch *d) {
  const int nr = st->nr, nt = st->nt, ne = st->ne, npa = st->npa;
  const double dts = st->dts, mdr = st->mdr, dphi = st->dphi, beta = st->betalim;
  const double qs = (double)d->qs;
  double *f = calloc((size_t)nr + 3, sizeof(double)); /* F(1..NR+2) at f[1..] */
  double *fbnd = calloc((size_t)nr + 1, sizeof(double));
  double *cr = calloc((size_t)nr * nt, sizeof(double));
  int *sgn = calloc((size_t)nr + 1, sizeof(int));
  double dtr = 100000.0; /* :115 */

  for (int i = 1; i <= nr; ++i) { /* :118-127 ExB radial drift */
    for (int j = 1; j <= nt; ++j) {
      const int j0 = (j == 1) ? nt - 1 : j - 1;
      const int j1 = (j == nt) ? 2 : j + 1;
      const double bsum = G2(st->bnes, i, j) + G2(st->bnes, i + 1, j);
      const double dv = G2(st->vt, i, j0) + G2(st->vt, i + 1, j0) - G2(st->vt, i, j1) -
                        G2(st->vt, i + 1, j1);
      cr[(i - 1) + (size_t)nr * (j - 1)] =
          d->vr[i - 1] * dv / bsum +
          (G2(st->eip, i, j) + G2(st->eip, i + 1, j)) / bsum * dts / mdr;
    }
  }

  for (int k = 1; k <= ne; ++k) {
    const double g = SK(st->grel, s, k);
    const double p4 = dts * V1(st->ekev, k) * 1000.0 * (g + 1.0) / g / dphi / mdr / qs; /* :131 */
    for (int l = 1; l <= npa; ++l) {
      for (int j = 1; j <= nt; ++j) {
        for (int i = 1; i <= nr; ++i) f[i] = F2(s, i, j, k, l);
        const int j0 = (j == 1) ? nt - 1 : j - 1;
        const int j1 = (j == nt) ? 2 : j + 1;
        for (int i = 1; i <= nr; ++i) { /* :139-153 */
          const double cgr1 = G3(st->fnis, i + 1, j1, l) + G3(st->fnis, i, j1, l) -
                              G3(st->fnis, i + 1, j0, l) - G3(st->fnis, i, j0, l);
          const double cgr2 = G2(st->bnes, i + 1, j1) + G2(st->bnes, i, j1) -
                              G2(st->bnes, i + 1, j0) - G2(st->bnes, i, j0);
          const double cgr3 =
              cgr1 + (G3(st->fnis, i + 1, j, l) + G3(st->fnis, i, j, l) -
                      2.0 * G3(st->fnhs, i + 1, j, l) - 2.0 * G3(st->fnhs, i, j, l)) *
                         cgr2 / 2.0 / (G2(st->bnes, i + 1, j) + G2(st->bnes, i, j));
          const double cgr = cgr3 / (G3(st->fnhs, i, j, l) + G3(st->fnhs, i + 1, j, l)) * p4 /
                             2.0 / (G2(st->bnes, i, j) + G2(st->bnes, i + 1, j)) /
                             (V1(st->rlz, i) + 0.5 * mdr);
          const double c = cr[(i - 1) + (size_t)nr * (j - 1)] + cgr;
          C4(d->cdr, i, j, k, l) = c;
          if (MGNP(i, j) == 0) {
            const double ctemp = dmax(fabs(c), 1E-10);
            dtr = dmin(dtr, st->fraccfl * dts / ctemp);
          }
          sgn[i] = (c < 0) ? -1 : 1;
        }
        int ur;
        if (sgn[nr] == 1) { /* :154-168 outflow / inflow at the outer edge */
          fbnd[1] = 0.0;
          fbnd[nr] = f[nr];
          ur = nr - 1;
        } else {
          fbnd[1] = f[2];
          ur = nr;
          if (MGNP(nr, j) == 1) {
            f[nr + 1] = 0.0;
            f[nr + 2] = 0.0;
          } else {
            f[nr + 1] = FGEOS(s, j, k, l) * st->conf1 * G3(st->fnhs, nr, j, l);
            f[nr + 2] = FGEOS(s, j, k, l) * st->conf2 * G3(st->fnhs, nr, j, l);
          }
        }
        for (int i = 2; i <= ur; ++i) { /* :169-183 limited face value */
          const double sg = (double)sgn[i];
          const double x = f[i + 1] - f[i];
          const double fup = 0.5 * (f[i] + f[i + 1] - sg * x);
          if (fabs(x) <= 1.E-27) {
            fbnd[i] = fup;
          } else {
            const int n = i + 1 - sgn[i];
            const double r = (f[n] - f[n - 1]) / x;
            if (r <= 0) {
              fbnd[i] = fup;
            } else {
              const double corr = -0.5 * (C4(d->cdr, i, j, k, l) - sg) * x;
              fbnd[i] = fup + limiter(r, beta) * corr;
            }
          }
        }
        for (int i = 2; i <= nr; ++i) { /* :185-191 conservative update */
          double v = F2(s, i, j, k, l) - C4(d->cdr, i, j, k, l) * fbnd[i] +
                     C4(d->cdr, i - 1, j, k, l) * fbnd[i - 1];
          if (v < 0) v = 1E-15;
          F2(s, i, j, k, l) = v;
        }
      }
    }
  }
  st->dtdrift_r[s - 1] = dtr;
  free(f); free(fbnd); free(cr); free(sgn);
}

/* DRIFTP, src/ModRamDrift.f90:204-279 (periodic in local time). */
void oracle_driftp(oracle_state *st, int s, oracle_drift_scratch *d) {
  const int nr = st->nr, nt = st->nt, ne = st->ne, npa = st->npa;
  const double dts = st->dts, mdr = st->mdr, dphi = st->dphi, beta = st->betalim;
  double *f = calloc((size_t)nt + 1, sizeof(double));
  double *fbnd = calloc((size_t)nt + 1, sizeof(double));
  double dtp = 100000.0; /* :223 */
  for (int l = 1; l <= npa; ++l) {
    for (int k = 1; k <= ne; ++k) {
      for (int i = 2; i <= nr; ++i) {
        const double rlz = V1(st->rlz, i);
        for (int j = 1; j <= nt; ++j) f[j] = F2(s, i, j, k, l);
        for (int j = 2; j <= nt; ++j) { /* :229-260 */
          const int j1 = (j == nt) ? 2 : j + 1;
          const double gpa1 =
              G3(st->fnis, i, j, l) + G3(st->fnis, i, j1, l) +
              (G3(st->fnis, i + 1, j1, l) + G3(st->fnis, i + 1, j, l) - G3(st->fnis, i - 1, j, l) -
               G3(st->fnis, i - 1, j1, l)) * rlz / 2.0 / mdr;
          const double gpa2 =
              rlz / 4.0 / mdr *
              (G3(st->fnis, i, j, l) + G3(st->fnis, i, j1, l) - 2.0 * G3(st->fnhs, i, j, l) -
               2.0 * G3(st->fnhs, i, j1, l)) *
              (G2(st->bnes, i + 1, j1) + G2(st->bnes, i + 1, j) - G2(st->bnes, i - 1, j) -
               G2(st->bnes, i - 1, j1)) /
              (G2(st->bnes, i, j) + G2(st->bnes, i, j1));
          const double c =
              ((G2(st->vt, i + 1, j) + G2(st->vt, i + 1, j1) - G2(st->vt, i - 1, j) -
                G2(st->vt, i - 1, j1)) * d->p1[i - 1] -
               RK(d->p2, i, k) * (gpa1 + gpa2) /
                   (G3(st->fnhs, i, j, l) + G3(st->fnhs, i, j1, l)) -
               (G2(st->eir, i, j1) + G2(st->eir, i, j)) / rlz * dts / dphi) /
                  (G2(st->bnes, i, j) + G2(st->bnes, i, j1)) +
              RAM_OME * dts / dphi;
          C4(d->cdp, i, j, k, l) = c;
          if (MGNP(i, j) == 0) {
            const double ctemp = dmax(fabs(c), 1E-10);
            dtp = dmin(dtp, st->fraccfl * dts / ctemp);
          }
          const int sgn = (c < 0) ? -1 : 1;
          const double sg = (double)sgn;
          const double x = f[j1] - f[j];
          const double fup = 0.5 * (f[j] + f[j1] - sg * x);
          if (fabs(x) <= 1.E-27) {
            fbnd[j] = fup;
          } else {
            int n = j + 1 - sgn;
            if (n > nt) n = n - nt + 1; /* :251 */
            const double r = (f[n] - f[n - 1]) / x;
            if (r <= 0) {
              fbnd[j] = fup;
            } else {
              const double corr = -0.5 * (c - sg) * x;
              fbnd[j] = fup + limiter(r, beta) * corr;
            }
          }
        }
        C4(d->cdp, i, 1, k, l) = C4(d->cdp, i, nt, k, l); /* :261-262 */
        fbnd[1] = fbnd[nt];
        for (int j = 2; j <= nt; ++j) { /* :263-271 */
          double v = F2(s, i, j, k, l) - C4(d->cdp, i, j, k, l) * fbnd[j] +
                     C4(d->cdp, i, j - 1, k, l) * fbnd[j - 1];
          if (v < 0) v = 1E-15;
          F2(s, i, j, k, l) = v;
        }
        F2(s, i, 1, k, l) = F2(s, i, nt, k, l); /* :272 */
      }
    }
  }
  st->dtdrift_p[s - 1] = dtp;
  free(f); free(fbnd);
}

/* DRIFTE, src/ModRamDrift.f90:285-376. */
void oracle_drifte(oracle_state *st, int s, oracle_drift_scratch *d) {
  const int nr = st->nr, nt = st->nt, ne = st->ne, npa = st->npa;
  const double dts = st->dts, mdr = st->mdr, dphi = st->dphi, beta = st->betalim;
  const double qs = (double)d->qs;
  double *fb = calloc((size_t)ne + 3, sizeof(double)); /* F(0..NE+2) */
  double *f = fb;                                       /* f[k] == F(k) */
  double *fbnd = calloc((size_t)ne + 1, sizeof(double));
  double dte = 10000.0; /* :308 */
  const double ezero = V1(st->ekev, 1) - V1(st->we, 1);                     /* :310 */
  const double grzero = 1.0 + ezero * 1000.0 * RAM_Q / st->rmas[s - 1] / RAM_CS / RAM_CS; /* :311 */
  const double g1 = SK(st->grel, s, 1), g2 = SK(st->grel, s, 2);
  f[ne + 1] = 0.0; /* :312-313 */
  f[ne + 2] = 0.0;
  for (int j = 1; j <= nt; ++j) {
    const int j0 = (j == 1) ? nt - 1 : j - 1;
    const int j2 = (j == nt) ? 2 : j + 1;
    for (int i = 2; i <= nr; ++i) {
      const double rlz = V1(st->rlz, i);
      const double bij = G2(st->bnes, i, j);
      const double drd1 = (G2(st->eip, i, j) * rlz - (G2(st->vt, i, j2) - G2(st->vt, i, j0)) / 2.0 / dphi) / bij;
      const double dpd1 = RAM_OME * rlz + ((G2(st->vt, i + 1, j) - G2(st->vt, i - 1, j)) / 2.0 / mdr - G2(st->eir, i, j)) / bij;
      for (int l = 1; l <= npa; ++l) {
        const double fnis = G3(st->fnis, i, j, l), fnhs = G3(st->fnhs, i, j, l);
        const double gpa = (1.0 - fnis / 2.0 / fnhs) / bij; /* :323-332 */
        const double gpr1 = gpa * (G2(st->bnes, i + 1, j) - G2(st->bnes, i - 1, j)) / 2.0 / mdr;
        const double gpr2 = -fnis / fnhs / rlz;
        const double gpr3 = -(G3(st->fnis, i + 1, j, l) - G3(st->fnis, i - 1, j, l)) / 2.0 / mdr / fnhs;
        const double gpp1 = gpa * (G2(st->bnes, i, j2) - G2(st->bnes, i, j0)) / 2.0 / dphi;
        const double gpp2 = -(G3(st->fnis, i, j2, l) - G3(st->fnis, i, j0, l)) / 2.0 / dphi / fnhs;
        const double drd2 = (G3(st->fnis, i, j2, l) - G3(st->fnis, i, j0, l)) / 2.0 / dphi +
                            (fnis - 2.0 * fnhs) * (G2(st->bnes, i, j2) - G2(st->bnes, i, j0)) / 4.0 / bij / dphi;
        const double dpd2 = fnis + (G3(st->fnis, i + 1, j, l) - G3(st->fnis, i - 1, j, l)) * rlz / 2.0 / mdr +
                            rlz * (fnis - 2.0 * fnhs) / 4.0 / mdr *
                                (G2(st->bnes, i + 1, j) - G2(st->bnes, i - 1, j)) / bij;
        for (int k = 1; k <= ne; ++k) f[k] = F2(s, i, j, k, l); /* :333-335 */
        f[1] = f[2] * g1 / g2 * sqrt((g2 * g2 - 1.0) / (g1 * g1 - 1.0));
        f[0] = f[1] * grzero / g1 * sqrt((g1 * g1 - 1.0) / (grzero * grzero - 1.0));
        for (int k = 1; k <= ne; ++k) { /* :336-362 */
          const double gb = SK(st->grbnd, s, k);
          const double edt1 = V1(st->ebnd, k) * 1e3 * (gb + 1.0) / 2.0 / gb / fnhs / rlz / bij / qs;
          const double drdt = drd1 + edt1 * drd2 * rlz;
          const double dpdt = dpd1 - edt1 * dpd2;
          const double dbdt1 = G2(st->dbdt, i, j) * (1.0 - fnis / 2.0 / fnhs) * rlz / bij;
          const double didt1 = -G3(st->didt, i, j, l) * rlz / fnhs;
          const double c = RK(d->edot, i, k) *
                           ((gpr1 + gpr2 + gpr3) * drdt + (gpp1 + gpp2) * dpdt + dbdt1 + didt1);
          C4(d->cde, i, j, k, l) = c;
          if (MGNP(i, j) == 0) {
            const double ctemp = dmax(fabs(c), 1E-10);
            dte = dmin(dte, st->fraccfl * dts * V1(st->de, k) / ctemp);
          }
          const int sgn = (c < 0) ? -1 : 1;
          const double sg = (double)sgn;
          const double x = f[k + 1] - f[k];
          const double fup = 0.5 * (f[k] + f[k + 1] - sg * x);
          if (fabs(x) <= 1.E-27) {
            fbnd[k] = fup;
          } else {
            const int n = k + 1 - sgn;
            const double r = (f[n] - f[n - 1]) / x;
            if (r <= 0) {
              fbnd[k] = fup;
            } else {
              const double corr = -0.5 * (c / V1(st->de, k) - sg) * x;
              fbnd[k] = fup + limiter(r, beta) * corr;
            }
          }
        }
        for (int k = 2; k <= ne; ++k) { /* :363-369 */
          const double wek = V1(st->we, k);
          double v = F2(s, i, j, k, l) - C4(d->cde, i, j, k, l) / wek * fbnd[k] +
                     C4(d->cde, i, j, k - 1, l) / wek * fbnd[k - 1];
          if (v < 0) v = 1E-15;
          F2(s, i, j, k, l) = v;
        }
      }
    }
  }
  st->dtdrift_e[s - 1] = dte;
  free(fb); free(fbnd);
}

/* DRIFTMU, src/ModRamDrift.f90:382-473. */
void oracle_driftmu(oracle_state *st, int s, oracle_drift_scratch *d) {
  const int nr = st->nr, nt = st->nt, ne = st->ne, npa = st->npa;
  const double dts = st->dts, mdr = st->mdr, dphi = st->dphi, beta = st->betalim;
  const double qs = (double)d->qs;
  double *f = calloc((size_t)npa + 1, sizeof(double));
  double *fbnd = calloc((size_t)npa + 1, sizeof(double));
  double dtmu = 10000.0; /* :404 */
  for (int k = 1; k <= ne; ++k) {
    const double g = SK(st->grel, s, k);
    for (int j = 1; j <= nt; ++j) {
      const int j0 = (j == 1) ? nt - 1 : j - 1;
      const int j1 = (j == nt) ? 2 : j + 1;
      for (int i = 2; i <= nr; ++i) {
        const double rlz = V1(st->rlz, i);
        const double bij = G2(st->bnes, i, j);
        for (int l = 1; l <= npa; ++l) f[l] = F2(s, i, j, k, l); /* :413-414 */
        f[1] = f[2];
        const double drm1 = (G2(st->eip, i, j) * rlz - (G2(st->vt, i, j1) - G2(st->vt, i, j0)) / 2.0 / dphi) / bij;
        const double dpm1 = RAM_OME * rlz + ((G2(st->vt, i + 1, j) - G2(st->vt, i - 1, j)) / 2.0 / mdr - G2(st->eir, i, j)) / bij;
        for (int l = 2; l <= npa; ++l) { /* :417-455 */
          const double bis = G3(st->bounis, i, j, l), bhs = G3(st->bounhs, i, j, l);
          const double cmudot = RK(d->mudot, i, l) * bis / bhs;
          const double gmr1 = (G2(st->bnes, i + 1, j) - G2(st->bnes, i - 1, j)) / 4.0 / mdr / bij;
          const double gmr2 = 1.0 / rlz;
          const double gmr3 = (G3(st->bounis, i + 1, j, l) - G3(st->bounis, i - 1, j, l)) / 2.0 / mdr / bis;
          const double gmp1 = (G2(st->bnes, i, j1) - G2(st->bnes, i, j0)) / 4.0 / dphi / bij;
          const double gmp2 = (G3(st->bounis, i, j1, l) - G3(st->bounis, i, j0, l)) / 2.0 / dphi / bis;
          const double edt = V1(st->ekev, k) * 1e3 * (g + 1.0) / 2.0 / g / bhs / rlz / bij / qs;
          const double drm2 = (G3(st->bounis, i, j1, l) - G3(st->bounis, i, j0, l)) / 2.0 / dphi +
                              (bis - 2.0 * bhs) * (G2(st->bnes, i, j1) - G2(st->bnes, i, j0)) / 4.0 / bij / dphi;
          const double dpm2 = bis + (G3(st->bounis, i + 1, j, l) - G3(st->bounis, i - 1, j, l)) * rlz / 2.0 / mdr +
                              (bis - 2.0 * bhs) * rlz / 4.0 / mdr *
                                  (G2(st->bnes, i + 1, j) - G2(st->bnes, i - 1, j)) / bij;
          const double drdm = drm1 + edt * drm2 * rlz;
          const double dpdm = dpm1 - edt * dpm2;
          const double dbdt2 = G2(st->dbdt, i, j) / 2.0 / bij * rlz;
          const double dibndt2 = G3(st->dibndt, i, j, l) * rlz / bis;
          const double c = -cmudot * ((gmr1 + gmr2 + gmr3) * drdm + (gmp1 + gmp2) * dpdm + dbdt2 + dibndt2);
          C4(d->cdmu, i, j, k, l) = c;
          if (MGNP(i, j) == 0) {
            const double ctemp = dmax(fabs(c), 1E-32);
            dtmu = dmin(dtmu, st->fraccfl * dts * V1(st->dmu, l) / ctemp);
          }
          const int isgm = (c < 0.0) ? -1 : 1;
          const double sg = (double)isgm;
          if (l <= npa - 2) {
            const double x = f[l + 1] - f[l];
            const double fup = 0.5 * (f[l] + f[l + 1] - sg * x);
            if (fabs(x) <= 1.E-27) {
              fbnd[l] = fup;
            } else {
              const int n = l + 1 - isgm;
              const double r = (f[n] - f[n - 1]) / x;
              if (r <= 0) {
                fbnd[l] = fup;
              } else {
                const double corr = -0.5 * (c / V1(st->dmu, l) - sg) * x;
                fbnd[l] = fup + limiter(r, beta) * corr;
              }
            }
          }
        }
        C4(d->cdmu, i, j, k, 1) = 0.0; /* :456-458 */
        fbnd[1] = 0.0;
        fbnd[npa - 1] = f[npa];
        for (int l = 2; l <= npa - 1; ++l) { /* :459-465 */
          const double wmul = V1(st->wmu, l);
          double v = F2(s, i, j, k, l) - C4(d->cdmu, i, j, k, l) / wmul * fbnd[l] +
                     C4(d->cdmu, i, j, k, l - 1) / wmul * fbnd[l - 1];
          if (v < 0) v = 1E-15;
          F2(s, i, j, k, l) = v;
        }
        F2(s, i, j, k, npa) = F2(s, i, j, k, npa - 1) * G3(st->fnhs, i, j, npa) * V1(st->mu, npa) /
                              G3(st->fnhs, i, j, npa - 1) / V1(st->mu, npa - 1); /* :466 */
      }
    }
  }
  st->dtdrift_mu[s - 1] = dtmu;
  free(f); free(fbnd);
}

/* CEPARA, src/ModRamLoss.f90:19-170: charge-exchange factors for the three
 * analytic cross-section fits (H+ :39-51, He+ :54-66, O+ :69-81) and the
 * atmospheric-loss factor ATLOS (:162-167).  The file-based branch
 * (:84-158) is not part of this path (no species here uses it). */
void oracle_cepara(const oracle_state *st, int s, oracle_loss_scratch *ls) {
  const int nr = st->nr, nt = st->nt, ne = st->ne, npa = st->npa;
  const double dts = st->dts;
  const size_t n4 = (size_t)nr * nt * ne * npa;
  for (size_t n = 0; n < n4; ++n) ls->charge[n] = 1.0; /* :34 */
  const int kind = st->cex_kind[s - 1];
  if (kind != ORACLE_CEX_NONE) {
    for (int l = 2; l <= npa; ++l)
      for (int k = 2; k <= ne; ++k)
        for (int i = 2; i <= nr; ++i)
          for (int j = 1; j <= nt; ++j) {
            double x = log10(V1(st->ekev, k));
            if (x < -2.0) x = -2.0;
            const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
            double y;
            if (kind == ORACLE_CEX_H)
              y = -18.767 - 0.11017 * x - 3.8173e-2 * x2 - 0.1232 * x3 - 5.0488e-2 * x4;
            else if (kind == ORACLE_CEX_HE)
              y = -20.789 + 0.92316 * x - 0.68017 * x2 + 0.66153 * x3 - 0.20998 * x4;
            else
              y = -18.987 - 0.10613 * x - 5.4841E-3 * x2 - 1.6262E-2 * x3 - 7.0554E-3 * x4;
            const double alpha = pow(10.0, y) * SK(st->v, s, k) * G3(st->hdns, i, j, l) * dts;
            C4(ls->charge, i, j, k, l) = exp(-alpha);
          }
  }
  for (int k = 2; k <= ne; ++k)
    for (int i = 2; i <= nr; ++i) {
      const double taub = 2.0 * V1(st->rlz, i) / SK(st->v, s, k);
      RK(ls->atlos, i, k) = exp(-dts / taub);
    }
}

/* CHAREXCHANGE, src/ModRamLoss.f90:457-478. */
void oracle_charexchange(oracle_state *st, int s, const oracle_loss_scratch *ls) {
  for (int k = 2; k <= st->ne; ++k)
    for (int j = 1; j <= st->nt; ++j)
      for (int i = 2; i <= st->nr; ++i)
        for (int l = 2; l <= st->npa; ++l)
          F2(s, i, j, k, l) = F2(s, i, j, k, l) * C4(ls->charge, i, j, k, l);
}

/* ATMOL, src/ModRamLoss.f90:485-507. */
void oracle_atmol(oracle_state *st, int s, const oracle_loss_scratch *ls) {
  for (int k = 2; k <= st->ne; ++k)
    for (int j = 1; j <= st->nt; ++j)
      for (int i = 2; i <= st->nr; ++i) {
        const int u = (int)V1(st->upa, i);
        for (int l = u; l <= st->npa; ++l)
          F2(s, i, j, k, l) = F2(s, i, j, k, l) * pow(RK(ls->atlos, i, k), 1.0 / G3(st->fnhs, i, j, l));
      }
}

/* WAVELO, src/ModRamWPI.f90:580-636: electron lifetimes inside (WALOS1
 * scaled by the wave amplitude Bw) and outside the plasmapause (WALOS2 with
 * the strong-diffusion WALOS3, energy corrections below 5 keV; 5 days / KP
 * above 1 MeV); F2 *= exp(-DTs/TAU) on K,I,L >= 2, all J. */
void oracle_wavelo(oracle_state *st, int s) {
  double bw = 30.;
  if (st->kp >= 4.0) bw = 100.;
  const int ns = st->ns, nr = st->nr;
#define WL(a, i, k) (a)[(s - 1) + (size_t)ns * (((i)-1) + (size_t)nr * ((k)-1))]
  for (int k = 2; k <= st->ne; ++k)
    for (int i = 2; i <= st->nr; ++i)
      for (int j = 1; j <= st->nt; ++j)
        for (int l = 2; l <= st->npa; ++l) {
          const double lz = V1(st->lz, i), ek = V1(st->ekev, k);
          double tau = 0.0;
          if (lz <= V1(st->rlpp, j)) {
            const double b = 10. / bw;
            tau = WL(st->walos1, i, k) * (b * b);
          } else if (lz > V1(st->rlpp, j)) {
            if (ek <= 1000.) {
              tau = WL(st->walos2, i, k) * (1 + WL(st->walos3, i, k) / WL(st->walos2, i, k));
              if (ek <= 1.1)
                tau = tau * 37.5813 * exp(-1.81255 * ek);
              else if (ek > 1.1 && ek <= 5.)
                tau = tau * (7.5 - 1.15 * ek);
            } else if (ek > 1000.) {
              tau = 5. * 3600 * 24 / st->kp;
            }
          }
          F2(s, i, j, k, l) = F2(s, i, j, k, l) * exp(-st->dts / tau);
        }
#undef WL
}

/* ---- Coulomb collisions, src/ModRamCoul.f90 ---------------------------- */

static const double RAM_PI = 3.1415926535897932384626433832795;
static const double RAM_MP = 1.673E-27;  /* src/ModRamConst.f90:18 */
static const double RAM_RE = 6.371E6;    /* src/ModRamConst.f90:14 */

/* FUNT and FUNI, Ejiri 1978 (src/ModRamFunctions.f90:90-143), in the
 * reference's expression order. */
static double coul_funt(double x) {
  const double y = sqrt(1 - x * x);
  const double alpha = 1. + log(2. + sqrt(3.)) / 2. / sqrt(3.);
  const double beta = alpha / 2. - RAM_PI * sqrt(2.) / 12.;
  const double a1 = 0.055, a2 = -0.037, a3 = -0.074, a4 = 0.056;
  return alpha - beta * (y + sqrt(y)) + a1 * pow(y, 1. / 3.) + a2 * pow(y, 2. / 3.) + a3 * y + a4 * pow(y, 4. / 3.);
}
static double coul_funi(double x) {
  double ylog = 0.0;
  const double y = sqrt(1 - x * x);
  if (y > 0) ylog = log(y);
  const double alpha = 1. + log(2. + sqrt(3.)) / 2. / sqrt(3.);
  const double beta = alpha / 2. - RAM_PI * sqrt(2.) / 12.;
  const double a1 = 0.055, a2 = -0.037, a3 = -0.074, a4 = 0.056;
  return 2. * alpha * (1. - y) + 2. * beta * y * ylog + 4. * beta * (y - sqrt(y)) +
         3. * a1 * (pow(y, 1. / 3.) - y) + 6. * a2 * (pow(y, 2. / 3.) - y) + 6. * a4 * (y - pow(y, 4. / 3.)) -
         2. * a3 * y * ylog;
}
/* Gcoul(x) = (erf(x) - x erf'(x)) / (2 x^2), src/ModRamFunctions.f90:72-87 */
static double coul_g(double x) {
  const double g1 = erf(x) - 2. * x / sqrt(RAM_PI) * exp(-x * x);
  return g1 / 2. / x / x;
}

/* The plasmasphere's species for the Coulomb terms: RAMSpecies(1..6) of
 * src/ModRamSpecies.f90:43-131 (mass in MP, charge, plasmasphereRatio);
 * those with a ratio below 1e-9 are skipped. */
static const double coul_ps_mass[6] = {5.4462E-4, 1.0, 4.0, 16.0, 14.0, 87.62};
static const double coul_ps_charge[6] = {-1, 1, 1, 1, 1, 1};
static const double coul_ps_ratio[6] = {1.0, 0.77, 0.2, 0.03, 0.0, 0.0};

#define CT(A, S, K, L) (A)[((S)-1) + (size_t)st->ns * (((K)-1) + (size_t)st->ne * ((L)-1))]

/* COULPARA(S), src/ModRamCoul.f90:17-125: COULE, COULI, CEDR, CIDR, ATA and
 * GTA of species S.  The plasmasphere sums CCE, CDE, EDRE, CCI, CDI, EDRI
 * start at 0 once and are not reset between energies, as in the
 * reference, so the K-th term holds the sums over K' <= K. */
void oracle_coulpara(oracle_state *st, int s) {
  const int ne = st->ne, npa = st->npa;
  double *coulde = (double *)calloc(npa, sizeof(double)), *couldi = (double *)calloc(npa, sizeof(double));
  double cce = 0, cde = 0, edre = 0, cci = 0, cdi = 0, edri = 0;
  const double eps = 8.854E-12, dln = 21.5;
  const double zt = (double)st->charge[s - 1];
  const double qe = (RAM_Q * RAM_Q / eps);
  const double gama = zt * zt * dln / 4. / RAM_PI * qe * 1E6 * qe;
  const double cco = gama / RAM_Q * st->dts / RAM_Q / 1E3;
  const double rm = st->rmas[s - 1];
  const double ccd = gama * st->dts / (rm * rm) / (RAM_CS * RAM_CS * RAM_CS);
  const double edrco = dln * qe * RAM_RE / rm * qe * 1E9 / RAM_Q;
  for (int k = 1; k <= ne; ++k) {
    for (int is = 0; is < 6; ++is) {
      const double ra = coul_ps_ratio[is];
      if (ra < 1e-9) continue;
      const double vf = sqrt(2. * RAM_Q / (RAM_MP * coul_ps_mass[is]));
      const double zb = coul_ps_charge[is];
      const double x = SK(st->vbnd, s, k) / vf, xd = SK(st->v, s, k) / vf;
      if (zb < 0.) {
        cce = cce + ra * coul_g(x);
        cde = cde + ra * (erf(xd) - coul_g(xd));
        edre = edre + ra * coul_g(xd);
      } else {
        cci = cci + ra * (zb * zb) * coul_g(x);
        cdi = cdi + ra * (zb * zb) * (erf(xd) - coul_g(xd));
        edri = edri + ra * coul_g(xd);
      }
    }
    const double vb = SK(st->vbnd, s, k), gb = SK(st->grbnd, s, k), gr = SK(st->grel, s, k), v = SK(st->v, s, k);
    CT(st->coule, s, k, 1) = -(cce * vb * cco * (gb * gb));
    CT(st->couli, s, k, 1) = -(cci * vb * cco * (gb * gb));
    CT(st->cedr, s, k, 1) = edrco * V1(st->ekev, k) * edre * (gr + 1) / (gr * gr) / (v * v);
    CT(st->cidr, s, k, 1) = edrco * V1(st->ekev, k) * edri * (gr + 1) / (gr * gr) / (v * v);
    const double ccde = ccd * cde * gr / pow(gr * gr - 1, 1.5);
    const double ccdi = ccd * cdi * gr / pow(gr * gr - 1, 1.5);
    for (int l = 2; l <= npa - 1; ++l) {
      const double mu = V1(st->mu, l), wmu = V1(st->wmu, l);
      const double bane = (1. - coul_funi(mu) / 2. / coul_funt(mu)) / (1. - mu * mu);
      CT(st->coule, s, k, l) = CT(st->coule, s, k, 1);
      CT(st->couli, s, k, l) = CT(st->couli, s, k, 1);
      CT(st->cedr, s, k, l) = CT(st->cedr, s, k, 1) * bane * coul_funt(mu) * mu * wmu;
      CT(st->cidr, s, k, l) = CT(st->cidr, s, k, 1) * bane * coul_funt(mu) * mu * wmu;
      const double muboun = mu + 0.5 * wmu;
      const double badif = (1. - muboun * muboun) / muboun / 2.;
      coulde[l - 1] = ccde * badif;
      const double afer = coulde[l - 1] / mu / V1(st->dmu, l) / wmu;
      const double asec = coulde[l - 2] / mu / V1(st->dmu, l - 1) / wmu;
      couldi[l - 1] = ccdi * badif;
      const double afir = couldi[l - 1] / mu / V1(st->dmu, l) / wmu;
      const double asic = couldi[l - 2] / mu / V1(st->dmu, l - 1) / wmu;
      CT(st->ata, s, k, l) = afir + afer;
      CT(st->gta, s, k, l) = asic + asec;
    }
    CT(st->cedr, s, k, npa) = CT(st->cedr, s, k, npa - 1);
    CT(st->cidr, s, k, npa) = CT(st->cidr, s, k, npa - 1);
    CT(st->ata, s, k, npa) = 0;
  }
  free(coulde);
  free(couldi);
}

/* COULEN(S), src/ModRamCoul.f90:133-221: the energy sweep of the Coulomb
 * drag, CccolE(K) = (COULE+COULI)(S,K,L) * NECR(I,J)*BANE(L) with BANE the
 * bounce factor (1-FNIS/2/FNHS)/(1-MU^2), its top 11 rows copied from row
 * NPA-11; ghosts F(1), F(0) and F(NE+1) = F(NE+2) = 0 as in DRIFTE; K=2..NE
 * updated and clamped to 1e-15. */
void oracle_coulen(oracle_state *st, int s) {
  const int ne = st->ne, npa = st->npa;
  double *f = (double *)calloc(ne + 3, sizeof(double)), *fbnd = (double *)calloc(ne + 1, sizeof(double));
  double *cc = (double *)calloc(ne + 1, sizeof(double)), *bane = (double *)calloc(npa + 1, sizeof(double));
  const double rm = st->rmas[s - 1];
  const double ezero = V1(st->ekev, 1) - V1(st->we, 1);
  const double grzero = 1. + ezero * 1000. * RAM_Q / rm / RAM_CS / RAM_CS;
  const double g1 = SK(st->grel, s, 1), g2 = SK(st->grel, s, 2);
  f[ne + 1] = 0.;
  f[ne + 2] = 0.;
  for (int j = 1; j <= st->nt; ++j)
    for (int i = 2; i <= st->nr; ++i) {
      for (int l = 2; l <= npa - 1; ++l) {
        const double mu = V1(st->mu, l);
        bane[l] = (1. - G3(st->fnis, i, j, l) / 2. / G3(st->fnhs, i, j, l)) / (1. - mu * mu);
      }
      for (int l = npa - 10; l <= npa; ++l) bane[l] = bane[l - 1];
      for (int l = 2; l <= npa; ++l) {
        const double xne = st->necr[(i - 1) + (size_t)st->nr * (j - 1)] * bane[l];
        for (int k = 2; k <= ne; ++k) f[k] = F2(s, i, j, k, l);
        f[1] = f[2] * g1 / g2 * sqrt((g1 * g1 - 1) / (g2 * g2 - 1));
        f[0] = f[1] * grzero / g1 * sqrt((grzero * grzero - 1) / (g1 * g1 - 1));
        for (int k = 1; k <= ne; ++k) {
          cc[k] = (CT(st->coule, s, k, l) + CT(st->couli, s, k, l)) * xne;
          const int isign = cc[k] < 0. ? -1 : 1;
          const double x = f[k + 1] - f[k];
          const double fup = 0.5 * (f[k] + f[k + 1] - isign * x);
          if (fabs(x) <= 1.E-27) fbnd[k] = fup;
          if (fabs(x) > 1.E-27) {
            const int n = k + 1 - isign;
            const double r = (f[n] - f[n - 1]) / x;
            if (r <= 0) fbnd[k] = fup;
            if (r > 0) {
              const double corr = -0.5 * (cc[k] / V1(st->de, k) - isign) * x;
              fbnd[k] = fup + limiter(r, st->betalim) * corr;
            }
          }
        }
        for (int k = 2; k <= ne; ++k) {
          F2(s, i, j, k, l) =
              F2(s, i, j, k, l) - cc[k] / V1(st->we, k) * fbnd[k] + cc[k - 1] / V1(st->we, k) * fbnd[k - 1];
          if (F2(s, i, j, k, l) < 0.) F2(s, i, j, k, l) = 1E-15;
        }
      }
    }
  free(f);
  free(fbnd);
  free(cc);
  free(bane);
}

/* COULMU(S), src/ModRamCoul.f90:229-296: implicit pitch-angle scattering
 * of every (I,J,K) line, a tridiagonal solve in the reference's recurrences
 * (RK, RL forward over L=2..NPA-1 with RK(1)=0, RL(1)=-1, then back
 * substitution), with BASCNE = NECR*BOUNIS/2/BOUNHS; F2 is solved for
 * F2/FNHS/MU and scaled back; negative cells are set to 1e-15 once
 * TimeRamElapsed > 0. */
void oracle_coulmu(oracle_state *st, int s) {
  const int npa = st->npa;
  double *rk = (double *)calloc(npa + 1, sizeof(double)), *rl = (double *)calloc(npa + 1, sizeof(double));
  double *bascne = (double *)calloc(npa + 1, sizeof(double));
  for (int j = 1; j <= st->nt; ++j)
    for (int i = 2; i <= st->nr; ++i) {
      const double xne = st->necr[(i - 1) + (size_t)st->nr * (j - 1)];
      for (int k = 2; k <= st->ne; ++k) {
        rk[1] = 0.;
        rl[1] = -1.;
        bascne[1] = xne * G3(st->bounis, i, j, 1) / 2. / G3(st->bounhs, i, j, 1);
        for (int l = 2; l <= npa - 1; ++l) {
          bascne[l] = xne * G3(st->bounis, i, j, l) / 2. / G3(st->bounhs, i, j, l);
          const double an = CT(st->ata, s, k, l) * bascne[l] / G3(st->fnhs, i, j, l) * G3(st->bounhs, i, j, l);
          const double gn =
              CT(st->gta, s, k, l) * bascne[l - 1] / G3(st->fnhs, i, j, l) * G3(st->bounhs, i, j, l - 1);
          const double bn = an + gn;
          const double rp = F2(s, i, j, k, l) / G3(st->fnhs, i, j, l) / V1(st->mu, l);
          const double denom = bn + gn * rl[l - 1] + 1;
          rk[l] = (rp + gn * rk[l - 1]) / denom;
          rl[l] = -an / denom;
        }
        F2(s, i, j, k, npa - 1) = rk[npa - 1] / (1 + rl[npa - 1]);
        for (int l = npa - 2; l >= 1; --l) F2(s, i, j, k, l) = rk[l] - rl[l] * F2(s, i, j, k, l + 1);
        F2(s, i, j, k, npa) = F2(s, i, j, k, npa - 1);
        for (int l = 1; l <= npa; ++l) {
          F2(s, i, j, k, l) = F2(s, i, j, k, l) * G3(st->fnhs, i, j, l) * V1(st->mu, l);
          if (st->time_elapsed > 0. && F2(s, i, j, k, l) < 0.) F2(s, i, j, k, l) = 1E-15;
        }
      }
    }
  free(rk);
  free(rl);
  free(bascne);
}
#undef CT

/* ---- implicit pitch-angle diffusion: WPADIF and FLCscatter -------------- */

/* The line solve shared by WPADIF(S) (src/ModRamWPI.f90:643-714) and
 * FLCscatter(S) (src/ModRamLoss.f90:513-575): for J=1..NT, I=2..NR,
 * K=2..NE, F = F2/FACMU with FACMU(L) = FNHS(I,J,L)*MU(L); RK(1)=0,
 * RL(1)=-1; for L=2..NPA-1 AN = D(L)/DMU(L)*DTs/FACMU(L)/WMU(L),
 * GN = D(L-1)/DMU(L-1)*DTs/FACMU(L)/WMU(L), the RK, RL recurrences; back
 * substitution, F2(NPA) = F2(NPA-1), F2 = F*FACMU.  D(I,J,K,L) is d1+d2
 * (d2 may be NULL), both (NR,NT,NE,NPA). */
#define D4(A, I, J, K, L) (A)[((I)-1) + (size_t)st->nr * (((J)-1) + (size_t)st->nt * (((K)-1) + (size_t)st->ne * ((L)-1)))]
static void pitch_diffusion(oracle_state *st, int s, const double *d1, const double *d2) {
  const int npa = st->npa;
  double *rk = (double *)calloc(npa + 1, sizeof(double)), *rl = (double *)calloc(npa + 1, sizeof(double));
  double *facmu = (double *)calloc(npa + 1, sizeof(double)), *f = (double *)calloc(npa + 1, sizeof(double));
  for (int j = 1; j <= st->nt; ++j)
    for (int i = 2; i <= st->nr; ++i)
      for (int k = 2; k <= st->ne; ++k) {
        for (int l = 2; l <= npa; ++l) {
          facmu[l] = G3(st->fnhs, i, j, l) * V1(st->mu, l);
          f[l] = F2(s, i, j, k, l) / facmu[l];
        }
        facmu[1] = G3(st->fnhs, i, j, 1) * V1(st->mu, 1);
        f[1] = f[2];
        rk[1] = 0.;
        rl[1] = -1.;
        for (int l = 2; l <= npa - 1; ++l) {
          const double dl = d2 ? D4(d1, i, j, k, l) + D4(d2, i, j, k, l) : D4(d1, i, j, k, l);
          const double dm = d2 ? D4(d1, i, j, k, l - 1) + D4(d2, i, j, k, l - 1) : D4(d1, i, j, k, l - 1);
          double an = dl / V1(st->dmu, l);
          double gn = dm / V1(st->dmu, l - 1);
          an = an * st->dts / facmu[l] / V1(st->wmu, l);
          gn = gn * st->dts / facmu[l] / V1(st->wmu, l);
          const double bn = an + gn;
          const double rp = f[l];
          const double denom = bn + gn * rl[l - 1] + 1;
          rk[l] = (rp + gn * rk[l - 1]) / denom;
          rl[l] = -an / denom;
        }
        F2(s, i, j, k, npa - 1) = rk[npa - 1] / (1 + rl[npa - 1]);
        for (int l = npa - 2; l >= 1; --l) F2(s, i, j, k, l) = rk[l] - rl[l] * F2(s, i, j, k, l + 1);
        F2(s, i, j, k, npa) = F2(s, i, j, k, npa - 1);
        for (int l = 1; l <= npa; ++l) F2(s, i, j, k, l) = F2(s, i, j, k, l) * facmu[l];
      }
  free(rk);
  free(rl);
  free(facmu);
  free(f);
}
#undef D4

/* WPADIF(S): d1, d2 = ATAW, ATAC for electrons, ATAW_emic_h, ATAW_emic_he
 * for ions with EMIC (src/ModRamWPI.f90:674-684). */
void oracle_wpadif(oracle_state *st, int s, const double *d1, const double *d2) {
  pitch_diffusion(st, s, d1, d2);
}

/* FLCscatter(S) with FLC_coef(S,:,:,:,:) as coef (NR,NT,NE,NPA); nothing
 * while TimeRamElapsed < Dt_bc (src/ModRamLoss.f90:526). */
void oracle_flcscatter(oracle_state *st, int s, const double *coef, double dt_bc) {
  if (st->time_elapsed < dt_bc) return;
  pitch_diffusion(st, s, coef, NULL);
}

/* SUMRC, src/ModRamRun.f90:231-259 (sequential sum in the I,K,L,J order). */
void oracle_sumrc(oracle_state *st, int s) {
  st->elorc[s - 1] = 0.0;
  const double enold = st->setrc[s - 1];
  double sum = 0.0;
  for (int i = 2; i <= st->nr; ++i)
    for (int k = 2; k <= st->ne; ++k)
      for (int l = 2; l <= st->npa; ++l)
        for (int j = 1; j <= st->nt - 1; ++j) {
          const double weight = F2(s, i, j, k, l) * V1(st->we, k) * V1(st->wmu, l);
          sum = sum + V1(st->ekev, k) * weight;
        }
  st->setrc[s - 1] = sum;
  st->elorc[s - 1] = enold - sum;
}

/* ANISCH(S), src/ModRamRun.f90:265-415, without the wave-particle branch
 * after it (:422-, DoUseWPI only) and without its file output.  For I=2..NR
 * and every J: F2(S,I,J,K,1)=F2(S,I,J,K,2) for K=2..NE, then the pitch-angle
 * sums over L=1..int(UPA(I)-1) of each K and the energy windows
 * K=2..6, 7..10, 11..25, 26..30, 31..NE (:297-298, khi(5)=NE), each window's
 * sums scaled by RFAC=4*PI/(CS*100) before they are added to the totals.
 * The reference indexes F2 past NE when NE < 30; here each window ends at
 * min(khi, NE), which is the reference's for NE >= 30.
 * ffactor (NS,NR,NE,NPA), ernh and epp (NS,NE); the outputs (NS,NR,NT) get
 * their (S, 2..NR, 1..NT) entries; edent and rnhtt (the EDENT and RNHTT of
 * the reference's wans file, :405) may be NULL. */
#define FF(A, S, I, K, L) \
  (A)[((S)-1) + (size_t)st->ns * (((I)-1) + (size_t)st->nr * (((K)-1) + (size_t)st->ne * ((L)-1)))]
#define SIJ(A, S, I, J) (A)[((S)-1) + (size_t)st->ns * (((I)-1) + (size_t)st->nr * ((J)-1))]
void oracle_anisch(oracle_state *st, int s, const double *ffactor, const double *ernh, const double *epp,
                   double *ppert, double *ppart, double *edent, double *rnhtt) {
  const double cv = 2.998E8 * 100;  /* CS*100 */
  const double rfac = 4 * 3.1415926535897932384626433832795 / cv;
  int khi[5] = {6, 10, 25, 30, st->ne};
  for (int i = 2; i <= st->nr; ++i)
    for (int j = 1; j <= st->nt; ++j) {
      int klo = 2;
      double ppt = 0.0, ppa = 0.0, rnhtt_ = 0.0, edent_ = 0.0;
      for (int iwa = 0; iwa < 5; ++iwa) {
        double pper = 0.0, ppar = 0.0, rnht = 0.0, eden = 0.0;
        const int kh = khi[iwa] < st->ne ? khi[iwa] : st->ne;
        for (int k = klo; k <= kh; ++k) {
          F2(s, i, j, k, 1) = F2(s, i, j, k, 2);
          double sume = 0.0, suma = 0.0, sumn = 0.0;
          const int u = (int)(V1(st->upa, i) - 1);
          for (int l = 1; l <= u; ++l) {
            const double ernm = V1(st->wmu, l) / FF(ffactor, s, i, k, l) / G3(st->fnhs, i, j, l);
            const double epma = ernm * V1(st->mu, l) * V1(st->mu, l);
            const double epme = ernm - epma;
            sume = sume + F2(s, i, j, k, l) * epme;
            suma = suma + F2(s, i, j, k, l) * epma;
            sumn = sumn + F2(s, i, j, k, l) * ernm;
          }
          pper = pper + SK(epp, s, k) * sume;
          ppar = ppar + SK(epp, s, k) * suma;
          rnht = rnht + SK(ernh, s, k) * sumn;
          eden = eden + SK(ernh, s, k) * V1(st->ekev, k) * sumn;
        }
        rnht = rnht * rfac;
        eden = eden * rfac;
        ppar = 2 * rfac * ppar;
        pper = rfac * pper;
        klo = khi[iwa] + 1;
        edent_ = edent_ + eden;
        ppt = ppt + pper;
        ppa = ppa + ppar;
        rnhtt_ = rnhtt_ + rnht;
      }
      SIJ(ppert, s, i, j) = ppt;
      SIJ(ppart, s, i, j) = ppa;
      if (edent) SIJ(edent, s, i, j) = edent_;
      if (rnhtt) SIJ(rnhtt, s, i, j) = rnhtt_;
    }
}

/* FLUX(S,I,J,K,L) = F2/FFACTOR/FNHS for I=2..NR, K=2..NE, L=2..NPA,
 * J=1..NT-1 (src/ModRamRun.f90:208-222); other entries are left as they are
 * (the reference's FLUX is zero there: it is zeroed at allocation,
 * src/ModRamInit.f90:78, and written only by this loop). */
void oracle_flux(const oracle_state *st, const double *ffactor, double *flux) {
  for (int s = 1; s <= st->ns; ++s)
    for (int i = 2; i <= st->nr; ++i)
      for (int k = 2; k <= st->ne; ++k)
        for (int l = 2; l <= st->npa; ++l)
          for (int j = 1; j <= st->nt - 1; ++j) {
            const size_t x = (size_t)(s - 1) + (size_t)st->ns * ((i - 1) + (size_t)st->nr *
                             ((j - 1) + (size_t)st->nt * ((k - 1) + (size_t)st->ne * (l - 1))));
            flux[x] = F2(s, i, j, k, l) / FF(ffactor, s, i, k, l) / G3(st->fnhs, i, j, l);
          }
}
#undef FF
#undef SIJ

/* The per-species body of ram_run, src/ModRamRun.f90:66-184, with the
 * optional processes as switched. */
static int ram_run_species(oracle_state *st, int s) {
  oracle_drift_scratch d;
  oracle_loss_scratch ls;
  if (oracle_drift_alloc(st, &d)) return -1;
  if (oracle_loss_alloc(st, &ls)) {
    oracle_drift_free(&d);
    return -1;
  }
  const int cex = st->cex_kind[s - 1] != ORACLE_CEX_NONE;
  const int wave = st->wpi && st->wpi[s - 1];              /* WAVELO */
  const int wpd = st->pd_wpi && st->pd_wpi[s - 1];         /* WPADIF in its place */
  const int emic = st->pd_emic && st->pd_emic[s - 1];      /* WPADIF for EMIC */
  const int flc = st->flc && st->flc[s - 1];
  const int wpi = wave || wpd;
  const int coul = st->use_coulomb;
  oracle_cepara(st, s, &ls);
  oracle_driftpara(st, s, &d);
  if (coul) oracle_coulpara(st, s);
  oracle_driftr(st, s, &d);
  oracle_driftp(st, s, &d);
  oracle_drifte(st, s, &d);
  oracle_driftmu(st, s, &d);
  oracle_sumrc(st, s);
  st->lsdr[s - 1] += st->elorc[s - 1];
  if (coul) {
    oracle_coulen(st, s);
    oracle_sumrc(st, s);
    st->lscoe[s - 1] += st->elorc[s - 1];
    oracle_coulmu(st, s);
    oracle_sumrc(st, s);
    st->lscsc[s - 1] += st->elorc[s - 1];
  }
  if (wpi) {
    if (wpd) oracle_wpadif(st, s, st->pd1[s - 1], st->pd2[s - 1]);
    else oracle_wavelo(st, s);
    oracle_sumrc(st, s);
    st->lswae[s - 1] += st->elorc[s - 1];
  }
  if (emic) {
    oracle_wpadif(st, s, st->pd1[s - 1], st->pd2[s - 1]);
    oracle_sumrc(st, s);
    st->lswae[s - 1] += st->elorc[s - 1];
  }
  if (cex) {
    oracle_charexchange(st, s, &ls);
    oracle_sumrc(st, s);
    st->lscha[s - 1] += st->elorc[s - 1];
  }
  if (flc) oracle_flcscatter(st, s, st->flc[s - 1], st->dt_bc);
  oracle_atmol(st, s, &ls);
  oracle_sumrc(st, s);
  st->lsatm[s - 1] += st->elorc[s - 1];
  oracle_atmol(st, s, &ls);
  oracle_sumrc(st, s);
  st->lsatm[s - 1] += st->elorc[s - 1];
  if (flc) oracle_flcscatter(st, s, st->flc[s - 1], st->dt_bc);
  if (cex) {
    oracle_charexchange(st, s, &ls);
    oracle_sumrc(st, s);
    st->lscha[s - 1] += st->elorc[s - 1];
  }
  if (emic) {
    oracle_wpadif(st, s, st->pd1[s - 1], st->pd2[s - 1]);
    oracle_sumrc(st, s);
    st->lswae[s - 1] += st->elorc[s - 1];
  }
  if (wpi) {
    if (wpd) oracle_wpadif(st, s, st->pd1[s - 1], st->pd2[s - 1]);
    else oracle_wavelo(st, s);
    oracle_sumrc(st, s);
    st->lswae[s - 1] += st->elorc[s - 1];
  }
  if (coul) {
    oracle_coulmu(st, s);
    oracle_sumrc(st, s);
    st->lscsc[s - 1] += st->elorc[s - 1];
    oracle_coulen(st, s);
    oracle_sumrc(st, s);
    st->lscoe[s - 1] += st->elorc[s - 1];
  }
  oracle_driftmu(st, s, &d);
  oracle_drifte(st, s, &d);
  oracle_driftp(st, s, &d);
  oracle_driftr(st, s, &d);
  oracle_sumrc(st, s);
  st->lsdr[s - 1] += st->elorc[s - 1];
  oracle_drift_free(&d);
  oracle_loss_free(&ls);
  return 0;
}

/* ram_run, src/ModRamRun.f90:64-205. */
int oracle_ram_run(oracle_state *st, int nthreads, double *dts_next) {
  int err = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static, 1) reduction(| : err)
#endif
  for (int s = 1; s <= st->ns; ++s) err |= ram_run_species(st, s) ? 1 : 0;
  if (err) return -1;
  /* :186 periodic copy F2(:,:,NT,:,:)=F2(:,:,1,:,:) */
  for (int l = 1; l <= st->npa; ++l)
    for (int k = 1; k <= st->ne; ++k)
      for (int i = 1; i <= st->nr; ++i)
        for (int s = 1; s <= st->ns; ++s) F2(s, i, st->nt, k, l) = F2(s, i, 1, k, l);
  /* :197-201 outside-magnetopause cells */
  for (int i = 1; i <= st->nr; ++i)
    for (int j = 1; j <= st->nt; ++j)
      if (MGNP(i, j) == 1)
        for (int l = 1; l <= st->npa; ++l)
          for (int k = 1; k <= st->ne; ++k)
            for (int s = 1; s <= st->ns; ++s) F2(s, i, j, k, l) = 1e-31;
  if (dts_next) { /* :204 (the DtsMin floor of :205 is the caller's) */
    double m = 1e300;
    for (int s = 0; s < st->ns; ++s) {
      m = dmin(m, st->dtdrift_r[s]);
      m = dmin(m, st->dtdrift_p[s]);
      m = dmin(m, st->dtdrift_e[s]);
      m = dmin(m, st->dtdrift_mu[s]);
    }
    *dts_next = m;
  }
  return 0;
}
