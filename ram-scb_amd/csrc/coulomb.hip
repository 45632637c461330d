// coulomb.hip -- Coulomb collisions of the split step (src/ModRamCoul.f90):
// the COULPARA tables and the C ABI of COULPARA, COULEN and COULMU.
//
// COULPARA's tables depend on the grid, the species, VBND and DTs only, so
// they are formed on the host once per change of those inputs (the
// reference recomputes them every step, src/ModRamRun.f90:69) with the C
// library's erf, exp, log, pow and sqrt, in the reference's expression
// order, and uploaded: COULEN and COULMU then run on the device in the plane
// kernels (k_plane<SEQ_COULEN|SEQ_COULMU|SEQ_STEPC>).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ctx.h"

using namespace ramgpu;

namespace {

// physical constants of src/ModRamConst.f90 beside Q, CS and RE
constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kMP = 1.673E-27;  // :18

// The bounce integrals of a dipole, Ejiri 1978 (src/ModRamFunctions.f90:
// 90-143), in the reference's expression order.
double ejiri_alpha() { return 1. + std::log(2. + std::sqrt(3.)) / 2. / std::sqrt(3.); }
double ejiri_beta(double alpha) { return alpha / 2. - kPi * std::sqrt(2.) / 12.; }
constexpr double kA1 = 0.055, kA2 = -0.037, kA3 = -0.074, kA4 = 0.056;

double funt(double x) {
  const double y = std::sqrt(1 - x * x), alpha = ejiri_alpha(), beta = ejiri_beta(alpha);
  return alpha - beta * (y + std::sqrt(y)) + kA1 * std::pow(y, 1. / 3.) + kA2 * std::pow(y, 2. / 3.) + kA3 * y +
         kA4 * std::pow(y, 4. / 3.);
}

double funi(double x) {
  const double y = std::sqrt(1 - x * x), alpha = ejiri_alpha(), beta = ejiri_beta(alpha);
  const double ylog = y > 0 ? std::log(y) : 0.0;
  return 2. * alpha * (1. - y) + 2. * beta * y * ylog + 4. * beta * (y - std::sqrt(y)) +
         3. * kA1 * (std::pow(y, 1. / 3.) - y) + 6. * kA2 * (std::pow(y, 2. / 3.) - y) +
         6. * kA4 * (y - std::pow(y, 4. / 3.)) - 2. * kA3 * y * ylog;
}

// (erf(x) - x erf'(x)) / (2 x^2), src/ModRamFunctions.f90:72-87
double gcoul(double x) {
  const double g1 = std::erf(x) - 2. * x / std::sqrt(kPi) * std::exp(-x * x);
  return g1 / 2. / x / x;
}

// The cold plasmasphere the test particles collide with: RAMSpecies(1..6)
// of src/ModRamSpecies.f90:43-136 (mass in proton masses, charge,
// plasmasphereRatio); ratios below 1e-9 are skipped (src/ModRamCoul.f90:69).
struct ColdSpecies {
  double mass, charge, ratio;
};
constexpr ColdSpecies kCold[6] = {{5.4462E-4, -1, 1.0}, {1.0, 1, 0.77}, {4.0, 1, 0.2},
                                  {16.0, 1, 0.03},     {14.0, 1, 0.0},  {87.62, 1, 0.0}};

enum CoulTable { CT_COULE, CT_COULI, CT_ATA, CT_GTA, CT_CEDR, CT_CIDR, CT_N };

// COULPARA(S), src/ModRamCoul.f90:17-125, for species s (0-based): the six
// tables as tab[CT_*][l*NE + k] (entries the reference never writes stay
// 0, as its zeroed allocation leaves them).  The plasmasphere sums start at
// 0 once per species and run on over K (:41-46, 64-84).
void coulpara_species(const ramgpu_ctx *c, int s, std::vector<double> (&tab)[CT_N]) {
  const int ne = c->ne, npa = c->npa;
  for (auto &v : tab) v.assign((size_t)ne * npa, 0.0);
  auto at = [&](int q, int k, int l) -> double & { return tab[q][(size_t)(l - 1) * ne + (k - 1)]; };
  auto sk = [&](const std::vector<double> &a, int k) { return a[(size_t)s + (size_t)c->ns * (k - 1)]; };
  std::vector<double> coulde(npa + 1, 0.0), couldi(npa + 1, 0.0);
  double cce = 0, cde = 0, edre = 0, cci = 0, cdi = 0, edri = 0;
  const double eps = 8.854E-12, dln = 21.5;
  const double zt = (double)c->t.charge[s];
  const double qe = (kQ * kQ / eps);
  const double gama = zt * zt * dln / 4. / kPi * qe * 1E6 * qe;
  const double cco = gama / kQ * c->t.dts / kQ / 1E3;
  const double rm = c->t.rmas[s];
  const double ccd = gama * c->t.dts / (rm * rm) / (kCS * kCS * kCS);
  const double edrco = dln * qe * kRE / rm * qe * 1E9 / kQ;
  const std::vector<double> &mu = c->h_mu, &wmu = c->h_wmu, &dmu = c->h_dmu;
  for (int k = 1; k <= ne; ++k) {
    const double vb = sk(c->h_vbnd, k), vk = sk(c->h_v, k), gb = sk(c->h_grbnd, k), gr = sk(c->h_grel, k);
    for (const ColdSpecies &cs : kCold) {
      const double ra = cs.ratio;
      if (ra < 1e-9) continue;
      const double vf = std::sqrt(2. * kQ / (kMP * cs.mass));
      const double zb = cs.charge, x = vb / vf, xd = vk / vf;
      if (zb < 0.) {
        cce = cce + ra * gcoul(x);
        cde = cde + ra * (std::erf(xd) - gcoul(xd));
        edre = edre + ra * gcoul(xd);
      } else {
        cci = cci + ra * (zb * zb) * gcoul(x);
        cdi = cdi + ra * (zb * zb) * (std::erf(xd) - gcoul(xd));
        edri = edri + ra * gcoul(xd);
      }
    }
    const double ek = c->h_ekev[k - 1];
    at(CT_COULE, k, 1) = -(cce * vb * cco * (gb * gb));
    at(CT_COULI, k, 1) = -(cci * vb * cco * (gb * gb));
    at(CT_CEDR, k, 1) = edrco * ek * edre * (gr + 1) / (gr * gr) / (vk * vk);
    at(CT_CIDR, k, 1) = edrco * ek * edri * (gr + 1) / (gr * gr) / (vk * vk);
    const double ccde = ccd * cde * gr / std::pow(gr * gr - 1, 1.5);
    const double ccdi = ccd * cdi * gr / std::pow(gr * gr - 1, 1.5);
    for (int l = 2; l <= npa - 1; ++l) {
      const double m = mu[l - 1], w = wmu[l - 1];
      const double bane = (1. - funi(m) / 2. / funt(m)) / (1. - m * m);
      at(CT_COULE, k, l) = at(CT_COULE, k, 1);
      at(CT_COULI, k, l) = at(CT_COULI, k, 1);
      at(CT_CEDR, k, l) = at(CT_CEDR, k, 1) * bane * funt(m) * m * w;
      at(CT_CIDR, k, l) = at(CT_CIDR, k, 1) * bane * funt(m) * m * w;
      const double muboun = m + 0.5 * w;
      const double badif = (1. - muboun * muboun) / muboun / 2.;
      coulde[l] = ccde * badif;
      couldi[l] = ccdi * badif;
      const double afer = coulde[l] / m / dmu[l - 1] / w, asec = coulde[l - 1] / m / dmu[l - 2] / w;
      const double afir = couldi[l] / m / dmu[l - 1] / w, asic = couldi[l - 1] / m / dmu[l - 2] / w;
      at(CT_ATA, k, l) = afir + afer;
      at(CT_GTA, k, l) = asic + asec;
    }
    at(CT_CEDR, k, npa) = at(CT_CEDR, k, npa - 1);
    at(CT_CIDR, k, npa) = at(CT_CIDR, k, npa - 1);
    at(CT_ATA, k, npa) = 0;
  }
}

int coul_alloc(ramgpu_ctx *c) {
  if (c->t.cc) return RAMGPU_OK;
  const size_t n = (size_t)c->ns * c->npa * c->ne;
  void *p[4] = {nullptr, nullptr, nullptr, nullptr};
  const size_t bytes[4] = {n * sizeof(double), n * sizeof(double2), (size_t)c->ns * 5 * sizeof(double),
                           (size_t)c->nr * c->nt * sizeof(double)};
  for (int q = 0; q < 4; ++q) {
    if (hipMalloc(&p[q], bytes[q]) != hipSuccess || hipMemset(p[q], 0, bytes[q]) != hipSuccess) {
      for (int r = 0; r <= q; ++r)
        if (p[r]) (void)hipFree(p[r]);
      return ramgpu_host::fail(RAMGPU_ENOMEM, "allocation of the Coulomb tables failed");
    }
  }
  for (void *q : p) c->allocs.push_back(q);
  c->t.cc = (double *)p[0];
  c->t.cag = (double2 *)p[1];
  c->t.crel = (double *)p[2];
  c->d_necr = (double *)p[3];
  c->t.necr = c->d_necr;
  c->tab_dirty = true;
  return RAMGPU_OK;
}

}  // namespace

namespace ramgpu_host {

// The COULPARA tables of every species and COULEN's ghost factors, formed
// and uploaded when an input they depend on has changed.
int coul_tables(ramgpu_ctx *c) {
  if (!c->coul_dirty) return RAMGPU_OK;
  if (!c->have_grid) return fail(RAMGPU_ESTATE, "ramgpu_set_grid has not been called");
  if (c->h_vbnd.empty()) return fail(RAMGPU_ESTATE, "ramgpu_set_coulomb has not been given VBND");
  for (int s = 0; s < c->ns; ++s)
    if (!c->have_species[s]) return fail(RAMGPU_ESTATE, "species %d not set", s + 1);
  int rc = coul_alloc(c);
  if (rc) return rc;
  const size_t P = (size_t)c->ne * c->npa;
  std::vector<double> &h = c->h_coul;
  h.assign((size_t)CT_N * c->ns * P, 0.0);
  std::vector<double> cc((size_t)c->ns * P), crel((size_t)c->ns * 5);
  std::vector<double2> cag((size_t)c->ns * P);
  std::vector<double> tab[CT_N];
  for (int s = 0; s < c->ns; ++s) {
    coulpara_species(c, s, tab);
    for (int q = 0; q < CT_N; ++q) std::copy(tab[q].begin(), tab[q].end(), h.begin() + ((size_t)q * c->ns + s) * P);
    for (size_t e = 0; e < P; ++e) {
      cc[s * P + e] = tab[CT_COULE][e] + tab[CT_COULI][e];  // CccolE's sum, src/ModRamCoul.f90:185
      cag[s * P + e] = make_double2(tab[CT_ATA][e], tab[CT_GTA][e]);
    }
    // COULEN's ghosts (src/ModRamCoul.f90:150-151,180-181)
    const double g1 = c->h_grel[s], g2 = c->h_grel[s + c->ns];
    const double ezero = c->h_ekev[0] - c->h_we[0];
    const double gz = 1. + ezero * 1000. * kQ / c->t.rmas[s] / kCS / kCS;
    crel[s * 5 + 0] = g1;
    crel[s * 5 + 1] = g2;
    crel[s * 5 + 2] = std::sqrt((g1 * g1 - 1) / (g2 * g2 - 1));
    crel[s * 5 + 3] = gz;
    crel[s * 5 + 4] = std::sqrt((gz * gz - 1) / (g1 * g1 - 1));
  }
  HIP_TRY(hipMemcpyAsync(c->t.cc, cc.data(), cc.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->t.cag, cag.data(), cag.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->t.crel, crel.data(), crel.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
  c->coul_dirty = false;
  return RAMGPU_OK;
}

int coul_check_lds(const ramgpu_ctx *c, int threads) {
  if (plane_coul_lds_bytes(c->ne, c->npa, threads) > c->lds_limit)
    return fail(RAMGPU_EINVAL, "a %dx%d plane with the Coulomb operators' tables exceeds the device LDS (%zu bytes)",
                c->ne, c->npa, c->lds_limit);
  return RAMGPU_OK;
}

}  // namespace ramgpu_host

using namespace ramgpu_host;

namespace {
// COULEN (op 0) or COULMU (op 1) of species s alone
int coul_op(ramgpu_ctx *c, int s, int op) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  if (!c->t.coul) return fail(RAMGPU_ESTATE, "Coulomb collisions are off (ramgpu_set_coulomb)");
  if ((rc = prepare(c))) return rc;
  const int clamp = (op == 1 && c->time_elapsed > 0.) ? 2 : 0;
  return run_plane(c, op == 0 ? SEQ_COULEN : SEQ_COULMU, clamp, s - 1, 1);
}
}  // namespace

extern "C" {

int ramgpu_set_coulomb(ramgpu_ctx *c, int use, const double *vbnd) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (use) {
    if (!vbnd) return fail(RAMGPU_EINVAL, "vbnd is null");
    if (c->npa < 13)
      return fail(RAMGPU_EINVAL, "Coulomb collisions need NPA >= 13 (COULEN's BANE(NPA-10..NPA) copy BANE(NPA-11))");
    if ((rc = coul_check_lds(c, c->plane.threads)) || (rc = coul_alloc(c))) return rc;
    c->h_vbnd.assign(vbnd, vbnd + (size_t)c->ns * c->ne);
  }
  c->t.coul = use ? 1 : 0;
  c->coul_dirty = true;
  c->tab_dirty = true;
  return RAMGPU_OK;
}

int ramgpu_set_necr(ramgpu_ctx *c, const double *necr) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!necr) return fail(RAMGPU_EINVAL, "necr is null");
  if ((rc = coul_alloc(c))) return rc;
  HIP_TRY(hipMemcpyAsync(c->d_necr, necr, (size_t)c->nr * c->nt * sizeof(double), hipMemcpyHostToDevice,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->have_necr = true;
  return RAMGPU_OK;
}

int ramgpu_set_time(ramgpu_ctx *c, double time_elapsed) {
  int rc = check_ctx(c);
  if (rc) return rc;
  c->time_elapsed = time_elapsed;
  return RAMGPU_OK;
}

int ramgpu_get_time(ramgpu_ctx *c, double *time_elapsed) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!time_elapsed) return fail(RAMGPU_EINVAL, "null output pointer");
  *time_elapsed = c->time_elapsed;
  return RAMGPU_OK;
}

int ramgpu_coulpara(ramgpu_ctx *c, int s) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  if (!c->t.coul) return fail(RAMGPU_ESTATE, "Coulomb collisions are off (ramgpu_set_coulomb)");
  return coul_tables(c);
}

int ramgpu_get_coulomb_tables(ramgpu_ctx *c, int s, double *coule, double *couli, double *ata, double *gta,
                              double *cedr, double *cidr) {
  int rc = check_ctx(c);
  if (rc || (rc = check_species(c, s))) return rc;
  if (!c->t.coul) return fail(RAMGPU_ESTATE, "Coulomb collisions are off (ramgpu_set_coulomb)");
  if ((rc = coul_tables(c))) return rc;
  double *out[CT_N] = {coule, couli, ata, gta, cedr, cidr};
  const size_t P = (size_t)c->ne * c->npa;
  for (int q = 0; q < CT_N; ++q)
    if (out[q]) {
      const double *src = c->h_coul.data() + ((size_t)q * c->ns + (s - 1)) * P;
      std::copy(src, src + P, out[q]);
    }
  return RAMGPU_OK;
}

int ramgpu_coulen(ramgpu_ctx *c, int s) { return coul_op(c, s, 0); }
int ramgpu_coulmu(ramgpu_ctx *c, int s) { return coul_op(c, s, 1); }

}  // extern "C"
