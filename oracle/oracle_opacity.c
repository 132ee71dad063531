/*
 * oracle_opacity.c -- TEST INFRASTRUCTURE: FP64 CPU restatement of the
 * per-zone opacity / emissivity build (transport::set_opacity and the
 * GasState / AtomicSpecies functions it calls).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* radioactive::decay_composition, src/transport/radioactive.cpp:58-113 */
void or_decay_composition(int n_el, const int32_t* Z, const int32_t* A, double* X, double t)
{
  double X_ni = 0, X_co = 0;
  for (int i = 0; i < n_el; i++) {
    if (Z[i] == 28 && A[i] == 56) X_ni = X[i];
    if (Z[i] == 27 && A[i] == 56) X_co = X[i];
  }
  double ni_f = exp(-t / OR_TAU_56NI);
  double co_f = OR_TAU_56CO / (OR_TAU_56NI - OR_TAU_56CO) * (exp(-t / OR_TAU_56NI) - exp(-t / OR_TAU_56CO));
  double fe_f = 1 - ni_f - co_f;
  double eco = exp(-t / OR_TAU_56CO);
  for (int i = 0; i < n_el; i++) {
    if (Z[i] == 28 && A[i] == 56) X[i] = X_ni * ni_f;
    if (Z[i] == 27 && A[i] == 56) X[i] = X_ni * co_f + X_co * eco;
    if (Z[i] == 26 && A[i] == 56) X[i] += X_ni * fe_f + X_co * (1 - eco);
  }
  double X_cr = 0, X_vn = 0;
  for (int i = 0; i < n_el; i++) {
    if (Z[i] == 24 && A[i] == 48) X_cr = X[i];
    if (Z[i] == 23 && A[i] == 48) X_vn = X[i];
  }
  double cr_f = exp(-t / OR_TAU_48CR);
  double vn_f = OR_TAU_48V / (OR_TAU_48CR - OR_TAU_48V) * (exp(-t / OR_TAU_48CR) - exp(-t / OR_TAU_48V));
  double ti_f = 1 - cr_f - vn_f;
  double evn = exp(-t / OR_TAU_48V);
  for (int i = 0; i < n_el; i++) {
    if (Z[i] == 24 && A[i] == 48) X[i] = X_cr * cr_f;
    if (Z[i] == 23 && A[i] == 48) X[i] = X_cr * vn_f + X_vn * evn;
    if (Z[i] == 22 && A[i] == 48) X[i] += X_cr * ti_f + X_vn * (1 - evn);
  }
}

/* r-process fit coefficients are not restated: the hot path's configs carry
 * no Z >= 58 element; radioactive::decay_energy_rate, radioactive.cpp:189-289 */
static double decay_energy_rate(int Z, int A, double t, double* gfrac)
{
  double total = 0, gtotal = 0;
  if (Z == 28 && A == 56) {
    double e_ni = exp(-t / OR_TAU_56NI);
    double e_co = exp(-t / OR_TAU_56CO);
    double ni56 = (e_ni / OR_TAU_56NI);
    double co56 = 1.0 / (OR_TAU_56NI - OR_TAU_56CO) * (e_ni - e_co);
    double ni_E = ni56 * (OR_E_56NI * OR_MEV_TO_ERGS);
    double co_E = co56 * (OR_E_56CO * OR_MEV_TO_ERGS);
    gtotal = ni_E + (1 - OR_CO_POS) * co_E;
    total = ni_E + co_E;
  }
  if (Z == 27 && A == 56) {
    double e_co = exp(-t / OR_TAU_56CO);
    double co56 = 1.0 / (OR_TAU_56CO) * e_co;
    double co_E = co56 * (OR_E_56CO * OR_MEV_TO_ERGS);
    gtotal = (1 - OR_CO_POS) * co_E;
    total = co_E;
  }
  if (Z == 24 && A == 48) {
    double e_cr = exp(-t / OR_TAU_48CR);
    double e_vn = exp(-t / OR_TAU_48V);
    double cr48 = (e_cr / OR_TAU_48CR);
    double vn48 = 1.0 / (OR_TAU_48CR - OR_TAU_48V) * (e_cr - e_vn);
    double cr_E = cr48 * (OR_E_48CR * OR_MEV_TO_ERGS);
    double vn_E = vn48 * (OR_E_48V * OR_MEV_TO_ERGS);
    gtotal += cr_E + vn_E;
    total += cr_E + vn_E;
  }
  *gfrac = gtotal / total;
  if (total == 0) *gfrac = 0;
  return total;
}

/* radioactive::rprocess_heating_rate, radioactive.cpp:116-142 */
static double rprocess_heating_rate(double t, double* gfrac)
{
  double td = t / 3600.0 / 24.0;
  double A1 = 8.4939E+09, alpha = 1.3642E+00, B1 = 8.3425E+09, beta1 = 3.6280E+00;
  double B2 = 8.8616E+08, beta2 = 1.0847E+01;
  double eps = A1 * pow(td, -1.0 * alpha) + B1 * exp(-1.0 * td / beta1) + B2 * exp(-1.0 * td / beta2);
  double af = 0.56, bf = 0.17, df = 0.74;
  double x = 2 * bf * pow(td, df);
  double f = 0.36 * (exp(-af * td) + log(1 + x) / x);
  *gfrac = 0;
  return eps * f;
}

/* radioactive::decay, src/transport/radioactive.cpp:151-184 */
double or_decay(int n_el, const int32_t* Z, const int32_t* A, const double* X, double t,
                double* gfrac, int force_rproc)
{
  int rproc = 0;
  for (int k = 0; k < n_el; k++)
    if (((Z[k] > 57) && (A[k] > 0)) || force_rproc) rproc = 1;
  if (rproc) return rprocess_heating_rate(t, gfrac);
  double total = 0, gtotal = 0, gf;
  for (int k = 0; k < n_el; k++) {
    double val = decay_energy_rate(Z[k], A[k], t, &gf);
    val = val * X[k] / (A[k] * OR_M_P);
    total += val;
    gtotal += val * gf;
  }
  *gfrac = gtotal / total;
  if (total == 0) *gfrac = 0;
  return total;
}

/* AtomicSpecies::solve_lte, src/opacity/AtomicSpecies_state.cpp:56-123.
 * ion_part/ion_frac index the atom's ions, lev_n its levels. */
void or_solve_lte(const sgpu_atoms* at, int a, double ne, double T,
                  double* ion_part, double* ion_frac, double* lev_n)
{
  int l0 = at->lev_off[a], nl = at->lev_off[a + 1] - l0;
  int i0 = at->ion_off[a], ni = at->ion_off[a + 1] - i0;
  for (int i = 0; i < ni; ++i) ion_part[i] = 0;
  for (int i = 0; i < nl; ++i) {
    double E = at->lev_E[l0 + i];
    int g = at->lev_g[l0 + i];
    int ion = at->lev_ion[l0 + i];
    lev_n[i] = g * exp(-E / OR_K_EV / T);
    ion_part[ion] += lev_n[i];
  }
  double lt = OR_H * OR_H / (2.0 * OR_PI * OR_M_E * OR_K * T);
  double fac = 2 / ne / pow(lt, 1.5);
  ion_frac[0] = 1.;
  double norm = 1.;
  for (int i = 1; i < ni; ++i) {
    double chi = at->ion_chi[i0 + i - 1];
    double saha = exp(-1.0 * chi / OR_K_EV / T);
    saha = saha * (ion_part[i] / ion_part[i - 1]) * fac;
    ion_frac[i] = saha * ion_frac[i - 1];
    if (ne < 1e-50) ion_frac[i] = 0;
    norm = 0;
    for (int j = 0; j <= i; j++) norm += ion_frac[j];
    for (int j = 0; j <= i; j++) ion_frac[j] = ion_frac[j] / norm;
  }
  for (int i = 0; i < nl; ++i) {
    double E = at->lev_E[l0 + i];
    int g = at->lev_g[l0 + i];
    int ion = at->lev_ion[l0 + i];
    double Z = ion_part[ion];
    double f = ion_frac[ion];
    lev_n[i] = f * g * exp(-E / OR_K_EV / T) / Z;
    if (lev_n[i] < 1e-30) lev_n[i] = 1e-30;   /* min_level_pop_, AtomicSpecies_init.cpp:36 */
  }
}

/* Brent state of one zone: per-atom arrays live in caller buffers */
typedef struct {
  const sgpu_atoms* at;
  int n_el;
  const int32_t* A;
  const double* mf;
  double dens, T;
  double* ion_part; double* ion_frac; double* lev_n;
  int n_evals;
  /* NLTE atoms (atom_nlte NULL: every atom in LTE): this zone's radiative
   * rates (global level / line index), flags and scratch */
  const int32_t* atom_nlte;
  const double* Pic; const double* Rci; const double* line_J;
  int use_collisions, no_ground_recomb;
  double* lev_lte; double* rates; double* lte_v; double* bvec;
} zone_gas;

/* GasState::charge_conservation, src/opacity/GasState.cpp:242-263 */
static double charge_conservation(zone_gas* z, double ne)
{
  double f = 0;
  z->n_evals++;
  for (int i = 0; i < z->n_el; ++i) {
    int l0 = z->at->lev_off[i], i0 = z->at->ion_off[i];
    int nl = z->at->lev_off[i + 1] - l0;
    or_solve_lte(z->at, i, ne, z->T, z->ion_part + i0, z->ion_frac + i0, z->lev_n + l0);
    if (z->atom_nlte && z->atom_nlte[i]) {
      /* AtomicSpecies::solve_nlte (AtomicSpecies_state.cpp:131-196): the LTE
       * populations just computed are lev_lte; rates, LU, populations */
      for (int l = 0; l < nl; ++l) z->lev_lte[l0 + l] = z->lev_n[l0 + l];
      int n = or_nlte_set_rates(z->at, i, ne, z->T, z->line_J, z->Pic, z->Rci, z->lev_lte, z->use_collisions,
                                z->no_ground_recomb, z->rates, z->lte_v);
      or_nlte_solve(n, z->rates, z->lte_v, z->bvec, NULL);
      or_nlte_populations(z->at, i, z->bvec, z->lte_v, z->lev_n + l0, z->ion_frac + i0);
    }
    double x = 0;   /* AtomicSpecies::get_net_ion_fraction, AtomicSpecies.h:125-131 */
    for (int l = 0; l < nl; ++l) x += z->lev_n[l0 + l] * z->at->lev_ion[l0 + l];
    f += z->dens * z->mf[i] / (z->A[i] * OR_M_P) * x;
  }
  return f - ne;
}

#define OR_SIGN(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a))

/* GasState::ne_brent_method, src/opacity/GasState.cpp:271-337 (Numerical
 * Recipes zbrent).  The atoms keep the state of the LAST evaluation. */
static double ne_brent(zone_gas* z, double x1, double x2, double tol, int* err)
{
  const int ITMAX = 100;
  const double EPS = 3.0e-8;
  double a = x1, b = x2, c = x2, d = 0, e = 0, min1, min2;
  double fa = charge_conservation(z, a);
  double fb = charge_conservation(z, b);
  double fc, p, q, r, s, tol1, xm;
  if ((fa > 0.0 && fb > 0.0) || (fa < 0.0 && fb < 0.0)) *err = 1;
  fc = fb;
  for (int iter = 1; iter <= ITMAX; iter++) {
    if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
      c = a; fc = fa; e = d = b - a;
    }
    if (fabs(fc) < fabs(fb)) {
      a = b; b = c; c = a;
      fa = fb; fb = fc; fc = fa;
    }
    tol1 = 2.0 * EPS * fabs(b) + 0.5 * tol;
    xm = 0.5 * (c - b);
    if (fabs(xm) <= tol1 || fb == 0.0) return b;
    if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
      s = fb / fa;
      if (a == c) {
        p = 2.0 * xm * s;
        q = 1.0 - s;
      } else {
        q = fa / fc;
        r = fb / fc;
        p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
        q = (q - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (p > 0.0) q = -q;
      p = fabs(p);
      min1 = 3.0 * xm * q - fabs(tol1 * q);
      min2 = fabs(e * q);
      if (2.0 * p < (min1 < min2 ? min1 : min2)) {
        e = d; d = p / q;
      } else {
        d = xm; e = d;
      }
    } else {
      d = xm; e = d;
    }
    a = b; fa = fb;
    if (fabs(d) > tol1) b += d;
    else b += OR_SIGN(tol1, xm);
    fb = charge_conservation(z, b);
  }
  *err = 2;
  return 0.0;
}

/* GasState::solve_state (LTE), src/opacity/GasState.cpp:208-233 */
int or_solve_state(const sgpu_atoms* at, int n_el, const int32_t* elem_A,
                   const double* mass_frac, double mu_I, double dens, double T,
                   double* n_elec, double* ion_part, double* ion_frac, double* lev_n,
                   int* n_evals)
{
  zone_gas z = {at, n_el, elem_A, mass_frac, dens, T, ion_part, ion_frac, lev_n, 0,
                NULL, NULL, NULL, NULL, 0, 0, NULL, NULL, NULL, NULL};
  double max_ne = 100 * dens / (mu_I * OR_M_P);
  double min_ne = 1e-10 * dens / (mu_I * OR_M_P);
  int err = 0;
  *n_elec = ne_brent(&z, min_ne, max_ne, 1e-3, &err);
  if (n_evals) *n_evals = z.n_evals;
  return err;
}

/* NLTE inputs of the opacity build (NULL: LTE) */
typedef struct {
  const double* J_nu;        /* [nz][n_nu] */
  const int32_t* atom_nlte;  /* [n_atoms] */
  double line_velocity_width;
  int use_collisions;
} nlte_in;

/* GasState::solve_state(J_nu) with use_nlte_ (GasState.cpp:208-233): the
 * radiative rates of every NLTE atom from this zone's J_nu once, then the
 * n_e Brent iteration with solve_nlte for those atoms. */
static int solve_state_nlte(const sgpu_atoms* at, const sgpu_locate* nu, int n_el, const int32_t* elem_A,
                            const double* mass_frac, double mu_I, double dens, double T, const nlte_in* nl,
                            const double* J, int no_ground_recomb, double* n_elec, double* ion_part,
                            double* ion_frac, double* lev_n)
{
  int n_lev = at->lev_off[at->n_atoms];
  size_t nbl = at->n_blines > 0 ? (size_t)at->n_blines : 1;
  int ml = at->max_levels > 0 ? at->max_levels : 1;
  double* Pic = (double*)calloc((size_t)n_lev, sizeof(double));
  double* Rci = (double*)calloc((size_t)n_lev, sizeof(double));
  double* line_J = (double*)calloc(nbl, sizeof(double));
  double* lev_lte = (double*)calloc((size_t)n_lev, sizeof(double));
  double* rates = (double*)malloc(sizeof(double) * (size_t)ml * ml);
  double* lte_v = (double*)malloc(sizeof(double) * (size_t)ml);
  double* bvec = (double*)malloc(sizeof(double) * (size_t)ml);
  for (int a = 0; a < n_el; ++a)
    if (nl->atom_nlte[a])
      or_nlte_radiative_rates(at, a, nu, J, T, elem_A[a], nl->line_velocity_width, Pic, Rci, line_J);
  zone_gas z = {at, n_el, elem_A, mass_frac, dens, T, ion_part, ion_frac, lev_n, 0,
                nl->atom_nlte, Pic, Rci, line_J, nl->use_collisions, no_ground_recomb,
                lev_lte, rates, lte_v, bvec};
  double max_ne = 100 * dens / (mu_I * OR_M_P);
  double min_ne = 1e-10 * dens / (mu_I * OR_M_P);
  int err = 0;
  *n_elec = ne_brent(&z, min_ne, max_ne, 1e-3, &err);
  free(Pic); free(Rci); free(line_J); free(lev_lte); free(rates); free(lte_v); free(bvec);
  return err;
}

/* transport::klein_nishina, src/transport/transport_opacity.cpp:306-316 */
double or_klein_nishina(double x)
{
  x = x / OR_M_E_MEV;
  double logfac = log(1 + 2 * x);
  double term1 = (1 + x) / x / x / x * (2 * x * (1 + x) / (1 + 2 * x) - logfac);
  double term2 = 1.0 / 2.0 / x * logfac;
  double term3 = -1.0 * (1 + 3 * x) / (1 + 2 * x) / (1 + 2 * x);
  return .75 * (term1 + term2 + term3);
}

/* transport::blackbody_nu, transport_opacity.cpp:322-326 */
double or_blackbody_nu(double T, double nu)
{
  double zeta = OR_H * nu / OR_K / T;
  return 2.0 * nu * nu * nu * OR_H / OR_C / OR_C / (exp(zeta) - 1);
}

/* transport::set_opacity zone loop, src/transport/transport_opacity.cpp:83-204,
 * with GasState::computeOpacity (GasState_opacities.cpp:11-136),
 * fuzz_expansion_opacity (:464-494), AtomicSpecies::fuzzline_expansion_opacity
 * (AtomicSpecies_opacities.cpp:292-324), get_planck_mean (:564-591),
 * get_rosseland_mean (:638-666) and the Fleck factor (:231-252). */
static int opacity_build(const sgpu_locate* nu, const sgpu_zone_state* zs, const sgpu_atoms* at,
                         const sgpu_opacity_params* p, const nlte_in* nlte, sgpu_opacity_out* out)
{
  int nz = zs->n_zones, ne_ = zs->n_elems, nn = nu->n;
  int n_lev = at && at->n_atoms ? at->lev_off[at->n_atoms] : 0;
  int n_ion = at && at->n_atoms ? at->ion_off[at->n_atoms] : 0;
  double* X_now = (double*)malloc(sizeof(double) * (ne_ > 0 ? ne_ : 1));
  double* mf = (double*)malloc(sizeof(double) * (ne_ > 0 ? ne_ : 1));
  double* ip = (double*)calloc(n_ion > 0 ? n_ion : 1, sizeof(double));
  double* ifr = (double*)calloc(n_ion > 0 ? n_ion : 1, sizeof(double));
  double* ln = (double*)calloc(n_lev > 0 ? n_lev : 1, sizeof(double));
  double* emis = (double*)malloc(sizeof(double) * nn);
  double* opac = (double*)malloc(sizeof(double) * nn);
  double* aopac = (double*)malloc(sizeof(double) * nn);
  double* atom_opac = (double*)malloc(sizeof(double) * nn);
  double* emis2 = (double*)malloc(sizeof(double) * nn);
  double* phifac = (double*)calloc(n_lev > 0 ? n_lev : 1, sizeof(double));
  int worst = 0;

  for (int i = 0; i < nz; i++) {
    double* abs = out->abs + (size_t)i * nn;
    double* scat = out->scat + (size_t)i * nn;
    double* cdf = out->emis_cdf + (size_t)i * nn;
    double dens = zs->rho[i];
    double temp = zs->T_gas[i];
    if (temp < p->temp_min) temp = p->temp_min;
    if (temp > p->temp_max) temp = p->temp_max;
    double time = p->t_now;
    for (int j = 0; j < ne_; j++) X_now[j] = zs->X[(size_t)i * ne_ + j];
    if (!p->omit_composition_decay) or_decay_composition(ne_, zs->elem_Z, zs->elem_A, X_now, p->t_now);
    /* GasState::set_mass_fractions, GasState.cpp:137-156 */
    double norm = 0.0;
    for (int j = 0; j < ne_; ++j) { mf[j] = X_now[j]; norm += X_now[j]; }
    for (int j = 0; j < ne_; ++j) mf[j] /= norm;
    double inv_mu = 0.;
    for (int j = 0; j < ne_; ++j) inv_mu += mf[j] / zs->elem_A[j];
    double mu_I = 1. / inv_mu;
    double grey = zs->grey_opacity ? zs->grey_opacity[i] : 0.0;
    if (p->first_step) temp = zs->T_gas[i];   /* transport_opacity.cpp:111-113 */
    int status = 0;
    double n_elec = 0;
    if (grey == 0) {
      /* transport_opacity.cpp:30-36, 114-136: LTE on the first step */
      if (nlte && !p->first_step)
        status = solve_state_nlte(at, nu, ne_, zs->elem_A, mf, mu_I, dens, temp, nlte,
                                  nlte->J_nu + (size_t)i * nn, p->no_ground_recomb, &n_elec, ip, ifr, ln);
      else
        status = or_solve_state(at, ne_, zs->elem_A, mf, mu_I, dens, temp, &n_elec, ip, ifr, ln, NULL);
    }
    if (status > worst) worst = status;
    if (out->zone_status) out->zone_status[i] = status;
    out->n_elec[i] = n_elec;
    if (out->ion_frac) for (int k = 0; k < n_ion; k++) out->ion_frac[(size_t)i * n_ion + k] = grey == 0 ? ifr[k] : 0;
    if (out->ion_part) for (int k = 0; k < n_ion; k++) out->ion_part[(size_t)i * n_ion + k] = grey == 0 ? ip[k] : 0;
    if (out->lev_n) for (int k = 0; k < n_lev; k++) out->lev_n[(size_t)i * n_lev + k] = grey == 0 ? ln[k] : 0;

    /* computeOpacity */
    for (int j = 0; j < nn; j++) { abs[j] = 0; scat[j] = 0; emis[j] = 0; }
    if (grey != 0) {
      double gopac = dens * grey;
      for (int j = 0; j < nn; j++) {
        abs[j] = gopac * p->epsilon;
        scat[j] = gopac * (1 - p->epsilon);
        double nuc = or_center(nu, j);
        double ezeta = exp(1.0 * OR_H * nuc / OR_K / temp);
        double bb = 2.0 * nuc * nuc * nuc * OR_H / OR_C / OR_C / (ezeta - 1);
        emis[j] += bb * abs[j];
      }
    } else {
      if (p->use_electron_scattering) {
        double es = OR_THOMSON * n_elec;
        for (int j = 0; j < nn; j++) {
          scat[j] += es;
          abs[j] += 1e-20 * p->epsilon * es;
        }
      }
      /* GasState::free_free_opacity (GasState_opacities.cpp:150-182), gaunt
       * factor 1: abs += fac nu^-3 (1 - e^{-h nu/kT}), emis += abs_ff B_nu */
      if (p->use_free_free) {
        double fac = 0;
        for (int a = 0; a < at->n_atoms; a++) {
          double Z_eff_sq = 0;
          int i0 = at->ion_off[a], ni = at->ion_off[a + 1] - i0;
          for (int j = 0; j < ni; j++) Z_eff_sq += ifr[i0 + j] * j * j;
          double n_ion = mf[a] * dens / (zs->elem_A[a] * OR_M_P);
          fac += n_ion * Z_eff_sq;
        }
        fac *= 3.7e8 * pow(temp, -0.5) * n_elec;
        for (int j = 0; j < nn; j++) {
          double nuc = or_center(nu, j);
          double ezeta = exp(-1.0 * OR_H * nuc / OR_K / temp);
          double bb = 2.0 * nuc * nuc * nuc * OR_H / OR_C / OR_C / (1.0 / ezeta - 1);
          double op = fac / nuc / nuc / nuc * (1 - ezeta);
          abs[j] += op;
          emis[j] += op * bb;
        }
      }
      /* GasState::bound_free_opacity (:352-372) with
       * AtomicSpecies::bound_free_opacity_general, coolheat 0
       * (AtomicSpecies_opacities.cpp:58-140): abs += sum over levels of
       * sigma (n_l - n_c phi e^{(E_ion - E)/kT} n_e), floored at 0; emis +=
       * n_e sum 2 h nu^3 / c^2 sigma n_c phi e^{(E_ion - E)/kT} */
      if (p->use_bound_free) {
        double kt_ev = OR_K_EV * temp;
        double lam_t = sqrt(OR_H * OR_H / (2 * OR_PI * OR_M_E * OR_K * temp));
        for (int j = 0; j < nn; j++) { opac[j] = 0; aopac[j] = 0; }   /* aopac: emissivity */
        for (int a = 0; a < at->n_atoms; a++) {
          int l0 = at->lev_off[a], nl = at->lev_off[a + 1] - l0, i0 = at->ion_off[a];
          double n_dens = dens * mf[a] / (zs->elem_A[a] * OR_M_P);
          for (int j = 0; j < nn; j++) { atom_opac[j] = 0; emis2[j] = 0; }
          for (int l = 0; l < nl; l++) {
            int k = l0 + l;
            int kc = or_continuum_level(at, a, k);
            phifac[l] = 0;
            if (kc == -1) continue;
            double nc = n_dens * ln[kc];
            double gl_o_gc = (1.0 * at->lev_g[k]) / (1.0 * at->lev_g[kc]);
            phifac[l] = nc * gl_o_gc / 2. * lam_t * lam_t * lam_t;
          }
          for (int j = 0; j < nn; j++) {
            double nuc = or_center(nu, j);
            double E = OR_H * nuc * 624150975230.2815;
            double emis_fac = 2. * OR_H * nuc * nuc * nuc / OR_C / OR_C;
            for (int l = 0; l < nl; l++) {
              int k = l0 + l;
              if (or_continuum_level(at, a, k) == -1) continue;
              double chi_ion = at->ion_chi[i0 + at->lev_ion[k]];
              double Eion = chi_ion - at->lev_E[k];
              if (E < Eion) continue;
              double zeta_net = (Eion - E) / kt_ev;
              double ezeta_net = exp(zeta_net);
              double sigma = or_photo_cs(Eion, chi_ion, at->lev_E[k], at->lev_ion[k], E);
              double opac_fac = n_dens * ln[k] - phifac[l] * n_elec * ezeta_net;
              if (opac_fac < 0) opac_fac = 0.;   /* kill maser */
              atom_opac[j] += sigma * opac_fac;
              if (at->lev_E[k] == 0 && p->no_ground_recomb) continue;
              emis2[j] += emis_fac * sigma * phifac[l] * ezeta_net;
            }
          }
          for (int j = 0; j < nn; j++) { opac[j] += atom_opac[j]; aopac[j] += emis2[j]; }
        }
        for (int j = 0; j < nn; j++) {
          abs[j] += opac[j];
          emis[j] += aopac[j] * n_elec;
        }
      }
      /* GasState::bound_bound_opacity (GasState_opacities.cpp:83-91,379-408) over
       * AtomicSpecies::bound_bound_opacity (AtomicSpecies_opacities.cpp:224-282):
       * each level-resolved line of the atom in stored order, Doppler width
       * from line_beta_dop (GasState.cpp:218-220), Voigt profile over the bins
       * [locate(nu0 - 5 dnu), locate(nu0 + 5 dnu)) at their centres */
      if (p->use_bound_bound && at->n_blines > 0) {
        for (int j = 0; j < nn; j++) { opac[j] = 0; aopac[j] = 0; }
        for (int a = 0; a < at->n_atoms; a++) {
          double n_dens = dens * mf[a] / (zs->elem_A[a] * OR_M_P);
          double vd = sqrt(2 * OR_K * temp / OR_M_P / zs->elem_A[a]);
          if (p->line_velocity_width > 0) vd = p->line_velocity_width;
          double beta_dop = vd / OR_C;
          for (int j = 0; j < nn; j++) { atom_opac[j] = 0; emis2[j] = 0; }
          const int64_t* off = at->bline_bin_off + (size_t)a * (nn + 1);
          for (int64_t l = off[0]; l < off[nn]; l++) {
            int ll = at->bline_l[l], lu = at->bline_u[l];
            double nl = ln[ll], nup = ln[lu];
            double gl = 1.0 * at->lev_g[ll], gu = 1.0 * at->lev_g[lu];
            double nu_0 = at->bline_nu[l];
            double dnu = beta_dop * nu_0;
            double A_ul = at->bline_A[l];
            double a_voigt = A_ul / 4 / OR_PI / dnu;
            if (nl == 0) continue;
            double alpha_0 = nl * n_dens * gu / gl * A_ul / (8 * OR_PI) * OR_C * OR_C;
            alpha_0 = alpha_0 * (1 - nup * gl / (nl * gu));
            if (alpha_0 <= 0) continue;
            /* AtomicSpecies_opacities.cpp:259 */
            if (alpha_0 / (nu_0 * nu_0 * dnu) < p->minimum_extinction) continue;
            int inu1 = or_locate_within_bounds(nu, nu_0 - dnu * 5);
            int inu2 = or_locate_within_bounds(nu, nu_0 + dnu * 5);
            double line_j = A_ul * nup * n_dens * OR_H / (4.0 * OR_PI);
            for (int j = inu1; j < inu2; ++j) {
              double nuc = or_center(nu, j);
              double x = (nu_0 - nuc) / dnu;
              double phi = or_voigt_profile(x, a_voigt) / dnu;
              atom_opac[j] += alpha_0 / nuc / nuc * phi;
              emis2[j] += line_j * nuc * phi;
            }
          }
          for (int j = 0; j < nn; j++) { opac[j] += atom_opac[j]; aopac[j] += emis2[j]; }
        }
        for (int j = 0; j < nn; j++) {
          abs[j] += opac[j];
          emis[j] += aopac[j];
        }
      }
      /* GasState::line_expansion_opacity (GasState_opacities.cpp:424-454) with
       * AtomicSpecies::line_expansion_opacity (AtomicSpecies_opacities.cpp:329-375) */
      if (p->use_line_expansion && at->n_blines > 0) {
        for (int j = 0; j < nn; j++) { opac[j] = 0; aopac[j] = 0; }
        for (int a = 0; a < at->n_atoms; a++) {
          double this_eps = at->atom_zero_eps && at->atom_zero_eps[a] ? 0 : p->epsilon;
          double n_dens = dens * mf[a] / (zs->elem_A[a] * OR_M_P);
          for (int j = 0; j < nn; j++) atom_opac[j] = 0;
          const int64_t* off = at->bline_bin_off + (size_t)a * (nn + 1);
          for (int64_t l = off[0]; l < off[nn]; ++l) {
            int ll = at->bline_l[l], lu = at->bline_u[l];
            double nl = ln[ll], nup = ln[lu];
            double gl = 1.0 * at->lev_g[ll], gu = 1.0 * at->lev_g[lu];
            if (nl < DBL_MIN) continue;
            double nu_0 = at->bline_nu[l];
            double lam = OR_C / nu_0;
            double tau = nl * n_dens * OR_SIGMA_TOT * at->bline_f[l] * time * lam;
            tau = tau * (1 - nup * gl / (nl * gu));
            if (nup * gl > nl * gu) tau = 0;   /* laser regime */
            double etau = exp(-tau);
            if (!isnan(etau)) atom_opac[at->bline_bin[l]] += (1 - etau);
          }
          for (int j = 0; j < nn; j++)
            atom_opac[j] = atom_opac[j] * or_center(nu, j) / or_delta(nu, j) / OR_C / time;
          for (int j = 0; j < nn; j++) {
            opac[j] += atom_opac[j];
            aopac[j] += atom_opac[j] * this_eps;
          }
        }
        for (int j = 0; j < nn; j++) {
          abs[j] += aopac[j];
          scat[j] += opac[j] - aopac[j];
          double nuc = or_center(nu, j);
          double ezeta = exp(1.0 * OR_H * nuc / OR_K / temp);
          double bb = 2.0 * nuc * nuc * nuc * OR_H / OR_C / OR_C / (ezeta - 1);
          emis[j] += bb * aopac[j];
        }
      }
      if (p->use_fuzz_expansion) {
        for (int j = 0; j < nn; j++) { opac[j] = 0; aopac[j] = 0; }
        for (int a = 0; a < at->n_atoms; a++) {
          double this_eps = at->atom_zero_eps && at->atom_zero_eps[a] ? 0 : p->epsilon;
          double n_dens = dens * mf[a] / (zs->elem_A[a] * OR_M_P);
          const double* ifa = ifr + at->ion_off[a];
          const double* ipa = ip + at->ion_off[a];
          for (int j = 0; j < nn; j++) atom_opac[j] = 0;
          const int64_t* off = at->line_bin_off + (size_t)a * (nn + 1);
          for (int64_t l = off[0]; l < off[nn]; ++l) {
            double lnu = at->line_nu[l];
            double gf = at->line_gf[l];
            double El = at->line_El[l];
            int ion = at->line_ion[l];
            int bin = at->line_bin[l];
            double nion = n_dens * ifa[ion];
            double nl = nion * exp(-1.0 * El / OR_K_EV / temp) / ipa[ion];
            double lam = OR_C / lnu;
            double stim_cor = (1 - exp(-OR_H * lnu / OR_K / temp));
            double tau = OR_SIGMA_TOT * lam * nl * gf * stim_cor * time;
            double etau = exp(-tau);
            atom_opac[bin] += (1 - etau);
          }
          for (int j = 0; j < nn; j++)
            atom_opac[j] = atom_opac[j] * or_center(nu, j) / or_delta(nu, j) / OR_C / time;
          for (int j = 0; j < nn; j++) {
            opac[j] += atom_opac[j];
            aopac[j] += atom_opac[j] * this_eps;
          }
        }
        for (int j = 0; j < nn; j++) {
          abs[j] += aopac[j];
          scat[j] += opac[j] - aopac[j];
          double nuc = or_center(nu, j);
          double ezeta = exp(1.0 * OR_H * nuc / OR_K / temp);
          double bb = 2.0 * nuc * nuc * nuc * OR_H / OR_C / OR_C / (ezeta - 1);
          emis[j] += bb * aopac[j];
        }
      }
    }

    /* emissivity CDF, L_thermal and clamps, transport_opacity.cpp:152-179 */
    double max_ext = p->maximum_opacity * dens;
    double L_th = 0;
    if (nn == 1) {
      double bb_int = OR_SB * pow(zs->T_gas[i], 4) / OR_PI;
      L_th += 4 * OR_PI * abs[0] * bb_int;
      cdf[0] = 1;
    } else {
      for (int j = 0; j < nn; j++) {
        double ednu = emis[j] * or_delta(nu, j);
        cdf[j] = j == 0 ? ednu : cdf[j - 1] + ednu;
        L_th += 4 * OR_PI * ednu;
        if (scat[j] > max_ext) scat[j] = max_ext;
        if (abs[j] > max_ext) abs[j] = max_ext;
      }
    }
    or_cdf_normalize(cdf, nn);
    out->L_thermal[i] = L_th;

    /* Planck / Rosseland means */
    if (nn == 1) {
      out->planck_mean[i] = abs[0];
      out->rosseland_mean[i] = abs[0] + scat[0];
    } else {
      double mean = 0, nrm = 0, rmean = 0, rnrm = 0;
      for (int j = 0; j < nn; j++) {
        double nuc = or_center(nu, j);
        double zeta = OR_H * nuc / OR_K / temp;
        double Bnu = 2.0 * nuc * nuc * nuc * OR_H / OR_C / OR_C / (exp(zeta) - 1);
        if (isnan(Bnu)) Bnu = 0;
        double W = Bnu * or_delta(nu, j);
        mean += W * (abs[j]);
        nrm += W;
        double ezeta = exp(zeta);
        double dBdT = pow(nuc, 4) * ezeta / (ezeta - 1) / (ezeta - 1);
        if (isnan(dBdT)) dBdT = 0;
        double Wr = dBdT * or_delta(nu, j);
        rmean += Wr / (abs[j] + scat[j]);
        rnrm += Wr;
      }
      out->planck_mean[i] = mean / nrm;
      out->rosseland_mean[i] = rnrm / rmean;
    }

    /* gamma-ray opacities, transport_opacity.cpp:190-203 (undecayed X) */
    double compton = 0, photoion = 0;
    for (int k = 0; k < ne_; k++) {
      double d = zs->X[(size_t)i * ne_ + k] * dens;
      double ndens = d / (OR_M_P * zs->elem_A[k]);
      compton += ndens * OR_THOMSON * zs->elem_Z[k];
      double photo = pow(OR_ALPHA_FS, 4.0) * 4.0 * sqrt(2.0);
      photo *= pow(1.0 * zs->elem_Z[k], 5.0);
      photo *= pow(OR_M_E_MEV, 3.5);
      photoion += ndens * 2.0 * OR_THOMSON * photo;
    }
    out->compton_opac[i] = compton;
    out->photoion_opac[i] = photoion;
  }

  /* Fleck factor, transport_opacity.cpp:231-252 */
  for (int i = 0; i < nz; i++) {
    if (p->radiative_eq) out->eps_imc[i] = 1.;
    else {
      double f_imc = 0.;
      if (p->fleck_alpha != 0) {
        double beta = 4.0 * OR_A_RAD * pow(zs->T_gas[i], 4) / (zs->e_gas[i] * zs->rho[i]);
        double tfac = OR_C * out->planck_mean[i] * p->dt;
        f_imc = p->fleck_alpha * beta * tfac;
      }
      out->eps_imc[i] = 1.0 / (1.0 + f_imc);
    }
  }
  free(X_now); free(mf); free(ip); free(ifr); free(ln);
  free(emis); free(opac); free(aopac); free(atom_opac); free(emis2); free(phifac);
  return worst;
}

/* transport::set_opacity zone loop, src/transport/transport_opacity.cpp:83-204 */
int or_opacity_build(const sgpu_locate* nu, const sgpu_zone_state* zs, const sgpu_atoms* at,
                     const sgpu_opacity_params* p, sgpu_opacity_out* out)
{
  return opacity_build(nu, zs, at, p, NULL, out);
}

/* the same with GasState::solve_state(J_nu) under use_nlte_ (GasState.cpp:208-263,
 * transport_opacity.cpp:133-136): atoms flagged in atom_nlte are solved by
 * solve_nlte at every n_e evaluation */
int or_opacity_build_nlte(const sgpu_locate* nu, const sgpu_zone_state* zs, const sgpu_atoms* at,
                          const sgpu_opacity_params* p, const double* J_nu, const int32_t* atom_nlte,
                          int use_collisions, sgpu_opacity_out* out)
{
  nlte_in nl = {J_nu, atom_nlte, p->line_velocity_width, use_collisions};
  return opacity_build(nu, zs, at, p, &nl, out);
}
