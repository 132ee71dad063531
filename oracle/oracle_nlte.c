/*
 * oracle_nlte.c -- TEST INFRASTRUCTURE: FP64 restatement of the NLTE
 * statistical-equilibrium solve of AtomicSpecies::solve_nlte
 * (src/opacity/AtomicSpecies_state.cpp:131-196).  The reference factors with
 * gsl_linalg_LU_decomp / gsl_linalg_LU_solve (GSL, version not pinned by the
 * reference's makefiles): Gaussian elimination with row partial pivoting, the
 * pivot being the entry of largest magnitude in the column (first one on
 * ties), followed by forward and back substitution.  That textbook algorithm
 * is restated here and checked against LAPACK getrf/getrs through numpy.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

int or_lu_decomp(int n, double* A, int* perm)
{
  int singular = 0;
  for (int i = 0; i < n; i++) perm[i] = i;
  for (int j = 0; j < n; j++) {
    int ip = j;
    double amax = fabs(A[(size_t)j * n + j]);
    for (int i = j + 1; i < n; i++) {
      double v = fabs(A[(size_t)i * n + j]);
      if (v > amax) { amax = v; ip = i; }
    }
    if (ip != j) {
      for (int k = 0; k < n; k++) {
        double t = A[(size_t)j * n + k];
        A[(size_t)j * n + k] = A[(size_t)ip * n + k];
        A[(size_t)ip * n + k] = t;
      }
      int t = perm[j]; perm[j] = perm[ip]; perm[ip] = t;
    }
    double ajj = A[(size_t)j * n + j];
    if (ajj == 0) { singular = 1; continue; }
    for (int i = j + 1; i < n; i++) {
      double lij = A[(size_t)i * n + j] / ajj;
      A[(size_t)i * n + j] = lij;
      for (int k = j + 1; k < n; k++) A[(size_t)i * n + k] -= lij * A[(size_t)j * n + k];
    }
  }
  return singular;
}

void or_lu_solve(int n, const double* LU, const int* perm, const double* rhs, double* x)
{
  for (int i = 0; i < n; i++) x[i] = rhs[perm[i]];
  for (int i = 0; i < n; i++) {
    double s = x[i];
    for (int k = 0; k < i; k++) s -= LU[(size_t)i * n + k] * x[k];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    for (int k = i + 1; k < n; k++) s -= LU[(size_t)i * n + k] * x[k];
    x[i] = s / LU[(size_t)i * n + i];
  }
}

/* Matrix assembly of solve_nlte (:146-164): M_ii = -sum_j R_ij,
 * M_ij = R_ji (i != j), last row = lte, rhs = e_{n-1}; then LU solve and
 * n_i = max(b_i lte_i, 1e-30) (:179-185).  Returns 1 when singular. */
int or_nlte_solve(int n, const double* R, const double* lte, double* b, double* pop)
{
  double* M = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* rhs = (double*)calloc((size_t)n, sizeof(double));
  int* perm = (int*)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    double Rout = 0.0;
    for (int j = 0; j < n; ++j) Rout += R[(size_t)i * n + j];
    M[(size_t)i * n + i] = -1 * Rout;
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (i != j) M[(size_t)i * n + j] = R[(size_t)j * n + i];
  for (int i = 0; i < n; ++i) M[(size_t)(n - 1) * n + i] = lte[i];
  rhs[n - 1] = 1.0;
  int sing = or_lu_decomp(n, M, perm);
  or_lu_solve(n, M, perm, rhs, b);
  if (pop)
    for (int i = 0; i < n; ++i) {
      double v = b[i] * lte[i];
      pop[i] = v < 1e-30 ? 1e-30 : v;
    }
  free(M); free(rhs); free(perm);
  return sing;
}

/* continuum level of level k (global index) of atom a: the ground (first)
 * level of the next ion stage, or -1 (AtomicData.cpp:598-602) */
int or_continuum_level(const sgpu_atoms* at, int a, int k)
{
  for (int j = at->lev_off[a]; j < at->lev_off[a + 1]; j++)
    if (at->lev_ion[j] == at->lev_ion[k] + 1) return j;
  return -1;
}

/* AtomicSpecies::set_rates, src/opacity/AtomicSpecies_state.cpp:294-432, for
 * atom a at (ne, T).  The atom's lines are those whose lower level is one of
 * its levels, taken in stored order.  line_J / Pic / Rci / lev_lte are
 * indexed by global line and level; rates (n x n, local level indices, row =
 * from-level) and lte (n) are written.  Returns n. */
int or_nlte_set_rates(const sgpu_atoms* at, int a, double ne, double T, const double* line_J,
                      const double* Pic, const double* Rci, const double* lev_lte,
                      int use_collisions, int no_ground_recomb, double* rates, double* lte)
{
  const int l0 = at->lev_off[a], n = at->lev_off[a + 1] - l0;
  for (int i = 0; i < n * n; i++) rates[i] = 0;
  /* radiative bound-bound (:303-321) */
  for (int64_t l = 0; l < at->n_blines; l++) {
    int ll = at->bline_l[l] - l0, lu = at->bline_u[l] - l0;
    if (ll < 0 || ll >= n) continue;
    double R_ul = at->bline_Bul[l] * line_J[l] + at->bline_A[l];
    double R_lu = at->bline_Blu[l] * line_J[l];
    if (at->bline_nu[l] == 0) { R_ul = 0; R_lu = 0; }
    rates[ll * n + lu] += R_lu;
    rates[lu * n + ll] += R_ul;
  }
  /* collisional bound-bound, van Regemorter (:326-380) */
  for (int64_t l = 0; l < at->n_blines; l++) {
    int ll = at->bline_l[l] - l0, lu = at->bline_u[l] - l0;
    if (ll < 0 || ll >= n) continue;
    int gl = at->lev_g[ll + l0], gu = at->lev_g[lu + l0];
    double El = at->lev_E[ll + l0], Eu = at->lev_E[lu + l0];
    double f_lu = at->bline_f[l];
    double dE = (Eu - El) * OR_EV_TO_ERGS;
    double zeta = dE / OR_K / T;
    double ezeta = exp(zeta);
    double eff_f = f_lu < 1.e-3 ? 1.e-3 : f_lu;
    if (use_collisions) {
      double C_up = 3.9 * pow(zeta, -1.) * pow(T, -1.5) / ezeta * ne * eff_f;
      if (zeta > 700) C_up = 0;
      double C_down = 3.9 * pow(zeta, -1.) * pow(T, -1.5) * ne * eff_f * gl / gu;
      rates[ll * n + lu] += C_up;
      rates[lu * n + ll] += C_down;
    }
  }
  /* bound-free (:386-423) */
  for (int i = 0; i < n; ++i) {
    int k = i + l0;
    int kc = or_continuum_level(at, a, k);
    if (kc == -1) continue;
    int ic = kc - l0;
    int istage = at->lev_ion[k];
    double chi = at->ion_chi[at->ion_off[a] + istage] - at->lev_E[k];
    double zeta = chi / OR_K_EV / T;
    if (use_collisions) {
      double C_ion = 2.7 / zeta / zeta * pow(T, -1.5) * exp(-zeta) * ne;
      rates[i * n + ic] += C_ion;
      int gi = at->lev_g[k], gc = at->lev_g[kc];
      double C_rec = 5.59080e-16 / zeta / zeta * pow(T, -3) * gi / gc * ne * ne;
      rates[ic * n + i] += C_rec;
    }
    double rci = Rci[k];
    if (no_ground_recomb && at->lev_E[k] == 0) rci = 0.;
    rates[ic * n + i] += rci * ne;
    rates[i * n + ic] += Pic[k];
  }
  /* times the LTE population of the source level (:429-431) */
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) rates[i * n + j] *= lev_lte[i + l0];
    lte[i] = lev_lte[i + l0];
  }
  return n;
}

/* end of AtomicSpecies::solve_nlte (:178-193): level populations and ion
 * fractions of atom a from departure coefficients b (local) */
void or_nlte_populations(const sgpu_atoms* at, int a, const double* b, const double* lte,
                         double* lev_n, double* ion_frac)
{
  const int l0 = at->lev_off[a], n = at->lev_off[a + 1] - l0;
  const int i0 = at->ion_off[a], ni = at->ion_off[a + 1] - i0;
  for (int i = 0; i < ni; i++) ion_frac[i] = 0;
  for (int i = 0; i < n; i++) {
    double v = b[i] * lte[i];
    if (v < 1e-30) v = 1e-30;
    lev_n[i] = v;
  }
  for (int i = 0; i < n; i++) ion_frac[at->lev_ion[i + l0]] += lev_n[i];
}

/* Voigt profile approximation of VoigtProfile::getProfile
 * (src/transport/VoigtProfile.cpp:37-55), with its own constants */
double or_voigt_profile(double x, double a)
{
  double sqrt_pi = 1.77245385091;
  double pi = 3.14159265359;
  double xsq = x * x;
  double c = (xsq - 0.855) / (xsq + 3.42);
  double q;
  if (c < 0) q = 0;
  else {
    double pic = 5.674 * c * c * c * c - 9.207 * c * c * c + 4.421 * c * c + 0.1117 * c;
    q = (1 + 21 / xsq) * a / pi / (xsq + 1) * pic;
  }
  double H = q * sqrt_pi + exp(-xsq);
  return H / sqrt_pi;
}

/* the hydrogenic photoionisation cross-section table of a level
 * (AtomicData.cpp:713-746: xy_array::init(E_ion, 10 E_ion, 9 E_ion / 100),
 * 100 points sigma = 6.3e-18 s_fac (E / E_ion)^-3, a 101st point left at 0
 * when init makes 101), read by xy_array::value_at_with_zero_edges
 * (xy_array.cpp:42-104: bisection locate, linear interpolation, 0 outside) */
double or_photo_cs(double E_ion, double chi_ion, double E_lev, int ion, double E_ev)
{
  const int npts = 100;
  const double fmax = 10;
  double E_max = E_ion * fmax;
  double dE = (E_max - E_ion) / npts;
  int n = (int)((E_max - E_ion) / dE) + 1;
  if (n <= 1) { n = 1; dE = E_max - E_ion; }
  if (dE < 0) dE = 0;
  double n_eff = pow(1 - (chi_ion - E_ion) / chi_ion, -0.5);
  double s_fac = n_eff / (ion + 1) / (ion + 1);
#define OR_PX(j) (E_ion + (j) * dE)
#define OR_PY(j) ((j) < npts ? 6.3e-18 * s_fac * pow(OR_PX(j) / E_ion, -3.0) : 0.0)
  if (E_ev < OR_PX(0)) return 0;
  if (E_ev > OR_PX(n - 1)) return 0;
  if (n == 1) return OR_PY(0);
  int ind;
  int bl = 0, bu = n - 1;
  if (E_ev >= OR_PX(bu)) ind = n;
  else if (E_ev <= OR_PX(bl)) ind = 0;
  else {
    while (bu - bl > 1) {
      int bm = (bu + bl) / 2;
      if (OR_PX(bm) <= E_ev) bl = bm;
      else bu = bm;
    }
    ind = bl;
  }
  int i1, i2;
  if (ind < n - 1) { i1 = ind; i2 = ind + 1; }
  else { i1 = n - 2; i2 = n - 1; ind = n - 1; }   /* E == x[n-1]: the last point */
  double slope = (OR_PY(i2) - OR_PY(i1)) / (OR_PX(i2) - OR_PX(i1));
  return OR_PY(ind) + slope * (E_ev - OR_PX(i1));
#undef OR_PX
#undef OR_PY
}

/* AtomicSpecies::calculate_radiative_rates (AtomicSpecies_state.cpp:202-287)
 * for atom a in a zone with mean intensity J_nu[nn] and gas temperature T,
 * Doppler width from GasState::solve_state (GasState.cpp:218-220).  Pic, Rci
 * (global level index) and line_J (global line index) receive this atom's
 * entries. */
void or_nlte_radiative_rates(const sgpu_atoms* at, int a, const sgpu_locate* nu, const double* J_nu,
                             double T, int elem_A, double line_velocity_width, double* Pic, double* Rci,
                             double* line_J)
{
  const int l0 = at->lev_off[a], n = at->lev_off[a + 1] - l0;
  const int i0 = at->ion_off[a];
  for (int j = 0; j < n; j++) { Pic[l0 + j] = 0; Rci[l0 + j] = 0; }
  const double fac1 = 2 / OR_C / OR_C;
  const int ng = nu->n;
  for (int i = 1; i < ng; ++i) {
    double nuc = or_center(nu, i);
    double E_ergs = OR_H * nuc;
    double E_ev = E_ergs * 624150975230.2815;
    double J = J_nu[i];
    double dnu = or_delta(nu, i);
    for (int j = 0; j < n; ++j) {
      int k = l0 + j;
      if (or_continuum_level(at, a, k) == -1) continue;
      double chi_ion = at->ion_chi[i0 + at->lev_ion[k]];
      double chi = chi_ion - at->lev_E[k];
      if (E_ev < chi) continue;
      double sigma = or_photo_cs(chi, chi_ion, at->lev_E[k], at->lev_ion[k], E_ev);
      double Jterm = sigma * J / E_ergs;
      Pic[k] += Jterm * dnu;
      Rci[k] += (sigma * fac1 * nuc * nuc + Jterm) * exp(-1.0 * (E_ev - chi) / OR_K_EV / T) * dnu;
    }
  }
  double lam_t = sqrt(OR_H * OR_H / (2 * OR_PI * OR_M_E * OR_K * T));
  double saha_fac = lam_t * lam_t * lam_t / 2.0;
  for (int j = 0; j < n; ++j) {
    int k = l0 + j;
    int kc = or_continuum_level(at, a, k);
    if (kc == -1) continue;
    double gl_o_gc = (1.0 * at->lev_g[k]) / (1.0 * at->lev_g[kc]);
    Pic[k] *= 4 * OR_PI;
    Rci[k] *= 4 * OR_PI * gl_o_gc * saha_fac;
  }
  double vd = sqrt(2 * OR_K * T / OR_M_P / elem_A);
  if (line_velocity_width > 0) vd = line_velocity_width;
  double beta = vd / OR_C;
  const double x_max = 5., dx = 0.05;
  for (int64_t l = 0; l < at->n_blines; l++) {
    int ll = at->bline_l[l] - l0;
    if (ll < 0 || ll >= n) continue;
    double nu0 = at->bline_nu[l];
    double sum = 0, J0 = 0;
    double nu_d = nu0 * beta;
    double gamma = at->bline_A[l];
    double a_voigt = gamma / 4 / OR_PI / nu_d;
    for (double x = -1 * x_max; x <= x_max; x += dx) {
      double phi = or_voigt_profile(x, a_voigt);
      double nx = nu0 + x * nu_d;
      double J1 = J_nu[or_locate_within_bounds(nu, nx)] * phi;   /* locate_array::value_at */
      sum += 0.5 * (J1 + J0) * dx;
      J0 = J1;
    }
    line_J[l] = sum;
  }
}
