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
static int continuum_level(const sgpu_atoms* at, int a, int k)
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
    int kc = continuum_level(at, a, k);
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
