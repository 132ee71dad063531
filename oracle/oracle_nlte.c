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
