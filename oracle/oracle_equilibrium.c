/* oracle_equilibrium.c -- CPU restatement (test infrastructure only) of the
 * radiative-equilibrium gas temperature of Sedona's transport step:
 * transport::solve_eq_temperature (src/transport/solve_equilibrium.cpp:66-94)
 * with the LTE balance rad_eq_function_LTE at solve_flag = 0 (:112-168) and
 * the Numerical Recipes Brent root finder temp_brent_method (:226-320). */
#include <math.h>
#include "oracle.h"

/* E_emitted - E_absorbed for zone c at temperature T, old opacities
 * (solve_flag 0): E_absorbed is the transport's e_abs tally */
static double rad_eq_lte(const sgpu_locate* nu, const double* abs, double e_abs, double T)
{
  double E_emitted = 0.;
  const int nn = nu->n;
  if (nn == 1) {
    E_emitted = 4.0 * OR_PI * abs[0] * OR_SB / OR_PI * pow(T, 4);
  } else {
    for (int i = 0; i < nn; i++) {
      double dnu = or_delta(nu, i);
      double nuc = or_center(nu, i);
      double B_nu = or_blackbody_nu(T, nuc);
      E_emitted += 4.0 * OR_PI * abs[i] * B_nu * dnu;
    }
  }
  return E_emitted - e_abs;
}

#define OR_SIGN(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a))

/* temp_brent_method (:226-320): tolerance 1e-2, EPS 3e-8, at most 100
 * iterations (then 0.0, as the reference returns); no bracketing check */
static double temp_brent(const sgpu_locate* nu, const double* abs, double e_abs, double tmin,
                         double tmax, int* n_eval)
{
  const double tol = 1.0e-2, EPS = 3.0e-8;
  double a = tmin, b = tmax, c = b, d = 0, e = 0, min1, min2;
  double fa = rad_eq_lte(nu, abs, e_abs, a);
  double fb = rad_eq_lte(nu, abs, e_abs, b);
  double fc, p, q, r, s, tol1, xm;
  int ne = 2;
  fc = fb;
  for (int iter = 1; iter <= 100; iter++) {
    if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { c = a; fc = fa; e = d = b - a; }
    if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    tol1 = 2.0 * EPS * fabs(b) + 0.5 * tol;
    xm = 0.5 * (c - b);
    if (fabs(xm) <= tol1 || fb == 0.0) { if (n_eval) *n_eval = ne; return b; }
    if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
      s = fb / fa;
      if (a == c) { p = 2.0 * xm * s; q = 1.0 - s; }
      else {
        q = fa / fc; r = fb / fc;
        p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
        q = (q - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (p > 0.0) q = -q;
      p = fabs(p);
      min1 = 3.0 * xm * q - fabs(tol1 * q);
      min2 = fabs(e * q);
      if (2.0 * p < (min1 < min2 ? min1 : min2)) { e = d; d = p / q; }
      else { d = xm; e = d; }
    } else { d = xm; e = d; }
    a = b; fa = fb;
    if (fabs(d) > tol1) b += d;
    else b += OR_SIGN(tol1, xm);
    fb = rad_eq_lte(nu, abs, e_abs, b);
    ne++;
  }
  if (n_eval) *n_eval = ne;
  return 0.0;
}

/* transport::solve_eq_temperature: every zone's T_gas from e_abs (or, with
 * set_Tgas_to_Trad, from e_rad: T = (e_rad / a)^(1/4)).  n_eval (may be
 * NULL) receives the rad_eq evaluations per zone. */
void or_solve_eq_temperature(const sgpu_locate* nu, int nz, const double* abs_opacity,
                             const double* e_abs, const double* e_rad, double tmin, double tmax,
                             int set_Tgas_to_Trad, double* T_gas, int* n_eval)
{
  for (int i = 0; i < nz; i++) {
    if (set_Tgas_to_Trad) { T_gas[i] = pow(e_rad[i] / OR_A_RAD, 0.25); continue; }
    T_gas[i] = temp_brent(nu, abs_opacity + (size_t)i * nu->n, e_abs[i], tmin, tmax,
                          n_eval ? n_eval + i : NULL);
  }
}
