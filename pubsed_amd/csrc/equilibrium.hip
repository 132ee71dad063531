// equilibrium.hip -- the radiative-equilibrium gas temperature of the
// transport step: transport::solve_eq_temperature
// (src/transport/solve_equilibrium.cpp:66-94) with the LTE balance at
// solve_flag = 0 (rad_eq_function_LTE, :112-168: Kirchhoff emission
// sum 4 pi kappa_abs B_nu(T) dnu against the e_abs tally) and the Numerical
// Recipes Brent root finder temp_brent_method (:226-320).
//
// One wave64 per zone: the Brent iteration is wave-uniform (every lane holds
// the same state; the butterfly reduction gives every lane the same sum), the
// emission sum of an evaluation is split over the lanes.
#include <hip/hip_runtime.h>
#include "sgpu_internal.h"
#include "sgpu_device.h"

namespace sgpu {

struct EqArgs {
  sgpu_locate nu;
  int nz;
  const double* abs;
  const double* e_abs;
  const double* e_rad;
  double tmin, tmax;
  int trad;
  double* T;
  int32_t* n_eval;
};

// E_emitted - E_absorbed of one zone at temperature T (old opacities)
__device__ double rad_eq(const EqArgs& A, const double* ab, double e_abs, double T, int lane)
{
  const int nn = A.nu.n;
  if (nn == 1) return 4.0 * PI * ab[0] * SB / PI * pow(T, 4) - e_abs;
  double part = 0;
  for (int i = lane; i < nn; i += WAVE) {
    const double dnu = loc_delta(A.nu, i);
    const double nu = loc_center(A.nu, i);
    const double zeta = H_PLANCK * nu / K_BOLTZ / T;   // transport::blackbody_nu
    const double B_nu = 2.0 * nu * nu * nu * H_PLANCK / C_LIGHT / C_LIGHT / (exp(zeta) - 1);
    part += 4.0 * PI * ab[i] * B_nu * dnu;
  }
  return wave_sum(part) - e_abs;
}

__global__ void __launch_bounds__(WAVE) eq_temperature_kernel(EqArgs A)
{
  const int z = blockIdx.x;
  const int lane = threadIdx.x;
  if (A.trad) {
    if (lane == 0) A.T[z] = pow(A.e_rad[z] / A_RAD, 0.25);
    return;
  }
  const double* ab = A.abs + (size_t)z * A.nu.n;
  const double eab = A.e_abs[z];
  const double tol = 1.0e-2, EPS = 3.0e-8;
  double a = A.tmin, b = A.tmax, c = b, d = 0, e = 0, min1, min2;
  double fa = rad_eq(A, ab, eab, a, lane);
  double fb = rad_eq(A, ab, eab, b, lane);
  double fc, p, q, r, s, tol1, xm;
  int ne = 2;
  double root = 0.0;
  fc = fb;
  for (int iter = 1; iter <= 100; iter++) {
    if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { c = a; fc = fa; e = d = b - a; }
    if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    tol1 = 2.0 * EPS * fabs(b) + 0.5 * tol;
    xm = 0.5 * (c - b);
    if (fabs(xm) <= tol1 || fb == 0.0) { root = b; break; }
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
    else b += (xm >= 0.0 ? fabs(tol1) : -fabs(tol1));
    fb = rad_eq(A, ab, eab, b, lane);
    ne++;
  }
  // 100 iterations without convergence leave root = 0.0, as the reference returns
  if (lane == 0) {
    A.T[z] = root;
    if (A.n_eval) A.n_eval[z] = ne;
  }
}

}  // namespace sgpu

using namespace sgpu;

extern "C" int sgpu_solve_eq_temperature(sgpu_ctx* ctx, const sgpu_locate* nu_grid, int32_t n_zones,
                                         const double* abs_opacity, const double* e_abs, const double* e_rad,
                                         double temp_min, double temp_max, int32_t set_Tgas_to_Trad,
                                         double* T_gas, int32_t* n_eval)
{
  SGPU_REQUIRE(ctx && nu_grid, "null argument");
  SGPU_REQUIRE(n_zones >= 0, "negative zone count");
  if (n_zones == 0) return SGPU_OK;
  SGPU_REQUIRE(T_gas, "T_gas missing");
  if (set_Tgas_to_Trad) {
    SGPU_REQUIRE(e_rad, "set_Tgas_to_Trad needs e_rad");
  } else {
    SGPU_REQUIRE(abs_opacity && e_abs, "abs_opacity and e_abs required");
    SGPU_REQUIRE(nu_grid->n >= 1 && nu_grid->x, "bad frequency grid");
    SGPU_REQUIRE(temp_min > 0 && temp_max > temp_min, "bad temperature range [%g, %g]", temp_min, temp_max);
  }
  EqArgs A{*nu_grid, n_zones, abs_opacity, e_abs, e_rad, temp_min, temp_max, set_Tgas_to_Trad, T_gas, n_eval};
  timer_begin(ctx, "eq_temperature");
  hipLaunchKernelGGL(eq_temperature_kernel, dim3(n_zones), dim3(WAVE), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "eq_temperature");
  return SGPU_OK;
}
