// sgpu_physics.h -- device restatements of the reference's small physics
// formulae used by the opacity build and the packet sources.
#pragma once
#include "sgpu_device.h"

namespace sgpu {

// radioactive::decay_composition, src/transport/radioactive.cpp:58-113;
// X has n_el entries (mass fractions), modified in place.
__device__ inline void decay_composition(int n_el, const int32_t* Z, const int32_t* A,
                                         double* X, double t)
{
  double X_ni = 0, X_co = 0;
  for (int i = 0; i < n_el; i++) {
    if (Z[i] == 28 && A[i] == 56) X_ni = X[i];
    if (Z[i] == 27 && A[i] == 56) X_co = X[i];
  }
  double ni_f = exp(-t / TAU_56NI);
  double co_f = TAU_56CO / (TAU_56NI - TAU_56CO) * (exp(-t / TAU_56NI) - exp(-t / TAU_56CO));
  double fe_f = 1 - ni_f - co_f;
  double eco = exp(-t / TAU_56CO);
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
  double cr_f = exp(-t / TAU_48CR);
  double vn_f = TAU_48V / (TAU_48CR - TAU_48V) * (exp(-t / TAU_48CR) - exp(-t / TAU_48V));
  double ti_f = 1 - cr_f - vn_f;
  double evn = exp(-t / TAU_48V);
  for (int i = 0; i < n_el; i++) {
    if (Z[i] == 24 && A[i] == 48) X[i] = X_cr * cr_f;
    if (Z[i] == 23 && A[i] == 48) X[i] = X_cr * vn_f + X_vn * evn;
    if (Z[i] == 22 && A[i] == 48) X[i] += X_cr * ti_f + X_vn * (1 - evn);
  }
}

// radioactive::rprocess_heating_rate, radioactive.cpp:116-142 (Lippuner &
// Roberts 2015 heating fit, Barnes et al. 2016 thermalisation); all of it is
// deposited locally, so the gamma-ray fraction is zero.
__device__ inline double rprocess_heating_rate(double t, double* gfrac)
{
  double td = t / 3600.0 / 24.0;
  double eps = 8.4939E+09 * pow(td, -1.0 * 1.3642E+00) + 8.3425E+09 * exp(-1.0 * td / 3.6280E+00) +
               8.8616E+08 * exp(-1.0 * td / 1.0847E+01);
  double x = 2 * 0.17 * pow(td, 0.74);
  double f = 0.36 * (exp(-0.56 * td) + log(1 + x) / x);
  *gfrac = 0;
  return eps * f;
}

// radioactive::decay_energy_rate for the 56Ni, 56Co and 48Cr chains,
// radioactive.cpp:189-242 (elements with Z >= 58 route through the
// r-process fit instead, as radioactive::decay does)
__device__ inline double decay_energy_rate(int Z, int A, double t, double* gfrac)
{
  double total = 0, gtotal = 0;
  if (Z == 28 && A == 56) {
    double e_ni = exp(-t / TAU_56NI);
    double e_co = exp(-t / TAU_56CO);
    double ni56 = (e_ni / TAU_56NI);
    double co56 = 1.0 / (TAU_56NI - TAU_56CO) * (e_ni - e_co);
    double ni_E = ni56 * (E_56NI * MEV_TO_ERGS);
    double co_E = co56 * (E_56CO * MEV_TO_ERGS);
    gtotal = ni_E + (1 - CO_POS) * co_E;
    total = ni_E + co_E;
  }
  if (Z == 27 && A == 56) {
    double e_co = exp(-t / TAU_56CO);
    double co56 = 1.0 / (TAU_56CO) * e_co;
    double co_E = co56 * (E_56CO * MEV_TO_ERGS);
    gtotal = (1 - CO_POS) * co_E;
    total = co_E;
  }
  if (Z == 24 && A == 48) {
    double e_cr = exp(-t / TAU_48CR);
    double e_vn = exp(-t / TAU_48V);
    double cr48 = (e_cr / TAU_48CR);
    double vn48 = 1.0 / (TAU_48CR - TAU_48V) * (e_cr - e_vn);
    double cr_E = cr48 * (E_48CR * MEV_TO_ERGS);
    double vn_E = vn48 * (E_48V * MEV_TO_ERGS);
    gtotal += cr_E + vn_E;
    total += cr_E + vn_E;
  }
  *gfrac = gtotal / total;
  if (total == 0) *gfrac = 0;
  return total;
}

// radioactive::decay, radioactive.cpp:151-184
__device__ inline double decay(int n_el, const int32_t* Z, const int32_t* A, const double* X,
                               double t, double* gfrac, int force_rproc)
{
  int rproc = 0;
  for (int k = 0; k < n_el; k++)
    if (((Z[k] > 57) && (A[k] > 0)) || force_rproc) rproc = 1;
  if (rproc) return rprocess_heating_rate(t, gfrac);
  double total = 0, gtotal = 0, gf;
  for (int k = 0; k < n_el; k++) {
    double val = decay_energy_rate(Z[k], A[k], t, &gf);
    val = val * X[k] / (A[k] * M_PROTON);
    total += val;
    gtotal += val * gf;
  }
  *gfrac = gtotal / total;
  if (total == 0) *gfrac = 0;
  return total;
}

// transport::klein_nishina, transport_opacity.cpp:306-316 (x in MeV)
template <class R>
__device__ __forceinline__ R klein_nishina(R x)
{
  x = x / (R)M_E_MEV;
  R logfac = log(1 + 2 * x);
  R term1 = (1 + x) / x / x / x * (2 * x * (1 + x) / (1 + 2 * x) - logfac);
  R term2 = (R)1.0 / (R)2.0 / x * logfac;
  R term3 = (R)-1.0 * (1 + 3 * x) / (1 + 2 * x) / (1 + 2 * x);
  return (R).75 * (term1 + term2 + term3);
}

// ground (first) level of the next ion stage of global level k of atom a, or -1
__device__ inline int continuum_level(const sgpu_atoms& at, int a, int k)
{
  const int ion = at.lev_ion[k] + 1;
  for (int j = at.lev_off[a]; j < at.lev_off[a + 1]; j++)
    if (at.lev_ion[j] == ion) return j;
  return -1;
}

// The hydrogenic photoionisation cross-section table of a level
// (AtomicData.cpp:713-746: nodes E_ion + j dE, dE = 9 E_ion / 100, y_j =
// 6.3e-18 s_fac (x_j / E_ion)^-3 for the first 100 nodes, 0 for a 101st),
// evaluated as xy_array::value_at_with_zero_edges (xy_array.cpp:42-104) on
// nodes computed where needed
__device__ inline double photo_cs(double E_ion, double chi_ion, int ion, double E_ev)
{
  const int npts = 100;
  const double fmax = 10;
  const double E_max = E_ion * fmax;
  double dE = (E_max - E_ion) / npts;
  int n = (int)((E_max - E_ion) / dE) + 1;
  if (n <= 1) { n = 1; dE = E_max - E_ion; }
  if (dE < 0) dE = 0;
  const double n_eff = pow(1 - (chi_ion - E_ion) / chi_ion, -0.5);
  const double s_fac = n_eff / (ion + 1) / (ion + 1);
  auto px = [&](int j) { return E_ion + j * dE; };
  auto py = [&](int j) { return j < npts ? 6.3e-18 * s_fac * pow(px(j) / E_ion, -3.0) : 0.0; };
  if (E_ev < px(0)) return 0;
  if (E_ev > px(n - 1)) return 0;
  if (n == 1) return py(0);
  int ind;
  int bl = 0, bu = n - 1;
  if (E_ev >= px(bu)) ind = n;
  else if (E_ev <= px(bl)) ind = 0;
  else {
    while (bu - bl > 1) {
      const int bm = (bu + bl) / 2;
      if (px(bm) <= E_ev) bl = bm;
      else bu = bm;
    }
    ind = bl;
  }
  int i1, i2;
  if (ind < n - 1) { i1 = ind; i2 = ind + 1; }
  else { i1 = n - 2; i2 = n - 1; ind = n - 1; }   // E == x[n-1]: the last node
  const double slope = (py(i2) - py(i1)) / (px(i2) - px(i1));
  return py(ind) + slope * (E_ev - px(i1));
}

// VoigtProfile::getProfile (src/transport/VoigtProfile.cpp:37-55), with its
// own constants
__device__ inline double voigt_profile(double x, double a)
{
  const double sqrt_pi = 1.77245385091;
  const double pi = 3.14159265359;
  const double xsq = x * x;
  const double c = (xsq - 0.855) / (xsq + 3.42);
  double q;
  if (c < 0) q = 0;
  else {
    const double pic = 5.674 * c * c * c * c - 9.207 * c * c * c + 4.421 * c * c + 0.1117 * c;
    q = (1 + 21 / xsq) * a / pi / (xsq + 1) * pic;
  }
  const double H = q * sqrt_pi + exp(-xsq);
  return H / sqrt_pi;
}

}  // namespace sgpu
