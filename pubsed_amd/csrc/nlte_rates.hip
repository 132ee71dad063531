// nlte_rates.hip -- rate-matrix assembly and level populations of the NLTE
// solve, batched over (zone, atom) systems.
//
//   set_rates_kernel    AtomicSpecies::set_rates (src/opacity/AtomicSpecies_state.cpp:294-432):
//                       one workgroup per system; matrix row r belongs to
//                       thread r % blockDim, which adds every term of that row
//                       in the reference's order (radiative lines in stored
//                       order, then van Regemorter collisions, then bound-free
//                       levels in order), so each entry is summed by exactly
//                       one thread, deterministically and without atomics.
//                       The matrices stream to HBM in the CSR layout that
//                       sgpu_nlte_batched reads.
//   populations_kernel  the end of AtomicSpecies::solve_nlte (:178-193).
//   radiative_rates_kernel  AtomicSpecies::calculate_radiative_rates (:202-287):
//                       photoionisation / recombination integrals over the
//                       MC J_nu per level and line J-bar by the Voigt profile.
#include <hip/hip_runtime.h>
#include "sgpu_internal.h"
#include "sgpu_device.h"
#include "sgpu_physics.h"

namespace sgpu {

constexpr int RATE_THREADS = 256;

struct RateArgs {
  sgpu_atoms at;
  const int32_t* sys_atom;
  const double* ne;
  const double* T;
  const double* line_J;    // [nsys][n_blines]
  const double* Pic;       // [nsys][n_levels]
  const double* Rci;       // [nsys][n_levels]
  const double* lev_lte;   // [nsys][n_levels]
  int use_collisions, no_ground_recomb;
  const int32_t* n;
  const int64_t* rate_off;
  const int64_t* vec_off;
  double* rates;
  double* lte;
  int32_t* status;         // set to 1 when n[s] differs from the atom's level count
};

__global__ void __launch_bounds__(RATE_THREADS) set_rates_kernel(RateArgs A)
{
  const int s = blockIdx.x;
  const int t = threadIdx.x, NT = blockDim.x;
  const sgpu_atoms& at = A.at;
  const int a = A.sys_atom[s];
  const int l0 = at.lev_off[a], n = at.lev_off[a + 1] - l0;
  if (n != A.n[s]) {
    if (t == 0) A.status[0] = 1;
    return;
  }
  const double ne = A.ne[s], T = A.T[s];
  double* R = A.rates + A.rate_off[s];
  const double* J = A.line_J + (size_t)s * at.n_blines;
  const double* Pic = A.Pic + (size_t)s * at.n_levels;
  const double* Rci = A.Rci + (size_t)s * at.n_levels;
  const double* lte = A.lev_lte + (size_t)s * at.n_levels;
  for (int r = t; r < n; r += NT)
    for (int j = 0; j < n; j++) R[(size_t)r * n + j] = 0;

  // radiative bound-bound (:303-321)
  for (int64_t l = 0; l < at.n_blines; l++) {
    const int ll = at.bline_l[l] - l0, lu = at.bline_u[l] - l0;
    if (ll < 0 || ll >= n) continue;
    const bool own_l = ll % NT == t, own_u = lu % NT == t;
    if (!own_l && !own_u) continue;
    double R_ul = at.bline_Bul[l] * J[l] + at.bline_A[l];
    double R_lu = at.bline_Blu[l] * J[l];
    if (at.bline_nu[l] == 0) { R_ul = 0; R_lu = 0; }
    if (own_l) R[(size_t)ll * n + lu] += R_lu;
    if (own_u) R[(size_t)lu * n + ll] += R_ul;
  }
  // collisional bound-bound, van Regemorter (:326-380)
  if (A.use_collisions) {
    for (int64_t l = 0; l < at.n_blines; l++) {
      const int ll = at.bline_l[l] - l0, lu = at.bline_u[l] - l0;
      if (ll < 0 || ll >= n) continue;
      const bool own_l = ll % NT == t, own_u = lu % NT == t;
      if (!own_l && !own_u) continue;
      const int gl = at.lev_g[ll + l0], gu = at.lev_g[lu + l0];
      const double dE = (at.lev_E[lu + l0] - at.lev_E[ll + l0]) * EV_TO_ERGS;
      const double zeta = dE / K_BOLTZ / T;
      const double ezeta = exp(zeta);
      const double f_lu = at.bline_f[l];
      const double eff_f = f_lu < 1.e-3 ? 1.e-3 : f_lu;
      if (own_l) {
        double C_up = 3.9 * pow(zeta, -1.) * pow(T, -1.5) / ezeta * ne * eff_f;
        if (zeta > 700) C_up = 0;
        R[(size_t)ll * n + lu] += C_up;
      }
      if (own_u) {
        const double C_down = 3.9 * pow(zeta, -1.) * pow(T, -1.5) * ne * eff_f * gl / gu;
        R[(size_t)lu * n + ll] += C_down;
      }
    }
  }
  // bound-free (:386-423)
  for (int i = 0; i < n; ++i) {
    const int k = i + l0;
    const bool own_i = i % NT == t;
    const int kc = continuum_level(at, a, k);
    if (kc == -1) continue;
    const int ic = kc - l0;
    const bool own_c = ic % NT == t;
    if (!own_i && !own_c) continue;
    const double chi = at.ion_chi[at.ion_off[a] + at.lev_ion[k]] - at.lev_E[k];
    const double zeta = chi / K_EV / T;
    if (A.use_collisions) {
      if (own_i) R[(size_t)i * n + ic] += 2.7 / zeta / zeta * pow(T, -1.5) * exp(-zeta) * ne;
      if (own_c) {
        const int gi = at.lev_g[k], gc = at.lev_g[kc];
        R[(size_t)ic * n + i] += 5.59080e-16 / zeta / zeta * pow(T, -3) * gi / gc * ne * ne;
      }
    }
    double rci = Rci[k];
    if (A.no_ground_recomb && at.lev_E[k] == 0) rci = 0.;
    if (own_c) R[(size_t)ic * n + i] += rci * ne;
    if (own_i) R[(size_t)i * n + ic] += Pic[k];
  }
  // times the LTE population of the source level (:429-431)
  double* lo = A.lte + A.vec_off[s];
  for (int r = t; r < n; r += NT) {
    const double w = lte[r + l0];
    for (int j = 0; j < n; j++) R[(size_t)r * n + j] *= w;
    lo[r] = w;
  }
}

struct PopArgs {
  sgpu_atoms at;
  const int32_t* sys_atom;
  const int64_t* vec_off;
  const double* b;
  const double* lte;
  double* lev_n;       // [nsys][n_levels]
  double* ion_frac;    // [nsys][n_ions]
};

__global__ void __launch_bounds__(RATE_THREADS) populations_kernel(PopArgs A)
{
  const int s = blockIdx.x;
  const sgpu_atoms& at = A.at;
  const int a = A.sys_atom[s];
  const int l0 = at.lev_off[a], n = at.lev_off[a + 1] - l0;
  const int i0 = at.ion_off[a], ni = at.ion_off[a + 1] - i0;
  const double* b = A.b + A.vec_off[s];
  const double* lte = A.lte + A.vec_off[s];
  double* ln = A.lev_n + (size_t)s * at.n_levels + l0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double v = b[i] * lte[i];
    ln[i] = v < 1e-30 ? 1e-30 : v;
  }
  __syncthreads();
  // one thread per ion, levels in order (the reference's summation order)
  for (int ion = threadIdx.x; ion < ni; ion += blockDim.x) {
    double acc = 0;
    for (int i = 0; i < n; i++)
      if (at.lev_ion[i + l0] == ion) acc += ln[i];
    A.ion_frac[(size_t)s * at.n_ions + i0 + ion] = acc;
  }
}

__global__ void zero_status_kernel(int32_t* st) { st[0] = 0; }

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

struct RadArgs {
  sgpu_atoms at;
  sgpu_locate nu;
  const int32_t* sys_atom;
  const int32_t* sys_zone;
  const double* T;
  const int32_t* elem_A;
  double line_velocity_width;
  const double* J_nu;     // [nz][n_nu]
  double* Pic;            // [nsys][n_levels]
  double* Rci;
  double* line_J;         // [nsys][n_blines]
};

// AtomicSpecies::calculate_radiative_rates (AtomicSpecies_state.cpp:202-287)
// for one (zone, atom) system per workgroup: a thread per level sums the
// bound-free integrand over the bins in order, a thread per line runs the
// 201-point trapezoid over the Voigt profile
__global__ void __launch_bounds__(RATE_THREADS) radiative_rates_kernel(RadArgs A)
{
  const int s = blockIdx.x;
  const sgpu_atoms& at = A.at;
  const int a = A.sys_atom[s];
  const int l0 = at.lev_off[a], n = at.lev_off[a + 1] - l0;
  const int i0 = at.ion_off[a];
  const double T = A.T[s];
  const int nn = A.nu.n;
  const double* J = A.J_nu + (size_t)A.sys_zone[s] * nn;
  double* Pic = A.Pic + (size_t)s * at.n_levels;
  double* Rci = A.Rci + (size_t)s * at.n_levels;
  const double fac1 = 2 / C_LIGHT / C_LIGHT;
  const double lam_t = sqrt(H_PLANCK * H_PLANCK / (2 * PI * M_ELECTRON * K_BOLTZ * T));
  const double saha_fac = lam_t * lam_t * lam_t / 2.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const int k = l0 + j;
    const int kc = continuum_level(at, a, k);
    double P = 0, Rr = 0;
    if (kc != -1) {
      const int ion = at.lev_ion[k];
      const double chi_ion = at.ion_chi[i0 + ion];
      const double chi = chi_ion - at.lev_E[k];
      for (int i = 1; i < nn; ++i) {
        const double nuc = loc_center(A.nu, i);
        const double E_ergs = H_PLANCK * nuc;
        const double E_ev = E_ergs * ERGS_TO_EV;
        if (E_ev < chi) continue;
        const double sigma = photo_cs(chi, chi_ion, ion, E_ev);
        const double Jterm = sigma * J[i] / E_ergs;
        const double dnu = loc_delta(A.nu, i);
        P += Jterm * dnu;
        Rr += (sigma * fac1 * nuc * nuc + Jterm) * exp(-1.0 * (E_ev - chi) / K_EV / T) * dnu;
      }
      const double gl_o_gc = (1.0 * at.lev_g[k]) / (1.0 * at.lev_g[kc]);
      P *= 4 * PI;
      Rr *= 4 * PI * gl_o_gc * saha_fac;
    }
    Pic[k] = P;
    Rci[k] = Rr;
  }
  double vd = sqrt(2 * K_BOLTZ * T / M_PROTON / A.elem_A[a]);
  if (A.line_velocity_width > 0) vd = A.line_velocity_width;
  const double beta = vd / C_LIGHT;
  const double x_max = 5., dx = 0.05;
  double* lJ = A.line_J + (size_t)s * at.n_blines;
  for (int64_t l = threadIdx.x; l < at.n_blines; l += blockDim.x) {
    const int ll = at.bline_l[l] - l0;
    if (ll < 0 || ll >= n) continue;
    const double nu0 = at.bline_nu[l];
    const double nu_d = nu0 * beta;
    const double a_voigt = at.bline_A[l] / 4 / PI / nu_d;
    double sum = 0, J0 = 0;
    for (double x = -1 * x_max; x <= x_max; x += dx) {
      const double phi = voigt_profile(x, a_voigt);
      const double J1 = J[loc_within(A.nu, nu0 + x * nu_d)] * phi;   // locate_array::value_at
      sum += 0.5 * (J1 + J0) * dx;
      J0 = J1;
    }
    lJ[l] = sum;
  }
}

}  // namespace sgpu

using namespace sgpu;

extern "C" int sgpu_nlte_set_rates(sgpu_ctx* ctx, const sgpu_atoms* atoms, int32_t nsys,
                                   const int32_t* sys_atom, const double* ne, const double* T,
                                   const double* line_J, const double* Pic, const double* Rci,
                                   const double* lev_lte, int32_t use_collisions, int32_t no_ground_recomb,
                                   const int32_t* n, const int64_t* rate_off, const int64_t* vec_off,
                                   double* rates, double* lte)
{
  SGPU_REQUIRE(ctx && atoms, "null argument");
  SGPU_REQUIRE(nsys >= 0, "negative system count");
  if (nsys == 0) return SGPU_OK;
  SGPU_REQUIRE(sys_atom && ne && T && Pic && Rci && lev_lte && n && rate_off && vec_off && rates && lte,
               "missing array");
  SGPU_REQUIRE(atoms->n_atoms >= 1 && atoms->lev_off && atoms->ion_off && atoms->lev_E && atoms->lev_g &&
               atoms->lev_ion && atoms->ion_chi, "atomic level data missing");
  SGPU_REQUIRE(atoms->n_blines == 0 || (line_J && atoms->bline_l && atoms->bline_u && atoms->bline_nu &&
                                        atoms->bline_f && atoms->bline_A && atoms->bline_Bul &&
                                        atoms->bline_Blu),
               "level-resolved line arrays (and line_J) missing");
  int st = 0;
  int32_t* flag = (int32_t*)scratch(ctx, "rates_status", sizeof(int32_t), &st);
  if (st) return st;
  hipLaunchKernelGGL(zero_status_kernel, dim3(1), dim3(1), 0, ctx->stream, flag);
  RateArgs A{*atoms, sys_atom, ne, T, line_J, Pic, Rci, lev_lte, use_collisions, no_ground_recomb,
             n, rate_off, vec_off, rates, lte, flag};
  timer_begin(ctx, "nlte_rates");
  hipLaunchKernelGGL(set_rates_kernel, dim3(nsys), dim3(RATE_THREADS), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "nlte_rates");
  int32_t h = 0;
  SGPU_HIP(hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  SGPU_REQUIRE(h == 0, "n[s] differs from the level count of atom sys_atom[s]");
  return SGPU_OK;
}

extern "C" int sgpu_nlte_populations(sgpu_ctx* ctx, const sgpu_atoms* atoms, int32_t nsys,
                                     const int32_t* sys_atom, const int64_t* vec_off, const double* b,
                                     const double* lte, double* lev_n, double* ion_frac)
{
  SGPU_REQUIRE(ctx && atoms, "null argument");
  SGPU_REQUIRE(nsys >= 0, "negative system count");
  if (nsys == 0) return SGPU_OK;
  SGPU_REQUIRE(sys_atom && vec_off && b && lte && lev_n && ion_frac, "missing array");
  SGPU_REQUIRE(atoms->n_atoms >= 1 && atoms->lev_off && atoms->ion_off && atoms->lev_ion,
               "atomic level data missing");
  PopArgs A{*atoms, sys_atom, vec_off, b, lte, lev_n, ion_frac};
  hipLaunchKernelGGL(populations_kernel, dim3(nsys), dim3(RATE_THREADS), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  return SGPU_OK;
}

extern "C" int sgpu_nlte_radiative_rates(sgpu_ctx* ctx, const sgpu_atoms* atoms, const sgpu_locate* nu_grid,
                                         int32_t nsys, const int32_t* sys_atom, const int32_t* sys_zone,
                                         const double* T, const int32_t* elem_A, double line_velocity_width,
                                         const double* J_nu, double* Pic, double* Rci, double* line_J)
{
  SGPU_REQUIRE(ctx && atoms && nu_grid, "null argument");
  SGPU_REQUIRE(nsys >= 0, "negative system count");
  if (nsys == 0) return SGPU_OK;
  SGPU_REQUIRE(sys_atom && sys_zone && T && elem_A && J_nu && Pic && Rci, "missing array");
  SGPU_REQUIRE(nu_grid->n >= 1 && nu_grid->x, "bad frequency grid");
  SGPU_REQUIRE(atoms->n_atoms >= 1 && atoms->lev_off && atoms->ion_off && atoms->lev_E && atoms->lev_g &&
               atoms->lev_ion && atoms->ion_chi, "atomic level data missing");
  SGPU_REQUIRE(atoms->n_blines == 0 || (line_J && atoms->bline_l && atoms->bline_nu && atoms->bline_A),
               "level-resolved line arrays (and line_J) missing");
  RadArgs A{*atoms, *nu_grid, sys_atom, sys_zone, T, elem_A, line_velocity_width, J_nu, Pic, Rci, line_J};
  timer_begin(ctx, "nlte_radrates");
  hipLaunchKernelGGL(radiative_rates_kernel, dim3(nsys), dim3(RATE_THREADS), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "nlte_radrates");
  return SGPU_OK;
}
