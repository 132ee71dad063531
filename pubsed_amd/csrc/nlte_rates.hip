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
#include <climits>
#include <hip/hip_runtime.h>
#include "sgpu_internal.h"
#include "sgpu_device.h"
#include "sgpu_physics.h"

namespace sgpu {

constexpr int RATE_CHUNK = 1024;   // lines (levels) staged in LDS per chunk
constexpr int RATE_MAX_STAGES = 256;   // ion stages of one atom held in LDS
constexpr int RATE_LDS_DOUBLES = 6400;  // matrices up to 80 x 80 are assembled in LDS
constexpr int RATE_MAX_VOIGT = 256;     // Voigt quadrature points (the reference's loop makes 201)
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
  // the matrix is accumulated in LDS when it fits (a row's terms are
  // read-modify-writes by its owner thread: in LDS they do not wait on
  // global-memory round trips), else in place in global memory
  __shared__ double R_lds[RATE_LDS_DOUBLES];
  double* const R_out = A.rates + A.rate_off[s];
  const bool in_lds = n * n <= RATE_LDS_DOUBLES;
  double* R = in_lds ? R_lds : R_out;
  const double* J = A.line_J + (size_t)s * at.n_blines;
  const double* Pic = A.Pic + (size_t)s * at.n_levels;
  const double* Rci = A.Rci + (size_t)s * at.n_levels;
  const double* lte = A.lev_lte + (size_t)s * at.n_levels;
  for (int r = t; r < n; r += NT)
    for (int j = 0; j < n; j++) R[(size_t)r * n + j] = 0;

  // The line list is read in chunks staged in LDS by the whole workgroup:
  // one thread per line computes the line's two terms (their transcendentals
  // run in parallel over the lines), then each row's owner walks the chunk
  // in stored order and adds the terms of its row, so every matrix element
  // still sums its terms in the reference's order: all radiative terms
  // (:303-321), then all collisional ones (van Regemorter, :326-380).
  __shared__ int32_t ch_l[RATE_CHUNK], ch_u[RATE_CHUNK];
  __shared__ double ch_lu[RATE_CHUNK], ch_ul[RATE_CHUNK];   // l -> u and u -> l terms
  __shared__ double bf_up[RATE_CHUNK], bf_down[RATE_CHUNK];
  for (int pass = 0; pass < (A.use_collisions ? 2 : 1); ++pass) {
    for (int64_t c0 = 0; c0 < at.n_blines; c0 += RATE_CHUNK) {
      const int m = (int)min((int64_t)RATE_CHUNK, at.n_blines - c0);
      __syncthreads();
      for (int q = t; q < m; q += NT) {
        const int64_t l = c0 + q;
        const int ll = at.bline_l[l] - l0, lu = at.bline_u[l] - l0;
        ch_l[q] = ll;
        ch_u[q] = lu;
        double t_lu = 0, t_ul = 0;
        if (ll >= 0 && ll < n) {
          if (pass == 0) {
            t_ul = at.bline_Bul[l] * J[l] + at.bline_A[l];
            t_lu = at.bline_Blu[l] * J[l];
            if (at.bline_nu[l] == 0) { t_ul = 0; t_lu = 0; }
          } else {
            const int gl = at.lev_g[ll + l0], gu = at.lev_g[lu + l0];
            const double dE = (at.lev_E[lu + l0] - at.lev_E[ll + l0]) * EV_TO_ERGS;
            const double zeta = dE / K_BOLTZ / T;
            const double ezeta = exp(zeta);
            const double f_lu = at.bline_f[l];
            const double eff_f = f_lu < 1.e-3 ? 1.e-3 : f_lu;
            t_lu = 3.9 * pow(zeta, -1.) * pow(T, -1.5) / ezeta * ne * eff_f;
            if (zeta > 700) t_lu = 0;
            t_ul = 3.9 * pow(zeta, -1.) * pow(T, -1.5) * ne * eff_f * gl / gu;
          }
        }
        ch_lu[q] = t_lu;
        ch_ul[q] = t_ul;
      }
      __syncthreads();
      for (int q = 0; q < m; q++) {
        const int ll = ch_l[q], lu = ch_u[q];
        if (ll < 0 || ll >= n) continue;
        if (ll % NT == t) R[(size_t)ll * n + lu] += ch_lu[q];
        if (lu % NT == t) R[(size_t)lu * n + ll] += ch_ul[q];
      }
    }
  }
  // bound-free (:386-423).  The continuum level of level k is the first
  // level of the next ion stage (continuum_level): the first level of every
  // stage is found once, in LDS.  One thread per level computes the level's
  // terms; the owners of the two rows then add them in level order.
  __shared__ int first_lev[RATE_MAX_STAGES];
  const int i0a = at.ion_off[a], n_st = at.ion_off[a + 1] - i0a;
  __syncthreads();
  for (int q = t; q < n_st && q < RATE_MAX_STAGES; q += NT) first_lev[q] = INT_MAX;
  __syncthreads();
  for (int i = t; i < n; i += NT) {
    const int st = at.lev_ion[i + l0];
    if (st >= 0 && st < n_st && st < RATE_MAX_STAGES) atomicMin(&first_lev[st], i + l0);
  }
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += RATE_CHUNK) {
    const int m = min(RATE_CHUNK, n - c0);
    __syncthreads();
    for (int q = t; q < m; q += NT) {
      const int i = c0 + q, k = i + l0;
      const int nxt = at.lev_ion[k] + 1;
      const int kc = (nxt < n_st && nxt < RATE_MAX_STAGES && first_lev[nxt] != INT_MAX) ? first_lev[nxt]
                     : (nxt < RATE_MAX_STAGES ? -1 : continuum_level(at, a, k));
      ch_l[q] = kc == -1 ? -1 : kc - l0;
      double up = 0, down = 0;   // i -> c (collisional ionisation + Pic), c -> i
      if (kc != -1) {
        const double chi = at.ion_chi[i0a + at.lev_ion[k]] - at.lev_E[k];
        const double zeta = chi / K_EV / T;
        double c_ion = 0, c_rec = 0;
        if (A.use_collisions) {
          c_ion = 2.7 / zeta / zeta * pow(T, -1.5) * exp(-zeta) * ne;
          const int gi = at.lev_g[k], gc = at.lev_g[kc];
          c_rec = 5.59080e-16 / zeta / zeta * pow(T, -3) * gi / gc * ne * ne;
        }
        double rci = Rci[k];
        if (A.no_ground_recomb && at.lev_E[k] == 0) rci = 0.;
        // two terms per element in the reference's order: collisional
        // first, then the radiative one
        ch_lu[q] = c_ion;
        ch_ul[q] = c_rec;
        up = Pic[k];
        down = rci * ne;
      } else {
        ch_lu[q] = 0;
        ch_ul[q] = 0;
      }
      bf_up[q] = up;
      bf_down[q] = down;
    }
    __syncthreads();
    for (int q = 0; q < m; q++) {
      const int ic = ch_l[q];
      if (ic < 0) continue;
      const int i = c0 + q;
      if (A.use_collisions) {
        if (i % NT == t) R[(size_t)i * n + ic] += ch_lu[q];
        if (ic % NT == t) R[(size_t)ic * n + i] += ch_ul[q];
      }
      if (ic % NT == t) R[(size_t)ic * n + i] += bf_down[q];
      if (i % NT == t) R[(size_t)i * n + ic] += bf_up[q];
    }
  }
  // times the LTE population of the source level (:429-431)
  double* lo = A.lte + A.vec_off[s];
  for (int r = t; r < n; r += NT) {
    const double w = lte[r + l0];
    for (int j = 0; j < n; j++) R[(size_t)r * n + j] *= w;
    lo[r] = w;
  }
  if (in_lds) {
    __syncthreads();
    for (int q = t; q < n * n; q += NT) R_out[q] = R_lds[q];
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
  const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE, n_wave = blockDim.x / WAVE;
  // the Voigt abscissae x = -5, -5 + dx, ... as the reference's loop forms
  // them, by repeated addition (its rounding fixes the points and their count)
  const double x_max = 5., dx = 0.05;
  __shared__ double xs[RATE_MAX_VOIGT];
  __shared__ int n_x;
  if (threadIdx.x == 0) {
    int m = 0;
    for (double x = -1 * x_max; x <= x_max && m < RATE_MAX_VOIGT; x += dx) xs[m++] = x;
    n_x = m;
  }
  __syncthreads();
  // one wave per level: the lanes share the level's frequency sum
  for (int j = wave; j < n; j += n_wave) {
    const int k = l0 + j;
    const int kc = continuum_level(at, a, k);
    double P = 0, Rr = 0;
    if (kc != -1) {
      const int ion = at.lev_ion[k];
      const double chi_ion = at.ion_chi[i0 + ion];
      const double chi = chi_ion - at.lev_E[k];
      for (int i = 1 + lane; i < nn; i += WAVE) {
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
      P = wave_sum(P);
      Rr = wave_sum(Rr);
      const double gl_o_gc = (1.0 * at.lev_g[k]) / (1.0 * at.lev_g[kc]);
      P *= 4 * PI;
      Rr *= 4 * PI * gl_o_gc * saha_fac;
    }
    if (lane == 0) {
      Pic[k] = P;
      Rci[k] = Rr;
    }
  }
  double vd = sqrt(2 * K_BOLTZ * T / M_PROTON / A.elem_A[a]);
  if (A.line_velocity_width > 0) vd = A.line_velocity_width;
  const double beta = vd / C_LIGHT;
  double* lJ = A.line_J + (size_t)s * at.n_blines;
  const int nx = n_x;
  // one wave per line: lane q takes the trapezoids (x_{q-1}, x_q), with
  // f_{-1} = 0 as the reference's J0 starts
  for (int64_t l = wave; l < at.n_blines; l += n_wave) {
    const int ll = at.bline_l[l] - l0;
    if (ll < 0 || ll >= n) continue;
    const double nu0 = at.bline_nu[l];
    const double nu_d = nu0 * beta;
    const double a_voigt = at.bline_A[l] / 4 / PI / nu_d;
    double sum = 0;
    for (int q = lane; q < nx; q += WAVE) {
      const double x1 = xs[q];
      const double J1 = J[loc_within(A.nu, nu0 + x1 * nu_d)] * voigt_profile(x1, a_voigt);   // value_at
      double J0 = 0;
      if (q > 0) {
        const double x0 = xs[q - 1];
        J0 = J[loc_within(A.nu, nu0 + x0 * nu_d)] * voigt_profile(x0, a_voigt);
      }
      sum += 0.5 * (J1 + J0) * dx;
    }
    sum = wave_sum(sum);
    if (lane == 0) lJ[l] = sum;
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
