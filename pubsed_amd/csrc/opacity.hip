// opacity.hip -- K1-K3: the per-zone opacity / emissivity build.
//
//   K1 state_kernel     one wave64 per zone: temperature clamp, composition
//                       decay, mass fractions, LTE Saha-Boltzmann state with
//                       the reference's Brent iteration for n_e; levels are
//                       spread over the lanes, partition functions and the
//                       net ion fraction are wave reductions.
//   K2 bins_kernel      one thread per (zone, frequency bin): electron
//                       scattering, free-free, bound-free (every level with
//                       its edge below the bin), level-resolved and fuzz-line
//                       expansion opacity summed over the bin's lines (lines
//                       sorted by bin, CSR index), Planck emissivity,
//                       max-extinction clamp; writes abs, scat and emis*dnu.
//   K3 finalize_kernel  one block per zone: emissivity CDF (block scan),
//                       normalisation, L_thermal, Planck / Rosseland means,
//                       gamma-ray coefficients, eps_imc.
//
// Reference: transport::set_opacity (src/transport/transport_opacity.cpp:21-260),
// GasState::solve_state / ne_brent_method (src/opacity/GasState.cpp:208-337),
// AtomicSpecies::solve_lte (src/opacity/AtomicSpecies_state.cpp:56-123),
// GasState::computeOpacity (src/opacity/GasState_opacities.cpp:11-136),
// fuzz_expansion_opacity (:464-494) and fuzzline_expansion_opacity
// (src/opacity/AtomicSpecies_opacities.cpp:292-324), get_planck_mean /
// get_rosseland_mean (GasState_opacities.cpp:564-666).
#include <cfloat>
#include <vector>
#include "sgpu_internal.h"
#include "sgpu_physics.h"

namespace sgpu {

constexpr int MAX_ELEMS = 64;        // elements per zone
constexpr int MAX_LEV = 2048;        // levels over all atoms (LDS per zone)
constexpr int MAX_ION = 256;         // ions over all atoms
constexpr int FIN_THREADS = 256;

struct StateArgs {
  sgpu_zone_state zs;
  sgpu_atoms at;
  sgpu_opacity_params p;
  int n_lev, n_ion;
  double* temp_used;   // [nz]
  double* nion;        // [nz][n_ion]  n_dens * ion_frac
  double* upart;       // [nz][n_ion]  partition function
  sgpu_opacity_out out;
  double* levn;        // [nz][n_lev]  level populations (fractions of the atom), or NULL
  double* ndens;       // [nz][n_atoms] atom number densities, or NULL
  double* fffac;       // [nz] free-free prefactor 3.7e8 T^-1/2 n_e sum n_ion Z^2, or NULL
  double* phifac;      // [nz][n_lev] bound-free n_c (g_l/g_c)/2 lambda_T^3, -1 without a continuum, or NULL
};

struct ZoneLds {
  double X[MAX_ELEMS];
  double mf[MAX_ELEMS];
  double bz[MAX_LEV];     // exp(-E/kT) per level
  double U[MAX_ION];      // partition function per ion
  double frac[MAX_ION];   // ion fractions of the current evaluation
  double mu_I;
};

// Saha chain of one atom (AtomicSpecies_state.cpp:69-97) into L.frac,
// evaluated by lane 0 (sequential over the atom's ions).
__device__ inline void saha_chain(ZoneLds& L, const sgpu_atoms& at, int a, double ne, double T)
{
  int i0 = at.ion_off[a], ni = at.ion_off[a + 1] - i0;
  double lt = H_PLANCK * H_PLANCK / (2.0 * PI * M_ELECTRON * K_BOLTZ * T);
  double fac = 2 / ne / pow(lt, 1.5);
  double* f = L.frac + i0;
  const double* U = L.U + i0;
  f[0] = 1.;
  for (int i = 1; i < ni; ++i) {
    double chi = at.ion_chi[i0 + i - 1];
    double saha = exp(-1.0 * chi / K_EV / T);
    saha = saha * (U[i] / U[i - 1]) * fac;
    f[i] = saha * f[i - 1];
    if (ne < 1e-50) f[i] = 0;
    double norm = 0;
    for (int j = 0; j <= i; j++) norm += f[j];
    for (int j = 0; j <= i; j++) f[j] = f[j] / norm;
  }
}

// GasState::charge_conservation (GasState.cpp:242-263) for one zone,
// executed by the whole wave (one wave per block).
__device__ double charge_conservation(ZoneLds& L, const StateArgs& A, double ne, double dens,
                                      double T, int lane)
{
  const sgpu_atoms& at = A.at;
  double f = 0;
  for (int a = 0; a < at.n_atoms; ++a) {
    if (lane == 0) saha_chain(L, at, a, ne, T);
    __syncthreads();
    int l0 = at.lev_off[a], l1 = at.lev_off[a + 1];
    int i0 = at.ion_off[a];
    double acc = 0;
    for (int l = l0 + lane; l < l1; l += WAVE) {
      int ion = at.lev_ion[l];
      double fr = L.frac[i0 + ion];
      double v = fr * (double)at.lev_g[l] * L.bz[l] / L.U[i0 + ion];
      if (v < 1e-30) v = 1e-30;
      acc += v * ion;
    }
    double x = wave_sum(acc);
    f += dens * L.mf[a] / (A.zs.elem_A[a] * M_PROTON) * x;
    __syncthreads();
  }
  return f - ne;
}

// The temperature-only part of a zone's gas state, shared by the LTE and
// NLTE solves: temperature clamp (transport_opacity.cpp:92-113), composition
// decay, mass fractions and mean ion mass (GasState::set_mass_fractions,
// GasState.cpp:137-156); for a gas zone also the Boltzmann factors and
// partition functions, which depend on T only and so serve every n_e
// evaluation (AtomicSpecies_state.cpp:59-67).  Returns false for a grey
// zone or a problem without atoms.  Executed by the whole wave.
__device__ bool zone_setup(ZoneLds& L, const sgpu_zone_state& zs, const sgpu_atoms& at,
                           const sgpu_opacity_params& p, int n_lev, int z, int lane, double* T_out)
{
  const int ne_ = zs.n_elems;
  double T = zs.T_gas[z];
  if (T < p.temp_min) T = p.temp_min;
  if (T > p.temp_max) T = p.temp_max;
  if (p.first_step) T = zs.T_gas[z];   // transport_opacity.cpp:111-113
  *T_out = T;
  if (lane == 0) {
    for (int j = 0; j < ne_; j++) L.X[j] = zs.X[(size_t)z * ne_ + j];
    if (!p.omit_composition_decay) decay_composition(ne_, zs.elem_Z, zs.elem_A, L.X, p.t_now);
    double norm = 0.0;
    for (int j = 0; j < ne_; ++j) { L.mf[j] = L.X[j]; norm += L.X[j]; }
    for (int j = 0; j < ne_; ++j) L.mf[j] /= norm;
    double inv = 0.;
    for (int j = 0; j < ne_; ++j) inv += L.mf[j] / zs.elem_A[j];
    L.mu_I = 1. / inv;
  }
  __syncthreads();
  double grey = zs.grey_opacity ? zs.grey_opacity[z] : 0.0;
  if (grey != 0 || at.n_atoms == 0) return false;
  for (int l = lane; l < n_lev; l += WAVE) L.bz[l] = exp(-at.lev_E[l] / K_EV / T);
  __syncthreads();
  for (int a = 0; a < at.n_atoms; ++a) {
    int l0 = at.lev_off[a], l1 = at.lev_off[a + 1];
    int i0 = at.ion_off[a], ni = at.ion_off[a + 1] - i0;
    for (int s = 0; s < ni; ++s) {
      double acc = 0;
      for (int l = l0 + lane; l < l1; l += WAVE)
        if (at.lev_ion[l] == s) acc += (double)at.lev_g[l] * L.bz[l];
      double u = wave_sum(acc);
      if (lane == 0) L.U[i0 + s] = u;
    }
  }
  __syncthreads();
  return true;
}

__global__ void __launch_bounds__(WAVE) state_kernel(StateArgs A)
{
  __shared__ ZoneLds L;
  const int z = blockIdx.x;
  const int lane = threadIdx.x;
  const sgpu_zone_state& zs = A.zs;
  const sgpu_atoms& at = A.at;
  double dens = zs.rho[z];
  double T;
  const bool gas = zone_setup(L, zs, at, A.p, A.n_lev, z, lane, &T);
  if (lane == 0) A.temp_used[z] = T;
  if (!gas) {
    for (int k = lane; k < A.n_ion; k += WAVE) {
      A.nion[(size_t)z * A.n_ion + k] = 0;
      A.upart[(size_t)z * A.n_ion + k] = 1;
      if (A.out.ion_frac) A.out.ion_frac[(size_t)z * A.n_ion + k] = 0;
      if (A.out.ion_part) A.out.ion_part[(size_t)z * A.n_ion + k] = 0;
    }
    if (A.out.lev_n)
      for (int k = lane; k < A.n_lev; k += WAVE) A.out.lev_n[(size_t)z * A.n_lev + k] = 0;
    if (lane == 0) {
      A.out.n_elec[z] = 0;
      if (A.out.zone_status) A.out.zone_status[z] = 0;
      if (A.fffac) A.fffac[z] = 0;
    }
    return;
  }

  // GasState::ne_brent_method (GasState.cpp:271-337), wave-uniform
  const double x1 = 1e-10 * dens / (L.mu_I * M_PROTON);
  const double x2 = 100 * dens / (L.mu_I * M_PROTON);
  const double tol = 1e-3, EPS = 3.0e-8;
  int err = 0;
  double a = x1, b = x2, c = x2, d = 0, e = 0, min1, min2;
  double fa = charge_conservation(L, A, a, dens, T, lane);
  double fb = charge_conservation(L, A, b, dens, T, lane);
  double fc, pp, q, r, s, tol1, xm;
  double root = 0.0;
  bool done = false;
  if ((fa > 0.0 && fb > 0.0) || (fa < 0.0 && fb < 0.0)) err = 1;
  fc = fb;
  for (int iter = 1; iter <= 100; iter++) {
    if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { c = a; fc = fa; e = d = b - a; }
    if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    tol1 = 2.0 * EPS * fabs(b) + 0.5 * tol;
    xm = 0.5 * (c - b);
    if (fabs(xm) <= tol1 || fb == 0.0) { root = b; done = true; break; }
    if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
      s = fb / fa;
      if (a == c) { pp = 2.0 * xm * s; q = 1.0 - s; }
      else {
        q = fa / fc; r = fb / fc;
        pp = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
        q = (q - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (pp > 0.0) q = -q;
      pp = fabs(pp);
      min1 = 3.0 * xm * q - fabs(tol1 * q);
      min2 = fabs(e * q);
      if (2.0 * pp < (min1 < min2 ? min1 : min2)) { e = d; d = pp / q; }
      else { d = xm; e = d; }
    } else { d = xm; e = d; }
    a = b; fa = fb;
    if (fabs(d) > tol1) b += d;
    else b += (xm >= 0.0 ? fabs(tol1) : -fabs(tol1));
    fb = charge_conservation(L, A, b, dens, T, lane);
  }
  if (!done) { err = 2; root = 0.0; }

  // outputs: the atoms keep the state of the LAST evaluation (L.frac)
  for (int at_i = 0; at_i < at.n_atoms; ++at_i) {
    double n_dens = dens * L.mf[at_i] / (zs.elem_A[at_i] * M_PROTON);
    for (int k = at.ion_off[at_i] + lane; k < at.ion_off[at_i + 1]; k += WAVE) {
      A.nion[(size_t)z * A.n_ion + k] = n_dens * L.frac[k];
      A.upart[(size_t)z * A.n_ion + k] = L.U[k];
      if (A.out.ion_frac) A.out.ion_frac[(size_t)z * A.n_ion + k] = L.frac[k];
      if (A.out.ion_part) A.out.ion_part[(size_t)z * A.n_ion + k] = L.U[k];
    }
  }
  if (A.out.lev_n || A.levn) {
    for (int at_i = 0; at_i < at.n_atoms; ++at_i) {
      int i0 = at.ion_off[at_i];
      for (int l = at.lev_off[at_i] + lane; l < at.lev_off[at_i + 1]; l += WAVE) {
        int ion = at.lev_ion[l];
        double v = L.frac[i0 + ion] * (double)at.lev_g[l] * L.bz[l] / L.U[i0 + ion];
        v = v < 1e-30 ? 1e-30 : v;
        if (A.out.lev_n) A.out.lev_n[(size_t)z * A.n_lev + l] = v;
        if (A.levn) A.levn[(size_t)z * A.n_lev + l] = v;
      }
    }
  }
  if (A.ndens)
    for (int at_i = lane; at_i < at.n_atoms; at_i += WAVE)
      A.ndens[(size_t)z * at.n_atoms + at_i] = dens * L.mf[at_i] / (zs.elem_A[at_i] * M_PROTON);
  // AtomicSpecies::bound_free_opacity_general (AtomicSpecies_opacities.cpp:94-104):
  // n_c g_l / g_c / 2 lambda_T^3 per level, with n_c the (floored) population
  // of the next ion's ground level
  if (A.phifac) {
    const double lam_t = sqrt(H_PLANCK * H_PLANCK / (2 * PI * M_ELECTRON * K_BOLTZ * T));
    for (int at_i = 0; at_i < at.n_atoms; ++at_i) {
      const int i0 = at.ion_off[at_i];
      const double n_dens = dens * L.mf[at_i] / (zs.elem_A[at_i] * M_PROTON);
      for (int l = at.lev_off[at_i] + lane; l < at.lev_off[at_i + 1]; l += WAVE) {
        const int kc = continuum_level(at, at_i, l);
        double ph = -1;
        if (kc != -1) {
          const int ion = at.lev_ion[kc];
          double v = L.frac[i0 + ion] * (double)at.lev_g[kc] * L.bz[kc] / L.U[i0 + ion];
          v = v < 1e-30 ? 1e-30 : v;
          const double nc = n_dens * v;
          const double gl_o_gc = (1.0 * at.lev_g[l]) / (1.0 * at.lev_g[kc]);
          ph = nc * gl_o_gc / 2. * lam_t * lam_t * lam_t;
        }
        A.phifac[(size_t)z * A.n_lev + l] = ph;
      }
    }
  }
  // GasState::free_free_opacity (GasState_opacities.cpp:160-173)
  if (A.fffac && lane == 0) {
    double fac = 0;
    for (int at_i = 0; at_i < at.n_atoms; ++at_i) {
      double Z_eff_sq = 0;
      const int i0 = at.ion_off[at_i], ni = at.ion_off[at_i + 1] - i0;
      for (int j = 0; j < ni; j++) Z_eff_sq += L.frac[i0 + j] * j * j;
      const double n_ion = L.mf[at_i] * dens / (zs.elem_A[at_i] * M_PROTON);
      fac += n_ion * Z_eff_sq;
    }
    A.fffac[z] = fac * (3.7e8 * pow(T, -0.5) * root);
  }
  if (lane == 0) {
    A.out.n_elec[z] = root;
    if (A.out.zone_status) A.out.zone_status[z] = err;
  }
}

struct BinArgs {
  sgpu_locate nu;
  sgpu_zone_state zs;
  sgpu_atoms at;
  sgpu_opacity_params p;
  int n_ion;
  const double* temp_used;
  const double* nion;
  const double* upart;
  double* abs;
  double* scat;
  double* ednu;    // emissivity * dnu, scanned in place by K3
  const double* n_elec;
  const double* levn;    // [nz][n_lev] (line expansion and bound-free only)
  const double* ndens;   // [nz][n_atoms]
  int n_lev;
  const double* fffac;   // [nz] (free-free only)
  const double* phifac;  // [nz][n_lev] (bound-free only)
};

// GasState::bound_bound_opacity (GasState_opacities.cpp:379-392) over
// AtomicSpecies::bound_bound_opacity (AtomicSpecies_opacities.cpp:224-282),
// gathered for one (zone, bin): the reference adds each line's Voigt profile
// into the bins [locate(nu_0 - 5 dnu), locate(nu_0 + 5 dnu)) at their
// centres, atom by atom; here the bin collects, atom by atom in element
// order, the lines of that atom whose window holds it, in the atom's line
// order.  Doppler width line_beta_dop nu_0 with line_beta_dop = v_D / c,
// v_D = sqrt(2 k T / (m_p A)) or line_velocity_width when set
// (GasState.cpp:218-220); Voigt damping A_ul / (4 pi dnu); extinction
// nl n g_u/g_l A_ul c^2 / (8 pi) corrected for stimulated emission; lines
// with alpha_0 / (nu_0^2 dnu) below opacity_minimum_extinction skipped.  A
// line can cover bin j only if nu_0 (1 + 5 beta) >= x[j-1] and
// nu_0 (1 - 5 beta) < x[j+1] (locate is upper_bound up to the log formula's
// unverified first bin), so the candidates are the lines of the atom's CSR
// bins holding nu_0 in [x[j-1] / (1 + 5 beta), x[j+1] / (1 - 5 beta)],
// widened by 1e-9 relative; each candidate is then tested with the same
// locate calls as the reference.  Every sum runs in a fixed order, so two
// builds are bit-identical.
__device__ inline void bound_bound_bin(const BinArgs& A, int z, int j, double T, double& opac, double& emis)
{
  const int nn = A.nu.n;
  const double nuc = loc_center(A.nu, j);
  const double* ln = A.levn + (size_t)z * A.n_lev;
  opac = 0;
  emis = 0;
  for (int a = 0; a < A.at.n_atoms; ++a) {
    const double n_dens = A.ndens[(size_t)z * A.at.n_atoms + a];
    double vd = sqrt(2 * K_BOLTZ * T / M_PROTON / A.zs.elem_A[a]);
    if (A.p.line_velocity_width > 0) vd = A.p.line_velocity_width;
    const double beta = vd / C_LIGHT;
    int b_lo = 0, b_hi = nn - 1;
    if (5 * beta < 0.5) {
      if (j >= 1) b_lo = upper_bound_dev(A.nu.x, nn, A.nu.x[j - 1] / (1 + 5 * beta) * (1 - 1e-9));
      if (j + 1 <= nn - 1) b_hi = upper_bound_dev(A.nu.x, nn, A.nu.x[j + 1] / (1 - 5 * beta) * (1 + 1e-9));
      b_lo = b_lo > 0 ? b_lo - 1 : 0;
      b_hi = b_hi + 1 < nn ? b_hi + 1 : nn - 1;
    }
    const int64_t* off = A.at.bline_bin_off + (size_t)a * (nn + 1);
    double atom_opac = 0, atom_emis = 0;
    for (int64_t l = off[b_lo]; l < off[b_hi + 1]; ++l) {
      const int ll = A.at.bline_l[l], lu = A.at.bline_u[l];
      const double nl = ln[ll], nup = ln[lu];
      if (nl == 0) continue;
      const double gl = 1.0 * A.at.lev_g[ll], gu = 1.0 * A.at.lev_g[lu];
      const double nu_0 = A.at.bline_nu[l];
      const double dnu = beta * nu_0;
      const double A_ul = A.at.bline_A[l];
      double alpha_0 = nl * n_dens * gu / gl * A_ul / (8 * PI) * C_LIGHT * C_LIGHT;
      alpha_0 = alpha_0 * (1 - nup * gl / (nl * gu));
      if (alpha_0 <= 0) continue;
      if (alpha_0 / (nu_0 * nu_0 * dnu) < A.p.minimum_extinction) continue;
      const int inu1 = loc_within(A.nu, nu_0 - dnu * 5);
      const int inu2 = loc_within(A.nu, nu_0 + dnu * 5);
      if (j < inu1 || j >= inu2) continue;
      const double a_voigt = A_ul / 4 / PI / dnu;
      const double line_j = A_ul * nup * n_dens * H_PLANCK / (4.0 * PI);
      const double x = (nu_0 - nuc) / dnu;
      const double phi = voigt_profile(x, a_voigt) / dnu;
      atom_opac += alpha_0 / nuc / nuc * phi;
      atom_emis += line_j * nuc * phi;
    }
    opac += atom_opac;
    emis += atom_emis;
  }
}

__global__ void __launch_bounds__(256) bins_kernel(BinArgs A)
{
  const int z = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int nn = A.nu.n;
  if (j >= nn) return;
  const double T = A.temp_used[z];
  const double dens = A.zs.rho[z];
  const double grey = A.zs.grey_opacity ? A.zs.grey_opacity[z] : 0.0;
  const double eps = A.p.epsilon;
  const double nuc = loc_center(A.nu, j);
  const double dnu = loc_delta(A.nu, j);
  double ab = 0, sc = 0, em = 0;
  if (grey != 0) {
    double gopac = dens * grey;
    ab = gopac * eps;
    sc = gopac * (1 - eps);
    double ezeta = exp(1.0 * H_PLANCK * nuc / K_BOLTZ / T);
    double bb = 2.0 * nuc * nuc * nuc * H_PLANCK / C_LIGHT / C_LIGHT / (ezeta - 1);
    em += bb * ab;
  } else {
    if (A.p.use_electron_scattering) {
      double es = THOMSON * A.n_elec[z];
      sc += es;
      ab += 1e-20 * eps * es;
    }
    // GasState::free_free_opacity (GasState_opacities.cpp:152-182), gaunt factor 1
    if (A.p.use_free_free) {
      const double fac = A.fffac[z];
      const double ezeta = exp(-1.0 * H_PLANCK * nuc / K_BOLTZ / T);
      const double bb = 2.0 * nuc * nuc * nuc * H_PLANCK / C_LIGHT / C_LIGHT / (1.0 / ezeta - 1);
      const double op = fac / nuc / nuc / nuc * (1 - ezeta);
      ab += op;
      em += op * bb;
    }
    // GasState::bound_free_opacity (GasState_opacities.cpp:352-372) over
    // AtomicSpecies::bound_free_opacity_general, coolheat 0
    // (AtomicSpecies_opacities.cpp:106-140): every level whose edge lies below
    // the bin centre; the emissivity sum is multiplied by n_e afterwards
    if (A.p.use_bound_free && A.at.n_atoms > 0) {
      const double kt_ev = K_EV * T;
      const double E = H_PLANCK * nuc * ERGS_TO_EV;
      const double emis_fac = 2. * H_PLANCK * nuc * nuc * nuc / C_LIGHT / C_LIGHT;
      const double ne = A.n_elec[z];
      const double* ln = A.levn + (size_t)z * A.n_lev;
      const double* ph = A.phifac + (size_t)z * A.n_lev;
      double opac = 0, emis = 0;
      for (int a = 0; a < A.at.n_atoms; ++a) {
        const double n_dens = A.ndens[(size_t)z * A.at.n_atoms + a];
        const int i0 = A.at.ion_off[a];
        double atom_opac = 0, atom_emis = 0;
        for (int k = A.at.lev_off[a]; k < A.at.lev_off[a + 1]; ++k) {
          const double phk = ph[k];
          if (phk < 0) continue;             // no continuum level
          const double E_lev = A.at.lev_E[k];
          const double chi_ion = A.at.ion_chi[i0 + A.at.lev_ion[k]];
          const double Eion = chi_ion - E_lev;
          if (E < Eion) continue;
          const double ezeta_net = exp((Eion - E) / kt_ev);
          const double sigma = photo_cs(Eion, chi_ion, A.at.lev_ion[k], E);
          double opac_fac = n_dens * ln[k] - phk * ne * ezeta_net;
          if (opac_fac < 0) opac_fac = 0.;   // kill maser
          atom_opac += sigma * opac_fac;
          if (E_lev == 0 && A.p.no_ground_recomb) continue;
          atom_emis += emis_fac * sigma * phk * ezeta_net;
        }
        opac += atom_opac;
        emis += atom_emis;
      }
      ab += opac;
      em += emis * ne;
    }
    // GasState::bound_bound_opacity (GasState_opacities.cpp:83-91)
    if (A.p.use_bound_bound && A.at.n_atoms > 0 && A.at.n_blines > 0) {
      double bo, be;
      bound_bound_bin(A, z, j, T, bo, be);
      ab += bo;
      em += be;
    }
    // GasState::line_expansion_opacity (GasState_opacities.cpp:424-454),
    // AtomicSpecies::line_expansion_opacity (AtomicSpecies_opacities.cpp:329-375):
    // the bin's level-resolved lines of each atom in their stored order
    if (A.p.use_line_expansion && A.at.n_blines > 0) {
      const double time = A.p.t_now;
      const double* ln = A.levn + (size_t)z * A.n_lev;
      double opac = 0, aopac = 0;
      for (int a = 0; a < A.at.n_atoms; ++a) {
        const double this_eps = (A.at.atom_zero_eps && A.at.atom_zero_eps[a]) ? 0 : eps;
        const double n_dens = A.ndens[(size_t)z * A.at.n_atoms + a];
        const int64_t* off = A.at.bline_bin_off + (size_t)a * (nn + 1);
        double atom_opac = 0;
        for (int64_t l = off[j]; l < off[j + 1]; ++l) {
          const int ll = A.at.bline_l[l], lu = A.at.bline_u[l];
          const double nl = ln[ll], nup = ln[lu];
          const double gl = 1.0 * A.at.lev_g[ll], gu = 1.0 * A.at.lev_g[lu];
          if (nl < DBL_MIN) continue;
          const double lam = C_LIGHT / A.at.bline_nu[l];
          double tau = nl * n_dens * SIGMA_TOT * A.at.bline_f[l] * time * lam;
          tau = tau * (1 - nup * gl / (nl * gu));
          if (nup * gl > nl * gu) tau = 0;   // laser regime
          const double etau = exp(-tau);
          if (!isnan(etau)) atom_opac += (1 - etau);
        }
        atom_opac = atom_opac * nuc / dnu / C_LIGHT / time;
        opac += atom_opac;
        aopac += atom_opac * this_eps;
      }
      ab += aopac;
      sc += opac - aopac;
      double ezeta = exp(1.0 * H_PLANCK * nuc / K_BOLTZ / T);
      double bb = 2.0 * nuc * nuc * nuc * H_PLANCK / C_LIGHT / C_LIGHT / (ezeta - 1);
      em += bb * aopac;
    }
    if (A.p.use_fuzz_expansion) {
      const double time = A.p.t_now;
      double opac = 0, aopac = 0;
      for (int a = 0; a < A.at.n_atoms; ++a) {
        const double this_eps = (A.at.atom_zero_eps && A.at.atom_zero_eps[a]) ? 0 : eps;
        const int64_t* off = A.at.line_bin_off + (size_t)a * (nn + 1);
        const int64_t l0 = off[j], l1 = off[j + 1];
        const int i0 = A.at.ion_off[a];
        const double* nion = A.nion + (size_t)z * A.n_ion + i0;
        const double* U = A.upart + (size_t)z * A.n_ion + i0;
        double atom_opac = 0;
        for (int64_t l = l0; l < l1; ++l) {
          const double lnu = A.at.line_nu[l];
          const int ion = A.at.line_ion[l];
          double nl = nion[ion] * exp(-1.0 * A.at.line_El[l] / K_EV / T) / U[ion];
          double lam = C_LIGHT / lnu;
          double stim_cor = (1 - exp(-H_PLANCK * lnu / K_BOLTZ / T));
          double tau = SIGMA_TOT * lam * nl * A.at.line_gf[l] * stim_cor * time;
          atom_opac += (1 - exp(-tau));
        }
        atom_opac = atom_opac * nuc / dnu / C_LIGHT / time;
        opac += atom_opac;
        aopac += atom_opac * this_eps;
      }
      ab += aopac;
      sc += opac - aopac;
      double ezeta = exp(1.0 * H_PLANCK * nuc / K_BOLTZ / T);
      double bb = 2.0 * nuc * nuc * nuc * H_PLANCK / C_LIGHT / C_LIGHT / (ezeta - 1);
      em += bb * aopac;
    }
  }
  double ed = em * dnu;
  if (nn > 1) {
    const double max_ext = A.p.maximum_opacity * dens;
    if (sc > max_ext) sc = max_ext;
    if (ab > max_ext) ab = max_ext;
  }
  const size_t k = (size_t)z * nn + j;
  A.abs[k] = ab;
  A.scat[k] = sc;
  A.ednu[k] = ed;
}

struct FinArgs {
  sgpu_locate nu;
  sgpu_zone_state zs;
  sgpu_opacity_params p;
  const double* temp_used;
  sgpu_opacity_out out;
};

template <int NT>
__device__ inline double block_sum(double v, double* red)
{
  v = wave_sum(v);
  const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0;
  for (int i = 0; i < NT / WAVE; i++) t += red[i];
  return t;
}

__global__ void __launch_bounds__(FIN_THREADS) finalize_kernel(FinArgs A)
{
  __shared__ double red[FIN_THREADS / WAVE];
  __shared__ double wsum[FIN_THREADS / WAVE];
  const int z = blockIdx.x;
  const int tid = threadIdx.x;
  const int nn = A.nu.n;
  const double T = A.temp_used[z];
  const double dens = A.zs.rho[z];
  double* cdf = A.out.emis_cdf + (size_t)z * nn;
  const double* ab = A.out.abs + (size_t)z * nn;
  const double* sc = A.out.scat + (size_t)z * nn;

  if (nn == 1) {
    if (tid == 0) {
      double Tg = A.zs.T_gas[z];
      double bb_int = SB * pow(Tg, 4) / PI;
      A.out.L_thermal[z] = 4 * PI * ab[0] * bb_int;
      cdf[0] = 1;
      A.out.planck_mean[z] = ab[0];
      A.out.rosseland_mean[z] = ab[0] + sc[0];
    }
  } else {
    // contiguous chunk of bins per thread; sequential inside the chunk
    const int per = (nn + FIN_THREADS - 1) / FIN_THREADS;
    const int j0 = tid * per, j1 = min(nn, j0 + per);
    double local = 0, lth = 0, pm = 0, pn = 0, rm = 0, rn = 0;
    for (int j = j0; j < j1; ++j) {
      double ed = cdf[j];
      local += ed;
      lth += 4 * PI * ed;
      double nuc = loc_center(A.nu, j), dnu = loc_delta(A.nu, j);
      double zeta = H_PLANCK * nuc / K_BOLTZ / T;
      double ezeta = exp(zeta);
      double Bnu = 2.0 * nuc * nuc * nuc * H_PLANCK / C_LIGHT / C_LIGHT / (ezeta - 1);
      if (isnan(Bnu)) Bnu = 0;
      double W = Bnu * dnu;
      pm += W * ab[j];
      pn += W;
      double dBdT = pow(nuc, 4) * ezeta / (ezeta - 1) / (ezeta - 1);
      if (isnan(dBdT)) dBdT = 0;
      double Wr = dBdT * dnu;
      rm += Wr / (ab[j] + sc[j]);
      rn += Wr;
    }
    // exclusive scan of chunk totals across the block (wave scan + LDS)
    const int lane = tid % WAVE, w = tid / WAVE;
    double incl = local;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      double t = __shfl_up(incl, o, WAVE);
      if (lane >= o) incl += t;
    }
    if (lane == WAVE - 1) wsum[w] = incl;
    __syncthreads();
    double base = 0;
    for (int i = 0; i < w; i++) base += wsum[i];
    double run = base + incl - local;
    for (int j = j0; j < j1; ++j) { run += cdf[j]; cdf[j] = run; }
    double L_th = block_sum<FIN_THREADS>(lth, red);
    double P_m = block_sum<FIN_THREADS>(pm, red);
    double P_n = block_sum<FIN_THREADS>(pn, red);
    double R_m = block_sum<FIN_THREADS>(rm, red);
    double R_n = block_sum<FIN_THREADS>(rn, red);
    __syncthreads();
    // cdf_array::normalize (cdf_array.h:62-75): NaNs propagate to the last
    // cumulative entry, and a NaN or zero total makes the CDF all ones
    const double last = cdf[nn - 1];
    const bool flat = isnan(last) || last == 0;
    __syncthreads();   // every thread has read the total before any rewrites it
    for (int j = j0; j < j1; ++j) cdf[j] = flat ? 1.0 : cdf[j] / last;
    if (tid == 0) {
      A.out.L_thermal[z] = L_th;
      A.out.planck_mean[z] = P_m / P_n;
      A.out.rosseland_mean[z] = R_n / R_m;
    }
  }
  if (tid == 0) {
    // gamma-ray opacities with the undecayed composition (transport_opacity.cpp:190-203)
    const int ne_ = A.zs.n_elems;
    double compton = 0, photoion = 0;
    for (int k = 0; k < ne_; k++) {
      double d = A.zs.X[(size_t)z * ne_ + k] * dens;
      double ndens = d / (M_PROTON * A.zs.elem_A[k]);
      compton += ndens * THOMSON * A.zs.elem_Z[k];
      double photo = pow(ALPHA_FS, 4.0) * 4.0 * sqrt(2.0);
      photo *= pow(1.0 * A.zs.elem_Z[k], 5.0);
      photo *= pow(M_E_MEV, 3.5);
      photoion += ndens * 2.0 * THOMSON * photo;
    }
    A.out.compton_opac[z] = compton;
    A.out.photoion_opac[z] = photoion;
    // Fleck factor (transport_opacity.cpp:231-252)
    double eps_imc;
    if (A.p.radiative_eq) eps_imc = 1.;
    else {
      double f_imc = 0.;
      if (A.p.fleck_alpha != 0) {
        double beta = 4.0 * A_RAD * pow(A.zs.T_gas[z], 4) / (A.zs.e_gas[z] * dens);
        double tfac = C_LIGHT * A.out.planck_mean[z] * A.p.dt;
        f_imc = A.p.fleck_alpha * beta * tfac;
      }
      eps_imc = 1.0 / (1.0 + f_imc);
    }
    A.out.eps_imc[z] = eps_imc;
  }
}

__global__ void max_status_kernel(const int32_t* st, int n, int32_t* worst)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && st[i] > 0) atomicMax(worst, st[i]);
}


// ------------------------------------------------------------------
// NLTE-coupled gas state (GasState::solve_state(J_nu) with use_nlte_,
// GasState.cpp:208-263): the n_e Brent iteration of every zone advances
// one evaluation per pass; an evaluation is the LTE state of every atom at
// the zone's trial n_e (K_a), set_rates + batched LU + populations for the
// NLTE atoms of all zones at once, and the charge balance (K_b).
// ------------------------------------------------------------------
struct BrentState {
  double a, b, c, d, e, fa, fb, fc;
  double x;       // the trial n_e of the next evaluation
  double root;
  int iter, phase, err, done;
};

struct NlteArgs {
  sgpu_zone_state zs;
  sgpu_atoms at;
  sgpu_opacity_params p;
  int n_lev, n_ion;
  const int32_t* sys_of;    // [nz][n_atoms]: system of an NLTE atom, -1 for LTE
  double* temp_used;        // [nz]
  BrentState* br;           // [nz]
  double* sys_ne;           // [nsys]
  double* lev_lte;          // [nsys][n_lev]   (global level index)
  const double* sys_levn;   // [nsys][n_lev]   NLTE populations of the last solve
  const double* sys_frac;   // [nsys][n_ion]
  double* levn;             // [nz][n_lev]     state of the zone's last evaluation
  double* frac;             // [nz][n_ion]
  double* nion;             // [nz][n_ion]
  double* upart;            // [nz][n_ion]
  double* ndens;            // [nz][n_atoms]
  double* x_lte;            // [nz][n_atoms]   net ion fraction of the LTE atoms
  double* fval;             // [nz]
  int32_t* active;          // zones still iterating
  sgpu_opacity_out out;
  double* fffac;
  double* phifac;
};

// set-up, the Brent bracket, and the outputs of grey zones
__global__ void __launch_bounds__(WAVE) nlte_init_kernel(NlteArgs A)
{
  __shared__ ZoneLds L;
  const int z = blockIdx.x, lane = threadIdx.x;
  const sgpu_atoms& at = A.at;
  double T;
  const bool gas = zone_setup(L, A.zs, at, A.p, A.n_lev, z, lane, &T);
  BrentState& B = A.br[z];
  if (lane == 0) {
    A.temp_used[z] = T;
    B = BrentState{};
    B.done = !gas;
  }
  if (!gas) {
    for (int k = lane; k < A.n_ion; k += WAVE) {
      A.nion[(size_t)z * A.n_ion + k] = 0;
      A.upart[(size_t)z * A.n_ion + k] = 1;
      A.frac[(size_t)z * A.n_ion + k] = 0;
    }
    for (int k = lane; k < A.n_lev; k += WAVE) A.levn[(size_t)z * A.n_lev + k] = 0;
    // the zone's NLTE systems are assembled and solved with the others but
    // never read: give them finite inputs
    for (int a = 0; a < at.n_atoms; ++a) {
      const int s = A.sys_of[(size_t)z * at.n_atoms + a];
      if (s < 0) continue;
      if (lane == 0) A.sys_ne[s] = 1.0;
      for (int l = at.lev_off[a] + lane; l < at.lev_off[a + 1]; l += WAVE) A.lev_lte[(size_t)s * A.n_lev + l] = 1.0;
    }
    return;
  }
  const double dens = A.zs.rho[z];
  for (int a = lane; a < at.n_atoms; a += WAVE)
    A.ndens[(size_t)z * at.n_atoms + a] = dens * L.mf[a] / (A.zs.elem_A[a] * M_PROTON);
  if (lane == 0) {
    const double x1 = 1e-10 * dens / (L.mu_I * M_PROTON);
    const double x2 = 100 * dens / (L.mu_I * M_PROTON);
    B.a = x1; B.b = x2; B.c = x2;
    B.x = x1;
  }
}

// K_a: the LTE state of every atom at the trial n_e (AtomicSpecies::solve_lte,
// AtomicSpecies_state.cpp:56-123); NLTE atoms hand their LTE populations
// and n_e to the rate assembly
__global__ void __launch_bounds__(WAVE) nlte_lte_kernel(NlteArgs A)
{
  __shared__ ZoneLds L;
  const int z = blockIdx.x, lane = threadIdx.x;
  if (A.br[z].done) return;
  const sgpu_atoms& at = A.at;
  double T;
  zone_setup(L, A.zs, at, A.p, A.n_lev, z, lane, &T);
  const double ne = A.br[z].x;
  for (int a = 0; a < at.n_atoms; ++a) {
    if (lane == 0) saha_chain(L, at, a, ne, T);
    __syncthreads();
    const int l0 = at.lev_off[a], l1 = at.lev_off[a + 1];
    const int i0 = at.ion_off[a], i1 = at.ion_off[a + 1];
    const int s = A.sys_of[(size_t)z * at.n_atoms + a];
    double acc = 0;
    for (int l = l0 + lane; l < l1; l += WAVE) {
      const int ion = at.lev_ion[l];
      double v = L.frac[i0 + ion] * (double)at.lev_g[l] * L.bz[l] / L.U[i0 + ion];
      v = v < 1e-30 ? 1e-30 : v;
      acc += v * ion;
      if (s >= 0) A.lev_lte[(size_t)s * A.n_lev + l] = v;
      else A.levn[(size_t)z * A.n_lev + l] = v;
    }
    const double x = wave_sum(acc);
    const double nd = A.ndens[(size_t)z * at.n_atoms + a];
    for (int k = i0 + lane; k < i1; k += WAVE) {
      A.upart[(size_t)z * A.n_ion + k] = L.U[k];
      if (s < 0) {
        A.frac[(size_t)z * A.n_ion + k] = L.frac[k];
        A.nion[(size_t)z * A.n_ion + k] = nd * L.frac[k];
      }
    }
    if (lane == 0) {
      A.x_lte[(size_t)z * at.n_atoms + a] = x;
      if (s >= 0) A.sys_ne[s] = ne;
    }
    __syncthreads();
  }
}

// K_b: GasState::charge_conservation (GasState.cpp:242-263) with the NLTE
// atoms' populations, which become the zone's state
__global__ void __launch_bounds__(WAVE) nlte_charge_kernel(NlteArgs A)
{
  const int z = blockIdx.x, lane = threadIdx.x;
  if (A.br[z].done) return;
  const sgpu_atoms& at = A.at;
  const double ne = A.br[z].x;
  double f = 0;
  for (int a = 0; a < at.n_atoms; ++a) {
    const int s = A.sys_of[(size_t)z * at.n_atoms + a];
    const double nd = A.ndens[(size_t)z * at.n_atoms + a];
    double x;
    if (s < 0) {
      x = A.x_lte[(size_t)z * at.n_atoms + a];
    } else {
      double acc = 0;
      for (int l = at.lev_off[a] + lane; l < at.lev_off[a + 1]; l += WAVE) {
        const double v = A.sys_levn[(size_t)s * A.n_lev + l];
        acc += v * at.lev_ion[l];
        A.levn[(size_t)z * A.n_lev + l] = v;
      }
      x = wave_sum(acc);
      for (int k = at.ion_off[a] + lane; k < at.ion_off[a + 1]; k += WAVE) {
        const double fr = A.sys_frac[(size_t)s * A.n_ion + k];
        A.frac[(size_t)z * A.n_ion + k] = fr;
        A.nion[(size_t)z * A.n_ion + k] = nd * fr;
      }
    }
    f += nd * x;
  }
  if (lane == 0) A.fval[z] = f - ne;
}

// GasState::ne_brent_method (GasState.cpp:271-337) as a state machine:
// phase 0 has f(x1), phase 1 f(x2), phase 2 f at the new b of an iteration
__device__ void brent_head(BrentState& B)
{
  const double tol = 1e-3, EPS = 3.0e-8;
  if ((B.fb > 0.0 && B.fc > 0.0) || (B.fb < 0.0 && B.fc < 0.0)) {
    B.c = B.a; B.fc = B.fa; B.e = B.d = B.b - B.a;
  }
  if (fabs(B.fc) < fabs(B.fb)) {
    B.a = B.b; B.b = B.c; B.c = B.a;
    B.fa = B.fb; B.fb = B.fc; B.fc = B.fa;
  }
  const double tol1 = 2.0 * EPS * fabs(B.b) + 0.5 * tol;
  const double xm = 0.5 * (B.c - B.b);
  if (fabs(xm) <= tol1 || B.fb == 0.0) { B.root = B.b; B.done = 1; return; }
  if (fabs(B.e) >= tol1 && fabs(B.fa) > fabs(B.fb)) {
    double pp, q, r;
    const double s = B.fb / B.fa;
    if (B.a == B.c) { pp = 2.0 * xm * s; q = 1.0 - s; }
    else {
      q = B.fa / B.fc; r = B.fb / B.fc;
      pp = s * (2.0 * xm * q * (q - r) - (B.b - B.a) * (r - 1.0));
      q = (q - 1.0) * (r - 1.0) * (s - 1.0);
    }
    if (pp > 0.0) q = -q;
    pp = fabs(pp);
    const double min1 = 3.0 * xm * q - fabs(tol1 * q);
    const double min2 = fabs(B.e * q);
    if (2.0 * pp < (min1 < min2 ? min1 : min2)) { B.e = B.d; B.d = pp / q; }
    else { B.d = xm; B.e = B.d; }
  } else { B.d = xm; B.e = B.d; }
  B.a = B.b; B.fa = B.fb;
  if (fabs(B.d) > tol1) B.b += B.d;
  else B.b += (xm >= 0.0 ? fabs(tol1) : -fabs(tol1));
  B.x = B.b;
  B.phase = 2;
}

__global__ void nlte_brent_kernel(BrentState* br, const double* fval, int nz, int32_t* active)
{
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  BrentState& B = br[z];
  if (B.done) return;
  const double f = fval[z];
  if (B.phase == 0) {
    B.fa = f;
    B.x = B.b;
    B.phase = 1;
  } else if (B.phase == 1) {
    B.fb = f;
    if ((B.fa > 0.0 && B.fb > 0.0) || (B.fa < 0.0 && B.fb < 0.0)) B.err = 1;
    B.fc = B.fb;
    B.iter = 1;
    brent_head(B);
  } else {
    B.fb = f;
    if (++B.iter > 100) { B.err = 2; B.root = 0.0; B.done = 1; }
    else brent_head(B);
  }
  if (!B.done) atomicAdd(active, 1);
}

// outputs of the converged state, and the free-free and bound-free factors
// of the bins kernel from it (the zone's last evaluation)
__global__ void __launch_bounds__(WAVE) nlte_finish_kernel(NlteArgs A)
{
  const int z = blockIdx.x, lane = threadIdx.x;
  const sgpu_atoms& at = A.at;
  const BrentState& B = A.br[z];
  const bool gas = !(A.zs.grey_opacity && A.zs.grey_opacity[z] != 0) && at.n_atoms > 0;
  for (int k = lane; k < A.n_ion; k += WAVE) {
    if (A.out.ion_frac) A.out.ion_frac[(size_t)z * A.n_ion + k] = gas ? A.frac[(size_t)z * A.n_ion + k] : 0;
    if (A.out.ion_part) A.out.ion_part[(size_t)z * A.n_ion + k] = gas ? A.upart[(size_t)z * A.n_ion + k] : 0;
  }
  if (A.out.lev_n)
    for (int k = lane; k < A.n_lev; k += WAVE) A.out.lev_n[(size_t)z * A.n_lev + k] = gas ? A.levn[(size_t)z * A.n_lev + k] : 0;
  const double T = A.temp_used[z];
  const double ne = gas ? B.root : 0.0;
  if (lane == 0) {
    A.out.n_elec[z] = ne;
    if (A.out.zone_status) A.out.zone_status[z] = gas ? B.err : 0;
  }
  if (!gas) {
    if (lane == 0 && A.fffac) A.fffac[z] = 0;
    return;
  }
  if (A.phifac) {
    const double lam_t = sqrt(H_PLANCK * H_PLANCK / (2 * PI * M_ELECTRON * K_BOLTZ * T));
    for (int a = 0; a < at.n_atoms; ++a) {
      const double n_dens = A.ndens[(size_t)z * at.n_atoms + a];
      for (int l = at.lev_off[a] + lane; l < at.lev_off[a + 1]; l += WAVE) {
        const int kc = continuum_level(at, a, l);
        double ph = -1;
        if (kc != -1) {
          const double nc = n_dens * A.levn[(size_t)z * A.n_lev + kc];
          const double gl_o_gc = (1.0 * at.lev_g[l]) / (1.0 * at.lev_g[kc]);
          ph = nc * gl_o_gc / 2. * lam_t * lam_t * lam_t;
        }
        A.phifac[(size_t)z * A.n_lev + l] = ph;
      }
    }
  }
  if (A.fffac && lane == 0) {
    double fac = 0;
    for (int a = 0; a < at.n_atoms; ++a) {
      double Z_eff_sq = 0;
      const int i0 = at.ion_off[a], ni = at.ion_off[a + 1] - i0;
      for (int j = 0; j < ni; j++) Z_eff_sq += A.frac[(size_t)z * A.n_ion + i0 + j] * j * j;
      fac += A.ndens[(size_t)z * at.n_atoms + a] * Z_eff_sq;
    }
    A.fffac[z] = fac * (3.7e8 * pow(T, -0.5) * ne);
  }
}

__global__ void gather_zone_kernel(const double* src, const int32_t* idx, int n, double* dst)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) dst[s] = src[idx[s]];
}

__global__ void zero_i32_kernel(int32_t* p) { p[0] = 0; }

}  // namespace sgpu

using namespace sgpu;

extern "C" int sgpu_opacity_build(sgpu_ctx* ctx, const sgpu_locate* nu_grid,
                                  const sgpu_zone_state* zones, const sgpu_atoms* atoms,
                                  const sgpu_opacity_params* params, sgpu_opacity_out* out)
{
  SGPU_REQUIRE(ctx && nu_grid && zones && params && out, "null argument");
  SGPU_REQUIRE(out->abs && out->scat && out->emis_cdf && out->L_thermal && out->planck_mean &&
               out->rosseland_mean && out->n_elec && out->compton_opac && out->photoion_opac &&
               out->eps_imc, "missing opacity output array");
  SGPU_REQUIRE(nu_grid->n >= 1 && nu_grid->x, "bad frequency grid");
  const int nz = zones->n_zones;
  SGPU_REQUIRE(nz >= 0, "negative zone count");
  if (nz == 0) return SGPU_OK;
  SGPU_REQUIRE(zones->n_elems >= 1 && zones->n_elems <= MAX_ELEMS, "n_elems %d outside [1, %d]",
               zones->n_elems, MAX_ELEMS);
  SGPU_REQUIRE(zones->rho && zones->T_gas && zones->X && zones->elem_Z && zones->elem_A,
               "missing zone arrays");
  SGPU_REQUIRE(params->fleck_alpha == 0 || params->radiative_eq || zones->e_gas,
               "fleck_alpha != 0 needs e_gas");
  sgpu_atoms at{};
  if (atoms) at = *atoms;
  if (at.n_atoms > 0) {
    SGPU_REQUIRE(at.n_atoms == zones->n_elems, "one atom per element required (%d atoms, %d elements)",
                 at.n_atoms, zones->n_elems);
    SGPU_REQUIRE(at.max_levels >= 1, "max_levels must be set");
  }
  int n_lev = 0, n_ion = 0;
  if (at.n_atoms > 0) {
    n_lev = at.n_levels; n_ion = at.n_ions;
    SGPU_REQUIRE(n_lev >= at.n_atoms && n_ion >= at.n_atoms, "n_levels / n_ions not set");
    SGPU_REQUIRE(n_lev <= MAX_LEV && n_ion <= MAX_ION, "atomic data too large: %d levels (max %d), %d ions (max %d)",
                 n_lev, MAX_LEV, n_ion, MAX_ION);
  }
  int st = 0;
  size_t nion_sz = (size_t)nz * (n_ion > 0 ? n_ion : 1);
  double* temp_used = (double*)scratch(ctx, "op_temp", sizeof(double) * nz, &st);
  if (st) return st;
  double* nion = (double*)scratch(ctx, "op_nion", sizeof(double) * nion_sz, &st);
  if (st) return st;
  double* upart = (double*)scratch(ctx, "op_upart", sizeof(double) * nion_sz, &st);
  if (st) return st;
  int32_t* worst = (int32_t*)scratch(ctx, "op_worst", sizeof(int32_t), &st);
  if (st) return st;
  const bool lexp = params->use_line_expansion && at.n_atoms > 0 && at.n_blines > 0;
  const bool bf = params->use_bound_free && at.n_atoms > 0;
  const bool bb = params->use_bound_bound && at.n_atoms > 0 && at.n_blines > 0;
  double* levn = nullptr;
  double* ndens = nullptr;
  double* fffac = nullptr;
  double* phifac = nullptr;
  if (lexp)
    SGPU_REQUIRE(at.bline_l && at.bline_u && at.bline_nu && at.bline_f && at.bline_bin && at.bline_bin_off,
                 "line expansion needs the level-resolved line arrays");
  if (bb)
    SGPU_REQUIRE(at.bline_l && at.bline_u && at.bline_nu && at.bline_A && at.bline_bin_off,
                 "bound-bound opacity needs the level-resolved line arrays");
  if (lexp || bf || bb) {
    levn = (double*)scratch(ctx, "op_levn", sizeof(double) * (size_t)nz * n_lev, &st);
    if (st) return st;
    ndens = (double*)scratch(ctx, "op_ndens", sizeof(double) * (size_t)nz * at.n_atoms, &st);
    if (st) return st;
  }
  if (bf) {
    phifac = (double*)scratch(ctx, "op_phifac", sizeof(double) * (size_t)nz * n_lev, &st);
    if (st) return st;
  }
  if (params->use_free_free) {
    fffac = (double*)scratch(ctx, "op_fffac", sizeof(double) * nz, &st);
    if (st) return st;
  }

  timer_begin(ctx, "opacity");
  StateArgs sa{*zones, at, *params, n_lev, n_ion, temp_used, nion, upart, *out, levn, ndens, fffac, phifac};
  hipLaunchKernelGGL(state_kernel, dim3(nz), dim3(WAVE), 0, ctx->stream, sa);
  SGPU_LAUNCH_CHECK();
  BinArgs ba{*nu_grid, *zones, at, *params, n_ion, temp_used, nion, upart,
             out->abs, out->scat, out->emis_cdf, out->n_elec, levn, ndens, n_lev, fffac, phifac};
  hipLaunchKernelGGL(bins_kernel, dim3((nu_grid->n + 255) / 256, nz), dim3(256), 0, ctx->stream, ba);
  SGPU_LAUNCH_CHECK();
  FinArgs fa{*nu_grid, *zones, *params, temp_used, *out};
  hipLaunchKernelGGL(finalize_kernel, dim3(nz), dim3(FIN_THREADS), 0, ctx->stream, fa);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "opacity");
  if (!out->zone_status) return SGPU_OK;
  SGPU_HIP(hipMemsetAsync(worst, 0, sizeof(int32_t), ctx->stream));
  hipLaunchKernelGGL(max_status_kernel, dim3((nz + 255) / 256), dim3(256), 0, ctx->stream,
                     out->zone_status, nz, worst);
  SGPU_LAUNCH_CHECK();
  int32_t hw = 0;
  SGPU_HIP(hipMemcpyAsync(&hw, worst, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  return hw;
}

extern "C" int sgpu_opacity_build_nlte(sgpu_ctx* ctx, const sgpu_locate* nu_grid, const sgpu_zone_state* zones,
                                       const sgpu_atoms* atoms, const sgpu_opacity_params* params,
                                       const double* J_nu, const int32_t* atom_nlte, int32_t use_collisions,
                                       int32_t lu_precision, sgpu_opacity_out* out, int32_t* n_evals)
{
  SGPU_REQUIRE(ctx && nu_grid && zones && params && out && atoms && atom_nlte, "null argument");
  SGPU_REQUIRE(lu_precision == SGPU_NLTE_FP64 || lu_precision == SGPU_NLTE_FP32_REFINE,
               "unknown LU precision %d", lu_precision);
  const sgpu_atoms& at = *atoms;
  int n_nlte = 0;
  if (at.n_atoms > 0)
    for (int a = 0; a < at.n_atoms; ++a) n_nlte += atom_nlte[a] != 0;
  if (n_evals) *n_evals = 0;
  // transport_opacity.cpp:30-36: LTE on the first step; no NLTE atom: LTE
  if (params->first_step || n_nlte == 0) return sgpu_opacity_build(ctx, nu_grid, zones, atoms, params, out);
  SGPU_REQUIRE(J_nu, "J_nu missing");
  SGPU_REQUIRE(out->abs && out->scat && out->emis_cdf && out->L_thermal && out->planck_mean &&
               out->rosseland_mean && out->n_elec && out->compton_opac && out->photoion_opac &&
               out->eps_imc, "missing opacity output array");
  SGPU_REQUIRE(nu_grid->n >= 1 && nu_grid->x, "bad frequency grid");
  const int nz = zones->n_zones;
  SGPU_REQUIRE(nz >= 0, "negative zone count");
  if (nz == 0) return SGPU_OK;
  SGPU_REQUIRE(zones->n_elems >= 1 && zones->n_elems <= MAX_ELEMS, "n_elems %d outside [1, %d]",
               zones->n_elems, MAX_ELEMS);
  SGPU_REQUIRE(zones->rho && zones->T_gas && zones->X && zones->elem_Z && zones->elem_A,
               "missing zone arrays");
  SGPU_REQUIRE(params->fleck_alpha == 0 || params->radiative_eq || zones->e_gas,
               "fleck_alpha != 0 needs e_gas");
  SGPU_REQUIRE(at.n_atoms == zones->n_elems, "one atom per element required (%d atoms, %d elements)",
               at.n_atoms, zones->n_elems);
  SGPU_REQUIRE(at.max_levels >= 1, "max_levels must be set");
  const int n_lev = at.n_levels, n_ion = at.n_ions;
  SGPU_REQUIRE(n_lev >= at.n_atoms && n_ion >= at.n_atoms, "n_levels / n_ions not set");
  SGPU_REQUIRE(n_lev <= MAX_LEV && n_ion <= MAX_ION, "atomic data too large: %d levels (max %d), %d ions (max %d)",
               n_lev, MAX_LEV, n_ion, MAX_ION);
  const bool lexp = params->use_line_expansion && at.n_blines > 0;
  if (lexp)
    SGPU_REQUIRE(at.bline_l && at.bline_u && at.bline_nu && at.bline_f && at.bline_bin && at.bline_bin_off,
                 "line expansion needs the level-resolved line arrays");
  if (params->use_bound_bound && at.n_blines > 0)
    SGPU_REQUIRE(at.bline_l && at.bline_u && at.bline_nu && at.bline_A && at.bline_bin_off,
                 "bound-bound opacity needs the level-resolved line arrays");

  // (zone, NLTE atom) systems, zone-major
  std::vector<int32_t> h_atom, h_zone, h_n, h_sys_of((size_t)nz * at.n_atoms, -1);
  std::vector<int64_t> h_roff(1, 0), h_voff(1, 0);
  std::vector<int32_t> lev_off(at.n_atoms + 1);
  SGPU_HIP(hipMemcpyAsync(lev_off.data(), at.lev_off, sizeof(int32_t) * (at.n_atoms + 1), hipMemcpyDeviceToHost,
                          ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  for (int z = 0; z < nz; ++z)
    for (int a = 0; a < at.n_atoms; ++a) {
      if (!atom_nlte[a]) continue;
      const int n = lev_off[a + 1] - lev_off[a];
      h_sys_of[(size_t)z * at.n_atoms + a] = (int32_t)h_atom.size();
      h_atom.push_back(a);
      h_zone.push_back(z);
      h_n.push_back(n);
      h_roff.push_back(h_roff.back() + (int64_t)n * n);
      h_voff.push_back(h_voff.back() + n);
    }
  const int nsys = (int)h_atom.size();
  const size_t nbl = at.n_blines > 0 ? (size_t)at.n_blines : 1;
  int st = 0;
  auto dbl = [&](const char* name, size_t n) { return (double*)scratch(ctx, name, sizeof(double) * (n ? n : 1), &st); };
  auto i32 = [&](const char* name, size_t n) { return (int32_t*)scratch(ctx, name, sizeof(int32_t) * (n ? n : 1), &st); };
  int32_t* d_atom = i32("opn_sys_atom", nsys);
  int32_t* d_zone = i32("opn_sys_zone", nsys);
  int32_t* d_n = i32("opn_sys_n", nsys);
  int32_t* d_sys_of = i32("opn_sys_of", (size_t)nz * at.n_atoms);
  int64_t* d_roff = (int64_t*)scratch(ctx, "opn_roff", sizeof(int64_t) * (nsys + 1), &st);
  int64_t* d_voff = (int64_t*)scratch(ctx, "opn_voff", sizeof(int64_t) * (nsys + 1), &st);
  double* temp_used = dbl("opn_temp", nz);
  double* nion = dbl("opn_nion", (size_t)nz * n_ion);
  double* upart = dbl("opn_upart", (size_t)nz * n_ion);
  double* frac = dbl("opn_frac", (size_t)nz * n_ion);
  double* levn = dbl("opn_levn", (size_t)nz * n_lev);
  double* ndens = dbl("opn_ndens", (size_t)nz * at.n_atoms);
  double* x_lte = dbl("opn_xlte", (size_t)nz * at.n_atoms);
  double* fval = dbl("opn_fval", nz);
  BrentState* br = (BrentState*)scratch(ctx, "opn_brent", sizeof(BrentState) * nz, &st);
  double* sys_T = dbl("opn_sys_T", nsys);
  double* sys_ne = dbl("opn_sys_ne", nsys);
  double* lev_lte = dbl("opn_lev_lte", (size_t)nsys * n_lev);
  double* Pic = dbl("opn_pic", (size_t)nsys * n_lev);
  double* Rci = dbl("opn_rci", (size_t)nsys * n_lev);
  double* line_J = dbl("opn_line_J", (size_t)nsys * nbl);
  double* rates = dbl("opn_rates", (size_t)h_roff.back());
  double* lte = dbl("opn_lte", (size_t)h_voff.back());
  double* bvec = dbl("opn_b", (size_t)h_voff.back());
  double* sys_levn = dbl("opn_sys_levn", (size_t)nsys * n_lev);
  double* sys_frac = dbl("opn_sys_frac", (size_t)nsys * n_ion);
  int32_t* active = i32("opn_active", 1);
  int32_t* worst = i32("opn_worst", 1);
  double* fffac = params->use_free_free ? dbl("opn_fffac", nz) : nullptr;
  double* phifac = params->use_bound_free ? dbl("opn_phifac", (size_t)nz * n_lev) : nullptr;
  if (st) return st;
  SGPU_HIP(hipMemcpyAsync(d_atom, h_atom.data(), sizeof(int32_t) * nsys, hipMemcpyHostToDevice, ctx->stream));
  SGPU_HIP(hipMemcpyAsync(d_zone, h_zone.data(), sizeof(int32_t) * nsys, hipMemcpyHostToDevice, ctx->stream));
  SGPU_HIP(hipMemcpyAsync(d_n, h_n.data(), sizeof(int32_t) * nsys, hipMemcpyHostToDevice, ctx->stream));
  SGPU_HIP(hipMemcpyAsync(d_sys_of, h_sys_of.data(), sizeof(int32_t) * h_sys_of.size(), hipMemcpyHostToDevice,
                          ctx->stream));
  SGPU_HIP(hipMemcpyAsync(d_roff, h_roff.data(), sizeof(int64_t) * (nsys + 1), hipMemcpyHostToDevice, ctx->stream));
  SGPU_HIP(hipMemcpyAsync(d_voff, h_voff.data(), sizeof(int64_t) * (nsys + 1), hipMemcpyHostToDevice, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));

  timer_begin(ctx, "opacity");
  NlteArgs A{*zones, at, *params, n_lev, n_ion, d_sys_of, temp_used, br, sys_ne, lev_lte, sys_levn, sys_frac,
             levn, frac, nion, upart, ndens, x_lte, fval, active, *out, fffac, phifac};
  hipLaunchKernelGGL(nlte_init_kernel, dim3(nz), dim3(WAVE), 0, ctx->stream, A);
  hipLaunchKernelGGL(gather_zone_kernel, dim3((nsys + 255) / 256), dim3(256), 0, ctx->stream, temp_used, d_zone,
                     nsys, sys_T);
  SGPU_LAUNCH_CHECK();
  // AtomicSpecies::calculate_radiative_rates once per zone (GasState.cpp:223)
  st = sgpu_nlte_radiative_rates(ctx, atoms, nu_grid, nsys, d_atom, d_zone, sys_T, zones->elem_A,
                                 params->line_velocity_width, J_nu, Pic, Rci, line_J);
  if (st < 0) return st;
  int evals = 0;
  for (; evals < 102; ++evals) {
    hipLaunchKernelGGL(nlte_lte_kernel, dim3(nz), dim3(WAVE), 0, ctx->stream, A);
    SGPU_LAUNCH_CHECK();
    st = sgpu_nlte_set_rates(ctx, atoms, nsys, d_atom, sys_ne, sys_T, line_J, Pic, Rci, lev_lte, use_collisions,
                             params->no_ground_recomb, d_n, d_roff, d_voff, rates, lte);
    if (st < 0) return st;
    st = sgpu_nlte_batched(ctx, nsys, d_n, d_roff, d_voff, rates, lte, bvec, nullptr, lu_precision, 1e-13, nullptr);
    if (st < 0) return st;
    st = sgpu_nlte_populations(ctx, atoms, nsys, d_atom, d_voff, bvec, lte, sys_levn, sys_frac);
    if (st < 0) return st;
    hipLaunchKernelGGL(nlte_charge_kernel, dim3(nz), dim3(WAVE), 0, ctx->stream, A);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, ctx->stream, active);
    hipLaunchKernelGGL(nlte_brent_kernel, dim3((nz + 255) / 256), dim3(256), 0, ctx->stream, br, fval, nz, active);
    SGPU_LAUNCH_CHECK();
    int32_t h_active = 0;
    SGPU_HIP(hipMemcpyAsync(&h_active, active, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    if (h_active == 0) { ++evals; break; }
  }
  if (n_evals) *n_evals = evals;
  hipLaunchKernelGGL(nlte_finish_kernel, dim3(nz), dim3(WAVE), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  BinArgs ba{*nu_grid, *zones, at, *params, n_ion, temp_used, nion, upart,
             out->abs, out->scat, out->emis_cdf, out->n_elec, levn, ndens, n_lev, fffac, phifac};
  hipLaunchKernelGGL(bins_kernel, dim3((nu_grid->n + 255) / 256, nz), dim3(256), 0, ctx->stream, ba);
  SGPU_LAUNCH_CHECK();
  FinArgs fa{*nu_grid, *zones, *params, temp_used, *out};
  hipLaunchKernelGGL(finalize_kernel, dim3(nz), dim3(FIN_THREADS), 0, ctx->stream, fa);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "opacity");
  if (!out->zone_status) return SGPU_OK;
  hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, ctx->stream, worst);
  hipLaunchKernelGGL(max_status_kernel, dim3((nz + 255) / 256), dim3(256), 0, ctx->stream,
                     out->zone_status, nz, worst);
  SGPU_LAUNCH_CHECK();
  int32_t hw = 0;
  SGPU_HIP(hipMemcpyAsync(&hw, worst, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  return hw;
}
