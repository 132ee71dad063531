// transport.hip -- K4/K5/K6/K7: packet emission, propagation, compaction,
// tally wipe / normalisation and the homologous expansion step.
//
// Propagation (K4) runs one packet per lane.  Three forms share the
// per-event code (sgpu_geometry.h):
//   SGPU_PREC_FP64   the plain drop-in: one thread per packet, FP64 geometry,
//                    every tally a global FP64 atomic, as the reference's
//                    omp-atomic loop (src/transport/transport.cpp:57-88,
//                    260-484).
//   SGPU_PREC_FP64_BATCHED  the same FP64 arithmetic (the same packet
//                    paths), launched as the mixed form: persistent lanes,
//                    zone tallies summed in registers over a zone visit and
//                    flushed to LDS, J_nu summed per (zone, bin) run.
//   SGPU_PREC_MIXED  persistent lanes that refill themselves with the next
//                    packet when theirs finishes (no lane idles on a
//                    long-lived neighbour), FP32 geometry and Doppler shifts,
//                    packed FP32 (kappa, eps) opacity table (one 8-byte
//                    gather per event), Walker alias tables for the
//                    emissivity, zone tallies summed in registers per zone
//                    visit into LDS, J_nu and spectra as FP64 global atomics.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>

#include "sgpu_internal.h"
#include "sgpu_geometry.h"

namespace sgpu {

constexpr int PROP_THREADS = 256;
constexpr int LDS_ZONE_MAX = 1024;   // zone tallies in LDS up to this many zones

struct Counts {
  uint32_t events, binshift, boundary, scatter, redist, escaped, absorbed, stopped, moving;
};

template <class R>
struct Tab;

template <>
struct Tab<double> {
  const double* abs;
  const double* scat;
  const double* cdf;
};

// one column of a zone's Walker alias table: column k keeps bin k with
// probability q, else gives bin a
struct __align__(8) AliasEntry {
  float q;
  int32_t a;
};

template <>
struct Tab<float> {
  const float2* ke;          // (kappa = abs + scat, eps = abs / kappa)
  const AliasEntry* alias;   // [nz][n_nu] alias tables of the emissivity distributions
};

template <class R>
struct PropArgs {
  GridView<R> g;
  Axis<R> nu;
  Tab<R> tab;
  const double* eps_imc;
  const double* compton;
  const double* photoion;
  sgpu_transport_params p;
  sgpu_packets pk;
  int64_t n;
  const int64_t* n_dev;   // when set: the packet count, read on the device
  sgpu_tallies ta;
  sgpu_counters* cnt;
  int lds_tallies;   // zone scalars accumulated in LDS
};

// The propagation kernel's argument block in kernarg memory (it is the
// kernel's only argument), through an opaque copy of the segment pointer:
// the compiler cannot hoist loads through it out of the event loop, so the
// cold fields (packet arrays, spectra, gamma-ray coefficients) are fetched
// with scalar loads where they are used instead of being held -- and
// spilled -- in SGPRs across the loop.
template <class R>
using KernargArgs = const __attribute__((address_space(4))) PropArgs<R>*;

template <class R>
__device__ __forceinline__ KernargArgs<R> cold_args()
{
  KernargArgs<R> p = (KernargArgs<R>)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return p;
}

// member-wise copies out of kernarg memory (the implicit copy constructors
// take generic references)
#define SGPU_KA __attribute__((address_space(4)))
__device__ __forceinline__ sgpu_packets ka_packets(const SGPU_KA sgpu_packets& k)
{
  sgpu_packets r;
  r.capacity = k.capacity;
  r.x = k.x; r.y = k.y; r.z = k.z; r.Dx = k.Dx; r.Dy = k.Dy; r.Dz = k.Dz;
  r.t = k.t; r.e = k.e; r.nu = k.nu; r.ind = k.ind; r.type = k.type; r.fate = k.fate;
  r.draw = k.draw; r.id = k.id;
  return r;
}

__device__ __forceinline__ sgpu_locate ka_locate(const SGPU_KA sgpu_locate& l)
{
  sgpu_locate r;
  r.n = l.n; r.type = l.type; r.min = l.min; r.del = l.del; r.x = l.x;
  return r;
}

__device__ __forceinline__ sgpu_spectrum ka_spectrum(const SGPU_KA sgpu_spectrum& s)
{
  sgpu_spectrum r;
  r.time_grid = ka_locate(s.time_grid); r.nu_grid = ka_locate(s.nu_grid);
  r.mu_grid = ka_locate(s.mu_grid); r.phi_grid = ka_locate(s.phi_grid);
  r.flux = s.flux; r.click = s.click;
  return r;
}

// dynamic LDS of the propagation and emission kernels
extern __shared__ double dyn_lds[];

// zone-scalar tally sink: LDS (per workgroup, ds_add_f64 on dyn_lds) or
// global FP64 atomics.  Fields are addressed by (field, zone).
enum { T_EABS = 0, T_FX, T_FY, T_FZ, T_FR, T_LDEP, T_NFIELDS };

// tally array of a field index known only at run time (a switch, so that the
// pointer table is not spilled to scratch memory for a dynamic index)
__device__ __forceinline__ double* tally_field(const sgpu_tallies& t, int f)
{
  switch (f) {
    case T_EABS: return t.e_abs;
    case T_FX: return t.fx_rad;
    case T_FY: return t.fy_rad;
    case T_FZ: return t.fz_rad;
    case T_FR: return t.fr_rad;
    default: return t.L_radio_dep;
  }
}

struct ZoneSink {
  double* g[T_NFIELDS];
  int nz;
  bool lds;
  // accumulate the per-event zone tallies in registers while the packet stays
  // in one zone and flush them when it leaves (the mixed form): a packet
  // random-walking through an opaque zone then costs one atomic per field
  // per zone visit instead of per event
  bool accum;
  int zone;
  double acc[5];
  // J_nu entry accumulated likewise while (zone, bin) is unchanged
  double* J;
  int64_t jkey;
  double jacc;
  // e_abs of an event is (its absorbed energy) * eps_imc[zone]: applied per
  // event in the direct form (the reference's operation order), once per
  // zone visit at the flush in the accumulating form
  const double* eps_imc;
  __device__ __forceinline__ void add(int field, int zone_, double v) const
  {
    if (lds) atomicAdd(&dyn_lds[field * nz + zone_], v);
    else atomic_add_f64(&g[field][zone_], v);
  }
  __device__ __forceinline__ void jnu(int64_t k, double v)
  {
    if (!accum) { atomic_add_f64(&J[k], v); return; }
    if (k != jkey) {
      if (jkey >= 0) atomic_add_f64(&J[jkey], jacc);
      jkey = k;
      jacc = 0;
    }
    jacc += v;
  }
  __device__ __forceinline__ void flush()
  {
    if (zone >= 0) {
      acc[T_EABS] *= eps_imc[zone];
#pragma unroll
      for (int f = 0; f < 5; f++) {
        if (acc[f] != 0) add(f, zone, acc[f]);
        acc[f] = 0;
      }
    }
    zone = -1;
    if (jkey >= 0) atomic_add_f64(&J[jkey], jacc);
    jkey = -1;
    jacc = 0;
  }
  __device__ __forceinline__ void event(int z, bool photon, double eabs, double fx, double fy, double fz,
                                        double fr)
  {
    if (!accum) {
      if (photon) add(T_EABS, z, eabs * eps_imc[z]);
      add(T_FX, z, fx);
      add(T_FY, z, fy);
      add(T_FZ, z, fz);
      add(T_FR, z, fr);
      return;
    }
    if (z != zone) {
      flush();
      zone = z;
    }
    if (photon) acc[T_EABS] += eabs;
    acc[T_FX] += fx;
    acc[T_FY] += fy;
    acc[T_FZ] += fz;
    acc[T_FR] += fr;
  }
};

// transport::get_opacity (transport_opacity.cpp:268-299)
__device__ __forceinline__ int get_opacity(const PropArgs<double>& A, const Pkt<double>& p,
                                           double dshift, double* opac, double* eps)
{
  double nu = p.nu * dshift;
  int i_nu = 0;
  if (p.type == SGPU_PHOTON) {
    i_nu = ax_within(A.nu, nu);
    size_t k = (size_t)p.ind * A.nu.n + i_nu;
    double a = A.tab.abs[k];
    double s = 0;
    if (!A.p.omit_scattering) s = A.tab.scat[k];
    *opac = a + s;
    *eps = (*opac == 0) ? 0 : a / *opac;
  } else {
    const double* compton = cold_args<double>()->compton;
    const double* photoion = cold_args<double>()->photoion;
    double co = compton[p.ind] * klein_nishina(p.nu);
    double po = photoion[p.ind] * pow(p.nu, -3.5);
    *opac = co + po;
    *eps = po / (co + po);
  }
  return i_nu;
}

__device__ __forceinline__ int get_opacity(const PropArgs<float>& A, const Pkt<float>& p,
                                           float dshift, float* opac, float* eps)
{
  float nu = p.nu * dshift;
  int i_nu = 0;
  if (p.type == SGPU_PHOTON) {
    i_nu = ax_within(A.nu, nu);
    float2 ke = A.tab.ke[(size_t)p.ind * A.nu.n + i_nu];
    *opac = ke.x;
    *eps = ke.y;
  } else {
    const double* compton = cold_args<float>()->compton;
    const double* photoion = cold_args<float>()->photoion;
    float co = (float)compton[p.ind] * klein_nishina(p.nu);
    float po = (float)photoion[p.ind] * __powf(p.nu, -3.5f);
    *opac = co + po;
    *eps = po / (co + po);
  }
  return i_nu;
}

// transport::sample_photon_frequency (emission.cpp:27-46)
__device__ __forceinline__ void sample_photon_frequency(const PropArgs<double>& A, Pkt<double>& p, Counts& c)
{
  if (p.type == SGPU_PHOTON) {
    const int nn = A.nu.n;
    int inu = min(cdf_sample(A.tab.cdf + (size_t)p.ind * nn, nn, p.rng.uniform()), nn - 1);
    p.nu = ax_sample(A.nu, inu, p.rng.uniform());
    c.redist++;
  } else {
    p.nu = 1;
  }
}

// FP32 form: the same distribution over bins -- p_k = cdf_k - cdf_{k-1}, the
// mass upper_bound(cdf, u) gives bin k -- sampled with a Walker alias table
// (Walker 1977; Vose 1991): a uniform column k and a uniform coin against q_k,
// one 8-byte load instead of a binary search over log2(n_nu) dependent loads
__device__ __forceinline__ void sample_photon_frequency(const PropArgs<float>& A, Pkt<float>& p, Counts& c)
{
  if (p.type == SGPU_PHOTON) {
    const int nn = A.nu.n;
    int k;
    float u_coin;
    AliasEntry e;
    if (p.pre.valid) {          // drawn and loaded when the event started
      k = p.pre.k;
      u_coin = p.pre.coin;
      e.q = p.pre.q;
      e.a = p.pre.a;
      p.pre.valid = false;
    } else {
      const float u_col = p.rng.uniform_f();
      u_coin = p.rng.uniform_f();
      k = min((int)(u_col * (float)nn), nn - 1);
      e = A.tab.alias[(size_t)p.ind * nn + k];
    }
    const int inu = u_coin < e.q ? k : e.a;
    p.nu = ax_sample(A.nu, inu, p.rng.uniform_f());
    c.redist++;
  } else {
    p.nu = 1;
  }
}

// grid_1D_sphere::sample_in_zone (grid_1D_sphere.cpp:681-700),
// grid_2D_cyln::sample_in_zone (grid_2D_cyln.cpp:528-541)
__device__ __forceinline__ void sample_in_zone(const GridView<double>& g, int i, const double* ran, double* r)
{
  if (g.kind == SGPU_GRID_1D_SPHERE) {
    double r_0 = (i == 0) ? g.r_out.min : g.r_out.x[i - 1];
    double ro = g.r_out.x[i];
    double r_samp = pow(r_0 * r_0 * r_0 + ran[0] * (ro * ro * ro - r_0 * r_0 * r_0), 1.0 / 3.0);
    double mu = 1 - 2.0 * ran[1];
    double phi = 2.0 * PI * ran[2];
    double sin_theta = sqrt(1 - mu * mu);
    r[0] = r_samp * sin_theta * cos(phi);
    r[1] = r_samp * sin_theta * sin(phi);
    r[2] = r_samp * mu;
  } else {
    const int nz = g.z_out.n;
    const int ix = i / nz, iz = i - ix * nz;
    double phi = 2 * PI * ran[1];
    double p_in = ax_left(g.x_out, ix);
    double p_out = g.x_out.x[ix];
    double p_samp = sqrt(p_in * p_in + ran[0] * (p_out * p_out - p_in * p_in));
    r[0] = p_samp * cos(phi);
    r[1] = p_samp * sin(phi);
    r[2] = ax_left(g.z_out, iz) + ran[2] * ax_delta(g.z_out, iz);
  }
}

// transport::isotropic_scatter (scatter.cpp:303-351); D divided by |D|
template <class R>
__device__ __forceinline__ void isotropic_scatter(const PropArgs<R>& A, Pkt<R>& p, int redist, Counts& c)
{
  const GridView<R>& g = A.g;
  const R cl = (R)C_LIGHT;
  R V[3] = {qdiv(p.x[0], g.t_now), qdiv(p.x[1], g.t_now), qdiv(p.x[2], g.t_now)};
  R vdotD = V[0] * p.D[0] + V[1] * p.D[1] + V[2] * p.D[2];
  R beta2 = qdiv(qdiv(V[0] * V[0] + V[1] * V[1] + V[2] * V[2], cl), cl);
  R gamma = inv_sqrt(1 - beta2);
  R dshift_in = gamma * (1 - qdiv(vdotD, cl));
  p.e = p.e * (double)dshift_in;
  p.nu = p.nu * dshift_in;
  R Dn[3];
  isotropic_dir(p, Dn);
  if (redist) sample_photon_frequency(A, p, c);
  for (int i = 0; i < 3; i++) V[i] = -1 * V[i];
  R vdp = (Dn[0] * V[0] + Dn[1] * V[1] + Dn[2] * V[2]);
  R vd_out = gamma * (1 - qdiv(vdp, cl));
#pragma unroll
  for (int k = 0; k < 3; k++)
    p.D[k] = qdiv((R)1, vd_out) *
             (Dn[k] - qdiv(gamma * V[k], cl) * (1 - qdiv(qdiv(gamma * vdp, cl), gamma + 1)));
  R norm = sqrt_(p.D[0] * p.D[0] + p.D[1] * p.D[1] + p.D[2] * p.D[2]);
  p.D[0] = qdiv(p.D[0], norm); p.D[1] = qdiv(p.D[1], norm); p.D[2] = qdiv(p.D[2], norm);
  p.e = p.e * (double)vd_out;
  p.nu = p.nu * vd_out;
}

// transport::compton_scatter (scatter.cpp:76-133)
template <class R>
__device__ __forceinline__ void compton_scatter(const PropArgs<R>& A, Pkt<R>& p, ZoneSink& zs,
                                                Counts& c)
{
  lorentz_transform(A.g, p, 0);
  R E_ratio, Dn[3];
  while (true) {
    isotropic_dir(p, Dn);
    R cost = p.D[0] * Dn[0] + p.D[1] * Dn[1] + p.D[2] * Dn[2];
    E_ratio = 1 / (1 + p.nu / (R)M_E_MEV * (1 - cost));
    R diff_cs = (R)0.5 * (E_ratio * E_ratio * (1 / E_ratio + E_ratio - 1 + cost * cost));
    if (p.rng.template u<R>() < diff_cs) break;
  }
  p.nu = p.nu * E_ratio;
  if (p.rng.template u<R>() > E_ratio) {
    zs.add(T_LDEP, p.ind, p.e);
    p.type = SGPU_PHOTON;
    isotropic_dir(p, Dn);
    sample_photon_frequency(A, p, c);
  }
  p.D[0] = Dn[0]; p.D[1] = Dn[1]; p.D[2] = Dn[2];
  lorentz_transform(A.g, p, 1);
}

// transport::do_scatter (scatter.cpp:12-70)
template <class R>
__device__ __forceinline__ int do_scatter(const PropArgs<R>& A, Pkt<R>& p, R eps, ZoneSink& zs,
                                          Counts& c)
{
  int fate = SGPU_MOVING;
  if (p.type == SGPU_PHOTON) {
    // one inlined scatter for both outcomes (lanes of a wavefront take
    // both), the draws in the reference's order
    int redist = 0;
    if (!(p.rng.template u<R>() > eps)) {
      R z2 = p.rng.template u<R>();
      if (A.p.radiative_eq || (z2 > (R)A.eps_imc[p.ind])) redist = 1;
      else fate = SGPU_ABSORBED;
    }
    if (fate == SGPU_MOVING) isotropic_scatter(A, p, redist, c);
  } else if (p.type == SGPU_GAMMARAY) {
    if (p.rng.template u<R>() > eps) compton_scatter(A, p, zs, c);
    else {
      zs.add(T_LDEP, p.ind, p.e);
      p.type = SGPU_PHOTON;
      isotropic_dir(p, p.D);
      sample_photon_frequency(A, p, c);
      lorentz_transform(A.g, p, 1);
    }
  }
  return fate;
}

// ---------------------------------------------------------------------------
// Discrete diffusion Monte Carlo (transport_use_ddmc = 2): 1D sphere, FP64
// form.  A photon in a zone of optical depth sigma dr above the threshold
// leaves the event loop and random-walks between zones with the leakage
// opacities of discrete_diffuse_DDMC; a Monte Carlo photon about to enter
// such a zone is converted at the interface with the asymptotic
// diffusion-limit probability or re-emitted into its own zone.
// ---------------------------------------------------------------------------
constexpr double DDMC_LAMBDA = 0.7104;     // discrete_diffusion.cpp:118
constexpr double DDMC_SML_PUSH = 1.0e-8;   // :119
// the inelastic-scattering loop of the walk, unbounded in the reference
// (:256-264), stops after this many redistributions
constexpr int DDMC_SCATTER_CAP = 100000;

// grid_1D_sphere::get_zone_size (grid_1D_sphere.h:65-68)
__device__ __forceinline__ double zone_size(const GridView<double>& g, int i)
{
  return i == 0 ? g.r_out.x[0] - g.r_out.min : g.r_out.x[i] - g.r_out.x[i - 1];
}

__device__ __forceinline__ double radius(const double* x) { return sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]); }

// the diffusion test of transport.cpp:219-231, 275-286, 298-309 for `zone`
__device__ __forceinline__ bool ddmc_zone(const PropArgs<double>& A, Pkt<double>& p, int zone, double dshift,
                                          double* sigma, double* dr)
{
  const int old = p.ind;
  double eps;
  p.ind = zone;
  get_opacity(A, p, dshift, sigma, &eps);
  p.ind = old;
  *dr = zone_size(A.g, zone);
  return (*sigma * *dr > A.p.ddmc_tau_threshold) && (p.type == SGPU_PHOTON);
}

// transport::sample_dir_from_blackbody_surface (discrete_diffusion.cpp:760-794)
__device__ __forceinline__ void sample_dir_bb_surface(Pkt<double>& p)
{
  const double mu = sqrt(p.rng.uniform());
  const double phi = 2.0 * PI * p.rng.uniform();
  const double smu = sqrt(1.0 - mu * mu);
  const double v1 = smu * cos(phi), v2 = smu * sin(phi), v3 = mu;
  const double r = radius(p.x);
  const double r1 = p.x[0] / r, r2 = p.x[1] / r, r3 = p.x[2] / r;
  const double costheta = r3, sintheta = sqrt(1 - costheta * costheta);
  const double v1p = v1 * costheta + r1 * v3 * sintheta - r2 * (1.0 - costheta) * (-r2 * v1 + r1 * v2);
  const double v2p = v2 * costheta + r2 * v3 * sintheta - r1 * (1.0 - costheta) * (-r2 * v1 + r1 * v2);
  const double v3p = v3 * costheta - (r1 * v1 + r2 * v2) * sintheta;
  const double norm = sqrt(v1p * v1p + v2p * v2p + v3p * v3p);
  p.D[0] = v1p / norm; p.D[1] = v2p / norm; p.D[2] = v3p / norm;
}

__device__ __forceinline__ void flip(Pkt<double>& p) { p.D[0] = -p.D[0]; p.D[1] = -p.D[1]; p.D[2] = -p.D[2]; }

// transport::discrete_diffuse_DDMC (discrete_diffusion.cpp:111-391)
__device__ __forceinline__ int ddmc_walk(const PropArgs<double>& A, Pkt<double>& p, double tstop, ZoneSink& zs,
                                      Counts& c)
{
  const GridView<double>& g = A.g;
  const int nz = g.n_zones, nn = A.nu.n;
  const double tau_thr = A.p.ddmc_tau_threshold;
  int fate = SGPU_MOVING;
  double dt_remaining = tstop - p.t;
  while (fate == SGPU_MOVING) {
    const int ii = p.ind;
    int ip = ii + 1, im = ii - 1;
    if (ip == nz) ip = ii;
    if (im < 0) im = 0;
    const double dx = zone_size(g, ii), dxp1 = zone_size(g, ip), dxm1 = zone_size(g, im);
    double sigma_i, sigma_ip1, sigma_im1, eps_i, eps_n;
    double dshift = dshift_lab_to_comoving(g, p);
    int i_nu = get_opacity(A, p, dshift, &sigma_i, &eps_i);
    p.ind = ip;
    get_opacity(A, p, dshift, &sigma_ip1, &eps_n);
    p.ind = im;
    get_opacity(A, p, dshift, &sigma_im1, &eps_n);
    p.ind = ii;
    if (sigma_i * dx <= tau_thr) return SGPU_MOVING;
    const double r_p = g.r_out.x[ii];
    const double r_m = ii == 0 ? g.r_out.min : g.r_out.x[im];
    const double r_0 = 0.5 * (r_p + r_m);
    const double Xi = 1.0 + dx * dx / (12.0 * r_0 * r_0);
    double geo = r_m * r_m / r_0 / r_0 / Xi;
    const bool on_im1 = sigma_im1 * dxm1 > tau_thr;
    double leak_left = on_im1 ? (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + sigma_im1 * dxm1)) * geo
                              : (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + 2.0 * DDMC_LAMBDA)) * geo;
    geo = r_p * r_p / r_0 / r_0 / Xi;
    const bool on_ip1 = sigma_ip1 * dxp1 > tau_thr;
    double leak_right = on_ip1 ? (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + sigma_ip1 * dxp1)) * geo
                               : (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + 2.0 * DDMC_LAMBDA)) * geo;
    if (ip == ii) leak_right = (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + 2.0 * DDMC_LAMBDA)) * geo;
    if (ii == 0) leak_left = 0.0;
    const double leak_tot = leak_left + leak_right;
    const double xi = p.rng.uniform();
    const double d_stay = C_LIGHT * dt_remaining;
    const double d_leak = -log(xi) / (leak_left + leak_right);
    double d_sc = INFINITY;
    if (nn > 1) {
      const double tau_r = -1.0 * log(p.rng.uniform());
      if (eps_i > 0.0) {
        double k_inel = sigma_i * eps_i;
        dshift = dshift_lab_to_comoving(g, p);
        i_nu = get_opacity(A, p, dshift, &sigma_i, &eps_i);
        const double elastic = A.tab.cdf[(size_t)p.ind * nn + i_nu];
        k_inel *= 1.0 - elastic;
        d_sc = tau_r / k_inel;
      }
    }
    int event;
    double this_d;
    if ((d_sc < d_leak) && (d_sc < d_stay)) { event = 0; this_d = d_sc; }
    else if (d_leak < d_stay) { event = 1; this_d = d_leak; }
    else { event = 2; this_d = d_stay; }
    const double this_dt = this_d / C_LIGHT;
    p.t += this_dt;
    dt_remaining -= this_dt;
    // bin 0 of the zone's J row, and e_abs without the eps_imc factor, as
    // the reference tallies them (:240-244)
    atomic_add_f64(&A.ta.J_nu[(size_t)p.ind * (A.p.store_Jnu ? nn : 1)], p.e * this_d);
    atomic_add_f64(&A.ta.e_abs[p.ind], (p.e * dshift) * this_d * sigma_i * eps_i);
    bool leaked2imc = false;
    if (event == 0) {
      int now_i_nu = i_nu;
      for (int k = 0; now_i_nu == i_nu && k < DDMC_SCATTER_CAP; k++) {
        fate = do_scatter(A, p, 1.0, zs, c);
        dshift = dshift_lab_to_comoving(g, p);
        now_i_nu = get_opacity(A, p, dshift, &sigma_i, &eps_i);
      }
    } else if (event == 2) {
      dt_remaining = -1.0;
      fate = SGPU_STOPPED;
    } else {
      const double P_left = leak_left / leak_tot;
      const double xi2 = p.rng.uniform();
      if (xi2 <= P_left) {
        p.ind--;
        const double rr = radius(p.x);
        double dr = rr - r_m;
        if (on_im1) dr += p.rng.uniform() * dxm1;
        else dr *= (1.0 + DDMC_SML_PUSH);
        p.x[0] -= p.x[0] / rr * dr;
        p.x[1] -= p.x[1] / rr * dr;
        p.x[2] -= p.x[2] / rr * dr;
        if (!on_im1) {
          leaked2imc = true;
          lorentz_transform(g, p, 0);
          sample_dir_bb_surface(p);
          const double mu = (p.x[0] * p.D[0] + p.x[1] * p.D[1] + p.x[2] * p.D[2]) / radius(p.x);
          if (mu > 0.0) flip(p);
          lorentz_transform(g, p, 1);
        }
      } else {
        p.ind++;
        const double rr = radius(p.x);
        double dr = r_p - rr;
        if (on_ip1) dr += p.rng.uniform() * dxp1;
        else dr *= (1.0 + DDMC_SML_PUSH);
        p.x[0] += p.x[0] / rr * dr;
        p.x[1] += p.x[1] / rr * dr;
        p.x[2] += p.x[2] / rr * dr;
        if (!on_ip1) {
          leaked2imc = true;
          lorentz_transform(g, p, 0);
          sample_dir_bb_surface(p);
          const double mu = (p.x[0] * p.D[0] + p.x[1] * p.D[1] + p.x[2] * p.D[2]) / radius(p.x);
          if (mu < 0.0) flip(p);
          lorentz_transform(g, p, 1);
        }
      }
    }
    // homologous zone velocity and dv/ds (grid_1D_sphere.cpp:707-716); the
    // adiabatic loss and the advection of :342-368
    const double tn = g.t_now;
    const double zv[3] = {p.x[0] / tn, p.x[1] / tn, p.x[2] / tn};
    const double dvds = 1.0 / tn;
    const double vel = sqrt(zv[0] * zv[0] + zv[1] * zv[1] + zv[2] * zv[2]);
    const double v_over_r = vel / radius(p.x);
    const double dlnrhodt = dvds + 2.0 * v_over_r;
    const double shift = (1.0 - dlnrhodt * this_dt / 3.0);
    lorentz_transform(g, p, 0);
    p.e *= shift;
    p.nu *= shift;
    lorentz_transform(g, p, 1);
    if (!leaked2imc) {
      p.x[0] += zv[0] * this_dt;
      p.x[1] += zv[1] * this_dt;
      p.x[2] += zv[2] * this_dt;
    }
    p.ind = get_zone(g, p.x);
    if (p.ind == -1) fate = SGPU_ABSORBED;
    if (p.ind == -2) fate = SGPU_ESCAPED;
  }
  return fate;
}

// transport::move_across_DDMC_interface (discrete_diffusion.cpp:398-509);
// true when converted (the packet is then left in the comoving frame, as
// the reference leaves it)
__device__ __forceinline__ bool ddmc_interface(const PropArgs<double>& A, Pkt<double>& p, int new_ind,
                                            double sigma_i, double dr)
{
  const GridView<double>& g = A.g;
  const int nz = g.n_zones;
  int ip = p.ind + 1, im = p.ind - 1;
  if (ip == nz) ip = p.ind;
  if (im < 0) im = 0;
  const double r_p = g.r_out.x[p.ind];
  const double r_m = p.ind == 0 ? g.r_out.min : g.r_out.x[im];
  const double r_interface = new_ind == ip ? r_p : r_m;
  lorentz_transform(g, p, 0);
  const double rr = radius(p.x);
  const double mu = fabs((p.x[0] * p.D[0] + p.x[1] * p.D[1] + p.x[2] * p.D[2]) / rr);
  const double p_convert = 4.0 * (1.0 + 1.5 * mu) / (3.0 * sigma_i * dr + 6.0 * 0.7104);
  const double xi = p.rng.uniform();
  double ran[3];
  ran[0] = p.rng.uniform(); ran[1] = p.rng.uniform(); ran[2] = p.rng.uniform();
  double new_r[3];
  if (xi <= p_convert) {
    p.ind = new_ind;
    sample_in_zone(g, p.ind, ran, new_r);
    p.x[0] = new_r[0]; p.x[1] = new_r[1]; p.x[2] = new_r[2];
    return true;
  }
  sample_in_zone(g, p.ind, ran, new_r);
  p.x[0] = new_r[0]; p.x[1] = new_r[1]; p.x[2] = new_r[2];
  // |x| is re-evaluated after each component moves, as in the reference
  if (new_ind == ip) {
    const double d = (r_interface - radius(p.x)) * (1.0 - DDMC_SML_PUSH);
    p.x[0] += p.x[0] / radius(p.x) * d;
    p.x[1] += p.x[1] / radius(p.x) * d;
    p.x[2] += p.x[2] / radius(p.x) * d;
  } else if (new_ind == im) {
    const double d = (radius(p.x) - r_interface) * (1.0 - DDMC_SML_PUSH);
    p.x[0] -= p.x[0] / radius(p.x) * d;
    p.x[1] -= p.x[1] / radius(p.x) * d;
    p.x[2] -= p.x[2] / radius(p.x) * d;
  }
  sample_dir_bb_surface(p);
  const double mu2 = (p.x[0] * p.D[0] + p.x[1] * p.D[1] + p.x[2] * p.D[2]) / radius(p.x);
  if (new_ind == ip && mu2 > 0.0) flip(p);
  else if (new_ind == im && mu2 < 0.0) flip(p);
  lorentz_transform(g, p, 1);
  return false;
}

// one iteration of transport::propagate_monte_carlo (transport.cpp:266-481)
template <class R, bool DDMC = false>
__device__ __forceinline__ int mc_event(const PropArgs<R>& A, Pkt<R>& p, double tstop,
                                        ZoneSink& zs, Counts& c)
{
  const int nn = A.nu.n;
  const R cl = (R)C_LIGHT;
  c.events++;
  if constexpr (sizeof(R) == 4) {
    // a scattering event keeps its zone, so the entry of its redistribution
    // is known now
    const float u_col = p.rng.uniform_f();
    p.pre.coin = p.rng.uniform_f();
    p.pre.k = min((int)(u_col * (float)nn), nn - 1);
    const AliasEntry e = A.tab.alias[(size_t)p.ind * nn + p.pre.k];
    p.pre.q = e.q;
    p.pre.a = e.a;
    p.pre.valid = p.type == SGPU_PHOTON;
  }
  // the opacity lookup before the boundary geometry (both are pure functions
  // of the packet), so that its table load is in flight during the geometry
  R dshift = dshift_lab_to_comoving(A.g, p);
  R kappa, eps;
  int i_nu = get_opacity(A, p, dshift, &kappa, &eps);
  R d_bn = 0;
  int new_ind = get_next_zone(A.g, p.x, p.D, p.ind, &d_bn);
  d_bn = floor_boundary_step(p, d_bn);
  bool new_cell_ddmc = false;
  R sig_n = 0, dr_n = 0;
  if constexpr (DDMC)
    if (new_ind >= 0) new_cell_ddmc = ddmc_zone(A, p, new_ind, dshift, &sig_n, &dr_n);
  R d_nu = qdiv(qdiv(ax_delta(A.nu, i_nu), p.nu) * cl, p.dvds);
  if (p.dvds == 0) d_nu = rinf<R>();
  if (d_nu < 0) d_nu = -1 * d_nu;
  bool binshift = false;
  if (d_nu < d_bn) { d_bn = d_nu; new_ind = p.ind; binshift = true; }
  R tot_lab = kappa * dshift;
  R tau_r = -1 * rlog<R>(1 - p.rng.template u<R>());
  R d_sc = qdiv(tau_r, tot_lab);
  if (tot_lab == 0) d_sc = rinf<R>();
  double d_tm = (tstop - p.t) * C_LIGHT;
  if (A.p.steady_state) d_tm = INFINITY;
  int event;
  R this_d;
  if ((d_sc < d_bn) && ((double)d_sc < d_tm)) { event = 0; this_d = d_sc; }
  else if ((double)d_bn < d_tm) { event = 1; this_d = d_bn; }
  else { event = 2; this_d = (R)d_tm; }
  const double this_E = p.e * (double)this_d;
  const int z = p.ind;
  const double ds = (double)dshift, kap = (double)kappa;
  const bool photon = p.type == SGPU_PHOTON;
  if (photon) zs.jnu((int64_t)z * (A.p.store_Jnu ? nn : 1) + (A.p.store_Jnu ? i_nu : 0), this_E);
  R rr = sqrt_(p.x[0] * p.x[0] + p.x[1] * p.x[1] + p.x[2] * p.x[2]);
  R xdotD = p.x[0] * p.D[0] + p.x[1] * p.D[1] + p.x[2] * p.D[2];
  zs.event(z, photon, photon ? this_E * ds * kap * (double)eps * ds : 0.0,
           this_E * ds * kap * (double)p.D[0] * ds, this_E * ds * kap * (double)p.D[1] * ds,
           this_E * ds * kap * (double)p.D[2] * ds, mul_ratio(this_E * ds * kap, xdotD, rr) * ds);
  p.x[0] += this_d * p.D[0];
  p.x[1] += this_d * p.D[1];
  p.x[2] += this_d * p.D[2];
  p.t = p.t + flight_time(this_d);
  int fate = SGPU_MOVING;
  if (event == 1) {
    c.boundary++;
    if (binshift) c.binshift++;
    if (new_ind == -1) {
      if (A.p.boundary_in_reflect) { p.D[0] *= -1; p.D[1] *= -1; p.D[2] *= -1; }
      else { p.ind = new_ind; fate = SGPU_ABSORBED; }
    } else if (new_ind == -2) {
      if (A.p.boundary_out_reflect) { p.D[0] *= -1; p.D[1] *= -1; p.D[2] *= -1; }
      else { p.ind = new_ind; fate = SGPU_ESCAPED; }
    } else if (DDMC && new_cell_ddmc && new_ind != p.ind) {
      if constexpr (DDMC)
        if (ddmc_interface(A, p, new_ind, sig_n, dr_n)) return SGPU_MOVING;
    } else {
      p.ind = new_ind;
    }
  } else if (event == 0) {
    c.scatter++;
    fate = do_scatter(A, p, eps, zs, c);
  } else {
    fate = SGPU_STOPPED;
  }
  return fate;
}

// spectrum_array::count (spectrum_array.cpp:215-246)
__device__ __forceinline__ void spectrum_count(const sgpu_spectrum& s, double t, double w, double E,
                                               double D0, double D1, double D2)
{
  if (!s.flux) return;
  double mu = D2;
  double phi = atan2(D1, D0);
  if (phi < 0) phi += 2. * PI;
  int t_bin = loc_locate(s.time_grid, t);
  int l_bin = loc_locate(s.nu_grid, w);
  int m_bin = loc_locate(s.mu_grid, mu);
  int p_bin = loc_locate(s.phi_grid, phi);
  if (l_bin < 0) l_bin = 0;
  if (l_bin >= s.nu_grid.n) l_bin = 0;
  if ((t_bin < 0) || (l_bin < 0) || (m_bin < 0) || (p_bin < 0)) return;
  if (t_bin >= s.time_grid.n) return;
  if (m_bin >= s.mu_grid.n) return;
  if (p_bin >= s.phi_grid.n) return;
  int a3 = s.phi_grid.n, a2 = s.mu_grid.n * a3, a1 = s.nu_grid.n * a2;
  size_t ind = (size_t)t_bin * a1 + (size_t)l_bin * a2 + (size_t)m_bin * a3 + p_bin;
  atomic_add_f64(&s.flux[ind], E);
  atomic_add_f64(&s.click[ind], 1.0);
}

__device__ __forceinline__ void init_rng(Rng& r, uint64_t seed, uint64_t id, uint32_t draw)
{
  r = make_rng(seed, id, draw);
}

__device__ __forceinline__ void init_rng(RngX& r, uint64_t seed, uint64_t id, uint32_t draw)
{
  r = make_rngx(seed, id, draw);
}

template <class R>
__device__ __forceinline__ void load_packet(const PropArgs<R>& A, int64_t i, Pkt<R>& p)
{
  const sgpu_packets k = ka_packets(cold_args<R>()->pk);
  p.x[0] = (R)k.x[i]; p.x[1] = (R)k.y[i]; p.x[2] = (R)k.z[i];
  p.D[0] = (R)k.Dx[i]; p.D[1] = (R)k.Dy[i]; p.D[2] = (R)k.Dz[i];
  p.t = k.t[i]; p.e = k.e[i]; p.nu = (R)k.nu[i];
  p.ind = k.ind[i]; p.type = k.type[i];
  init_rng(p.rng, cold_args<R>()->p.seed, k.id[i], k.draw[i]);
  p.dvds = 0;
  if constexpr (sizeof(R) == 4) p.pre.valid = false;
}

template <class R>
__device__ __forceinline__ void finish_packet(const PropArgs<R>& A, int64_t i, Pkt<R>& p, int fate, Counts& c)
{
  const sgpu_packets k = ka_packets(cold_args<R>()->pk);
  if (fate == SGPU_ESCAPED) {
    double x0 = p.x[0], x1 = p.x[1], x2 = p.x[2], D0 = p.D[0], D1 = p.D[1], D2 = p.D[2];
    double t_obs = p.t - (x0 * D0 + x1 * D1 + x2 * D2) / C_LIGHT;
    if (p.type == SGPU_PHOTON) {
      const sgpu_spectrum sp = ka_spectrum(cold_args<R>()->ta.optical);
      spectrum_count(sp, t_obs, (double)p.nu, p.e, D0, D1, D2);
    } else if (p.type == SGPU_GAMMARAY) {
      const sgpu_spectrum sp = ka_spectrum(cold_args<R>()->ta.gamma);
      spectrum_count(sp, t_obs, (double)p.nu, p.e, D0, D1, D2);
    }
    p.t = t_obs;
    c.escaped++;
  } else if (fate == SGPU_ABSORBED) c.absorbed++;
  else if (fate == SGPU_STOPPED) c.stopped++;
  else c.moving++;
  k.x[i] = p.x[0]; k.y[i] = p.x[1]; k.z[i] = p.x[2];
  k.Dx[i] = p.D[0]; k.Dy[i] = p.D[1]; k.Dz[i] = p.D[2];
  k.t[i] = p.t; k.e[i] = p.e; k.nu[i] = p.nu;
  k.ind[i] = p.ind; k.type[i] = p.type; k.fate[i] = fate;
  k.draw[i] = p.rng.draw;
}

__device__ __forceinline__ void flush_counts(sgpu_counters* cnt, const Counts& c, uint64_t npk)
{
  if (!cnt) return;
  unsigned long long v[10] = {npk, c.events, c.binshift, c.boundary, c.scatter, c.redist,
                              c.escaped, c.absorbed, c.stopped, c.moving};
  unsigned long long* dst = (unsigned long long*)cnt;
  const int lane = threadIdx.x % WAVE;
#pragma unroll
  for (int f = 0; f < 10; f++) {
    unsigned long long s = wave_sum(v[f]);
    if (lane == 0 && s) atomicAdd(&dst[f], s);
  }
}

// copy an FP32 axis' edges into LDS at `ax` and point `dst` at them.  The
// pointer is replaced unconditionally (an empty axis gets an empty range),
// so every edge access provably addresses LDS and compiles to ds_read
// rather than a flat load.
__device__ __forceinline__ void stage_axis(Axis<float>& dst, const Axis<float>& src, float*& ax)
{
  const int n = (src.n > 0 && src.x) ? src.n : 0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) ax[k] = src.x[k];
  dst.x = ax;
  ax += n;
}

// PERSIST = false: one thread per packet (the plain drop-in).
// PERSIST = true : lanes stride through the packets and refill on finish.
// OCC: minimum waves per SIMD the register allocation must allow (1: no bound).
// KIND: the grid kind when known at compile time (0: read from the arguments);
// a fixed kind removes the other geometry from the loop body.
template <class R, bool PERSIST, bool ACCUM, int OCC = 1, int KIND = 0, bool DDMC = false>
__global__ void __launch_bounds__(PROP_THREADS) __attribute__((amdgpu_waves_per_eu(OCC)))
propagate_kernel(PropArgs<R> A0)
{
  PropArgs<R> A = A0;
  if (A0.n_dev) A.n = *A0.n_dev;
  if constexpr (KIND != 0) A.g.kind = KIND;
  const int nz = A.g.n_zones;
  const bool use_lds = A.lds_tallies;
  ZoneSink zs = {{A.ta.e_abs, A.ta.fx_rad, A.ta.fy_rad, A.ta.fz_rad, A.ta.fr_rad, A.ta.L_radio_dep},
                 nz, use_lds, ACCUM, -1, {0, 0, 0, 0, 0}, A.ta.J_nu, -1, 0.0, A.eps_imc};
  if (use_lds)
    for (int k = threadIdx.x; k < T_NFIELDS * nz; k += blockDim.x) dyn_lds[k] = 0;
  if constexpr (sizeof(R) == 4) {
    // the FP32 form reads the frequency and zone edges from LDS: every event
    // locates and samples its bin and measures its zone through them
    float* ax = (float*)(dyn_lds + (use_lds ? T_NFIELDS * nz : 0));
    stage_axis(A.nu, A0.nu, ax);
    stage_axis(A.g.r_out, A0.g.r_out, ax);
    stage_axis(A.g.x_out, A0.g.x_out, ax);
    stage_axis(A.g.z_out, A0.g.z_out, ax);
  }
  __syncthreads();
  Counts c = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double tstop = A.p.t_now + A.p.dt;
  const int64_t stride = PERSIST ? (int64_t)gridDim.x * blockDim.x : A.n;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t mine = 0;
  bool have = false;
  int n_ev = 0;
  Pkt<R> p;
  while (true) {
    if (!have) {
      if (i >= A.n) break;
      load_packet(A, i, p);
      mine++;
      p.ind = get_zone(A.g, p.x);
      n_ev = 0;
      if (p.ind < 0) {
        finish_packet(A, i, p, p.ind == -1 ? SGPU_ABSORBED : SGPU_ESCAPED, c);
        i += stride;
        continue;
      }
      have = true;
    }
    int fate;
    if constexpr (DDMC) {
      // transport::propagate (transport.cpp:216-248): diffuse while the
      // photon sits in an optically thick zone
      R sig, dr;
      if (ddmc_zone(A, p, p.ind, dshift_lab_to_comoving(A.g, p), &sig, &dr)) fate = ddmc_walk(A, p, tstop, zs, c);
      else fate = mc_event<R, true>(A, p, tstop, zs, c);
    } else {
      fate = mc_event(A, p, tstop, zs, c);
    }
    n_ev++;
    if (fate != SGPU_MOVING || (A.p.max_events > 0 && n_ev >= A.p.max_events)) {
      finish_packet(A, i, p, fate, c);
      have = false;
      i += stride;
    }
  }
  zs.flush();
  flush_counts(A.cnt, c, mine);
  if (use_lds) {
    __syncthreads();
    for (int k = threadIdx.x; k < T_NFIELDS * nz; k += blockDim.x) {
      double v = dyn_lds[k];
      if (v != 0) atomic_add_f64(tally_field(A.ta, k / nz) + k % nz, v);
    }
  }
}

// FP32 (kappa, eps) table of the mixed form
__global__ void pack_tables_kernel(const double* abs, const double* scat, int omit_scat, size_t n, float2* ke)
{
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double a = abs[k];
  double s = omit_scat ? 0.0 : scat[k];
  double op = a + s;
  ke[k] = make_float2((float)op, op == 0 ? 0.f : (float)(a / op));
}

// Alias table of one zone's emissivity distribution (Walker 1977), in FP64,
// one wave per zone with the scaled probabilities in LDS (nn <= ALIAS_LDS_MAX).
// The wave classifies the columns and compacts the small (p n < 1) and large
// ones into two index-ordered lists (ballot prefix sums); lane 0 then sweeps
// both lists once (the stack-free form of Vose's pairing, as in
// Huebschle-Schneider & Sanders 2019): the current large column fills the
// next small one, r <- (r + p_small) - 1 as Vose updates it, and when r
// drops below 1 the large column is itself filled from the next large one.
// Every column ends with its own threshold and alias, so the table samples
// the distribution exactly; leftovers (rounding) get threshold 1.  The
// distribution is the one upper_bound over the normalised CDF samples,
// including the last bin taking the remainder when cdf[nn-1] < 1 (the
// min(upper_bound, nn - 1) clamp); a zone with no emission sends every
// column to the last bin, as upper_bound on an all-zero CDF does.
constexpr int ALIAS_LDS_MAX = 4096;   // columns (the lists hold 16-bit indices)

__global__ void __launch_bounds__(64) alias_kernel(const double* cdf, int nn, AliasEntry* tab)
{
  extern __shared__ double alds[];
  double* sc = alds;                              // scaled probabilities
  uint16_t* small = (uint16_t*)(alds + nn);       // small columns, ascending
  uint16_t* large = small + nn;                   // large columns, ascending
  const int z = blockIdx.x, lane = threadIdx.x;
  const double* c = cdf + (size_t)z * nn;
  AliasEntry* t = tab + (size_t)z * nn;
  double part = 0;
  for (int k = lane; k < nn; k += WAVE) {
    double lo = k == 0 ? 0.0 : c[k - 1];
    double pk = (k == nn - 1 ? 1.0 : c[k]) - lo;
    if (!(pk > 0)) pk = 0;   // also a NaN CDF
    sc[k] = pk;
    part += pk;
  }
  const double tot = wave_sum(part);
  if (!(tot > 0)) {
    for (int k = lane; k < nn; k += WAVE) t[k] = AliasEntry{0.f, nn - 1};
    return;
  }
  const double scale = (double)nn / tot;
  int ns = 0, nl = 0;
  for (int k0 = 0; k0 < nn; k0 += WAVE) {
    const int k = k0 + lane;
    bool is_small = false, valid = k < nn;
    if (valid) {
      const double v = sc[k] * scale;
      sc[k] = v;
      is_small = v < 1.0;
    }
    const uint64_t bs = __ballot(valid && is_small), bl = __ballot(valid && !is_small);
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (valid && is_small) small[ns + __popcll(bs & below)] = (uint16_t)k;
    if (valid && !is_small) large[nl + __popcll(bl & below)] = (uint16_t)k;
    ns += __popcll(bs);
    nl += __popcll(bl);
  }
  __syncthreads();
  if (lane != 0) return;
  int is = 0, il = 0;
  if (nl > 0) {
    int g = large[0];
    double r = sc[g];
    // the next small and large columns and their weights are read one step
    // ahead, so the LDS loads overlap the dependent update of r
    int l = ns > 0 ? small[0] : 0;
    double pl = sc[l];
    int g2 = nl > 1 ? large[1] : g;
    double p2 = sc[g2];
    while (is < ns) {
      const int l_next = is + 1 < ns ? small[is + 1] : l;
      const double pl_next = sc[l_next];
      t[l] = AliasEntry{(float)pl, g};
      r = (r + pl) - 1.0;
      ++is;
      // the large column has turned small: the next large one fills it
      while (r < 1.0 && il + 1 < nl) {
        ++il;
        const int g3 = il + 1 < nl ? large[il + 1] : g2;
        const double p3 = sc[g3];
        t[g] = AliasEntry{(float)r, g2};
        r = (p2 + r) - 1.0;
        g = g2;
        g2 = g3;
        p2 = p3;
      }
      if (r < 1.0) break;            // no large column left to fill from
      l = l_next;
      pl = pl_next;
    }
    t[g] = AliasEntry{1.f, g};       // the last large column (or a leftover)
    for (int k = il + 1; k < nl; k++) t[large[k]] = AliasEntry{1.f, large[k]};
  }
  for (; is < ns; is++) t[small[is]] = AliasEntry{1.f, small[is]};
}

__global__ void to_f32_kernel(const double* src, float* dst, int n)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = (float)src[k];
}

// ------------------------------------------------------------------
// K5: radioactive emission (emission.cpp:165-243, :54-105)
// ------------------------------------------------------------------
struct EmitPrep {
  sgpu_grid g;
  sgpu_zone_state zs;
  double t_now;
  int force_rproc;
  double* cdf;       // [nz] zone emission CDF
  double* gfrac;     // [nz]
  double* L_emit;    // [nz] tallies.L_radio_emit
  double* L_tot;     // [1]
};

// single workgroup: per-zone decay luminosity, CDF (block scan), L_tot
__global__ void __launch_bounds__(1024) emit_prep_kernel(EmitPrep P)
{
  __shared__ double wsum[1024 / WAVE];
  const int nz = P.g.n_zones, tid = threadIdx.x, NT = blockDim.x;
  const int per = (nz + NT - 1) / NT;
  const int j0 = tid * per, j1 = min(nz, j0 + per);
  double local = 0;
  for (int i = j0; i < j1; ++i) {
    double gf;
    double L = decay(P.zs.n_elems, P.zs.elem_Z, P.zs.elem_A, P.zs.X + (size_t)i * P.zs.n_elems,
                     P.t_now, &gf, P.force_rproc);
    L = P.zs.rho[i] * L * P.g.vol[i];
    P.L_emit[i] = L;
    P.gfrac[i] = gf;
    P.cdf[i] = L;
    local += L;
  }
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
  for (int i = j0; i < j1; ++i) { run += P.cdf[i]; P.cdf[i] = run; }
  __syncthreads();
  const double last = P.cdf[nz - 1];
  const bool flat = isnan(last) || last == 0;
  __syncthreads();
  for (int i = j0; i < j1; ++i) P.cdf[i] = flat ? 1.0 : P.cdf[i] / last;
  if (tid == 0) {
    double tot = 0;
    for (int i = 0; i < NT / WAVE; i++) tot += wsum[i];
    *P.L_tot = tot;
  }
}

struct EmitArgs {
  GridView<double> g;
  Axis<double> nu;
  const double* emis_cdf;
  const double* zcdf;
  const double* gfrac;
  const double* L_tot;
  double t_now, dt;
  uint64_t seed, id_base;
  int64_t n_emit, offset;
  const int64_t* offset_dev;   // when set: the offset, read on the device; nothing is
                               // emitted if the packets would pass the capacity
  sgpu_packets pk;
  double* L_dep;
  int lds_dep;
};


__global__ void __launch_bounds__(256) emit_kernel(EmitArgs A)
{
  const int nz = A.g.n_zones;
  if (A.lds_dep) {
    for (int k = threadIdx.x; k < nz; k += blockDim.x) dyn_lds[k] = 0;
    __syncthreads();
  }
  const double L_tot = *A.L_tot;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t offset = A.offset;
  bool fits = true;
  if (A.offset_dev) {
    offset = *A.offset_dev;
    fits = offset + A.n_emit <= A.pk.capacity;
  }
  if (q < A.n_emit && L_tot != 0 && fits) {
    const double E_p = L_tot * A.dt / (1.0 * A.n_emit);
    Pkt<double> p;
    p.rng = make_rng(A.seed, A.id_base + (uint64_t)q, 0);
    int i = min(cdf_sample(A.zcdf, nz, p.rng.uniform()), nz - 1);
    double t = A.t_now + A.dt * p.rng.uniform();
    if (p.rng.uniform() < A.gfrac[i]) p.type = SGPU_GAMMARAY;
    else {
      p.type = SGPU_PHOTON;
      if (A.lds_dep) atomicAdd(&dyn_lds[i], E_p);
      else atomic_add_f64(&A.L_dep[i], E_p);
    }
    p.ind = i;
    double ran[3];
    ran[0] = p.rng.uniform(); ran[1] = p.rng.uniform(); ran[2] = p.rng.uniform();
    sample_in_zone(A.g, i, ran, p.x);
    isotropic_dir(p, p.D);
    if (p.type == SGPU_PHOTON) {
      const int nn = A.nu.n;
      int inu = min(cdf_sample(A.emis_cdf + (size_t)i * nn, nn, p.rng.uniform()), nn - 1);
      p.nu = ax_sample(A.nu, inu, p.rng.uniform());
    } else {
      p.nu = 1;
    }
    p.e = E_p;
    lorentz_transform(A.g, p, 1);
    p.t = t;
    const int64_t o = offset + q;
    const sgpu_packets& k = A.pk;
    k.x[o] = p.x[0]; k.y[o] = p.x[1]; k.z[o] = p.x[2];
    k.Dx[o] = p.D[0]; k.Dy[o] = p.D[1]; k.Dz[o] = p.D[2];
    k.t[o] = p.t; k.e[o] = p.e; k.nu[o] = p.nu;
    k.ind[o] = p.ind; k.type[o] = p.type; k.fate[o] = SGPU_MOVING;
    k.draw[o] = p.rng.draw; k.id[o] = A.id_base + (uint64_t)q;
  }
  if (A.lds_dep) {
    __syncthreads();
    for (int k = threadIdx.x; k < nz; k += blockDim.x)
      if (dyn_lds[k] != 0) atomic_add_f64(&A.L_dep[k], dyn_lds[k]);
  }
}

// ------------------------------------------------------------------
// initial radiation field (transport::initialize_particles,
// emission.cpp:112-159, with create_isotropic_particle :54-105)
// ------------------------------------------------------------------
// y[0..n) := cdf_array::set_value's running sum then normalize
// (cdf_array.h:53-75) by one workgroup of NT threads: each thread scans a
// contiguous chunk, the chunk totals are scanned across the block; a NaN or
// zero total makes the CDF all ones.  Returns the total to every thread.
template <int NT>
__device__ double block_cdf(double* y, int n, double* wsum)
{
  const int tid = threadIdx.x;
  const int per = (n + NT - 1) / NT;
  const int j0 = tid * per, j1 = min(n, j0 + per);
  double local = 0;
  for (int j = j0; j < j1; ++j) local += y[j];
  const int lane = tid % WAVE, w = tid / WAVE;
  double incl = local;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    double t = __shfl_up(incl, o, WAVE);
    if (lane >= o) incl += t;
  }
  if (lane == WAVE - 1) wsum[w] = incl;
  __syncthreads();
  double base = 0, total = 0;
  for (int i = 0; i < NT / WAVE; i++) {
    if (i < w) base += wsum[i];
    total += wsum[i];
  }
  double run = base + incl - local;
  for (int j = j0; j < j1; ++j) { run += y[j]; y[j] = run; }
  __syncthreads();
  const double last = y[n - 1];
  const bool flat = isnan(last) || last == 0;
  __syncthreads();   // every thread has read the total before any rewrites it
  for (int j = j0; j < j1; ++j) y[j] = flat ? 1.0 : y[j] / last;
  return total;
}

constexpr int INIT_THREADS = 256;

// zone CDF over e_rad V and E_sum (one workgroup)
__global__ void __launch_bounds__(1024) init_zone_cdf_kernel(const double* e_rad, const double* vol, int nz,
                                                             double* cdf, double* E_sum)
{
  __shared__ double wsum[1024 / WAVE];
  for (int i = threadIdx.x; i < nz; i += blockDim.x) cdf[i] = e_rad[i] * vol[i];
  __syncthreads();
  const double tot = block_cdf<1024>(cdf, nz, wsum);
  if (threadIdx.x == 0) *E_sum = tot;
}

// every zone's emissivity CDF from B_nu(T_gas) dnu at the bin centres
// (transport::blackbody_nu, transport_opacity.cpp:322-326); one workgroup per zone
__global__ void __launch_bounds__(INIT_THREADS) init_blackbody_cdf_kernel(sgpu_locate nu, const double* T_gas,
                                                                          double* emis_cdf)
{
  __shared__ double wsum[INIT_THREADS / WAVE];
  const int z = blockIdx.x, nn = nu.n;
  const double T = T_gas[z];
  double* y = emis_cdf + (size_t)z * nn;
  for (int j = threadIdx.x; j < nn; j += INIT_THREADS) {
    const double nu_m = loc_center(nu, j);
    const double zeta = H_PLANCK * nu_m / K_BOLTZ / T;
    y[j] = 2.0 * nu_m * nu_m * nu_m * H_PLANCK / C_LIGHT / C_LIGHT / (exp(zeta) - 1) * loc_delta(nu, j);
  }
  __syncthreads();
  block_cdf<INIT_THREADS>(y, nn, wsum);
}

struct InitArgs {
  GridView<double> g;
  Axis<double> nu;
  const double* emis_cdf;
  const double* zcdf;
  const double* E_sum;
  double t_now;
  uint64_t seed, id_base;
  int64_t n_init, offset;
  sgpu_packets pk;
};

// one photon per thread: zone, 3 position, 2 direction and 2 frequency draws
__global__ void __launch_bounds__(256) init_particles_kernel(InitArgs A)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.n_init) return;
  const int nz = A.g.n_zones;
  Pkt<double> p;
  p.rng = make_rng(A.seed, A.id_base + (uint64_t)q, 0);
  const int i = min(cdf_sample(A.zcdf, nz, p.rng.uniform()), nz - 1);
  p.ind = i;
  p.type = SGPU_PHOTON;
  double ran[3];
  ran[0] = p.rng.uniform(); ran[1] = p.rng.uniform(); ran[2] = p.rng.uniform();
  sample_in_zone(A.g, i, ran, p.x);
  isotropic_dir(p, p.D);
  const int nn = A.nu.n;
  const int inu = min(cdf_sample(A.emis_cdf + (size_t)i * nn, nn, p.rng.uniform()), nn - 1);
  p.nu = ax_sample(A.nu, inu, p.rng.uniform());
  p.e = *A.E_sum / (1.0 * A.n_init);
  lorentz_transform(A.g, p, 1);
  p.t = A.t_now;
  const int64_t o = A.offset + q;
  const sgpu_packets& k = A.pk;
  k.x[o] = p.x[0]; k.y[o] = p.x[1]; k.z[o] = p.x[2];
  k.Dx[o] = p.D[0]; k.Dy[o] = p.D[1]; k.Dz[o] = p.D[2];
  k.t[o] = p.t; k.e[o] = p.e; k.nu[o] = p.nu;
  k.ind[o] = p.ind; k.type[o] = p.type; k.fate[o] = SGPU_MOVING;
  k.draw[o] = p.rng.draw; k.id[o] = A.id_base + (uint64_t)q;
}

// ------------------------------------------------------------------
// inner boundary source (transport::setup_core_emission,
// transport_init.cpp:465-548; transport::emit_inner_source, emission.cpp:370-480)
// ------------------------------------------------------------------
// the core spectrum's CDF over the nu grid and the packet energy
// L_core dt / n_emit (L_core from the spectrum when not given); one workgroup
__global__ void __launch_bounds__(INIT_THREADS) core_prep_kernel(sgpu_locate nu, double r_core, double T_core,
                                                                 double L_core, double dt, int64_t n_emit,
                                                                 double* cdf, double* Ep)
{
  __shared__ double wsum[INIT_THREADS / WAVE];
  const int nn = nu.n;
  for (int j = threadIdx.x; j < nn; j += INIT_THREADS) {
    const double nu_m = loc_center(nu, j);
    double bb = 1;
    if (T_core > 0) {
      const double zeta = H_PLANCK * nu_m / K_BOLTZ / T_core;
      bb = 2.0 * nu_m * nu_m * nu_m * H_PLANCK / C_LIGHT / C_LIGHT / (exp(zeta) - 1);
    }
    cdf[j] = bb * loc_delta(nu, j);
  }
  __syncthreads();
  const double tot = block_cdf<INIT_THREADS>(cdf, nn, wsum);
  if (threadIdx.x == 0) {
    // L_sum = sum 4 pi r^2 pi B dnu, in the reference's factor order
    const double L = L_core != 0 ? L_core : 4.0 * PI * r_core * r_core * PI * tot;
    *Ep = L * dt / (double)n_emit;
  }
}

struct CoreArgs {
  GridView<double> g;
  Axis<double> nu;
  const double* cdf;
  const double* Ep;
  double r_core, core_frequency, t_now, dt;
  uint64_t seed, id_base;
  int64_t n_emit, offset;
  sgpu_packets pk;
};

__global__ void __launch_bounds__(256) core_emit_kernel(CoreArgs A)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.n_emit) return;
  Pkt<double> p;
  p.rng = make_rng(A.seed, A.id_base + (uint64_t)q, 0);
  if (A.r_core == 0) {
    p.x[0] = p.x[1] = p.x[2] = 0;
    isotropic_dir(p, p.D);
  } else {
    // a uniform point of the photosphere, a cos-theta outward direction
    // about its normal, rotated into the grid frame
    const double phi_core = 2 * PI * p.rng.uniform();
    const double cosp_core = cos(phi_core), sinp_core = sin(phi_core);
    const double cost_core = 1 - 2.0 * p.rng.uniform();
    const double sint_core = sqrt(1 - cost_core * cost_core);
    const double a_phot = A.r_core + A.r_core * 1e-10;
    p.x[0] = a_phot * sint_core * cosp_core;
    p.x[1] = a_phot * sint_core * sinp_core;
    p.x[2] = a_phot * cost_core;
    const double phi_loc = 2 * PI * p.rng.uniform();
    const double cost_loc = sqrt(p.rng.uniform());
    const double sint_loc = sqrt(1 - cost_loc * cost_loc);
    const double D_xl = sint_loc * cos(phi_loc), D_yl = sint_loc * sin(phi_loc), D_zl = cost_loc;
    p.D[0] = cost_core * cosp_core * D_xl - sinp_core * D_yl + sint_core * cosp_core * D_zl;
    p.D[1] = cost_core * sinp_core * D_xl + cosp_core * D_yl + sint_core * sinp_core * D_zl;
    p.D[2] = -sint_core * D_xl + cost_core * D_zl;
  }
  p.e = *A.Ep;
  if (A.core_frequency > 0) p.nu = A.core_frequency;
  else {
    const int nn = A.nu.n;
    const int inu = min(cdf_sample(A.cdf, nn, p.rng.uniform()), nn - 1);
    p.nu = ax_sample(A.nu, inu, p.rng.uniform());
  }
  p.ind = get_zone(A.g, p.x);
  lorentz_transform(A.g, p, 1);
  p.t = A.t_now + p.rng.uniform() * A.dt;
  p.type = SGPU_PHOTON;
  const int64_t o = A.offset + q;
  const sgpu_packets& k = A.pk;
  k.x[o] = p.x[0]; k.y[o] = p.x[1]; k.z[o] = p.x[2];
  k.Dx[o] = p.D[0]; k.Dy[o] = p.D[1]; k.Dz[o] = p.D[2];
  k.t[o] = p.t; k.e[o] = p.e; k.nu[o] = p.nu;
  k.ind[o] = p.ind; k.type[o] = p.type; k.fate[o] = SGPU_MOVING;
  k.draw[o] = p.rng.draw; k.id[o] = A.id_base + (uint64_t)q;
}

// the device packet count after an emission: + n_emit when packets were
// emitted (non-zero luminosity, room in the buffer); one thread
__global__ void emit_count_kernel(const double* L_tot, int64_t n_emit, int64_t capacity, int64_t* count)
{
  const int64_t c = *count;
  if (*L_tot != 0 && c + n_emit <= capacity) *count = c + n_emit;
}

// ------------------------------------------------------------------
// K6: stream compaction (clean_up_particle_vector), order-keeping
// ------------------------------------------------------------------
constexpr int CMP_THREADS = 256;

__global__ void keep_count_kernel(const int32_t* fate, int64_t n, int32_t* blk_keep, int32_t* blk_esc,
                                  const int64_t* n_dev)
{
  __shared__ int ks[CMP_THREADS / WAVE], es[CMP_THREADS / WAVE];
  if (n_dev) n = *n_dev;
  int64_t i = (int64_t)blockIdx.x * CMP_THREADS + threadIdx.x;
  int f = i < n ? fate[i] : SGPU_ESCAPED + 100;
  int keep = (i < n) && f != SGPU_ESCAPED && f != SGPU_ABSORBED;
  int esc = (i < n) && f == SGPU_ESCAPED;
  int k = wave_sum(keep), e = wave_sum(esc);
  if (threadIdx.x % WAVE == 0) { ks[threadIdx.x / WAVE] = k; es[threadIdx.x / WAVE] = e; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int tk = 0, te = 0;
    for (int w = 0; w < CMP_THREADS / WAVE; w++) { tk += ks[w]; te += es[w]; }
    blk_keep[blockIdx.x] = tk;
    blk_esc[blockIdx.x] = te;
  }
}

// single workgroup exclusive scan of block counts; totals to out[0..1]
__global__ void __launch_bounds__(1024) scan_blocks_kernel(int32_t* blk_keep, const int32_t* blk_esc,
                                                           int nb, long long* totals)
{
  __shared__ long long wsum[1024 / WAVE];
  __shared__ long long esum[1024 / WAVE];
  const int tid = threadIdx.x, NT = blockDim.x;
  const int per = (nb + NT - 1) / NT;
  const int j0 = tid * per, j1 = min(nb, j0 + per);
  long long local = 0, le = 0;
  for (int j = j0; j < j1; ++j) { local += blk_keep[j]; le += blk_esc[j]; }
  const int lane = tid % WAVE, w = tid / WAVE;
  long long incl = local;
  for (int o = 1; o < WAVE; o <<= 1) {
    long long t = __shfl_up(incl, o, WAVE);
    if (lane >= o) incl += t;
  }
  long long es = le;
  for (int o = 32; o > 0; o >>= 1) es += __shfl_xor(es, o, WAVE);
  if (lane == WAVE - 1) wsum[w] = incl;
  if (lane == 0) esum[w] = es;
  __syncthreads();
  long long base = 0;
  for (int i = 0; i < w; i++) base += wsum[i];
  long long run = base + incl - local;
  for (int j = j0; j < j1; ++j) { int v = blk_keep[j]; blk_keep[j] = (int32_t)run; run += v; }
  if (tid == 0) {
    long long tk = 0, te = 0;
    for (int i = 0; i < NT / WAVE; i++) { tk += wsum[i]; te += esum[i]; }
    totals[0] = tk;
    totals[1] = te;
  }
}

__global__ void scatter_keep_kernel(sgpu_packets src, sgpu_packets dst, int64_t n, const int32_t* blk_off,
                                    const int64_t* n_dev)
{
  __shared__ int woff[CMP_THREADS / WAVE];
  if (n_dev) n = *n_dev;
  int64_t i = (int64_t)blockIdx.x * CMP_THREADS + threadIdx.x;
  int f = i < n ? src.fate[i] : SGPU_ESCAPED;
  int keep = (i < n) && f != SGPU_ESCAPED && f != SGPU_ABSORBED;
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  unsigned long long mask = __ballot(keep);
  int before = __popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) woff[w] = __popcll(mask);
  __syncthreads();
  int base = blk_off[blockIdx.x];
  for (int k = 0; k < w; k++) base += woff[k];
  if (keep) {
    int64_t o = base + before;
    dst.x[o] = src.x[i]; dst.y[o] = src.y[i]; dst.z[o] = src.z[i];
    dst.Dx[o] = src.Dx[i]; dst.Dy[o] = src.Dy[i]; dst.Dz[o] = src.Dz[i];
    dst.t[o] = src.t[i]; dst.e[o] = src.e[i]; dst.nu[o] = src.nu[i];
    dst.ind[o] = src.ind[i]; dst.type[o] = src.type[i]; dst.fate[o] = src.fate[i];
    dst.draw[o] = src.draw[i]; dst.id[o] = src.id[i];
  }
}

template <class T>
__global__ void copy_kernel(const T* __restrict__ s, T* __restrict__ d, int64_t n)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

// the first tot[0] packets of src into dst, every field (one launch)
__global__ void copy_packets_kernel(sgpu_packets src, sgpu_packets dst, const long long* tot)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *tot) return;
  dst.x[i] = src.x[i]; dst.y[i] = src.y[i]; dst.z[i] = src.z[i];
  dst.Dx[i] = src.Dx[i]; dst.Dy[i] = src.Dy[i]; dst.Dz[i] = src.Dz[i];
  dst.t[i] = src.t[i]; dst.e[i] = src.e[i]; dst.nu[i] = src.nu[i];
  dst.ind[i] = src.ind[i]; dst.type[i] = src.type[i]; dst.fate[i] = src.fate[i];
  dst.draw[i] = src.draw[i]; dst.id[i] = src.id[i];
}

// the device counts after a compaction
__global__ void compact_counts_kernel(const long long* tot, int64_t* n_alive, int64_t* n_escaped)
{
  *n_alive = tot[0];
  if (n_escaped) *n_escaped = tot[1];
}

__global__ void fill_kernel(double* p, int64_t n, double v)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------
// K7: tally normalisation (transport_mpi.cpp:415-441)
// ------------------------------------------------------------------
__global__ void __launch_bounds__(256) normalize_kernel(sgpu_locate nu, const double* vol, double dt,
                                                        int store_Jnu, sgpu_tallies ta)
{
  __shared__ double red[256 / WAVE];
  const int i = blockIdx.x, tid = threadIdx.x, nn = nu.n;
  const double v = vol[i];
  if (tid == 0) {
    ta.e_abs[i] /= v * dt;
    ta.L_radio_dep[i] /= v * dt;
    ta.L_radio_emit[i] /= v;
    ta.fr_rad[i] /= v * C_LIGHT * dt;
  }
  if (nn == 1 || !store_Jnu) {
    if (tid == 0) ta.e_rad[i] = ta.J_nu[(size_t)i * (store_Jnu ? nn : 1)] / (v * dt * C_LIGHT);
    return;
  }
  double es = 0;
  for (int j = tid; j < nn; j += blockDim.x) {
    double* J = &ta.J_nu[(size_t)i * nn + j];
    double dnu = loc_delta(nu, j);
    *J /= v * dt * 4 * PI * dnu;
    es += *J * dnu * 4 * PI / C_LIGHT;
  }
  es = wave_sum(es);
  if (tid % WAVE == 0) red[tid / WAVE] = es;
  __syncthreads();
  if (tid == 0) {
    double t = 0;
    for (int w = 0; w < 256 / WAVE; w++) t += red[w];
    ta.e_rad[i] = t;
  }
}

// ------------------------------------------------------------------
// homologous step (hydro_homologous.cpp:27-61, grid expand)
// ------------------------------------------------------------------
__global__ void homologous_kernel(int kind, int nz, double* rho, double* r_out, int nr, double* x_out,
                                  int nx, double* z_out, int nzz, double* vol, double r_min_new, double e)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nz) rho[i] = rho[i] / e / e / e;
  if (kind == SGPU_GRID_1D_SPHERE) {
    if (i < nr) r_out[i] *= e;
  } else {
    if (i < nx) x_out[i] *= e;
    if (i < nzz) z_out[i] *= e;
    if (i < nz) vol[i] = vol[i] * e * e * e;
  }
}

__global__ void shell_volume_kernel(int nz, const double* r_out, double r_min, double* vol)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nz) return;
  double r0 = (i == 0) ? r_min : r_out[i - 1];
  vol[i] = 4.0 * PI / 3.0 * (r_out[i] * r_out[i] * r_out[i] - r0 * r0 * r0);
}

}  // namespace sgpu

using namespace sgpu;

namespace {

template <class R>
Axis<R> make_axis(const sgpu_locate& l, const R* x)
{
  Axis<R> a;
  a.n = l.n; a.type = l.type; a.min = (R)l.min; a.del = (R)l.del;
  a.inv_log1p = (R)(1.0 / log1p(l.del));
  a.x = x;
  return a;
}

template <class R>
GridView<R> make_grid(const sgpu_grid& g, const R* r, const R* xo, const R* zo)
{
  GridView<R> v;
  v.kind = g.kind; v.n_zones = g.n_zones;
  v.t_now = (R)g.t_now; v.inv_t = (R)(1.0 / g.t_now); v.r_core = (R)g.r_core;
  v.r_out = make_axis<R>(g.r_out, r);
  v.x_out = make_axis<R>(g.x_out, xo);
  v.z_out = make_axis<R>(g.z_out, zo);
  return v;
}

int check_grid(const sgpu_grid* g)
{
  SGPU_REQUIRE(g, "null grid");
  SGPU_REQUIRE(g->kind == SGPU_GRID_1D_SPHERE || g->kind == SGPU_GRID_2D_CYLN, "unknown grid kind %d", g->kind);
  SGPU_REQUIRE(g->t_now > 0, "grid t_now must be > 0");
  SGPU_REQUIRE(g->vol, "grid volumes missing");
  if (g->kind == SGPU_GRID_1D_SPHERE)
    SGPU_REQUIRE(g->r_out.x && g->r_out.n == g->n_zones && g->n_zones > 0, "1D grid needs r_out with n_zones bins");
  else
    SGPU_REQUIRE(g->x_out.x && g->z_out.x && (int64_t)g->x_out.n * g->z_out.n == g->n_zones,
                 "2D grid needs x_out, z_out with nx*nz = n_zones");
  return SGPU_OK;
}

int check_packets(const sgpu_packets* p)
{
  SGPU_REQUIRE(p && p->x && p->y && p->z && p->Dx && p->Dy && p->Dz && p->t && p->e && p->nu &&
               p->ind && p->type && p->fate && p->draw && p->id, "packet arrays missing");
  return SGPU_OK;
}

unsigned grid_blocks(int64_t n, int threads)
{
  return (unsigned)((n + threads - 1) / threads);
}

}  // namespace

namespace {
// the propagation of packets [0, n), or of [0, *n_dev) read on the device
// when n_dev is set (the grids are then sized for n, the buffer's capacity)
int propagate_impl(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid, const sgpu_tables* tables,
                   const sgpu_transport_params* params, sgpu_packets* packets, int64_t n, const int64_t* n_dev,
                   sgpu_tallies* tallies, sgpu_counters* counters)
{
  SGPU_REQUIRE(ctx && nu_grid && tables && params && tallies, "null argument");
  int st = check_grid(grid);
  if (st) return st;
  st = check_packets(packets);
  if (st) return st;
  SGPU_REQUIRE(n >= 0 && n <= packets->capacity, "n = %lld outside [0, capacity = %lld]", (long long)n,
               (long long)packets->capacity);
  SGPU_REQUIRE(nu_grid->n >= 1 && nu_grid->x, "bad frequency grid");
  SGPU_REQUIRE(tables->abs && tables->emis_cdf && tables->eps_imc && tables->compton_opac &&
               tables->photoion_opac && (tables->scat || params->omit_scattering), "opacity tables missing");
  SGPU_REQUIRE(tallies->e_abs && tallies->J_nu && tallies->fx_rad && tallies->fy_rad && tallies->fz_rad &&
               tallies->fr_rad && tallies->L_radio_dep, "tally arrays missing");
  SGPU_REQUIRE(params->precision == SGPU_PREC_FP64 || params->precision == SGPU_PREC_MIXED ||
               params->precision == SGPU_PREC_FP64_BATCHED, "unknown precision %d", params->precision);
  SGPU_REQUIRE(params->use_ddmc == 0 || params->use_ddmc == 2,
               "transport_use_ddmc = %d: only 0 (off) and 2 (DDMC) are supported", params->use_ddmc);
  if (params->use_ddmc == 2)
    SGPU_REQUIRE(grid->kind == SGPU_GRID_1D_SPHERE && params->precision != SGPU_PREC_MIXED &&
                 params->ddmc_tau_threshold >= 0,
                 "DDMC runs on 1D spherical grids in the FP64 form with a threshold >= 0");
  if (n == 0) return SGPU_OK;
  const int nz = grid->n_zones, nn = nu_grid->n;
  if (params->precision == SGPU_PREC_FP64_BATCHED && params->use_ddmc != 2) {
    PropArgs<double> A;
    A.g = make_grid<double>(*grid, grid->r_out.x, grid->x_out.x, grid->z_out.x);
    A.nu = make_axis<double>(*nu_grid, nu_grid->x);
    A.tab = {tables->abs, tables->scat, tables->emis_cdf};
    A.eps_imc = tables->eps_imc; A.compton = tables->compton_opac; A.photoion = tables->photoion_opac;
    A.p = *params; A.pk = *packets; A.n = n; A.n_dev = n_dev; A.ta = *tallies; A.cnt = counters;
    A.lds_tallies = nz <= LDS_ZONE_MAX;
    const size_t lds = A.lds_tallies ? sizeof(double) * T_NFIELDS * nz : 0;
    const char* ew = getenv("SGPU_PROP_WG_PER_CU");
    const int wg_per_cu = ew ? std::max(1, atoi(ew)) : 12;
    const int64_t want = (int64_t)ctx->n_cu * wg_per_cu, need = grid_blocks(n, PROP_THREADS);
    const dim3 gb((unsigned)(want < need ? want : need)), tb(PROP_THREADS);
    timer_begin(ctx, "propagate");
    if (grid->kind == SGPU_GRID_1D_SPHERE)
      hipLaunchKernelGGL((propagate_kernel<double, true, true, 1, SGPU_GRID_1D_SPHERE>), gb, tb, lds, ctx->stream, A);
    else
      hipLaunchKernelGGL((propagate_kernel<double, true, true, 1, SGPU_GRID_2D_CYLN>), gb, tb, lds, ctx->stream, A);
    SGPU_LAUNCH_CHECK();
  } else if (params->precision != SGPU_PREC_MIXED) {
    PropArgs<double> A;
    A.g = make_grid<double>(*grid, grid->r_out.x, grid->x_out.x, grid->z_out.x);
    A.nu = make_axis<double>(*nu_grid, nu_grid->x);
    A.tab = {tables->abs, tables->scat, tables->emis_cdf};
    A.eps_imc = tables->eps_imc; A.compton = tables->compton_opac; A.photoion = tables->photoion_opac;
    A.p = *params; A.pk = *packets; A.n = n; A.n_dev = n_dev; A.ta = *tallies; A.cnt = counters;
    A.lds_tallies = 0;
    timer_begin(ctx, "propagate");
    if (params->use_ddmc == 2)
      hipLaunchKernelGGL((propagate_kernel<double, false, false, 1, 0, true>), dim3(grid_blocks(n, PROP_THREADS)),
                         dim3(PROP_THREADS), 0, ctx->stream, A);
    else
      hipLaunchKernelGGL((propagate_kernel<double, false, false>), dim3(grid_blocks(n, PROP_THREADS)),
                         dim3(PROP_THREADS), 0, ctx->stream, A);
    SGPU_LAUNCH_CHECK();
  } else {
    const size_t ntab = (size_t)nz * nn;
    float2* ke = (float2*)scratch(ctx, "prop_ke", sizeof(float2) * ntab, &st);
    if (st) return st;
    SGPU_REQUIRE(nn <= ALIAS_LDS_MAX, "the mixed form samples frequencies through alias tables built "
                 "in LDS: n_nu = %d exceeds %d", nn, ALIAS_LDS_MAX);
    AliasEntry* alias = (AliasEntry*)scratch(ctx, "prop_alias", sizeof(AliasEntry) * ntab, &st);
    if (st) return st;
    const int n_edges = nn + grid->r_out.n + grid->x_out.n + grid->z_out.n;
    float* e32 = (float*)scratch(ctx, "prop_edges32", sizeof(float) * n_edges, &st);
    if (st) return st;
    hipLaunchKernelGGL(pack_tables_kernel, dim3(grid_blocks((int64_t)ntab, 256)), dim3(256), 0, ctx->stream,
                       tables->abs, tables->scat, params->omit_scattering, ntab, ke);
    hipLaunchKernelGGL(alias_kernel, dim3(nz), dim3(64), (sizeof(double) + 2 * sizeof(uint16_t)) * nn,
                       ctx->stream, tables->emis_cdf, nn, alias);
    SGPU_LAUNCH_CHECK();
    float* nu32 = e32;
    float* r32 = nu32 + nn;
    float* x32 = r32 + grid->r_out.n;
    float* z32 = x32 + grid->x_out.n;
    hipLaunchKernelGGL(to_f32_kernel, dim3(grid_blocks(nn, 256)), dim3(256), 0, ctx->stream, nu_grid->x, nu32, nn);
    if (grid->r_out.n > 0 && grid->r_out.x)
      hipLaunchKernelGGL(to_f32_kernel, dim3(grid_blocks(grid->r_out.n, 256)), dim3(256), 0, ctx->stream,
                         grid->r_out.x, r32, grid->r_out.n);
    if (grid->x_out.n > 0 && grid->x_out.x)
      hipLaunchKernelGGL(to_f32_kernel, dim3(grid_blocks(grid->x_out.n, 256)), dim3(256), 0, ctx->stream,
                         grid->x_out.x, x32, grid->x_out.n);
    if (grid->z_out.n > 0 && grid->z_out.x)
      hipLaunchKernelGGL(to_f32_kernel, dim3(grid_blocks(grid->z_out.n, 256)), dim3(256), 0, ctx->stream,
                         grid->z_out.x, z32, grid->z_out.n);
    SGPU_LAUNCH_CHECK();
    PropArgs<float> A;
    A.g = make_grid<float>(*grid, r32, x32, z32);
    A.nu = make_axis<float>(*nu_grid, nu32);
    A.tab = {ke, alias};
    A.eps_imc = tables->eps_imc; A.compton = tables->compton_opac; A.photoion = tables->photoion_opac;
    A.p = *params; A.pk = *packets; A.n = n; A.n_dev = n_dev; A.ta = *tallies; A.cnt = counters;
    // launch shape (tuning knobs for sweeps, defaults below):
    //   SGPU_PROP_SHAPE  0 persistent lanes + register tallies (default)
    //                    1 one packet per thread + register tallies
    //                    2 one packet per thread, per-event atomics
    //   SGPU_PROP_WG_PER_CU  persistent workgroups per CU (default 12)
    //   SGPU_PROP_LDS    0: zone tallies as global atomics instead of LDS
    //   SGPU_PROP_OCC    waves per SIMD the register allocation must allow:
    //                    4 (default: 128 VGPRs, 32 spilled; 1% faster than 3
    //                    at 1e7 packets, 6% at 1e6), 3 (168 VGPRs), 1 (no bound)
    const char* es = getenv("SGPU_PROP_SHAPE");
    const int shape = es ? atoi(es) : 0;
    const char* ew = getenv("SGPU_PROP_WG_PER_CU");
    const int wg_per_cu = ew ? std::max(1, atoi(ew)) : 12;
    const char* el = getenv("SGPU_PROP_LDS");
    A.lds_tallies = nz <= LDS_ZONE_MAX && shape != 2 && !(el && atoi(el) == 0);
    const size_t lds = (A.lds_tallies ? sizeof(double) * T_NFIELDS * nz : 0) + sizeof(float) * n_edges;
    SGPU_REQUIRE(lds <= 96 * 1024, "mixed-form LDS (zone tallies + %d FP32 edges) exceeds 96 KiB", n_edges);
    int64_t want = (int64_t)ctx->n_cu * wg_per_cu;
    int64_t need = grid_blocks(n, PROP_THREADS);
    unsigned blocks = (unsigned)(want < need ? want : need);
    const char* eo = getenv("SGPU_PROP_OCC");
    const int occ = eo ? atoi(eo) : 4;
    const bool d1 = grid->kind == SGPU_GRID_1D_SPHERE;
    timer_begin(ctx, "propagate");
    const dim3 gb(blocks), tb(PROP_THREADS);
    if (shape == 1)
      hipLaunchKernelGGL((propagate_kernel<float, false, true>), dim3((unsigned)need), tb, lds, ctx->stream, A);
    else if (shape == 2)
      hipLaunchKernelGGL((propagate_kernel<float, false, false>), dim3((unsigned)need), tb, lds, ctx->stream, A);
    else if (occ == 4 && d1)
      hipLaunchKernelGGL((propagate_kernel<float, true, true, 4, SGPU_GRID_1D_SPHERE>), gb, tb, lds, ctx->stream, A);
    else if (occ == 4)
      hipLaunchKernelGGL((propagate_kernel<float, true, true, 4, SGPU_GRID_2D_CYLN>), gb, tb, lds, ctx->stream, A);
    else if (occ == 3 && d1)
      hipLaunchKernelGGL((propagate_kernel<float, true, true, 3, SGPU_GRID_1D_SPHERE>), gb, tb, lds, ctx->stream, A);
    else if (occ == 3)
      hipLaunchKernelGGL((propagate_kernel<float, true, true, 3, SGPU_GRID_2D_CYLN>), gb, tb, lds, ctx->stream, A);
    else if (d1)
      hipLaunchKernelGGL((propagate_kernel<float, true, true, 1, SGPU_GRID_1D_SPHERE>), gb, tb, lds, ctx->stream, A);
    else
      hipLaunchKernelGGL((propagate_kernel<float, true, true, 1, SGPU_GRID_2D_CYLN>), gb, tb, lds, ctx->stream, A);
    SGPU_LAUNCH_CHECK();
  }
  timer_end(ctx, "propagate");
  return SGPU_OK;
}
}  // namespace

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}

// `per` FP64 atomic additions of 1.0 per thread at hash-scattered indices
__global__ void __launch_bounds__(256) atomic_rate_kernel(double* buf, uint32_t n, int per)
{
  uint32_t h = mix32(blockIdx.x * blockDim.x + threadIdx.x);
  for (int k = 0; k < per; ++k) {
    h = mix32(h + (uint32_t)k);
    atomic_add_f64(&buf[h % n], 1.0);
  }
}

}  // namespace

extern "C" int sgpu_atomic_rate(sgpu_ctx* ctx, double* buf, int64_t n, int64_t adds, double* ms)
{
  SGPU_REQUIRE(ctx && buf && ms, "null argument");
  SGPU_REQUIRE(n >= 1 && n <= 0xffffffffLL, "n = %lld outside [1, 2^32)", (long long)n);
  const int per = 256;
  const int64_t threads = (adds + per - 1) / per;
  SGPU_REQUIRE(adds >= 1 && threads % 256 == 0 && threads / 256 <= 0x7fffffff,
               "adds = %lld must be a positive multiple of 65536", (long long)adds);
  hipEvent_t e0, e1;
  SGPU_HIP(hipEventCreate(&e0));
  SGPU_HIP(hipEventCreate(&e1));
  SGPU_HIP(hipEventRecord(e0, ctx->stream));
  hipLaunchKernelGGL(atomic_rate_kernel, dim3((unsigned)(threads / 256)), dim3(256), 0, ctx->stream, buf,
                     (uint32_t)n, per);
  SGPU_LAUNCH_CHECK();
  SGPU_HIP(hipEventRecord(e1, ctx->stream));
  SGPU_HIP(hipEventSynchronize(e1));
  float t = 0;
  SGPU_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return SGPU_OK;
}

extern "C" int sgpu_propagate(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                              const sgpu_tables* tables, const sgpu_transport_params* params,
                              sgpu_packets* packets, int64_t n, sgpu_tallies* tallies,
                              sgpu_counters* counters)
{
  return propagate_impl(ctx, grid, nu_grid, tables, params, packets, n, nullptr, tallies, counters);
}

extern "C" int sgpu_propagate_dev(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                  const sgpu_tables* tables, const sgpu_transport_params* params,
                                  sgpu_packets* packets, const int64_t* d_n, sgpu_tallies* tallies,
                                  sgpu_counters* counters)
{
  SGPU_REQUIRE(d_n && packets, "null argument");
  return propagate_impl(ctx, grid, nu_grid, tables, params, packets, packets->capacity, d_n, tallies, counters);
}

namespace {
// the radioactive emission's kernels; the offset from the host (d_count
// null) or from the device count, which the emission then advances
int emit_radioactive_impl(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                          const sgpu_zone_state* zones, const sgpu_tables* tables,
                          const sgpu_transport_params* params, int64_t n_emit, int32_t force_rproc,
                          uint64_t id_base, sgpu_packets* packets, int64_t offset, int64_t* d_count,
                          sgpu_tallies* tallies, double** ltot_out)
{
  int st = 0;
  const int nz = grid->n_zones;
  double* zcdf = (double*)scratch(ctx, "emit_zcdf", sizeof(double) * (2 * (size_t)nz + 1), &st);
  if (st) return st;
  double* gfr = zcdf + nz;
  double* ltot = gfr + nz;
  timer_begin(ctx, "emit");
  EmitPrep P{*grid, *zones, params->t_now, force_rproc, zcdf, gfr, tallies->L_radio_emit, ltot};
  hipLaunchKernelGGL(emit_prep_kernel, dim3(1), dim3(1024), 0, ctx->stream, P);
  SGPU_LAUNCH_CHECK();
  if (n_emit > 0) {
    EmitArgs A;
    A.g = make_grid<double>(*grid, grid->r_out.x, grid->x_out.x, grid->z_out.x);
    A.nu = make_axis<double>(*nu_grid, nu_grid->x);
    A.emis_cdf = tables->emis_cdf; A.zcdf = zcdf; A.gfrac = gfr; A.L_tot = ltot;
    A.t_now = params->t_now; A.dt = params->dt; A.seed = params->seed; A.id_base = id_base;
    A.n_emit = n_emit; A.offset = offset; A.offset_dev = d_count; A.pk = *packets;
    A.L_dep = tallies->L_radio_dep;
    A.lds_dep = nz <= 4096;
    hipLaunchKernelGGL(emit_kernel, dim3(grid_blocks(n_emit, 256)), dim3(256),
                       A.lds_dep ? sizeof(double) * nz : 0, ctx->stream, A);
    SGPU_LAUNCH_CHECK();
    if (d_count) {
      hipLaunchKernelGGL(emit_count_kernel, dim3(1), dim3(1), 0, ctx->stream, ltot, n_emit, packets->capacity,
                         d_count);
      SGPU_LAUNCH_CHECK();
    }
  }
  timer_end(ctx, "emit");
  *ltot_out = ltot;
  return SGPU_OK;
}

int emit_args(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid, const sgpu_zone_state* zones,
              const sgpu_tables* tables, const sgpu_transport_params* params, const sgpu_packets* packets,
              const sgpu_tallies* tallies, int64_t n_emit)
{
  SGPU_REQUIRE(ctx && nu_grid && zones && tables && params && tallies, "null argument");
  int st = check_grid(grid);
  if (st) return st;
  st = check_packets(packets);
  if (st) return st;
  SGPU_REQUIRE(zones->n_zones == grid->n_zones, "zone count mismatch");
  SGPU_REQUIRE(n_emit >= 0, "negative packet count");
  SGPU_REQUIRE(tables->emis_cdf && tallies->L_radio_dep && tallies->L_radio_emit, "missing arrays");
  return SGPU_OK;
}
}  // namespace

extern "C" int sgpu_emit_radioactive_dev(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                         const sgpu_zone_state* zones, const sgpu_tables* tables,
                                         const sgpu_transport_params* params, int64_t n_emit,
                                         int32_t force_rproc, uint64_t id_base, sgpu_packets* packets,
                                         int64_t* d_count, sgpu_tallies* tallies)
{
  int st = emit_args(ctx, grid, nu_grid, zones, tables, params, packets, tallies, n_emit);
  if (st) return st;
  SGPU_REQUIRE(d_count, "device count missing");
  double* ltot;
  return emit_radioactive_impl(ctx, grid, nu_grid, zones, tables, params, n_emit, force_rproc, id_base, packets,
                               0, d_count, (sgpu_tallies*)tallies, &ltot);
}

extern "C" int sgpu_emit_radioactive(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                     const sgpu_zone_state* zones, const sgpu_tables* tables,
                                     const sgpu_transport_params* params, int64_t n_emit,
                                     int32_t force_rproc, uint64_t id_base, sgpu_packets* packets,
                                     int64_t offset, sgpu_tallies* tallies, int64_t* n_out)
{
  int st = emit_args(ctx, grid, nu_grid, zones, tables, params, packets, tallies, n_emit);
  if (st) return st;
  SGPU_REQUIRE(n_out, "null argument");
  SGPU_REQUIRE(offset >= 0 && offset + n_emit <= packets->capacity,
               "emission range [%lld, %lld) exceeds capacity %lld", (long long)offset,
               (long long)(offset + n_emit), (long long)packets->capacity);
  *n_out = 0;
  double* ltot;
  if ((st = emit_radioactive_impl(ctx, grid, nu_grid, zones, tables, params, n_emit, force_rproc, id_base, packets,
                                  offset, nullptr, tallies, &ltot)))
    return st;
  double hl = 0;
  SGPU_HIP(hipMemcpyAsync(&hl, ltot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  *n_out = (hl == 0) ? 0 : n_emit;
  return SGPU_OK;
}

extern "C" int sgpu_initialize_particles(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                         const sgpu_zone_state* zones, const double* e_rad, double* emis_cdf,
                                         const sgpu_transport_params* params, int64_t n_init, uint64_t id_base,
                                         sgpu_packets* packets, int64_t offset, int64_t* n_out)
{
  SGPU_REQUIRE(ctx && nu_grid && zones && params && n_out, "null argument");
  int st = check_grid(grid);
  if (st) return st;
  st = check_packets(packets);
  if (st) return st;
  SGPU_REQUIRE(zones->n_zones == grid->n_zones, "zone count mismatch");
  SGPU_REQUIRE(e_rad && emis_cdf && zones->T_gas, "missing arrays");
  SGPU_REQUIRE(nu_grid->n >= 1 && nu_grid->x, "bad frequency grid");
  SGPU_REQUIRE(n_init >= 0 && offset >= 0, "negative count or offset");
  *n_out = 0;
  // emission.cpp:116-129: nothing to do, or no room -- return before the
  // emissivity is rewritten
  if (n_init == 0 || offset + n_init > packets->capacity) return SGPU_OK;
  const int nz = grid->n_zones;
  double* zcdf = (double*)scratch(ctx, "init_zcdf", sizeof(double) * ((size_t)nz + 1), &st);
  if (st) return st;
  double* esum = zcdf + nz;
  timer_begin(ctx, "emit");
  hipLaunchKernelGGL(init_zone_cdf_kernel, dim3(1), dim3(1024), 0, ctx->stream, e_rad, grid->vol, nz, zcdf, esum);
  hipLaunchKernelGGL(init_blackbody_cdf_kernel, dim3(nz), dim3(INIT_THREADS), 0, ctx->stream, *nu_grid,
                     zones->T_gas, emis_cdf);
  SGPU_LAUNCH_CHECK();
  InitArgs A;
  A.g = make_grid<double>(*grid, grid->r_out.x, grid->x_out.x, grid->z_out.x);
  A.nu = make_axis<double>(*nu_grid, nu_grid->x);
  A.emis_cdf = emis_cdf; A.zcdf = zcdf; A.E_sum = esum;
  A.t_now = params->t_now; A.seed = params->seed; A.id_base = id_base;
  A.n_init = n_init; A.offset = offset; A.pk = *packets;
  hipLaunchKernelGGL(init_particles_kernel, dim3(grid_blocks(n_init, 256)), dim3(256), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "emit");
  *n_out = n_init;
  return SGPU_OK;
}

namespace {
// the three compaction kernels and the copy back of the survivors, with the
// per-block counts blk [2 * nb] and the two totals tot in caller-chosen device
// memory; *n_alive / *n_escaped come back through a synchronous copy
int compact_with(sgpu_ctx* ctx, sgpu_packets* packets, int64_t n, sgpu_packets* scr, int32_t* blk,
                 long long* tot, int64_t* n_alive, int64_t* n_escaped)
{
  const unsigned nb = grid_blocks(n, CMP_THREADS);
  hipLaunchKernelGGL(keep_count_kernel, dim3(nb), dim3(CMP_THREADS), 0, ctx->stream, packets->fate, n, blk,
                     blk + nb, nullptr);
  hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, ctx->stream, blk, blk + nb, (int)nb, tot);
  hipLaunchKernelGGL(scatter_keep_kernel, dim3(nb), dim3(CMP_THREADS), 0, ctx->stream, *packets, *scr, n, blk,
                     nullptr);
  SGPU_LAUNCH_CHECK();
  long long ht[2];
  SGPU_HIP(hipMemcpyAsync(ht, tot, sizeof(ht), hipMemcpyDeviceToHost, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  const int64_t k = ht[0];
  if (k > 0) {
    const unsigned b = grid_blocks(k, 256);
    const sgpu_packets& s = *scr;
    const sgpu_packets& d = *packets;
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.x, d.x, k);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.y, d.y, k);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.z, d.z, k);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.Dx, d.Dx, k);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.Dy, d.Dy, k);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.Dz, d.Dz, k);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.t, d.t, k);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.e, d.e, k);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(b), dim3(256), 0, ctx->stream, s.nu, d.nu, k);
    hipLaunchKernelGGL(copy_kernel<int32_t>, dim3(b), dim3(256), 0, ctx->stream, s.ind, d.ind, k);
    hipLaunchKernelGGL(copy_kernel<int32_t>, dim3(b), dim3(256), 0, ctx->stream, s.type, d.type, k);
    hipLaunchKernelGGL(copy_kernel<int32_t>, dim3(b), dim3(256), 0, ctx->stream, s.fate, d.fate, k);
    hipLaunchKernelGGL(copy_kernel<uint32_t>, dim3(b), dim3(256), 0, ctx->stream, s.draw, d.draw, k);
    hipLaunchKernelGGL(copy_kernel<uint64_t>, dim3(b), dim3(256), 0, ctx->stream, s.id, d.id, k);
    SGPU_LAUNCH_CHECK();
  }
  *n_alive = k;
  *n_escaped = ht[1];
  return SGPU_OK;
}

int compact_args(sgpu_ctx* ctx, sgpu_packets* packets, int64_t n, sgpu_packets* scr, int64_t* n_alive,
                 int64_t* n_escaped)
{
  SGPU_REQUIRE(ctx && n_alive && n_escaped, "null argument");
  int st = check_packets(packets);
  if (st) return st;
  st = check_packets(scr);
  if (st) return st;
  SGPU_REQUIRE(n >= 0 && n <= packets->capacity && n <= scr->capacity, "n exceeds capacity");
  *n_alive = 0; *n_escaped = 0;
  return SGPU_OK;
}
}  // namespace

extern "C" int sgpu_emit_inner_source(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                      const sgpu_core_source* core, const sgpu_transport_params* params,
                                      int64_t n_emit, uint64_t id_base, sgpu_packets* packets, int64_t offset,
                                      int64_t* n_out)
{
  SGPU_REQUIRE(ctx && nu_grid && core && params && n_out, "null argument");
  int st = check_grid(grid);
  if (st) return st;
  st = check_packets(packets);
  if (st) return st;
  SGPU_REQUIRE(nu_grid->n >= 1 && nu_grid->x, "bad frequency grid");
  SGPU_REQUIRE(n_emit >= 0 && offset >= 0, "negative count or offset");
  SGPU_REQUIRE(core->r_core >= 0, "core radius must be >= 0");
  *n_out = 0;
  // emission.cpp:375, 391-392: nothing to emit, or no room
  if (n_emit == 0 || offset + n_emit > packets->capacity) return SGPU_OK;
  double T_core = core->T_core;
  // transport_init.cpp:478-479: the blackbody of L and R
  if (core->L_core != 0 && core->r_core != 0 && T_core == 0)
    T_core = pow(core->L_core / (4.0 * PI * core->r_core * core->r_core * SB), 0.25);
  const int nn = nu_grid->n;
  double* cdf = (double*)scratch(ctx, "core_cdf", sizeof(double) * ((size_t)nn + 1), &st);
  if (st) return st;
  double* ep = cdf + nn;
  timer_begin(ctx, "emit");
  hipLaunchKernelGGL(core_prep_kernel, dim3(1), dim3(INIT_THREADS), 0, ctx->stream, *nu_grid, core->r_core, T_core,
                     core->L_core, params->dt, n_emit, cdf, ep);
  SGPU_LAUNCH_CHECK();
  CoreArgs A;
  A.g = make_grid<double>(*grid, grid->r_out.x, grid->x_out.x, grid->z_out.x);
  A.nu = make_axis<double>(*nu_grid, nu_grid->x);
  A.cdf = cdf; A.Ep = ep; A.r_core = core->r_core; A.core_frequency = core->core_frequency;
  A.t_now = params->t_now; A.dt = params->dt; A.seed = params->seed; A.id_base = id_base;
  A.n_emit = n_emit; A.offset = offset; A.pk = *packets;
  hipLaunchKernelGGL(core_emit_kernel, dim3(grid_blocks(n_emit, 256)), dim3(256), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "emit");
  *n_out = n_emit;
  return SGPU_OK;
}

extern "C" int sgpu_compact_dev(sgpu_ctx* ctx, sgpu_packets* packets, int64_t* d_count, sgpu_packets* scr,
                                int64_t* d_escaped)
{
  SGPU_REQUIRE(ctx && d_count, "null argument");
  int st = check_packets(packets);
  if (st) return st;
  st = check_packets(scr);
  if (st) return st;
  SGPU_REQUIRE(scr->capacity >= packets->capacity, "staging buffer smaller than the packet buffer");
  const int64_t cap = packets->capacity;
  if (cap == 0) return SGPU_OK;
  const unsigned nb = grid_blocks(cap, CMP_THREADS);
  int32_t* blk = (int32_t*)scratch(ctx, "cmpd_blk", sizeof(int32_t) * 2 * (size_t)nb + 64, &st);
  if (st) return st;
  long long* tot = (long long*)scratch(ctx, "cmpd_tot", sizeof(long long) * 2, &st);
  if (st) return st;
  hipLaunchKernelGGL(keep_count_kernel, dim3(nb), dim3(CMP_THREADS), 0, ctx->stream, packets->fate, cap, blk,
                     blk + nb, (const int64_t*)d_count);
  hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, ctx->stream, blk, blk + nb, (int)nb, tot);
  hipLaunchKernelGGL(scatter_keep_kernel, dim3(nb), dim3(CMP_THREADS), 0, ctx->stream, *packets, *scr, cap, blk,
                     (const int64_t*)d_count);
  hipLaunchKernelGGL(copy_packets_kernel, dim3(grid_blocks(cap, 256)), dim3(256), 0, ctx->stream, *scr, *packets,
                     tot);
  hipLaunchKernelGGL(compact_counts_kernel, dim3(1), dim3(1), 0, ctx->stream, tot, d_count, d_escaped);
  SGPU_LAUNCH_CHECK();
  return SGPU_OK;
}

extern "C" int sgpu_compact(sgpu_ctx* ctx, sgpu_packets* packets, int64_t n, sgpu_packets* scr,
                            int64_t* n_alive, int64_t* n_escaped)
{
  int st = compact_args(ctx, packets, n, scr, n_alive, n_escaped);
  if (st || n == 0) return st;
  const unsigned nb = grid_blocks(n, CMP_THREADS);
  int32_t* blk = (int32_t*)scratch(ctx, "cmp_blk", sizeof(int32_t) * 2 * (size_t)nb + 64, &st);
  if (st) return st;
  long long* tot = (long long*)scratch(ctx, "cmp_tot", sizeof(long long) * 2, &st);
  if (st) return st;
  return compact_with(ctx, packets, n, scr, blk, tot, n_alive, n_escaped);
}

// ------------------------------------------------------------------
// sgpu_compact_noalloc: the same compaction with its per-block counts and
// totals in a fixed file-scope device array, so that the call allocates
// nothing.  Every GPU holds its own copy of the array, at its own address,
// until the process ends; the address is looked up once per device with
// hipGetSymbolAddress and only that address is handed to kernels and copies.
// Calls that overlap on one GPU share the array and race.
// ------------------------------------------------------------------
constexpr int64_t NOALLOC_MAX_PACKETS = (int64_t)1 << 28;                   // 268,435,456
constexpr size_t NOALLOC_MAX_BLOCKS = (size_t)(NOALLOC_MAX_PACKETS / CMP_THREADS);   // 2^20
// int32 words: keep and escape counts of every block, then 256 B holding the
// two 64-bit totals (8 MiB + 256 B in all)
constexpr size_t NOALLOC_WORDS = 2 * NOALLOC_MAX_BLOCKS + 64;
static __device__ __align__(256) int32_t compact_scratch[NOALLOC_WORDS];

namespace {
std::mutex noalloc_mu;
std::map<int, int32_t*> noalloc_addr;   // device -> device address of compact_scratch

int noalloc_scratch(sgpu_ctx* ctx, int32_t** out)
{
  std::lock_guard<std::mutex> lock(noalloc_mu);
  auto it = noalloc_addr.find(ctx->device);
  if (it != noalloc_addr.end()) { *out = it->second; return SGPU_OK; }
  int cur = -1;
  SGPU_HIP(hipGetDevice(&cur));
  if (cur != ctx->device) SGPU_HIP(hipSetDevice(ctx->device));
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(compact_scratch));
  if (cur != ctx->device) SGPU_HIP(hipSetDevice(cur));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(compact_scratch)", __FILE__, __LINE__);
  SGPU_REQUIRE(p != nullptr, "no device address for the compaction scratch array");
  noalloc_addr[ctx->device] = (int32_t*)p;
  *out = (int32_t*)p;
  return SGPU_OK;
}
}  // namespace

extern "C" int sgpu_compact_noalloc(sgpu_ctx* ctx, sgpu_packets* packets, int64_t n, sgpu_packets* scr,
                                    int64_t* n_alive, int64_t* n_escaped)
{
  int st = compact_args(ctx, packets, n, scr, n_alive, n_escaped);
  if (st) return st;
  if (n > NOALLOC_MAX_PACKETS) {
    set_error("sgpu_compact_noalloc: %lld packets exceed its fixed scratch limit of %lld",
              (long long)n, (long long)NOALLOC_MAX_PACKETS);
    return SGPU_ERR_SIZE;
  }
  if (n == 0) return SGPU_OK;
  int32_t* base = nullptr;
  if ((st = noalloc_scratch(ctx, &base))) return st;
  // every word the kernels read is written first in the same call (block
  // counts by keep_count_kernel, totals by scan_blocks_kernel): nothing to clear
  long long* tot = (long long*)(base + 2 * NOALLOC_MAX_BLOCKS);
  return compact_with(ctx, packets, n, scr, base, tot, n_alive, n_escaped);
}

static int fill(sgpu_ctx* ctx, double* p, int64_t n)
{
  if (!p || n <= 0) return SGPU_OK;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_blocks(n, 256)), dim3(256), 0, ctx->stream, p, n, 0.0);
  SGPU_LAUNCH_CHECK();
  return SGPU_OK;
}

static int64_t spectrum_size(const sgpu_spectrum& s)
{
  return (int64_t)s.time_grid.n * s.nu_grid.n * s.mu_grid.n * s.phi_grid.n;
}

extern "C" int sgpu_wipe_radiation(sgpu_ctx* ctx, int32_t nz, int32_t nn, sgpu_tallies* ta, int32_t wipe_spectra)
{
  SGPU_REQUIRE(ctx && ta && nz >= 0 && nn >= 1, "bad argument");
  int st;
  double* zs[] = {ta->e_abs, ta->fx_rad, ta->fy_rad, ta->fz_rad, ta->fr_rad, ta->L_radio_dep,
                  ta->L_radio_emit, ta->e_rad};
  for (double* p : zs) if ((st = fill(ctx, p, nz))) return st;
  if ((st = fill(ctx, ta->J_nu, (int64_t)nz * nn))) return st;
  if (wipe_spectra) {
    if ((st = fill(ctx, ta->optical.flux, spectrum_size(ta->optical)))) return st;
    if ((st = fill(ctx, ta->optical.click, spectrum_size(ta->optical)))) return st;
    if ((st = fill(ctx, ta->gamma.flux, spectrum_size(ta->gamma)))) return st;
    if ((st = fill(ctx, ta->gamma.click, spectrum_size(ta->gamma)))) return st;
  }
  return SGPU_OK;
}

extern "C" int sgpu_normalize_tallies(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                      const sgpu_transport_params* params, sgpu_tallies* ta)
{
  SGPU_REQUIRE(ctx && nu_grid && params && ta, "null argument");
  int st = check_grid(grid);
  if (st) return st;
  SGPU_REQUIRE(ta->e_abs && ta->J_nu && ta->fr_rad && ta->L_radio_dep && ta->L_radio_emit && ta->e_rad,
               "tally arrays missing");
  hipLaunchKernelGGL(normalize_kernel, dim3(grid->n_zones), dim3(256), 0, ctx->stream, *nu_grid, grid->vol,
                     params->dt, params->store_Jnu, *ta);
  SGPU_LAUNCH_CHECK();
  return SGPU_OK;
}

extern "C" int sgpu_alias_tables(sgpu_ctx* ctx, const double* emis_cdf, int32_t n_zones, int32_t n_nu,
                                 sgpu_alias_entry* table)
{
  SGPU_REQUIRE(ctx && emis_cdf && table, "null argument");
  SGPU_REQUIRE(n_zones >= 0 && n_nu >= 1 && n_nu <= ALIAS_LDS_MAX, "n_nu = %d outside [1, %d]", n_nu,
               ALIAS_LDS_MAX);
  static_assert(sizeof(sgpu_alias_entry) == sizeof(AliasEntry), "alias entry layout");
  if (n_zones == 0) return SGPU_OK;
  hipLaunchKernelGGL(alias_kernel, dim3(n_zones), dim3(64), (sizeof(double) + 2 * sizeof(uint16_t)) * n_nu,
                     ctx->stream, emis_cdf, n_nu, (AliasEntry*)table);
  SGPU_LAUNCH_CHECK();
  return SGPU_OK;
}

extern "C" int sgpu_homologous_step(sgpu_ctx* ctx, sgpu_grid* grid, double* rho, double* r_out,
                                    double* x_out, double* z_out, double* vol, double e)
{
  SGPU_REQUIRE(ctx && grid && rho && vol && e > 0, "bad argument");
  const int nz = grid->n_zones;
  int nr = 0, nx = 0, nzz = 0;
  if (grid->kind == SGPU_GRID_1D_SPHERE) {
    SGPU_REQUIRE(r_out, "1D grid needs r_out");
    nr = grid->r_out.n;
    grid->r_out.min *= e;
    if (grid->r_out.type == SGPU_LOC_LIN) grid->r_out.del *= e;
  } else {
    SGPU_REQUIRE(x_out && z_out, "2D grid needs x_out and z_out");
    nx = grid->x_out.n; nzz = grid->z_out.n;
    grid->x_out.min *= e;
    grid->z_out.min *= e;
    if (grid->x_out.type == SGPU_LOC_LIN) grid->x_out.del *= e;
    if (grid->z_out.type == SGPU_LOC_LIN) grid->z_out.del *= e;
  }
  int m = nz;
  if (nr > m) m = nr;
  if (nx > m) m = nx;
  if (nzz > m) m = nzz;
  hipLaunchKernelGGL(homologous_kernel, dim3(grid_blocks(m, 256)), dim3(256), 0, ctx->stream, grid->kind, nz,
                     rho, r_out, nr, x_out, nx, z_out, nzz, vol, 0.0, e);
  SGPU_LAUNCH_CHECK();
  if (grid->kind == SGPU_GRID_1D_SPHERE) {
    hipLaunchKernelGGL(shell_volume_kernel, dim3(grid_blocks(nz, 256)), dim3(256), 0, ctx->stream, nz, r_out,
                       grid->r_out.min, vol);
    SGPU_LAUNCH_CHECK();
  }
  return SGPU_OK;
}
