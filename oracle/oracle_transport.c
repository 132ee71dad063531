/*
 * oracle_transport.c -- TEST INFRASTRUCTURE: FP64 CPU restatement of packet
 * emission, propagation, interactions, escape spectrum, tally normalisation,
 * compaction and homologous expansion.  Every packet draws from its own
 * Philox stream, so the result does not depend on the thread count.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

typedef struct {
  double x[3], D[3], t, e, nu;
  int ind, type;
  uint64_t id;
  uint32_t draw;
  double gamma, dshift, dvds;
} pkt_t;

typedef struct {
  const sgpu_grid* g;
  const sgpu_locate* nu;
  const sgpu_tables* tab;
  const sgpu_transport_params* p;
  sgpu_tallies* ta;
  uint64_t n_events, n_binshift, n_boundary, n_scatter, n_redist;
} tctx_t;

static inline double urand(const tctx_t* c, pkt_t* p)
{
  return or_uniform(c->p->seed, p->id, p->draw++);
}

static inline void atomic_add(double* a, double v)
{
#pragma omp atomic
  *a += v;
}

/* grid_1D_sphere::get_zone (grid_1D_sphere.cpp:494-505),
 * grid_2D_cyln::get_zone (grid_2D_cyln.cpp:329-354) */
int or_get_zone(const sgpu_grid* g, const double* x)
{
  if (g->kind == SGPU_GRID_1D_SPHERE) {
    double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (r < g->r_out.min) return -1;
    if (r >= g->r_out.x[g->r_out.n - 1]) return -2;
    return or_locate_within_bounds(&g->r_out, r);
  }
  double p = sqrt(x[0] * x[0] + x[1] * x[1]);
  int nx = g->x_out.n, nz = g->z_out.n;
  if (p < g->x_out.min) return -2;
  if (p > g->x_out.x[nx - 1]) return -2;
  if (x[2] < g->z_out.min) return -2;
  if (x[2] > g->z_out.x[nz - 1]) return -2;
  int i = or_locate_within_bounds(&g->x_out, p);
  int k = or_locate_within_bounds(&g->z_out, x[2]);
  return i * nz + k;
}

/* grid_1D_sphere::get_next_zone, grid_1D_sphere.cpp:510-590 */
static int next_zone_1d(const sgpu_grid* g, const double* x, const double* D, int i,
                        double r_core, double* l)
{
  double rsq = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  double xdotD = (D[0] * x[0] + D[1] * x[1] + D[2] * x[2]);
  double r_o = g->r_out.x[i];
  double l_out = -1 * xdotD + sqrt(xdotD * xdotD + r_o * r_o - rsq);
  double r_in;
  int ind_in;
  if (i != 0) { r_in = g->r_out.x[i - 1]; ind_in = i - 1; }
  else { r_in = g->r_out.min; ind_in = -1; }
  if (r_core >= r_in) { r_in = r_core; ind_in = -1; }
  double l_in;
  if ((i == 0) && (r_in == 0)) l_in = -1;
  else {
    double rad = xdotD * xdotD + r_in * r_in - rsq;
    if (rad < 0) l_in = -1;
    else l_in = -1 * xdotD - sqrt(rad);
  }
  int ind;
  double tiny_offset = 1 + 1e-6;
  if ((l_out < l_in) || (l_in < 0)) {
    ind = i + 1;
    if (ind == g->n_zones) { ind = -2; *l = l_out / tiny_offset; }
    else *l = l_out * tiny_offset;
  } else {
    ind = ind_in;
    if (ind_in == -1) *l = l_in / tiny_offset;
    else *l = l_in * tiny_offset;
  }
  return ind;
}

/* grid_2D_cyln::get_next_zone, grid_2D_cyln.cpp:360-465 */
static int next_zone_2d(const sgpu_grid* g, const double* x, const double* D, int i, double* l)
{
  int nz = g->z_out.n, nx = g->x_out.n;
  double p = sqrt(x[0] * x[0] + x[1] * x[1]);
  double z = x[2];
  double Dz = D[2];
  int ix = i / nz, iz = i % nz;
  double lp, lz;
  int d_ip = 0, d_iz = 0;
  double tiny = 1e-10;
  double dz = or_delta(&g->z_out, iz), dx = or_delta(&g->x_out, ix);
  if (Dz > 0) {
    double zt = g->z_out.x[iz] + dz * tiny;
    lz = (zt - z) / Dz;
    d_iz = 1;
  } else {
    double zt = or_left(&g->z_out, iz) - dz * tiny;
    lz = (zt - z) / Dz;
    if (Dz == 0) lz = INFINITY;
    d_iz = -1;
  }
  double pt, c, lp_out, lp_in, det;
  double a = D[0] * D[0] + D[1] * D[1];
  double b = 2 * (x[0] * D[0] + x[1] * D[1]);
  if (a == 0) lp = INFINITY;
  else {
    pt = g->x_out.x[ix] + dx * tiny;
    c = p * p - pt * pt;
    det = b * b - 4 * a * c;
    if (det < 0) lp_out = INFINITY;
    else lp_out = (-1.0 * b + sqrt(det)) / (2 * a);
    if (lp_out < 0) lp_out = INFINITY;
    pt = or_left(&g->x_out, ix) - dx * tiny;
    c = p * p - pt * pt;
    det = b * b - 4 * a * c;
    if (det < 0) lp_in = INFINITY;
    else lp_in = (-1.0 * b - sqrt(det)) / (2 * a);
    if (lp_in < 0) lp_in = INFINITY;
    if (ix == 0) lp_in = INFINITY;
    if (lp_in < lp_out) { lp = lp_in; d_ip = -1; }
    else { lp = lp_out; d_ip = 1; }
  }
  int new_iz = iz, new_ip = ix;
  if (lz < lp) { *l = lz; new_iz += d_iz; }
  else { *l = lp; new_ip += d_ip; }
  if ((new_iz < 0) || (new_iz >= nz) || (new_ip >= nx)) return -2;
  return new_ip * nz + new_iz;
}

int or_get_next_zone(const sgpu_grid* g, const double* x, const double* D, int i,
                     double r_core, double* l)
{
  if (g->kind == SGPU_GRID_1D_SPHERE) return next_zone_1d(g, x, D, i, r_core, l);
  return next_zone_2d(g, x, D, i, l);
}

/* homologous get_velocity (grid_1D_sphere.cpp:707-716, grid_2D_cyln.cpp:546-554)
 * and transport::do_dshift (relativity.cpp:14-43) */
static double do_dshift(const tctx_t* c, pkt_t* p, int tolab)
{
  double tn = c->g->t_now;
  double v[3] = {p->x[0] / tn, p->x[1] / tn, p->x[2] / tn};
  double dvds = 1.0 / tn;
  if (tolab) { v[0] *= -1; v[1] *= -1; v[2] *= -1; }
  double beta2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / OR_C / OR_C;
  double vdd = v[0] * p->D[0] + v[1] * p->D[1] + v[2] * p->D[2];
  double gamma = 1.0 / sqrt(1 - beta2);
  double dshift = gamma * (1 - vdd / OR_C);
  p->gamma = gamma; p->dshift = dshift; p->dvds = dvds;
  return dshift;
}

/* transport::lorentz_transform, relativity.cpp:84-127 (D divided by |D|^2) */
static void lorentz_transform(const tctx_t* c, pkt_t* p, int tolab)
{
  double tn = c->g->t_now;
  double v[3] = {p->x[0] / tn, p->x[1] / tn, p->x[2] / tn};
  double dvds = 1.0 / tn;
  if (tolab) { v[0] *= -1; v[1] *= -1; v[2] *= -1; }
  double beta2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / OR_C / OR_C;
  double vdd = v[0] * p->D[0] + v[1] * p->D[1] + v[2] * p->D[2];
  double gamma = 1.0 / sqrt(1 - beta2);
  double dshift = gamma * (1 - vdd / OR_C);
  p->gamma = gamma; p->dshift = dshift; p->dvds = dvds;
  p->e *= dshift;
  p->nu *= dshift;
  for (int k = 0; k < 3; k++)
    p->D[k] = 1.0 / dshift * (p->D[k] - gamma * v[k] / OR_C * (1 - gamma * vdd / OR_C / (gamma + 1)));
  double norm = p->D[0] * p->D[0] + p->D[1] * p->D[1] + p->D[2] * p->D[2];
  p->D[0] /= norm; p->D[1] /= norm; p->D[2] /= norm;
}

/* transport::get_opacity, transport_opacity.cpp:268-299 */
static int get_opacity(const tctx_t* c, const pkt_t* p, double dshift, double* opac, double* eps)
{
  double nu = p->nu * dshift;
  int i_nu = 0;
  int nn = c->nu->n;
  if (p->type == SGPU_PHOTON) {
    i_nu = or_locate_within_bounds(c->nu, nu);
    size_t k = (size_t)p->ind * nn + i_nu;
    double a = c->tab->abs[k];
    double s = 0;
    if (!c->p->omit_scattering) s = c->tab->scat[k];
    *opac = a + s;
    *eps = (*opac == 0) ? 0 : a / *opac;
  } else {
    double co = c->tab->compton_opac[p->ind] * or_klein_nishina(p->nu);
    double po = c->tab->photoion_opac[p->ind] * pow(p->nu, -3.5);
    *opac = co + po;
    *eps = po / (co + po);
  }
  return i_nu;
}

/* transport::sample_photon_frequency, emission.cpp:27-46 */
static void sample_photon_frequency(tctx_t* c, pkt_t* p)
{
  if (p->type == SGPU_PHOTON) {
    int nn = c->nu->n;
    int inu = or_cdf_sample(c->tab->emis_cdf + (size_t)p->ind * nn, nn, urand(c, p));
    p->nu = or_sample(c->nu, inu, urand(c, p));
    c->n_redist++;
  } else {
    p->nu = 1;
  }
}

/* transport::isotropic_scatter, scatter.cpp:303-351 (D divided by |D|) */
static void isotropic_scatter(tctx_t* c, pkt_t* p, int redist)
{
  double tn = c->g->t_now;
  double V[3] = {p->x[0] / tn, p->x[1] / tn, p->x[2] / tn};
  double vdotD = V[0] * p->D[0] + V[1] * p->D[1] + V[2] * p->D[2];
  double beta2 = (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) / OR_C / OR_C;
  double gamma = 1.0 / sqrt(1 - beta2);
  double dshift_in = gamma * (1 - vdotD / OR_C);
  p->e = p->e * dshift_in;
  p->nu = p->nu * dshift_in;
  double mu = 1 - 2.0 * urand(c, p);
  double phi = 2.0 * OR_PI * urand(c, p);
  double smu = sqrt(1 - mu * mu);
  double Dn[3] = {smu * cos(phi), smu * sin(phi), mu};
  if (redist) sample_photon_frequency(c, p);
  for (int i = 0; i < 3; i++) V[i] = -1 * V[i];
  double vdp = (Dn[0] * V[0] + Dn[1] * V[1] + Dn[2] * V[2]);
  double vd_out = gamma * (1 - vdp / OR_C);
  for (int k = 0; k < 3; k++)
    p->D[k] = 1.0 / vd_out * (Dn[k] - gamma * V[k] / OR_C * (1 - gamma * vdp / OR_C / (gamma + 1)));
  double norm = sqrt(p->D[0] * p->D[0] + p->D[1] * p->D[1] + p->D[2] * p->D[2]);
  p->D[0] = p->D[0] / norm; p->D[1] = p->D[1] / norm; p->D[2] = p->D[2] / norm;
  p->e = p->e * vd_out;
  p->nu = p->nu * vd_out;
}

/* transport::compton_scatter, scatter.cpp:76-133 */
static void compton_scatter(tctx_t* c, pkt_t* p)
{
  lorentz_transform(c, p, 0);
  double E_ratio, Dn[3];
  while (1) {
    double mu = 1 - 2.0 * urand(c, p);
    double phi = 2.0 * OR_PI * urand(c, p);
    double smu = sqrt(1 - mu * mu);
    Dn[0] = smu * cos(phi); Dn[1] = smu * sin(phi); Dn[2] = mu;
    double cost = p->D[0] * Dn[0] + p->D[1] * Dn[1] + p->D[2] * Dn[2];
    E_ratio = 1 / (1 + p->nu / OR_M_E_MEV * (1 - cost));
    double diff_cs = 0.5 * (E_ratio * E_ratio * (1 / E_ratio + E_ratio - 1 + cost * cost));
    if (urand(c, p) < diff_cs) break;
  }
  p->nu = p->nu * E_ratio;
  if (urand(c, p) > E_ratio) {
    atomic_add(&c->ta->L_radio_dep[p->ind], p->e);
    p->type = SGPU_PHOTON;
    double mu = 1 - 2.0 * urand(c, p);
    double phi = 2.0 * OR_PI * urand(c, p);
    double smu = sqrt(1 - mu * mu);
    Dn[0] = smu * cos(phi); Dn[1] = smu * sin(phi); Dn[2] = mu;
    sample_photon_frequency(c, p);
  }
  p->D[0] = Dn[0]; p->D[1] = Dn[1]; p->D[2] = Dn[2];
  lorentz_transform(c, p, 1);
}

/* transport::do_scatter, scatter.cpp:12-70 */
static int do_scatter(tctx_t* c, pkt_t* p, double eps)
{
  int fate = SGPU_MOVING;
  if (p->type == SGPU_PHOTON) {
    if (urand(c, p) > eps) isotropic_scatter(c, p, 0);
    else {
      double z2 = urand(c, p);
      if ((z2 > c->tab->eps_imc[p->ind]) || (c->p->radiative_eq)) isotropic_scatter(c, p, 1);
      else fate = SGPU_ABSORBED;
    }
  } else if (p->type == SGPU_GAMMARAY) {
    if (urand(c, p) > eps) compton_scatter(c, p);
    else {
      atomic_add(&c->ta->L_radio_dep[p->ind], p->e);
      p->type = SGPU_PHOTON;
      double mu = 1 - 2.0 * urand(c, p);
      double phi = 2.0 * OR_PI * urand(c, p);
      double smu = sqrt(1 - mu * mu);
      p->D[0] = smu * cos(phi); p->D[1] = smu * sin(phi); p->D[2] = mu;
      sample_photon_frequency(c, p);
      lorentz_transform(c, p, 1);
    }
  }
  return fate;
}


/* ---------------------------------------------------------------------------
 * Discrete diffusion Monte Carlo (transport_use_ddmc = 2), 1D sphere
 * ------------------------------------------------------------------------- */
#define DDMC_LAMBDA 0.7104     /* discrete_diffusion.cpp:118 */
#define DDMC_SML_PUSH 1.0e-8   /* :119 */
#define DDMC_SCATTER_CAP 100000

static void sample_in_zone(const sgpu_grid* g, int i, const double* ran, double* r);

/* grid_1D_sphere::get_zone_size (grid_1D_sphere.h:65-68): r_out.delta(i) */
static double zone_size(const sgpu_grid* g, int i)
{
  return i == 0 ? g->r_out.x[0] - g->r_out.min : g->r_out.x[i] - g->r_out.x[i - 1];
}

static double radius(const double* x) { return sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]); }

/* photon in a zone whose optical depth sigma dr exceeds the threshold
 * (transport.cpp:219-231, 275-286, 298-309) */
static int ddmc_zone(tctx_t* c, pkt_t* p, int zone, double dshift, double* sigma, double* dr)
{
  int old = p->ind;
  double eps;
  p->ind = zone;
  get_opacity(c, p, dshift, sigma, &eps);
  p->ind = old;
  *dr = zone_size(c->g, zone);
  return (*sigma * *dr > c->p->ddmc_tau_threshold) && (p->type == SGPU_PHOTON);
}

/* transport::sample_dir_from_blackbody_surface, discrete_diffusion.cpp:760-794 */
static void sample_dir_bb_surface(tctx_t* c, pkt_t* p)
{
  double mu = sqrt(urand(c, p));
  double phi = 2.0 * OR_PI * urand(c, p);
  double smu = sqrt(1.0 - mu * mu);
  double v1 = smu * cos(phi), v2 = smu * sin(phi), v3 = mu;
  double r1 = p->x[0] / radius(p->x), r2 = p->x[1] / radius(p->x), r3 = p->x[2] / radius(p->x);
  double costheta = r3, sintheta = sqrt(1 - costheta * costheta);
  double v1p = v1 * costheta + r1 * v3 * sintheta - r2 * (1.0 - costheta) * (-r2 * v1 + r1 * v2);
  double v2p = v2 * costheta + r2 * v3 * sintheta - r1 * (1.0 - costheta) * (-r2 * v1 + r1 * v2);
  double v3p = v3 * costheta - (r1 * v1 + r2 * v2) * sintheta;
  double norm = sqrt(v1p * v1p + v2p * v2p + v3p * v3p);
  p->D[0] = v1p / norm; p->D[1] = v2p / norm; p->D[2] = v3p / norm;
}

/* transport::discrete_diffuse_DDMC, discrete_diffusion.cpp:111-391: leakage
 * between zones with the spherical geometry factors, effective inelastic
 * scattering on a frequency grid, adiabatic loss and advection.  The
 * inelastic-scattering loop, unbounded in the reference, stops after
 * DDMC_SCATTER_CAP redistributions. */
static int ddmc_walk(tctx_t* c, pkt_t* p, double tstop)
{
  const sgpu_grid* g = c->g;
  const int nz = g->n_zones, nn = c->nu->n;
  const double tau_thr = c->p->ddmc_tau_threshold;
  int fate = SGPU_MOVING;
  double dt_remaining = tstop - p->t;
  while (fate == SGPU_MOVING) {
    int ii = p->ind, ip = ii + 1, im = ii - 1;
    if (ip == nz) ip = ii;
    if (im < 0) im = 0;
    double dx = zone_size(g, ii), dxp1 = zone_size(g, ip), dxm1 = zone_size(g, im);
    double sigma_i, sigma_ip1, sigma_im1, eps_i, eps_ip1, eps_im1;
    double dshift = do_dshift(c, p, 0);
    int i_nu = get_opacity(c, p, dshift, &sigma_i, &eps_i);
    p->ind = ip;
    get_opacity(c, p, dshift, &sigma_ip1, &eps_ip1);
    p->ind = im;
    get_opacity(c, p, dshift, &sigma_im1, &eps_im1);
    p->ind = ii;
    if (sigma_i * dx <= tau_thr) return SGPU_MOVING;
    double r_p = g->r_out.x[ii];
    double r_m = g->r_out.x[im];
    if (ii == 0) r_m = g->r_out.min;
    double r_0 = 0.5 * (r_p + r_m);
    double Xi = 1.0 + dx * dx / (12.0 * r_0 * r_0);
    double geo = r_m * r_m / r_0 / r_0 / Xi;
    int on_im1 = sigma_im1 * dxm1 > tau_thr;
    double leak_left = on_im1 ? (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + sigma_im1 * dxm1)) * geo
                              : (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + 2.0 * DDMC_LAMBDA)) * geo;
    geo = r_p * r_p / r_0 / r_0 / Xi;
    int on_ip1 = sigma_ip1 * dxp1 > tau_thr;
    double leak_right = on_ip1 ? (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + sigma_ip1 * dxp1)) * geo
                               : (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + 2.0 * DDMC_LAMBDA)) * geo;
    if (ip == ii) leak_right = (2.0 / 3.0 / dx) * (1.0 / (sigma_i * dx + 2.0 * DDMC_LAMBDA)) * geo;
    if (ii == 0) leak_left = 0.0;
    double leak_tot = leak_left + leak_right;
    double xi = urand(c, p);
    double d_stay = OR_C * dt_remaining;
    double d_leak = -log(xi) / (leak_left + leak_right);
    double d_sc = INFINITY;
    if (nn > 1) {
      double tau_r = -1.0 * log(urand(c, p));
      if (eps_i > 0.0) {
        double k_inel = sigma_i * eps_i;
        dshift = do_dshift(c, p, 0);
        i_nu = get_opacity(c, p, dshift, &sigma_i, &eps_i);
        double elastic = c->tab->emis_cdf[(size_t)p->ind * nn + i_nu];
        k_inel *= 1.0 - elastic;
        d_sc = tau_r / k_inel;
      }
    }
    int event;
    double this_d;
    if ((d_sc < d_leak) && (d_sc < d_stay)) { event = 0; this_d = d_sc; }
    else if (d_leak < d_stay) { event = 1; this_d = d_leak; }
    else { event = 2; this_d = d_stay; }
    double this_dt = this_d / OR_C;
    p->t += this_dt;
    dt_remaining -= this_dt;
    atomic_add(&c->ta->J_nu[(size_t)p->ind * (c->p->store_Jnu ? nn : 1)], p->e * this_d);
    atomic_add(&c->ta->e_abs[p->ind], (p->e * dshift) * this_d * sigma_i * eps_i);
    int leaked2imc = 0;
    if (event == 0) {
      int now_i_nu = i_nu;
      for (int k = 0; now_i_nu == i_nu && k < DDMC_SCATTER_CAP; k++) {
        fate = do_scatter(c, p, 1.0);
        dshift = do_dshift(c, p, 0);
        now_i_nu = get_opacity(c, p, dshift, &sigma_i, &eps_i);
      }
    } else if (event == 2) {
      dt_remaining = -1.0;
      fate = SGPU_STOPPED;
    } else {
      double P_left = leak_left / leak_tot;
      double xi2 = urand(c, p);
      if (xi2 <= P_left) {
        p->ind--;
        double rr = radius(p->x);
        double dr = rr - r_m;
        if (on_im1) dr += urand(c, p) * dxm1;
        else dr *= (1.0 + DDMC_SML_PUSH);
        p->x[0] -= p->x[0] / rr * dr;
        p->x[1] -= p->x[1] / rr * dr;
        p->x[2] -= p->x[2] / rr * dr;
        if (!on_im1) {
          leaked2imc = 1;
          lorentz_transform(c, p, 0);
          sample_dir_bb_surface(c, p);
          double mu = (p->x[0] * p->D[0] + p->x[1] * p->D[1] + p->x[2] * p->D[2]) / radius(p->x);
          if (mu > 0.0) { p->D[0] = -p->D[0]; p->D[1] = -p->D[1]; p->D[2] = -p->D[2]; }
          lorentz_transform(c, p, 1);
        }
      } else {
        p->ind++;
        double rr = radius(p->x);
        double dr = r_p - rr;
        if (on_ip1) dr += urand(c, p) * dxp1;
        else dr *= (1.0 + DDMC_SML_PUSH);
        p->x[0] += p->x[0] / rr * dr;
        p->x[1] += p->x[1] / rr * dr;
        p->x[2] += p->x[2] / rr * dr;
        if (!on_ip1) {
          leaked2imc = 1;
          lorentz_transform(c, p, 0);
          sample_dir_bb_surface(c, p);
          double mu = (p->x[0] * p->D[0] + p->x[1] * p->D[1] + p->x[2] * p->D[2]) / radius(p->x);
          if (mu < 0.0) { p->D[0] = -p->D[0]; p->D[1] = -p->D[1]; p->D[2] = -p->D[2]; }
          lorentz_transform(c, p, 1);
        }
      }
    }
    /* homologous zone velocity and dv/ds (grid_1D_sphere.cpp:707-716) */
    double tn = g->t_now;
    double zv[3] = {p->x[0] / tn, p->x[1] / tn, p->x[2] / tn};
    double dvds = 1.0 / tn;
    double vel = sqrt(zv[0] * zv[0] + zv[1] * zv[1] + zv[2] * zv[2]);
    double v_over_r = vel / radius(p->x);
    double dlnrhodt = dvds + 2.0 * v_over_r;
    double shift = (1.0 - dlnrhodt * this_dt / 3.0);
    lorentz_transform(c, p, 0);
    p->e *= shift;
    p->nu *= shift;
    lorentz_transform(c, p, 1);
    if (!leaked2imc) {
      p->x[0] += zv[0] * this_dt;
      p->x[1] += zv[1] * this_dt;
      p->x[2] += zv[2] * this_dt;
    }
    p->ind = or_get_zone(g, p->x);
    if (p->ind == -1) fate = SGPU_ABSORBED;
    if (p->ind == -2) fate = SGPU_ESCAPED;
  }
  return fate;
}

/* transport::move_across_DDMC_interface, discrete_diffusion.cpp:398-509:
 * converts the photon into the diffusing neighbour with the asymptotic
 * diffusion-limit probability, else re-emits it from the interface back into
 * its own zone.  Returns 1 when converted (the packet is then left in the
 * comoving frame, as the reference leaves it). */
static int ddmc_interface(tctx_t* c, pkt_t* p, int new_ind, double sigma_i, double dr)
{
  const sgpu_grid* g = c->g;
  int nz = g->n_zones;
  int ip = p->ind + 1, im = p->ind - 1;
  if (ip == nz) ip = p->ind;
  if (im < 0) im = 0;
  double r_p = g->r_out.x[p->ind];
  double r_m = g->r_out.x[im];
  if (p->ind == 0) r_m = g->r_out.min;
  double r_interface = new_ind == ip ? r_p : r_m;
  lorentz_transform(c, p, 0);
  double rr = radius(p->x);
  double mu = fabs((p->x[0] * p->D[0] + p->x[1] * p->D[1] + p->x[2] * p->D[2]) / rr);
  double p_convert = 4.0 * (1.0 + 1.5 * mu) / (3.0 * sigma_i * dr + 6.0 * 0.7104);
  double xi = urand(c, p);
  double ran[3];
  ran[0] = urand(c, p); ran[1] = urand(c, p); ran[2] = urand(c, p);
  double new_r[3];
  if (xi <= p_convert) {
    p->ind = new_ind;
    sample_in_zone(g, p->ind, ran, new_r);
    p->x[0] = new_r[0]; p->x[1] = new_r[1]; p->x[2] = new_r[2];
    return 1;
  }
  sample_in_zone(g, p->ind, ran, new_r);
  p->x[0] = new_r[0]; p->x[1] = new_r[1]; p->x[2] = new_r[2];
  /* the reference re-evaluates |x| after each component moves */
  if (new_ind == ip) {
    double d = (r_interface - radius(p->x)) * (1.0 - DDMC_SML_PUSH);
    p->x[0] += p->x[0] / radius(p->x) * d;
    p->x[1] += p->x[1] / radius(p->x) * d;
    p->x[2] += p->x[2] / radius(p->x) * d;
  } else if (new_ind == im) {
    double d = (radius(p->x) - r_interface) * (1.0 - DDMC_SML_PUSH);
    p->x[0] -= p->x[0] / radius(p->x) * d;
    p->x[1] -= p->x[1] / radius(p->x) * d;
    p->x[2] -= p->x[2] / radius(p->x) * d;
  }
  sample_dir_bb_surface(c, p);
  double mu2 = (p->x[0] * p->D[0] + p->x[1] * p->D[1] + p->x[2] * p->D[2]) / radius(p->x);
  if (new_ind == ip && mu2 > 0.0) { p->D[0] = -p->D[0]; p->D[1] = -p->D[1]; p->D[2] = -p->D[2]; }
  else if (new_ind == im && mu2 < 0.0) { p->D[0] = -p->D[0]; p->D[1] = -p->D[1]; p->D[2] = -p->D[2]; }
  lorentz_transform(c, p, 1);
  return 0;
}

/* transport::propagate_monte_carlo, transport.cpp:260-484, with the DDMC
 * checks of :272-309 and :441-450 when use_ddmc = 2 */
static int propagate_mc(tctx_t* c, pkt_t* p, double tstop, int max_events, int* n_ev)
{
  int fate = SGPU_MOVING;
  int nn = c->nu->n;
  const int ddmc = c->p->use_ddmc == 2;
  while (fate == SGPU_MOVING) {
    if (max_events > 0 && *n_ev >= max_events) break;
    double sig_n = 0, dr_n = 0;
    if (ddmc && ddmc_zone(c, p, p->ind, do_dshift(c, p, 0), &sig_n, &dr_n)) return SGPU_MOVING;
    (*n_ev)++;
    c->n_events++;
    double d_bn = 0;
    int new_ind = or_get_next_zone(c->g, p->x, p->D, p->ind, c->g->r_core, &d_bn);
    double dshift = do_dshift(c, p, 0);
    int new_cell_ddmc = 0;
    if (ddmc && new_ind >= 0) new_cell_ddmc = ddmc_zone(c, p, new_ind, dshift, &sig_n, &dr_n);
    double kappa, eps;
    int i_nu = get_opacity(c, p, dshift, &kappa, &eps);
    double d_nu = or_delta(c->nu, i_nu) / p->nu * OR_C / p->dvds;
    if (p->dvds == 0) d_nu = INFINITY;
    if (d_nu < 0) d_nu = -1 * d_nu;
    int binshift = 0;
    if (d_nu < d_bn) { d_bn = d_nu; new_ind = p->ind; binshift = 1; }
    double tot_lab = kappa * dshift;
    double tau_r = -1.0 * log(1 - urand(c, p));
    double d_sc = tau_r / tot_lab;
    if (tot_lab == 0) d_sc = INFINITY;
    double d_tm = (tstop - p->t) * OR_C;
    if (c->p->steady_state) d_tm = INFINITY;
    int event;
    double this_d;
    if ((d_sc < d_bn) && (d_sc < d_tm)) { event = 0; this_d = d_sc; }
    else if (d_bn < d_tm) { event = 1; this_d = d_bn; }
    else { event = 2; this_d = d_tm; }
    double this_E = p->e * this_d;
    int z = p->ind;
    if (p->type == SGPU_PHOTON) {
      atomic_add(&c->ta->e_abs[z], this_E * dshift * (kappa) * eps * dshift * c->tab->eps_imc[z]);
      if (c->p->store_Jnu) atomic_add(&c->ta->J_nu[(size_t)z * nn + i_nu], this_E);
      else atomic_add(&c->ta->J_nu[(size_t)z * (c->p->store_Jnu ? nn : 1)], this_E);
    }
    atomic_add(&c->ta->fx_rad[z], this_E * dshift * kappa * p->D[0] * dshift);
    atomic_add(&c->ta->fy_rad[z], this_E * dshift * kappa * p->D[1] * dshift);
    atomic_add(&c->ta->fz_rad[z], this_E * dshift * kappa * p->D[2] * dshift);
    double rr = sqrt(p->x[0] * p->x[0] + p->x[1] * p->x[1] + p->x[2] * p->x[2]);
    double xdotD = p->x[0] * p->D[0] + p->x[1] * p->D[1] + p->x[2] * p->D[2];
    atomic_add(&c->ta->fr_rad[z], this_E * dshift * kappa * xdotD / rr * dshift);
    p->x[0] += this_d * p->D[0];
    p->x[1] += this_d * p->D[1];
    p->x[2] += this_d * p->D[2];
    p->t = p->t + this_d / OR_C;
    if (event == 1) {
      c->n_boundary++;
      if (binshift) c->n_binshift++;
      if (new_ind == -1) {
        if (c->p->boundary_in_reflect) { p->D[0] *= -1; p->D[1] *= -1; p->D[2] *= -1; }
        else { p->ind = new_ind; fate = SGPU_ABSORBED; }
      } else if (new_ind == -2) {
        if (c->p->boundary_out_reflect) { p->D[0] *= -1; p->D[1] *= -1; p->D[2] *= -1; }
        else { p->ind = new_ind; fate = SGPU_ESCAPED; }
      } else if (new_cell_ddmc && new_ind != p->ind) {
        if (ddmc_interface(c, p, new_ind, sig_n, dr_n)) return SGPU_MOVING;
      } else {
        p->ind = new_ind;
      }
    } else if (event == 0) {
      c->n_scatter++;
      fate = do_scatter(c, p, eps);
    } else {
      fate = SGPU_STOPPED;
    }
  }
  return fate;
}

/* spectrum_array::count, spectrum_array.cpp:215-246 */
static void spectrum_count(sgpu_spectrum* s, double t, double w, double E, const double* D)
{
  double mu = D[2];
  double phi = atan2(D[1], D[0]);
  if (phi < 0) phi += 2. * OR_PI;
  int t_bin = or_locate(&s->time_grid, t);
  int l_bin = or_locate(&s->nu_grid, w);
  int m_bin = or_locate(&s->mu_grid, mu);
  int p_bin = or_locate(&s->phi_grid, phi);
  if (l_bin < 0) l_bin = 0;
  if (l_bin >= s->nu_grid.n) l_bin = 0;
  if ((t_bin < 0) || (l_bin < 0) || (m_bin < 0) || (p_bin < 0)) return;
  if (t_bin >= s->time_grid.n) return;
  if (m_bin >= s->mu_grid.n) return;
  if (p_bin >= s->phi_grid.n) return;
  int a3 = s->phi_grid.n, a2 = s->mu_grid.n * a3, a1 = s->nu_grid.n * a2;
  size_t ind = (size_t)t_bin * a1 + (size_t)l_bin * a2 + (size_t)m_bin * a3 + p_bin;
  atomic_add(&s->flux[ind], E);
  atomic_add(&s->click[ind], 1);
}

static void load_pkt(const sgpu_packets* pk, int64_t i, pkt_t* p)
{
  p->x[0] = pk->x[i]; p->x[1] = pk->y[i]; p->x[2] = pk->z[i];
  p->D[0] = pk->Dx[i]; p->D[1] = pk->Dy[i]; p->D[2] = pk->Dz[i];
  p->t = pk->t[i]; p->e = pk->e[i]; p->nu = pk->nu[i];
  p->ind = pk->ind[i]; p->type = pk->type[i];
  p->id = pk->id[i]; p->draw = pk->draw[i];
}

static void store_pkt(sgpu_packets* pk, int64_t i, const pkt_t* p, int fate)
{
  pk->x[i] = p->x[0]; pk->y[i] = p->x[1]; pk->z[i] = p->x[2];
  pk->Dx[i] = p->D[0]; pk->Dy[i] = p->D[1]; pk->Dz[i] = p->D[2];
  pk->t[i] = p->t; pk->e[i] = p->e; pk->nu[i] = p->nu;
  pk->ind[i] = p->ind; pk->type[i] = p->type; pk->fate[i] = fate;
  pk->id[i] = p->id; pk->draw[i] = p->draw;
}

/* packet loop of transport::step (transport.cpp:57-88) with
 * transport::propagate (:202-254) */
int or_propagate(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                 const sgpu_tables* tab, const sgpu_transport_params* prm,
                 sgpu_packets* pk, int64_t n, sgpu_tallies* ta,
                 sgpu_counters* cnt, int n_threads)
{
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
  uint64_t tot[10] = {0};
#pragma omp parallel
  {
    tctx_t c = {grid, nu_grid, tab, prm, ta, 0, 0, 0, 0, 0};
    uint64_t esc = 0, absb = 0, stp = 0, mov = 0;
#pragma omp for schedule(guided)
    for (int64_t i = 0; i < n; i++) {
      pkt_t p;
      load_pkt(pk, i, &p);
      int fate;
      int n_ev = 0;
      p.ind = or_get_zone(grid, p.x);
      if (p.ind == -1) fate = SGPU_ABSORBED;
      else if (p.ind == -2) fate = SGPU_ESCAPED;
      else {
        double tstop = prm->t_now + prm->dt;
        fate = SGPU_MOVING;
        while (fate == SGPU_MOVING) {
          double sig, dr;
          /* transport::propagate, transport.cpp:216-248 */
          if (prm->use_ddmc == 2 && ddmc_zone(&c, &p, p.ind, do_dshift(&c, &p, 0), &sig, &dr))
            fate = ddmc_walk(&c, &p, tstop);
          else
            fate = propagate_mc(&c, &p, tstop, prm->max_events, &n_ev);
          if (prm->max_events > 0 && n_ev >= prm->max_events) break;
        }
      }
      if (fate == SGPU_ESCAPED) {
        double t_obs = p.t - (p.x[0] * p.D[0] + p.x[1] * p.D[1] + p.x[2] * p.D[2]) / OR_C;
        if (p.type == SGPU_PHOTON && ta->optical.flux) spectrum_count(&ta->optical, t_obs, p.nu, p.e, p.D);
        if (p.type == SGPU_GAMMARAY && ta->gamma.flux) spectrum_count(&ta->gamma, t_obs, p.nu, p.e, p.D);
        p.t = t_obs;
        esc++;
      } else if (fate == SGPU_ABSORBED) absb++;
      else if (fate == SGPU_STOPPED) stp++;
      else mov++;
      store_pkt(pk, i, &p, fate);
    }
#pragma omp critical
    {
      tot[0] += c.n_events; tot[1] += c.n_binshift; tot[2] += c.n_boundary;
      tot[3] += c.n_scatter; tot[4] += c.n_redist; tot[5] += esc; tot[6] += absb;
      tot[7] += stp; tot[8] += mov;
    }
  }
  if (cnt) {
    cnt->n_packets += (uint64_t)n;
    cnt->n_events += tot[0]; cnt->n_binshift += tot[1]; cnt->n_boundary += tot[2];
    cnt->n_scatter += tot[3]; cnt->n_redist += tot[4]; cnt->n_escaped += tot[5];
    cnt->n_absorbed += tot[6]; cnt->n_stopped += tot[7]; cnt->n_moving += tot[8];
  }
  return 0;
}

/* sample_in_zone, grid_1D_sphere.cpp:681-700 and grid_2D_cyln.cpp:528-541 */
static void sample_in_zone(const sgpu_grid* g, int i, const double* ran, double* r)
{
  if (g->kind == SGPU_GRID_1D_SPHERE) {
    double r_0 = (i == 0) ? g->r_out.min : g->r_out.x[i - 1];
    double ro = g->r_out.x[i];
    double r_samp = pow(r_0 * r_0 * r_0 + ran[0] * (ro * ro * ro - r_0 * r_0 * r_0), 1.0 / 3.0);
    double mu = 1 - 2.0 * ran[1];
    double phi = 2.0 * OR_PI * ran[2];
    double sin_theta = sqrt(1 - mu * mu);
    r[0] = r_samp * sin_theta * cos(phi);
    r[1] = r_samp * sin_theta * sin(phi);
    r[2] = r_samp * mu;
  } else {
    int nz = g->z_out.n;
    int ix = i / nz, iz = i % nz;
    double phi = 2 * OR_PI * ran[1];
    double p_in = or_left(&g->x_out, ix);
    double p_out = g->x_out.x[ix];
    double p_samp = sqrt(p_in * p_in + ran[0] * (p_out * p_out - p_in * p_in));
    r[0] = p_samp * cos(phi);
    r[1] = p_samp * sin(phi);
    r[2] = or_left(&g->z_out, iz) + ran[2] * or_delta(&g->z_out, iz);
  }
}

/* transport::emit_radioactive (emission.cpp:165-243) with
 * create_isotropic_particle (:54-105).  Packet q draws, in order: zone,
 * time, type, 3 position, 2 direction, [2 frequency for photons]. */
int or_emit_radioactive(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                        const sgpu_zone_state* zs, const sgpu_tables* tab,
                        const sgpu_transport_params* prm, int64_t n_emit,
                        int force_rproc, uint64_t id_base,
                        sgpu_packets* pk, int64_t offset, sgpu_tallies* ta,
                        int64_t* n_out)
{
  int nz = grid->n_zones;
  double* cdf = (double*)malloc(sizeof(double) * nz);
  double* gfr = (double*)malloc(sizeof(double) * nz);
  double L_tot = 0;
  *n_out = 0;
  for (int i = 0; i < nz; i++) {
    double vol = grid->vol[i];
    double gfrac;
    double L = or_decay(zs->n_elems, zs->elem_Z, zs->elem_A, zs->X + (size_t)i * zs->n_elems,
                        prm->t_now, &gfrac, force_rproc);
    L = zs->rho[i] * L * vol;
    ta->L_radio_emit[i] = L;
    gfr[i] = gfrac;
    L_tot += L;
    cdf[i] = i == 0 ? L : cdf[i - 1] + L;
  }
  or_cdf_normalize(cdf, nz);
  if (L_tot == 0 || n_emit <= 0) { free(cdf); free(gfr); return 0; }
  double E_p = L_tot * prm->dt / (1.0 * n_emit);
  tctx_t c = {grid, nu_grid, tab, prm, ta, 0, 0, 0, 0, 0};
  for (int64_t q = 0; q < n_emit; q++) {
    pkt_t p;
    memset(&p, 0, sizeof(p));
    p.id = id_base + (uint64_t)q;
    p.draw = 0;
    int i = or_cdf_sample(cdf, nz, urand(&c, &p));
    double t = prm->t_now + prm->dt * urand(&c, &p);
    int type;
    if (urand(&c, &p) < gfr[i]) type = SGPU_GAMMARAY;
    else { ta->L_radio_dep[i] += E_p; type = SGPU_PHOTON; }
    p.ind = i;
    p.type = type;
    double ran[3];
    ran[0] = urand(&c, &p); ran[1] = urand(&c, &p); ran[2] = urand(&c, &p);
    sample_in_zone(grid, i, ran, p.x);
    double mu = 1 - 2.0 * urand(&c, &p);
    double phi = 2.0 * OR_PI * urand(&c, &p);
    double smu = sqrt(1 - mu * mu);
    p.D[0] = smu * cos(phi); p.D[1] = smu * sin(phi); p.D[2] = mu;
    sample_photon_frequency(&c, &p);
    p.e = E_p;
    lorentz_transform(&c, &p, 1);
    p.t = t;
    store_pkt(pk, offset + q, &p, SGPU_MOVING);
  }
  *n_out = n_emit;
  free(cdf); free(gfr);
  return 0;
}

/* transport::initialize_particles (emission.cpp:112-159) with
 * create_isotropic_particle (:54-105): the zone CDF over e_rad V, each
 * zone's emissivity overwritten with its normalised B_nu(T_gas) dnu at the
 * bin centres (written into emis_cdf, as the reference overwrites
 * emissivity_), E_p = sum(e_rad V) / n, photons at t_now.  Packet q draws,
 * in order: zone, 3 position, 2 direction, 2 frequency.  Nothing is done
 * (and emis_cdf is kept) for n_init = 0 or when the packets do not fit, as
 * the reference returns early in both cases. */
int or_initialize_particles(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                            const sgpu_zone_state* zs, const double* e_rad, double* emis_cdf,
                            const sgpu_transport_params* prm, int64_t n_init, uint64_t id_base,
                            sgpu_packets* pk, int64_t offset, int64_t* n_out)
{
  *n_out = 0;
  if (n_init <= 0 || offset + n_init > pk->capacity) return 0;
  int nz = grid->n_zones, nn = nu_grid->n;
  double* cdf = (double*)malloc(sizeof(double) * nz);
  double E_sum = 0;
  for (int i = 0; i < nz; i++) {
    double E_zone = e_rad[i] * grid->vol[i];
    cdf[i] = i == 0 ? E_zone : cdf[i - 1] + E_zone;
    E_sum += E_zone;
    double T = zs->T_gas[i];
    double* em = emis_cdf + (size_t)i * nn;
    for (int j = 0; j < nn; j++) {
      double nu_m = or_center(nu_grid, j);
      double e = or_blackbody_nu(T, nu_m) * or_delta(nu_grid, j);
      em[j] = j == 0 ? e : em[j - 1] + e;
    }
    or_cdf_normalize(em, nn);
  }
  or_cdf_normalize(cdf, nz);
  double Ep = E_sum / (1.0 * n_init);
  sgpu_tables tab = {0};
  tab.emis_cdf = emis_cdf;
  tctx_t c = {grid, nu_grid, &tab, prm, 0, 0, 0, 0, 0, 0};
  for (int64_t q = 0; q < n_init; q++) {
    pkt_t p;
    memset(&p, 0, sizeof(p));
    p.id = id_base + (uint64_t)q;
    p.draw = 0;
    int i = or_cdf_sample(cdf, nz, urand(&c, &p));
    p.ind = i;
    p.type = SGPU_PHOTON;
    double ran[3];
    ran[0] = urand(&c, &p); ran[1] = urand(&c, &p); ran[2] = urand(&c, &p);
    sample_in_zone(grid, i, ran, p.x);
    double mu = 1 - 2.0 * urand(&c, &p);
    double phi = 2.0 * OR_PI * urand(&c, &p);
    double smu = sqrt(1 - mu * mu);
    p.D[0] = smu * cos(phi); p.D[1] = smu * sin(phi); p.D[2] = mu;
    sample_photon_frequency(&c, &p);
    p.e = Ep;
    lorentz_transform(&c, &p, 1);
    p.t = prm->t_now;
    store_pkt(pk, offset + q, &p, SGPU_MOVING);
  }
  *n_out = n_init;
  free(cdf);
  return 0;
}

/* transport::setup_core_emission (transport_init.cpp:465-548) and
 * transport::emit_inner_source (emission.cpp:370-480).  The core spectrum
 * is B_nu(T_core) dnu at the bin centres (1 dnu for T_core <= 0), L_core
 * defaults to its 4 pi r^2 pi sum; every emissivity weight is 1
 * (transport_init.cpp:321-332).  Packet q draws, in order: for r_core > 0
 * phi_core, cos theta_core, phi_loc, cos theta_loc = sqrt(u); for r_core = 0
 * mu, phi; then (without core_frequency) 2 frequency draws; then the time. */
int or_emit_inner_source(const sgpu_grid* grid, const sgpu_locate* nu_grid, const sgpu_core_source* core,
                         const sgpu_transport_params* prm, int64_t n_emit, uint64_t id_base,
                         sgpu_packets* pk, int64_t offset, int64_t* n_out)
{
  *n_out = 0;
  if (n_emit <= 0 || offset + n_emit > pk->capacity) return 0;
  int nn = nu_grid->n;
  double r_core = core->r_core, T_core = core->T_core, L_core = core->L_core;
  if (L_core != 0 && r_core != 0 && T_core == 0) T_core = pow(L_core / (4.0 * OR_PI * r_core * r_core * OR_SB), 0.25);
  double* cdf = (double*)malloc(sizeof(double) * nn);
  double L_sum = 0;
  for (int j = 0; j < nn; j++) {
    double nu = or_center(nu_grid, j), dnu = or_delta(nu_grid, j);
    double bb = T_core <= 0 ? 1 : or_blackbody_nu(T_core, nu);
    cdf[j] = (j == 0 ? 0 : cdf[j - 1]) + bb * dnu;
    L_sum += 4.0 * OR_PI * r_core * r_core * OR_PI * bb * dnu;
  }
  or_cdf_normalize(cdf, nn);
  if (L_core == 0) L_core = L_sum;
  double Ep = L_core * prm->dt / n_emit;
  tctx_t c = {grid, nu_grid, 0, prm, 0, 0, 0, 0, 0, 0};
  for (int64_t q = 0; q < n_emit; q++) {
    pkt_t p;
    memset(&p, 0, sizeof(p));
    p.id = id_base + (uint64_t)q;
    p.draw = 0;
    if (r_core == 0) {
      double mu = 1 - 2.0 * urand(&c, &p);
      double phi = 2.0 * OR_PI * urand(&c, &p);
      double smu = sqrt(1 - mu * mu);
      p.D[0] = smu * cos(phi); p.D[1] = smu * sin(phi); p.D[2] = mu;
    } else {
      double phi_core = 2 * OR_PI * urand(&c, &p);
      double cosp_core = cos(phi_core), sinp_core = sin(phi_core);
      double cost_core = 1 - 2.0 * urand(&c, &p);
      double sint_core = sqrt(1 - cost_core * cost_core);
      double a_phot = r_core + r_core * 1e-10;
      p.x[0] = a_phot * sint_core * cosp_core;
      p.x[1] = a_phot * sint_core * sinp_core;
      p.x[2] = a_phot * cost_core;
      double phi_loc = 2 * OR_PI * urand(&c, &p);
      double cost_loc = sqrt(urand(&c, &p));
      double sint_loc = sqrt(1 - cost_loc * cost_loc);
      double D_xl = sint_loc * cos(phi_loc), D_yl = sint_loc * sin(phi_loc), D_zl = cost_loc;
      p.D[0] = cost_core * cosp_core * D_xl - sinp_core * D_yl + sint_core * cosp_core * D_zl;
      p.D[1] = cost_core * sinp_core * D_xl + cosp_core * D_yl + sint_core * sinp_core * D_zl;
      p.D[2] = -sint_core * D_xl + cost_core * D_zl;
    }
    p.e = Ep;
    if (core->core_frequency > 0) p.nu = core->core_frequency;
    else {
      int inu = or_cdf_sample(cdf, nn, urand(&c, &p));
      p.nu = or_sample(nu_grid, inu, urand(&c, &p));
    }
    p.ind = or_get_zone(grid, p.x);
    lorentz_transform(&c, &p, 1);
    p.t = prm->t_now + urand(&c, &p) * prm->dt;
    p.type = SGPU_PHOTON;
    store_pkt(pk, offset + q, &p, SGPU_MOVING);
  }
  *n_out = n_emit;
  free(cdf);
  return 0;
}

/* transport::reduce_radiation normalisation, transport_mpi.cpp:415-441 */
int or_normalize_tallies(const sgpu_grid* grid, const sgpu_locate* nu,
                         const sgpu_transport_params* prm, sgpu_tallies* ta)
{
  int nn = nu->n;
  double dt = prm->dt;
  for (int i = 0; i < grid->n_zones; i++) {
    double vol = grid->vol[i];
    ta->e_abs[i] /= vol * dt;
    ta->L_radio_dep[i] /= vol * dt;
    ta->L_radio_emit[i] /= vol;
    ta->fr_rad[i] /= vol * OR_C * dt;
    if (nn == 1 || !prm->store_Jnu) {
      ta->e_rad[i] = ta->J_nu[(size_t)i * (prm->store_Jnu ? nn : 1)] / (vol * dt * OR_C);
    } else {
      double esum = 0;
      for (int j = 0; j < nn; j++) {
        double* J = &ta->J_nu[(size_t)i * nn + j];
        *J /= vol * dt * 4 * OR_PI * or_delta(nu, j);
        esum += *J * or_delta(nu, j) * 4 * OR_PI / OR_C;
      }
      ta->e_rad[i] = esum;
    }
  }
  return 0;
}

/* transport::clean_up_particle_vector, transport.cpp:164-194 (order-keeping) */
int64_t or_compact(sgpu_packets* pk, int64_t n, int64_t* n_escaped)
{
  int64_t k = 0, esc = 0;
  for (int64_t i = 0; i < n; i++) {
    int f = pk->fate[i];
    if (f == SGPU_ESCAPED) esc++;
    if (f == SGPU_ESCAPED || f == SGPU_ABSORBED) continue;
    if (k != i) {
      pk->x[k] = pk->x[i]; pk->y[k] = pk->y[i]; pk->z[k] = pk->z[i];
      pk->Dx[k] = pk->Dx[i]; pk->Dy[k] = pk->Dy[i]; pk->Dz[k] = pk->Dz[i];
      pk->t[k] = pk->t[i]; pk->e[k] = pk->e[i]; pk->nu[k] = pk->nu[i];
      pk->ind[k] = pk->ind[i]; pk->type[k] = pk->type[i]; pk->fate[k] = pk->fate[i];
      pk->draw[k] = pk->draw[i]; pk->id[k] = pk->id[i];
    }
    k++;
  }
  *n_escaped = esc;
  return k;
}

/* hydro_homologous::step (hydro_homologous.cpp:27-61) with
 * grid_1D_sphere::expand (grid_1D_sphere.cpp:413-426) and
 * grid_2D_cyln::expand (grid_2D_cyln.cpp:315-324); the caller advances t_now. */
void or_homologous_step(sgpu_grid* g, double* rho, double* r_out, double* x_out,
                        double* z_out, double* vol, double e)
{
  for (int i = 0; i < g->n_zones; i++) rho[i] = rho[i] / e / e / e;
  if (g->kind == SGPU_GRID_1D_SPHERE) {
    for (int i = 0; i < g->r_out.n; i++) r_out[i] *= e;
    if (g->r_out.type == SGPU_LOC_LIN) g->r_out.del *= e;
    g->r_out.min *= e;
    for (int i = 0; i < g->n_zones; i++) {
      double r0 = (i == 0) ? g->r_out.min : r_out[i - 1];
      vol[i] = 4.0 * OR_PI / 3.0 * (r_out[i] * r_out[i] * r_out[i] - r0 * r0 * r0);
    }
  } else {
    for (int i = 0; i < g->x_out.n; i++) x_out[i] *= e;
    for (int i = 0; i < g->z_out.n; i++) z_out[i] *= e;
    if (g->x_out.type == SGPU_LOC_LIN) g->x_out.del *= e;
    if (g->z_out.type == SGPU_LOC_LIN) g->z_out.del *= e;
    g->x_out.min *= e;
    g->z_out.min *= e;
    for (int i = 0; i < g->n_zones; i++) vol[i] = vol[i] * e * e * e;
  }
}

/* test hooks: the static per-event functions above, callable on one packet */
void or_lorentz(const sgpu_grid* grid, double* x, double* D, double* e, double* nu, int tolab)
{
  tctx_t c = {grid, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  pkt_t p;
  memset(&p, 0, sizeof(p));
  for (int k = 0; k < 3; k++) { p.x[k] = x[k]; p.D[k] = D[k]; }
  p.e = *e; p.nu = *nu;
  lorentz_transform(&c, &p, tolab);
  for (int k = 0; k < 3; k++) D[k] = p.D[k];
  *e = p.e; *nu = p.nu;
}

/* Klein-Nishina rejection sampling of compton_scatter (scatter.cpp:82-103)
 * in the comoving frame for a photon moving along +z with energy nu (MeV):
 * returns n samples of cos(theta) and E_ratio. */
void or_compton_sample(double nu, int64_t n, uint64_t seed, double* cost_out, double* eratio_out)
{
  sgpu_transport_params prm;
  memset(&prm, 0, sizeof(prm));
  prm.seed = seed;
  tctx_t c = {0, 0, 0, &prm, 0, 0, 0, 0, 0, 0};
  for (int64_t i = 0; i < n; i++) {
    pkt_t p;
    memset(&p, 0, sizeof(p));
    p.id = (uint64_t)i;
    p.D[2] = 1;
    p.nu = nu;
    double E_ratio, cost;
    while (1) {
      double mu = 1 - 2.0 * urand(&c, &p);
      double phi = 2.0 * OR_PI * urand(&c, &p);
      double smu = sqrt(1 - mu * mu);
      double Dn[3] = {smu * cos(phi), smu * sin(phi), mu};
      cost = p.D[0] * Dn[0] + p.D[1] * Dn[1] + p.D[2] * Dn[2];
      E_ratio = 1 / (1 + p.nu / OR_M_E_MEV * (1 - cost));
      double diff_cs = 0.5 * (E_ratio * E_ratio * (1 / E_ratio + E_ratio - 1 + cost * cost));
      if (urand(&c, &p) < diff_cs) break;
    }
    cost_out[i] = cost;
    eratio_out[i] = E_ratio;
  }
}
