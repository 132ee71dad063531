/*
 * oracle_adjoint.c -- TEST INFRASTRUCTURE: FP64 CPU reference of the
 * differentiable transport estimator and of its gradient.
 *
 * Continuous absorption: a packet's weight is w = exp(-sum alpha dshift ds)
 * with alpha = rho kappa(zone, comoving bin); scattering (table scat) and the
 * zone / frequency-bin steps of propagate_monte_carlo (transport.cpp:288-359)
 * decide the path, so the path does not depend on kappa and
 *   dF_b/dkappa_zk = -sum_{packets escaping in b} e w sum_{segments in (z,k)} rho dshift ds.
 * No reference counterpart exists (the reference absorbs discretely,
 * scatter.cpp:26-41); the pins are central finite differences with common
 * random numbers and the GPU kernel's own replay.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static double urand_p(uint64_t seed, uint64_t id, uint32_t* draw)
{
  return or_uniform(seed, id, (*draw)++);
}

typedef struct {
  double x[3], D[3], e, nu;
  int ind;
  uint64_t id;
  uint32_t draw;
} apk_t;

static double dshift_of(const sgpu_grid* g, const apk_t* p)
{
  double tn = g->t_now;
  double v[3] = {p->x[0] / tn, p->x[1] / tn, p->x[2] / tn};
  double beta2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / OR_C / OR_C;
  double vdd = v[0] * p->D[0] + v[1] * p->D[1] + v[2] * p->D[2];
  double gamma = 1.0 / sqrt(1 - beta2);
  return gamma * (1 - vdd / OR_C);
}

static void coherent_scatter(const sgpu_grid* g, apk_t* p, uint64_t seed)
{
  double tn = g->t_now;
  double V[3] = {p->x[0] / tn, p->x[1] / tn, p->x[2] / tn};
  double vdotD = V[0] * p->D[0] + V[1] * p->D[1] + V[2] * p->D[2];
  double beta2 = (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) / OR_C / OR_C;
  double gamma = 1.0 / sqrt(1 - beta2);
  double din = gamma * (1 - vdotD / OR_C);
  p->e *= din;
  p->nu *= din;
  double mu = 1 - 2.0 * urand_p(seed, p->id, &p->draw);
  double phi = 2.0 * OR_PI * urand_p(seed, p->id, &p->draw);
  double smu = sqrt(1 - mu * mu);
  double Dn[3] = {smu * cos(phi), smu * sin(phi), mu};
  for (int i = 0; i < 3; i++) V[i] = -V[i];
  double vdp = Dn[0] * V[0] + Dn[1] * V[1] + Dn[2] * V[2];
  double vd_out = gamma * (1 - vdp / OR_C);
  for (int k = 0; k < 3; k++)
    p->D[k] = 1.0 / vd_out * (Dn[k] - gamma * V[k] / OR_C * (1 - gamma * vdp / OR_C / (gamma + 1)));
  double norm = sqrt(p->D[0] * p->D[0] + p->D[1] * p->D[1] + p->D[2] * p->D[2]);
  for (int k = 0; k < 3; k++) p->D[k] /= norm;
  p->e *= vd_out;
  p->nu *= vd_out;
}

/* one packet; mode 0 forward (returns bin, sets *ew = e w), mode 1 backward */
static int replay(const sgpu_grid* g, const sgpu_locate* nu, const sgpu_locate* spec,
                  const float* kappa, const double* rho, const double* scat,
                  const sgpu_transport_params* prm, const sgpu_packets* pk, int64_t i,
                  int mode, double coef, double* grad, double* ew)
{
  apk_t p;
  p.x[0] = pk->x[i]; p.x[1] = pk->y[i]; p.x[2] = pk->z[i];
  p.D[0] = pk->Dx[i]; p.D[1] = pk->Dy[i]; p.D[2] = pk->Dz[i];
  p.e = pk->e[i]; p.nu = pk->nu[i]; p.id = pk->id[i]; p.draw = pk->draw[i];
  int nn = nu->n;
  p.ind = or_get_zone(g, p.x);
  int fate = p.ind == -2 ? SGPU_ESCAPED : (p.ind == -1 ? SGPU_ABSORBED : SGPU_MOVING);
  double tau = 0;
  int n_ev = 0;
  while (fate == SGPU_MOVING) {
    if (prm->max_events > 0 && n_ev >= prm->max_events) break;
    n_ev++;
    double d_bn = 0;
    int new_ind = or_get_next_zone(g, p.x, p.D, p.ind, g->r_core, &d_bn);
    double ds = dshift_of(g, &p);
    int b = or_locate_within_bounds(nu, p.nu * ds);
    size_t k = (size_t)p.ind * nn + b;
    double d_nu = or_delta(nu, b) / p.nu * OR_C / (1.0 / g->t_now);
    if (d_nu < 0) d_nu = -d_nu;
    if (d_nu < d_bn) { d_bn = d_nu; new_ind = p.ind; }
    double tot = scat[k] * ds;
    double tau_r = -1.0 * log(1 - urand_p(prm->seed, p.id, &p.draw));
    double d_sc = tot == 0 ? INFINITY : tau_r / tot;
    int ev;
    double d;
    if (d_sc < d_bn) { ev = 0; d = d_sc; } else { ev = 1; d = d_bn; }
    double seg = rho[p.ind] * ds * d;
    if (mode == 1) grad[k] += coef * seg;
    else tau += (double)kappa[k] * seg;
    for (int q = 0; q < 3; q++) p.x[q] += d * p.D[q];
    if (ev == 1) {
      if (new_ind == -1) fate = SGPU_ABSORBED;
      else if (new_ind == -2) fate = SGPU_ESCAPED;
      else p.ind = new_ind;
    } else {
      coherent_scatter(g, &p, prm->seed);
    }
  }
  if (mode == 1) return 0;
  *ew = 0;
  if (fate != SGPU_ESCAPED) return -1;
  int l = or_locate(spec, p.nu);
  if (l < 0 || l >= spec->n) l = 0;
  *ew = p.e * exp(-tau);
  return l;
}

int or_adjoint_forward(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                       const sgpu_locate* spec_grid, const float* kappa_abs,
                       const double* rho, const double* scat,
                       const sgpu_transport_params* p, const sgpu_packets* pk,
                       int64_t n, double* F, int32_t* rec_bin, double* rec_w,
                       double* dF_dkappa_dir, const double* dir)
{
  (void)dF_dkappa_dir; (void)dir;
  for (int64_t i = 0; i < n; i++) {
    double ew;
    int b = replay(grid, nu_grid, spec_grid, kappa_abs, rho, scat, p, pk, i, 0, 0, 0, &ew);
    if (b >= 0) F[b] += ew;
    if (rec_bin) rec_bin[i] = b;
    if (rec_w) rec_w[i] = ew;
  }
  return 0;
}

int or_adjoint_backward(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                        const sgpu_locate* spec_grid, const float* kappa_abs,
                        const double* rho, const double* scat,
                        const sgpu_transport_params* p, const sgpu_packets* pk,
                        int64_t n, const double* dL_dF, const int32_t* rec_bin,
                        const double* rec_w, double* grad)
{
  for (int64_t i = 0; i < n; i++) {
    if (rec_bin[i] < 0) continue;
    double coef = -dL_dF[rec_bin[i]] * rec_w[i];
    replay(grid, nu_grid, spec_grid, kappa_abs, rho, scat, p, pk, i, 1, coef, grad, 0);
  }
  return 0;
}
