// adjoint.hip -- K9: differentiable transport (continuous-absorption
// estimator) and its adjoint replay.
//
// The reference samples absorption discretely (scatter.cpp:26-41), which has
// no derivative with respect to the opacity.  Here absorption acts as a
// packet weight w = exp(-sum alpha dshift ds) with alpha(z, b) = rho_z
// kappa(z, b) read in the comoving bin b, while scattering (the table scat,
// sampled as in propagate_monte_carlo, transport.cpp:336-341) and the
// frequency-bin / zone-boundary steps decide the path.  The path is then
// independent of kappa, so
//   F_nu        = sum over packets escaping in bin nu of e w
//   dF_nu/dk_zb = - sum over those packets of e w sum_{segments in (z,b)} rho_z dshift ds
// exactly, and the backward kernel replays every path from its Philox
// counters (the only per-packet memory kept between the passes is the escape
// bin and final weight: 12 bytes; the path itself is recomputed -- the
// checkpoint is the packet's initial state plus its counter).
#include <hip/hip_runtime.h>
#include "sgpu_internal.h"
#include "sgpu_geometry.h"

namespace sgpu {

struct AdjArgs {
  GridView<double> g;
  Axis<double> nu;
  sgpu_locate spec;
  const float* kappa;     // [nz][nn] cm^2/g
  const double* rho;      // [nz]
  const double* scat;     // [nz][nn] cm^-1
  sgpu_transport_params p;
  sgpu_packets pk;        // initial packet states (read only)
  int64_t n;
  double* F;              // forward: [n_spec]
  int32_t* rec_bin;       // forward out / backward in
  double* rec_w;
  const double* dLdF;     // backward
  double* grad;           // backward: [nz][nn]
};

// isotropic comoving re-emission without redistribution (scatter.cpp:303-351)
__device__ __forceinline__ void coherent_scatter(const GridView<double>& g, Pkt<double>& p)
{
  const double cl = C_LIGHT;
  double V[3] = {p.x[0] / g.t_now, p.x[1] / g.t_now, p.x[2] / g.t_now};
  double vdotD = V[0] * p.D[0] + V[1] * p.D[1] + V[2] * p.D[2];
  double beta2 = (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) / cl / cl;
  double gamma = 1.0 / sqrt(1 - beta2);
  double dshift_in = gamma * (1 - vdotD / cl);
  p.e *= dshift_in;
  p.nu *= dshift_in;
  double Dn[3];
  isotropic_dir(p, Dn);
  for (int i = 0; i < 3; i++) V[i] = -V[i];
  double vdp = Dn[0] * V[0] + Dn[1] * V[1] + Dn[2] * V[2];
  double vd_out = gamma * (1 - vdp / cl);
  for (int k = 0; k < 3; k++)
    p.D[k] = 1.0 / vd_out * (Dn[k] - gamma * V[k] / cl * (1 - gamma * vdp / cl / (gamma + 1)));
  double norm = sqrt(p.D[0] * p.D[0] + p.D[1] * p.D[1] + p.D[2] * p.D[2]);
  p.D[0] /= norm; p.D[1] /= norm; p.D[2] /= norm;
  p.e *= vd_out;
  p.nu *= vd_out;
}

template <bool BACKWARD>
__global__ void __launch_bounds__(256) adjoint_kernel(AdjArgs A)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const sgpu_packets& k = A.pk;
  Pkt<double> p;
  p.x[0] = k.x[i]; p.x[1] = k.y[i]; p.x[2] = k.z[i];
  p.D[0] = k.Dx[i]; p.D[1] = k.Dy[i]; p.D[2] = k.Dz[i];
  p.t = k.t[i]; p.e = k.e[i]; p.nu = k.nu[i];
  p.type = SGPU_PHOTON;
  p.rng = make_rng(A.p.seed, k.id[i], k.draw[i]);
  p.dvds = 0;
  const int nn = A.nu.n;
  // backward: every segment of a packet that escaped into bin b contributes
  // -dL/dF_b (e_f w) rho dshift ds; the forward record rec_w holds e_f w
  double coef = 0;
  if (BACKWARD) {
    const int b = A.rec_bin[i];
    if (b < 0) return;
    coef = -A.dLdF[b] * A.rec_w[i];
  }
  p.ind = get_zone(A.g, p.x);
  double tau = 0;
  int fate = p.ind == -2 ? SGPU_ESCAPED : (p.ind == -1 ? SGPU_ABSORBED : SGPU_MOVING);
  int n_ev = 0;
  while (fate == SGPU_MOVING) {
    if (A.p.max_events > 0 && n_ev >= A.p.max_events) break;
    n_ev++;
    double d_bn = 0;
    int new_ind = get_next_zone(A.g, p.x, p.D, p.ind, &d_bn);
    double dshift = dshift_lab_to_comoving(A.g, p);
    double nuc = p.nu * dshift;
    int b = ax_within(A.nu, nuc);
    size_t kz = (size_t)p.ind * nn + b;
    double sig = A.scat[kz];
    double d_nu = ax_delta(A.nu, b) / p.nu * C_LIGHT / p.dvds;
    if (d_nu < 0) d_nu = -d_nu;
    if (d_nu < d_bn) { d_bn = d_nu; new_ind = p.ind; }
    double tot_lab = sig * dshift;
    double tau_r = -1.0 * log(1 - p.rng.uniform());
    double d_sc = tot_lab == 0 ? INFINITY : tau_r / tot_lab;
    int event;
    double d;
    if (d_sc < d_bn) { event = 0; d = d_sc; }
    else { event = 1; d = d_bn; }
    const double seg = A.rho[p.ind] * dshift * d;
    if (BACKWARD) atomic_add_f64(&A.grad[kz], coef * seg);
    else tau += (double)A.kappa[kz] * seg;
    p.x[0] += d * p.D[0]; p.x[1] += d * p.D[1]; p.x[2] += d * p.D[2];
    p.t += d / C_LIGHT;
    if (event == 1) {
      if (new_ind == -1) fate = SGPU_ABSORBED;
      else if (new_ind == -2) fate = SGPU_ESCAPED;
      else p.ind = new_ind;
    } else {
      coherent_scatter(A.g, p);
    }
  }
  if (!BACKWARD) {
    int bin = -1;
    double w = exp(-tau);
    if (fate == SGPU_ESCAPED) {
      int l = loc_locate(A.spec, p.nu);
      if (l < 0 || l >= A.spec.n) l = 0;
      bin = l;
      atomic_add_f64(&A.F[bin], p.e * w);
    }
    if (A.rec_bin) A.rec_bin[i] = bin;
    // the record keeps e_final * w: the backward coefficient of every segment
    if (A.rec_w) A.rec_w[i] = p.e * w;
  }
}

}  // namespace sgpu

using namespace sgpu;

namespace {

int prep(const sgpu_grid* grid, const sgpu_locate* nu, const sgpu_locate* spec, const float* kappa,
         const double* rho, const double* scat, const sgpu_transport_params* params,
         const sgpu_packets* pk, int64_t n, AdjArgs* A)
{
  SGPU_REQUIRE(grid && nu && spec && kappa && rho && scat && params && pk, "null argument");
  SGPU_REQUIRE(grid->kind == SGPU_GRID_1D_SPHERE && grid->r_out.x && grid->n_zones == grid->r_out.n,
               "the differentiable form runs on the 1D spherical grid");
  SGPU_REQUIRE(params->steady_state, "the differentiable form is steady state");
  SGPU_REQUIRE(pk->x && pk->y && pk->z && pk->Dx && pk->Dy && pk->Dz && pk->t && pk->e && pk->nu &&
               pk->draw && pk->id && n >= 0 && n <= pk->capacity, "bad packets");
  SGPU_REQUIRE(nu->n >= 1 && nu->x && spec->n >= 1 && spec->x, "bad frequency grids");
  A->g.kind = grid->kind; A->g.n_zones = grid->n_zones; A->g.t_now = grid->t_now;
  A->g.inv_t = 1.0 / grid->t_now; A->g.r_core = grid->r_core;
  A->g.r_out = {grid->r_out.n, grid->r_out.type, grid->r_out.min, grid->r_out.del, 0.0, grid->r_out.x};
  A->g.x_out = {0, 0, 0, 0, 0, nullptr};
  A->g.z_out = {0, 0, 0, 0, 0, nullptr};
  A->nu = {nu->n, nu->type, nu->min, nu->del, 0.0, nu->x};
  A->spec = *spec;
  A->kappa = kappa; A->rho = rho; A->scat = scat; A->p = *params; A->pk = *pk; A->n = n;
  return SGPU_OK;
}

}  // namespace

extern "C" int sgpu_adjoint_forward(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                    const sgpu_locate* spec_grid, const float* kappa_abs, const double* rho,
                                    const double* scat, const sgpu_transport_params* params,
                                    const sgpu_packets* packets, int64_t n, double* F, int32_t* rec_bin,
                                    double* rec_w)
{
  SGPU_REQUIRE(ctx && F, "null argument");
  AdjArgs A{};
  int st = prep(grid, nu_grid, spec_grid, kappa_abs, rho, scat, params, packets, n, &A);
  if (st) return st;
  A.F = F; A.rec_bin = rec_bin; A.rec_w = rec_w;
  if (n == 0) return SGPU_OK;
  timer_begin(ctx, "adjoint_fwd");
  hipLaunchKernelGGL(adjoint_kernel<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "adjoint_fwd");
  return SGPU_OK;
}

extern "C" int sgpu_adjoint_backward(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                     const sgpu_locate* spec_grid, const float* kappa_abs, const double* rho,
                                     const double* scat, const sgpu_transport_params* params,
                                     const sgpu_packets* packets, int64_t n, const double* dL_dF,
                                     const int32_t* rec_bin, const double* rec_w, double* dL_dkappa)
{
  SGPU_REQUIRE(ctx && dL_dF && rec_bin && rec_w && dL_dkappa, "null argument");
  AdjArgs A{};
  int st = prep(grid, nu_grid, spec_grid, kappa_abs, rho, scat, params, packets, n, &A);
  if (st) return st;
  A.dLdF = dL_dF; A.rec_bin = const_cast<int32_t*>(rec_bin); A.rec_w = const_cast<double*>(rec_w);
  A.grad = dL_dkappa;
  if (n == 0) return SGPU_OK;
  timer_begin(ctx, "adjoint_bwd");
  hipLaunchKernelGGL(adjoint_kernel<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, A);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "adjoint_bwd");
  return SGPU_OK;
}
