// sgpu_geometry.h -- packet state and per-event physics of the Monte Carlo
// transport, templated on the geometry type R (double: the FP64 drop-in;
// float: the FP32-geometry form).  Energies, times and every tally stay FP64.
#pragma once
#include "sgpu_physics.h"

namespace sgpu {

// binned axis in precision R (FP32 copies of the edges for the mixed form)
template <class R>
struct Axis {
  int n, type;
  R min, del;
  R inv_log1p;   // 1/log(1+del) (FP32 form only)
  const R* x;
};

template <class R>
__device__ __forceinline__ R ax_left(const Axis<R>& g, int i) { return i == 0 ? g.min : g.x[i - 1]; }
template <class R>
__device__ __forceinline__ R ax_delta(const Axis<R>& g, int i)
{
  return i == 0 ? g.x[0] - g.min : g.x[i] - g.x[i - 1];
}

// locate_array::locate for double (the reference's exact expression) ...
__device__ __forceinline__ int ax_locate(const Axis<double>& g, double xv)
{
  int n = g.n;
  if (n == 1) return 0;
  if (xv >= g.x[n - 1]) return n;
  if (xv < g.min) return 0;
  int ind;
  if (g.type == SGPU_LOC_LIN) ind = (int)floor((xv - g.min) / g.del);
  else if (g.type == SGPU_LOC_LOG) ind = (int)floor(log(xv / g.min) / log(1 + g.del));
  else if (g.type == SGPU_LOC_NONE) ind = 0;
  else ind = upper_bound_dev(g.x, n, xv);
  if (ind < 0) ind = 0;        // a NaN coordinate stays inside the table
  if (ind > n) ind = n;
  if (g.type != SGPU_LOC_NONE && ind != 0 && ind != n && (ax_left(g, ind) > xv || g.x[ind] <= xv))
    ind = upper_bound_dev(g.x, n, xv);
  return ind;
}

// ... and for float (precomputed 1/log(1+del), same verify-and-fall-back)
__device__ __forceinline__ int ax_locate(const Axis<float>& g, float xv)
{
  int n = g.n;
  if (n == 1) return 0;
  if (xv >= g.x[n - 1]) return n;
  if (xv < g.min) return 0;
  int ind;
  if (g.type == SGPU_LOC_LIN) ind = (int)floorf((xv - g.min) / g.del);
  else if (g.type == SGPU_LOC_LOG) ind = (int)floorf(__logf(xv * __builtin_amdgcn_rcpf(g.min)) * g.inv_log1p);
  else if (g.type == SGPU_LOC_NONE) ind = 0;
  else ind = upper_bound_dev(g.x, n, xv);
  if (ind < 0) ind = 0;
  if (ind > n) ind = n;
  if (g.type != SGPU_LOC_NONE && ind != 0 && ind != n && (ax_left(g, ind) > xv || g.x[ind] <= xv))
    ind = upper_bound_dev(g.x, n, xv);
  return ind;
}

template <class R>
__device__ __forceinline__ int ax_within(const Axis<R>& g, R xv)
{
  int i = ax_locate(g, xv);
  return i == g.n ? g.n - 1 : i;
}

template <class R>
__device__ __forceinline__ R ax_sample(const Axis<R>& g, int i, R r)
{
  if (g.n == 1) return 0;
  if (i == 0) return g.min + (g.x[0] - g.min) * r;
  return g.x[i - 1] + (g.x[i] - g.x[i - 1]) * r;
}

template <class R>
struct GridView {
  int kind, n_zones;
  R t_now, inv_t, r_core;
  Axis<R> r_out, x_out, z_out;
};

// The mixed form draws the alias column and coin of an event's possible
// frequency redistribution when the event starts and loads the table entry
// then, so that the load is in flight during the event's geometry; an event
// that does not redistribute discards them.  The FP64 form draws in the
// reference's order and carries nothing.
struct AliasPre { float q; int32_t a; int k; float coin; bool valid; };
struct NoPre {};
template <class R> struct PreOf { typedef NoPre type; };
template <> struct PreOf<float> { typedef AliasPre type; };

template <class R>
struct Pkt {
  R x[3], D[3], nu;
  double t, e;
  int ind, type;
  R dvds;
  typename RngOf<R>::type rng;
  typename PreOf<R>::type pre;
};

template <class R> __device__ __forceinline__ R sqrt_(R v);
template <> __device__ __forceinline__ double sqrt_<double>(double v) { return sqrt(v); }
template <> __device__ __forceinline__ float sqrt_<float>(float v) { return __builtin_amdgcn_sqrtf(v); }   // v_sqrt_f32, 1 ulp
template <class R> __device__ __forceinline__ R rlog(R v);
template <> __device__ __forceinline__ double rlog<double>(double v) { return log(v); }
template <> __device__ __forceinline__ float rlog<float>(float v) { return __logf(v); }
template <class R> __device__ __forceinline__ void rsincos(R v, R* s, R* c);
template <> __device__ __forceinline__ void rsincos<double>(double v, double* s, double* c)
{
  *s = sin(v); *c = cos(v);
}
template <> __device__ __forceinline__ void rsincos<float>(float v, float* s, float* c)
{
  __sincosf(v, s, c);   // hardware sine / cosine (v in [0, 2 pi): ~2^-21 absolute)
}

template <class R> __device__ __forceinline__ R rinf() { return (R)INFINITY; }

// Division, 1/sqrt and the time of flight.  The FP64 form keeps IEEE
// division and square root in the reference's operation order (it follows
// the CPU path packet for packet); the FP32 form uses the hardware
// reciprocal and reciprocal square root (1-2 ulp), far inside what its
// statistical parity resolves, and multiplies by 1/c instead of dividing.
template <class R> __device__ __forceinline__ R qdiv(R a, R b);
template <> __device__ __forceinline__ double qdiv<double>(double a, double b) { return a / b; }
template <> __device__ __forceinline__ float qdiv<float>(float a, float b)
{
  return a * __builtin_amdgcn_rcpf(b);   // v_rcp_f32 (1 ulp); HIP's __fdividef divides exactly
}
template <class R> __device__ __forceinline__ R inv_sqrt(R v);
template <> __device__ __forceinline__ double inv_sqrt<double>(double v) { return 1 / sqrt(v); }
template <> __device__ __forceinline__ float inv_sqrt<float>(float v) { return __frsqrt_rn(v); }
// ds / c in seconds, accumulated into the FP64 packet time
__device__ __forceinline__ double flight_time(double ds) { return ds / C_LIGHT; }
__device__ __forceinline__ double flight_time(float ds) { return (double)ds * (1.0 / C_LIGHT); }
// acc * num / den for an FP64 tally term
__device__ __forceinline__ double mul_ratio(double acc, double num, double den) { return acc * num / den; }
__device__ __forceinline__ double mul_ratio(double acc, float num, float den)
{
  return acc * (double)(num * __builtin_amdgcn_rcpf(den));
}

// grid_1D_sphere::get_zone / grid_2D_cyln::get_zone
template <class R>
__device__ __forceinline__ int get_zone(const GridView<R>& g, const R* x)
{
  if (g.kind == SGPU_GRID_1D_SPHERE) {
    R r = sqrt_(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (r < g.r_out.min) return -1;
    if (r >= g.r_out.x[g.r_out.n - 1]) return -2;
    return ax_within(g.r_out, r);
  }
  R p = sqrt_(x[0] * x[0] + x[1] * x[1]);
  int nx = g.x_out.n, nz = g.z_out.n;
  if (p < g.x_out.min) return -2;
  if (p > g.x_out.x[nx - 1]) return -2;
  if (x[2] < g.z_out.min) return -2;
  if (x[2] > g.z_out.x[nz - 1]) return -2;
  int i = ax_within(g.x_out, p);
  int k = ax_within(g.z_out, x[2]);
  return i * nz + k;
}

// grid_1D_sphere::get_next_zone, grid_1D_sphere.cpp:510-590
template <class R>
__device__ __forceinline__ int next_zone_1d(const GridView<R>& g, const R* x, const R* D, int i, R* l)
{
  R rsq = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  R xdotD = (D[0] * x[0] + D[1] * x[1] + D[2] * x[2]);
  R r_o = g.r_out.x[i];
  R l_out = -1 * xdotD + sqrt_(xdotD * xdotD + r_o * r_o - rsq);
  R r_in;
  int ind_in;
  if (i != 0) { r_in = g.r_out.x[i - 1]; ind_in = i - 1; }
  else { r_in = g.r_out.min; ind_in = -1; }
  if (g.r_core >= r_in) { r_in = g.r_core; ind_in = -1; }
  R l_in;
  if ((i == 0) && (r_in == 0)) l_in = -1;
  else {
    R rad = xdotD * xdotD + r_in * r_in - rsq;
    if (rad < 0) l_in = -1;
    else l_in = -1 * xdotD - sqrt_(rad);
  }
  int ind;
  const R tiny_offset = (R)(1 + 1e-6);
  if ((l_out < l_in) || (l_in < 0)) {
    ind = i + 1;
    if (ind == g.n_zones) { ind = -2; *l = qdiv(l_out, tiny_offset); }
    else *l = l_out * tiny_offset;
  } else {
    ind = ind_in;
    if (ind_in == -1) *l = qdiv(l_in, tiny_offset);
    else *l = l_in * tiny_offset;
  }
  return ind;
}

// grid_2D_cyln::get_next_zone, grid_2D_cyln.cpp:360-465
template <class R>
__device__ __forceinline__ int next_zone_2d(const GridView<R>& g, const R* x, const R* D, int i, R* l)
{
  const int nz = g.z_out.n, nx = g.x_out.n;
  R p = sqrt_(x[0] * x[0] + x[1] * x[1]);
  R z = x[2];
  R Dz = D[2];
  int ix = i / nz, iz = i - ix * nz;
  R lp, lz;
  int d_ip = 0, d_iz = 0;
  const R tiny = (R)1e-10;
  R dz = ax_delta(g.z_out, iz), dx = ax_delta(g.x_out, ix);
  if (Dz > 0) {
    R zt = g.z_out.x[iz] + dz * tiny;
    lz = (zt - z) / Dz;
    d_iz = 1;
  } else {
    R zt = ax_left(g.z_out, iz) - dz * tiny;
    lz = (zt - z) / Dz;
    if (Dz == 0) lz = rinf<R>();
    d_iz = -1;
  }
  R a = D[0] * D[0] + D[1] * D[1];
  R b = 2 * (x[0] * D[0] + x[1] * D[1]);
  if (a == 0) lp = rinf<R>();
  else {
    R pt = g.x_out.x[ix] + dx * tiny;
    R c = p * p - pt * pt;
    R det = b * b - 4 * a * c;
    R lp_out, lp_in;
    if (det < 0) lp_out = rinf<R>();
    else lp_out = (-1 * b + sqrt_(det)) / (2 * a);
    if (lp_out < 0) lp_out = rinf<R>();
    pt = ax_left(g.x_out, ix) - dx * tiny;
    c = p * p - pt * pt;
    det = b * b - 4 * a * c;
    if (det < 0) lp_in = rinf<R>();
    else lp_in = (-1 * b - sqrt_(det)) / (2 * a);
    if (lp_in < 0) lp_in = rinf<R>();
    if (ix == 0) lp_in = rinf<R>();
    if (lp_in < lp_out) { lp = lp_in; d_ip = -1; }
    else { lp = lp_out; d_ip = 1; }
  }
  int new_iz = iz, new_ip = ix;
  if (lz < lp) { *l = lz; new_iz += d_iz; }
  else { *l = lp; new_ip += d_ip; }
  if ((new_iz < 0) || (new_iz >= nz) || (new_ip >= nx)) return -2;
  return new_ip * nz + new_iz;
}

template <class R>
__device__ __forceinline__ int get_next_zone(const GridView<R>& g, const R* x, const R* D, int i, R* l)
{
  if (g.kind == SGPU_GRID_1D_SPHERE) return next_zone_1d(g, x, D, i, l);
  return next_zone_2d(g, x, D, i, l);
}

// Shortest boundary step.  An FP32 position resolves about 2^-24 |x|, and
// r_o^2 - r^2 of a packet on a shell cancels to zero (or a few ulps, or a
// negative value whose square root is NaN).  A packet skimming a shell then
// gets boundary steps that do not move it and crosses back and forth between
// the two zones without advancing, for up to millions of events.  The FP32
// form therefore steps at least 2^-20 |x|_max across a boundary: x.D grows by
// the step length with every step, so successive crossings make progress.
// The FP64 form keeps the reference's step unchanged.
__device__ __forceinline__ double floor_boundary_step(const Pkt<double>&, double l) { return l; }
__device__ __forceinline__ float floor_boundary_step(const Pkt<float>& p, float l)
{
  const float m = fmaxf(fabsf(p.x[0]), fmaxf(fabsf(p.x[1]), fabsf(p.x[2])));
  return fmaxf(l, m * 9.5367431640625e-7f);   // fmaxf also replaces a NaN step
}

// transport::do_dshift with homologous velocity (relativity.cpp:14-43)
template <class R>
__device__ __forceinline__ R dshift_lab_to_comoving(const GridView<R>& g, Pkt<R>& p)
{
  R v0 = qdiv(p.x[0], g.t_now), v1 = qdiv(p.x[1], g.t_now), v2 = qdiv(p.x[2], g.t_now);
  const R c = (R)C_LIGHT;
  R beta2 = qdiv(qdiv(v0 * v0 + v1 * v1 + v2 * v2, c), c);
  R vdd = v0 * p.D[0] + v1 * p.D[1] + v2 * p.D[2];
  R gamma = inv_sqrt(1 - beta2);
  p.dvds = qdiv((R)1, g.t_now);
  return gamma * (1 - qdiv(vdd, c));
}

// transport::lorentz_transform (relativity.cpp:84-127); D divided by |D|^2
template <class R>
__device__ __forceinline__ void lorentz_transform(const GridView<R>& g, Pkt<R>& p, int tolab)
{
  R v[3] = {qdiv(p.x[0], g.t_now), qdiv(p.x[1], g.t_now), qdiv(p.x[2], g.t_now)};
  p.dvds = qdiv((R)1, g.t_now);
  if (tolab) { v[0] *= -1; v[1] *= -1; v[2] *= -1; }
  const R c = (R)C_LIGHT;
  R beta2 = qdiv(qdiv(v[0] * v[0] + v[1] * v[1] + v[2] * v[2], c), c);
  R vdd = v[0] * p.D[0] + v[1] * p.D[1] + v[2] * p.D[2];
  R gamma = inv_sqrt(1 - beta2);
  R dshift = gamma * (1 - qdiv(vdd, c));
  p.e *= (double)dshift;
  p.nu *= dshift;
#pragma unroll
  for (int k = 0; k < 3; k++)
    p.D[k] = qdiv((R)1, dshift) *
             (p.D[k] - qdiv(gamma * v[k], c) * (1 - qdiv(qdiv(gamma * vdd, c), gamma + 1)));
  R norm = p.D[0] * p.D[0] + p.D[1] * p.D[1] + p.D[2] * p.D[2];
  p.D[0] = qdiv(p.D[0], norm); p.D[1] = qdiv(p.D[1], norm); p.D[2] = qdiv(p.D[2], norm);
}

template <class R>
__device__ __forceinline__ void isotropic_dir(Pkt<R>& p, R* Dn)
{
  R mu = 1 - 2 * p.rng.template u<R>();
  R phi = 2 * (R)PI * p.rng.template u<R>();
  R smu = sqrt_(1 - mu * mu);
  R s, c;
  rsincos(phi, &s, &c);
  Dn[0] = smu * c; Dn[1] = smu * s; Dn[2] = mu;
}

}  // namespace sgpu
