// sgpu_device.h -- device-side building blocks of the transport hot path
// (gfx950, wave64): reference constants, Philox-4x32-10, locate_array
// lookups, homologous-grid geometry.  Geometry and lookups are templated on
// the arithmetic type R so that the FP64 drop-in and the FP32-geometry form
// share one source; tallies always accumulate in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/sedona_gpu.h"

namespace sgpu {

// physical constants: the values of src/utils/physical_constants.h:9-33,
// including its truncated pi (3.14159) and alpha_fs (0.007)
constexpr double PI = 3.14159;
constexpr double C_LIGHT = 2.99792458e10;
constexpr double H_PLANCK = 6.6260755e-27;
constexpr double K_BOLTZ = 1.380658e-16;
constexpr double K_EV = 8.6173324e-5;
constexpr double M_PROTON = 1.67262158e-24;
constexpr double M_ELECTRON = 9.10938188e-28;
constexpr double SB = 5.6704e-5;
constexpr double A_RAD = 7.5657e-15;
constexpr double M_E_MEV = 0.510998910511;
constexpr double THOMSON = 0.66523e-24;
constexpr double SIGMA_TOT = 0.0265400193567;
constexpr double ALPHA_FS = 0.007;
constexpr double MEV_TO_ERGS = 1.60217646e-6;
constexpr double EV_TO_ERGS = 1.60217646e-12;
constexpr double ERGS_TO_EV = 624150975230.2815;
// radioactive decay constants, src/transport/radioactive.h:11-35
constexpr double TAU_56NI = 757241.8;
constexpr double TAU_56CO = 9627379.;
constexpr double E_56NI = 1.725;
constexpr double E_56CO = 3.730;
constexpr double CO_POS = 0.0324;
constexpr double TAU_48CR = 111976.;
constexpr double TAU_48V = 1.99107843e6;
constexpr double E_48CR = 0.442;
constexpr double E_48V = 3.066;

constexpr int WAVE = 64;

// ------------------------------------------------------------------
// Philox-4x32-10 (Salmon et al., SC'11).  Counter {draw, 0, id_lo, id_hi},
// key {seed_lo, seed_hi}.  Identical bit-for-bit to the CPU oracle's.
// ------------------------------------------------------------------
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32->64 product per multiplier (v_mad_u64_u32: both halves in
    // one quarter-rate instruction, where mul_lo + mul_hi take two)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    uint32_t n0 = hi1 ^ c[1] ^ k0;
    uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

struct Rng {
  uint32_t k0, k1, id_lo, id_hi, draw;
  __device__ __forceinline__ double uniform()
  {
    uint32_t c[4] = {draw++, 0u, id_lo, id_hi};
    philox4x32(c, k0, k1);
    uint64_t u = ((uint64_t)c[1] << 32) | c[0];
    return (double)(u >> 11) * (1.0 / 9007199254740992.0);
  }
  __device__ __forceinline__ float uniform_f()
  {
    uint32_t c[4] = {draw++, 0u, id_lo, id_hi};
    philox4x32(c, k0, k1);
    return (float)(c[0] >> 8) * (1.0f / 16777216.0f);
  }
  template <class R> __device__ __forceinline__ R u();
};
template <> __device__ __forceinline__ double Rng::u<double>() { return uniform(); }
template <> __device__ __forceinline__ float Rng::u<float>() { return uniform_f(); }

__device__ __forceinline__ Rng make_rng(uint64_t seed, uint64_t id, uint32_t draw)
{
  Rng r;
  r.k0 = (uint32_t)seed; r.k1 = (uint32_t)(seed >> 32);
  r.id_lo = (uint32_t)id; r.id_hi = (uint32_t)(id >> 32);
  r.draw = draw;
  return r;
}

// FP32 draws of the mixed form: xoshiro128** (Blackman and Vigna,
// "Scrambled linear pseudorandom number generators", ACM TOMS 47 (2021)),
// a 128-bit state in registers and a dozen integer operations per draw,
// seeded for each packet and propagation call by one Philox block (counter
// {draw, 0, id}).  `draw` then advances by one, so a packet resumed from its
// stored counter seeds a fresh stream from a block no stream of it has used.
// Philox blocks of four draws, every event restarting its block to keep the
// refills of a wave's lanes together, ran the event loop about 3% slower
// (20 rounds of 32-bit products per redistribution event).
// The FP64 form keeps one Philox block per draw (Rng above), which the
// oracle replays draw for draw.
struct RngX {
  uint32_t s0, s1, s2, s3, draw;
  __device__ __forceinline__ float uniform_f()
  {
    const uint32_t m = s1 * 5u;
    const uint32_t r = ((m << 7) | (m >> 25)) * 9u;
    const uint32_t t = s1 << 9;
    s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3; s2 ^= t;
    s3 = (s3 << 11) | (s3 >> 21);
    return (float)(r >> 8) * (1.0f / 16777216.0f);
  }
  template <class R> __device__ __forceinline__ R u() { return uniform_f(); }
};

__device__ __forceinline__ RngX make_rngx(uint64_t seed, uint64_t id, uint32_t draw)
{
  uint32_t c[4] = {draw, 0u, (uint32_t)id, (uint32_t)(id >> 32)};
  philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  RngX r;
  r.s0 = c[0] | (uint32_t)((c[0] | c[1] | c[2] | c[3]) == 0u);   // the all-zero state is a fixed point
  r.s1 = c[1]; r.s2 = c[2]; r.s3 = c[3];
  r.draw = draw + 1;
  return r;
}

template <class R> struct RngOf { typedef Rng type; };
template <> struct RngOf<float> { typedef RngX type; };

// ------------------------------------------------------------------
// locate_array (src/utils/locate_array.cpp:191-242, locate_array.h:54-72)
// ------------------------------------------------------------------
template <class T>
__device__ __forceinline__ int upper_bound_dev(const T* x, int n, T v)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (x[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double loc_left(const sgpu_locate& g, int i)
{
  return i == 0 ? g.min : g.x[i - 1];
}

// locate(): closed-left bins, floor(log(x/min)/log(1+del)) for log grids
// with the reference's verify-and-fall-back-to-upper_bound step
__device__ __forceinline__ int loc_locate(const sgpu_locate& g, double xv)
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
  if (g.type != SGPU_LOC_NONE && ind != 0 && ind != n && (loc_left(g, ind) > xv || g.x[ind] <= xv))
    ind = upper_bound_dev(g.x, n, xv);
  return ind;
}

__device__ __forceinline__ int loc_within(const sgpu_locate& g, double xv)
{
  int ind = loc_locate(g, xv);
  return ind == g.n ? g.n - 1 : ind;
}

__device__ __forceinline__ double loc_center(const sgpu_locate& g, int i)
{
  return i == 0 ? 0.5 * (g.min + g.x[0]) : 0.5 * (g.x[i - 1] + g.x[i]);
}

__device__ __forceinline__ double loc_delta(const sgpu_locate& g, int i)
{
  return i == 0 ? g.x[0] - g.min : g.x[i] - g.x[i - 1];
}

__device__ __forceinline__ double loc_sample(const sgpu_locate& g, int i, double r)
{
  if (g.n == 1) return 0;
  if (i == 0) return g.min + (g.x[0] - g.min) * r;
  return g.x[i - 1] + (g.x[i] - g.x[i - 1]) * r;
}

// cdf_array::sample: first entry larger than yval (cdf_array.h:84-89)
template <class T>
__device__ __forceinline__ int cdf_sample(const T* y, int n, T yval)
{
  if (n == 1) return 0;
  return upper_bound_dev(y, n, yval);
}

// ------------------------------------------------------------------
// wave64 reductions (DPP/permute lowered by the compiler from shfl_xor)
// ------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ void atomic_add_f64(double* p, double v)
{
  unsafeAtomicAdd(p, v);   // hardware global_atomic_add_f64 on gfx950
}

}  // namespace sgpu
