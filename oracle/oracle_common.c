/*
 * oracle_common.c -- TEST INFRASTRUCTURE: Philox RNG and locate_array /
 * cdf_array restatements used by the CPU oracle.
 */
#include <math.h>
#include <float.h>
#include "oracle.h"

/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
 * published multipliers and Weyl key increments.  The reference draws from
 * one GSL mt19937 stream per thread (src/transport/thread_RNG.cpp:65-73);
 * this build gives every packet its own counter-addressed stream instead. */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform double in [0,1): the top 53 bits of the first two output words of
 * block {draw, 0, id_lo, id_hi} under key {seed_lo, seed_hi}. */
double or_uniform(uint64_t seed, uint64_t id, uint32_t draw)
{
  uint32_t ctr[4] = {draw, 0u, (uint32_t)id, (uint32_t)(id >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  or_philox4x32(ctr, key, o);
  uint64_t u = ((uint64_t)o[1] << 32) | o[0];
  return (double)(u >> 11) * (1.0 / 9007199254740992.0);
}

float or_uniform_f32(uint64_t seed, uint64_t id, uint32_t draw)
{
  uint32_t ctr[4] = {draw, 0u, (uint32_t)id, (uint32_t)(id >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  or_philox4x32(ctr, key, o);
  return (float)(o[0] >> 8) * (1.0f / 16777216.0f);
}

/* locate_array::log_init, src/utils/locate_array.cpp:54-80: right edges by
 * repeated multiplication v <- v + v*del while v < stop, then stop itself.
 * Returns the bin count (or -1 when max_n is too small). */
int or_log_init(double start, double stop, double del, double* x, int max_n, double* min_out)
{
  int n = 0;
  *min_out = start;
  double v = start + start * del;
  while (v < stop) {
    if (n >= max_n) return -1;
    x[n++] = v;
    v = v + v * del;
  }
  if (n >= max_n) return -1;
  x[n++] = stop;
  return n;
}

/* locate_array::init(start, stop, del), src/utils/locate_array.cpp:25-46 */
int or_lin_init(double start, double stop, double del, double* x, int max_n, double* min_out)
{
  int n = (int)ceil((stop - start) / del);
  if (n < 1) n = 1;
  if (n > max_n) return -1;
  *min_out = start;
  for (int i = 0; i < n - 1; i++) x[i] = start + (i + 1) * del;
  x[n - 1] = stop;
  return n;
}

static int upper_bound(const double* x, int n, double v)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (x[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

double or_left(const sgpu_locate* g, int i) { return i == 0 ? g->min : g->x[i - 1]; }

/* locate_array::locate, src/utils/locate_array.cpp:191-218 */
int or_locate(const sgpu_locate* g, double xval)
{
  int n = g->n;
  if (n == 1) return 0;
  if (xval >= g->x[n - 1]) return n;
  if (xval < g->min) return 0;
  int ind;
  if (g->type == SGPU_LOC_LIN) ind = (int)floor((xval - g->min) / g->del);
  else if (g->type == SGPU_LOC_LOG) ind = (int)floor(log(xval / g->min) / log(1 + g->del));
  else if (g->type == SGPU_LOC_NONE) ind = 0;
  else ind = upper_bound(g->x, n, xval);
  if (g->type != SGPU_LOC_NONE && ind != 0 && ind != n &&
      (or_left(g, ind) > xval || g->x[ind] <= xval))
    ind = upper_bound(g->x, n, xval);
  return ind;
}

/* locate_array::locate_within_bounds, :225-230 */
int or_locate_within_bounds(const sgpu_locate* g, double xval)
{
  int ind = or_locate(g, xval);
  if (ind == g->n) return g->n - 1;
  return ind;
}

/* locate_array::center / delta, src/utils/locate_array.h:54-72 */
double or_center(const sgpu_locate* g, int i)
{
  return i == 0 ? 0.5 * (g->min + g->x[0]) : 0.5 * (g->x[i - 1] + g->x[i]);
}

double or_delta(const sgpu_locate* g, int i)
{
  return i == 0 ? g->x[0] - g->min : g->x[i] - g->x[i - 1];
}

/* locate_array::sample, src/utils/locate_array.cpp:237-242 */
double or_sample(const sgpu_locate* g, int i, double r)
{
  if (g->n == 1) return 0;
  if (i == 0) return g->min + (g->x[0] - g->min) * r;
  return g->x[i - 1] + (g->x[i] - g->x[i - 1]) * r;
}

/* cdf_array::sample, src/transport/cdf_array.h:84-89 */
int or_cdf_sample(const double* y, int n, double yval)
{
  if (n == 1) return 0;
  return upper_bound(y, n, yval);
}

/* cdf_array::normalize, src/transport/cdf_array.h:62-75 */
void or_cdf_normalize(double* y, int n)
{
  for (int i = 0; i < n; i++) if (isnan(y[i])) y[i] = 0;
  if (y[n - 1] == 0) for (int i = 0; i < n; i++) y[i] = 1.0;
  double N = y[n - 1];
  for (int i = 0; i < n; i++) y[i] /= N;
}
