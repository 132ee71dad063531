// nlte.hip -- K8: NLTE statistical-equilibrium solves.
//
// The reference (AtomicSpecies::solve_nlte, src/opacity/AtomicSpecies_state.cpp:131-196)
// builds, from rates R_ij (level i -> j, already multiplied by the LTE
// population of level i, :429-431), the matrix M_ii = -sum_j R_ij,
// M_ij = R_ji, replaces the last row by the LTE populations (number
// conservation for departure coefficients), and solves M b = e_{n-1} with
// GSL's partial-pivot LU (:174-175); populations are max(b lte, 1e-30).
//
// Forms:
//   sgpu_nlte_solve     one system per call, FP64 LU in device memory by one
//                       workgroup: the drop-in for one solve_nlte call.
//   sgpu_nlte_batched   every system in one persistent launch.  Workgroups
//                       pull systems (largest first) from a device queue.
//     FP64 mode         FP64 LU in a per-workgroup device-memory slot.
//     FP32_REFINE mode  row/column power-of-two equilibration, FP32 LU in LDS
//                       (n <= 202: 163 KB), trailing updates on MFMA
//                       (v_mfma_f32_32x32x2f32) for the bulk of the matrix,
//                       FP64 residuals r = e - M b straight from the FP64
//                       rates, FP32 correction solves, until
//                       max|db| <= tol max|b|; systems that do not converge
//                       (or are too large for LDS) are re-solved in FP64.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include "sgpu_internal.h"
#include "sgpu_device.h"

namespace sgpu {

constexpr int NLTE_THREADS = 256;
constexpr int NLTE_LDS_MAX_N = 202;          // 202*202*4 = 163,216 B of LDS
constexpr int NLTE_MAX_ITERS = 12;

// ------------------------------------------------------------------
// block-wide helpers (one workgroup per system)
// ------------------------------------------------------------------
struct ArgMax {
  double v;
  int i;
};

__device__ __forceinline__ ArgMax better(ArgMax a, ArgMax b)
{
  // larger magnitude wins; the smaller row index wins ties (first maximum)
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

__device__ ArgMax block_argmax(ArgMax m, ArgMax* red)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ArgMax t;
    t.v = __shfl_xor(m.v, o, WAVE);
    t.i = __shfl_xor(m.i, o, WAVE);
    m = better(m, t);
  }
  const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  __syncthreads();
  if (lane == 0) red[w] = m;
  __syncthreads();
  ArgMax r = red[0];
  for (int k = 1; k < (int)(blockDim.x / WAVE); k++) r = better(r, red[k]);
  return r;
}

__device__ double block_max(double v, double* red)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = red[0];
  for (int k = 1; k < (int)(blockDim.x / WAVE); k++) r = fmax(r, red[k]);
  return r;
}

// Assemble M (row-major, n x n) from rates R (row-major) and lte into dst.
template <class T>
__device__ void assemble(int n, const double* R, const double* lte, T* M, int ld)
{
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double Rout = 0.0;
    for (int j = 0; j < n; ++j) Rout += R[(size_t)i * n + j];
    if (i != n - 1) M[(size_t)i * ld + i] = (T)(-1 * Rout);
  }
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    int i = idx / n, j = idx - i * n;
    if (i == n - 1) M[(size_t)i * ld + j] = (T)lte[j];
    else if (i != j) M[(size_t)i * ld + j] = (T)R[(size_t)j * n + i];
  }
  __syncthreads();
}

// Right-looking partial-pivot LU of an n x n FP64 matrix in device memory,
// then forward / back substitution of rhs e_{n-1}.  x receives the solution.
__device__ void lu_solve_f64(int n, double* M, int* perm, double* x, ArgMax* red, int* singular)
{
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int i = tid; i < n; i += NT) perm[i] = i;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    ArgMax m = {-1.0, n};
    for (int i = j + tid; i < n; i += NT) {
      ArgMax t = {fabs(M[(size_t)i * n + j]), i};
      m = better(m, t);
    }
    ArgMax p = block_argmax(m, red);
    const int ip = p.i;
    if (ip != j) {
      for (int k = tid; k < n; k += NT) {
        double t = M[(size_t)j * n + k];
        M[(size_t)j * n + k] = M[(size_t)ip * n + k];
        M[(size_t)ip * n + k] = t;
      }
      if (tid == 0) { int t = perm[j]; perm[j] = perm[ip]; perm[ip] = t; }
    }
    __syncthreads();
    const double ajj = M[(size_t)j * n + j];
    if (ajj == 0) { if (tid == 0) *singular = 1; __syncthreads(); continue; }
    for (int i = j + 1 + tid; i < n; i += NT) M[(size_t)i * n + j] /= ajj;
    __syncthreads();
    const int m_ = n - j - 1;
    for (int idx = tid; idx < m_ * m_; idx += NT) {
      int i = j + 1 + idx / m_, k = j + 1 + idx % m_;
      M[(size_t)i * n + k] -= M[(size_t)i * n + j] * M[(size_t)j * n + k];
    }
    __syncthreads();
  }
  // x = P e_{n-1}
  for (int i = tid; i < n; i += NT) x[i] = perm[i] == n - 1 ? 1.0 : 0.0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {          // forward, unit lower
    const double xk = x[k];
    for (int i = k + 1 + tid; i < n; i += NT) x[i] -= M[(size_t)i * n + k] * xk;
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {     // backward
    if (tid == 0) x[k] = x[k] / M[(size_t)k * n + k];
    __syncthreads();
    const double xk = x[k];
    for (int i = tid; i < k; i += NT) x[i] -= M[(size_t)i * n + k] * xk;
    __syncthreads();
  }
}

__device__ void write_pops(int n, const double* x, const double* lte, double* b, double* pop)
{
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    b[i] = x[i];
    if (pop) {
      double v = x[i] * lte[i];
      pop[i] = v < 1e-30 ? 1e-30 : v;
    }
  }
}

// ------------------------------------------------------------------
// single system (drop-in)
// ------------------------------------------------------------------
__global__ void __launch_bounds__(NLTE_THREADS) nlte_single_kernel(int n, const double* R, const double* lte,
                                                                   double* b, double* pop, double* work)
{
  __shared__ ArgMax red[NLTE_THREADS / WAVE];
  __shared__ int perm[1024];
  __shared__ int singular;
  if (threadIdx.x == 0) singular = 0;
  double* M = work;
  double* x = work + (size_t)n * n;
  assemble<double>(n, R, lte, M, n);
  lu_solve_f64(n, M, perm, x, red, &singular);
  write_pops(n, x, lte, b, pop);
}

// ------------------------------------------------------------------
// batched: device queue of systems, largest first
// ------------------------------------------------------------------
// smallest and largest system size (single workgroup): mm[0] = min, mm[1] = max
__global__ void __launch_bounds__(1024) size_range_kernel(const int32_t* n, int nsys, int32_t* mm)
{
  __shared__ int lo[1024 / WAVE], hi[1024 / WAVE];
  int a = 1 << 30, b = -(1 << 30);
  for (int s = threadIdx.x; s < nsys; s += blockDim.x) { a = min(a, n[s]); b = max(b, n[s]); }
  for (int o = 32; o > 0; o >>= 1) { a = min(a, __shfl_xor(a, o, WAVE)); b = max(b, __shfl_xor(b, o, WAVE)); }
  if (threadIdx.x % WAVE == 0) { lo[threadIdx.x / WAVE] = a; hi[threadIdx.x / WAVE] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x / WAVE); w++) { a = min(a, lo[w]); b = max(b, hi[w]); }
    a = min(a, lo[0]); b = max(b, hi[0]);
    mm[0] = a; mm[1] = b;
  }
}

// counting sort of system indices by size, descending (single workgroup)
__global__ void __launch_bounds__(1024) order_kernel(const int32_t* n, int nsys, int32_t* order, int32_t* counter)
{
  __shared__ int hist[1025];
  for (int k = threadIdx.x; k < 1025; k += blockDim.x) hist[k] = 0;
  __syncthreads();
  for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
    int v = n[s];
    v = v < 0 ? 0 : (v > 1024 ? 1024 : v);
    atomicAdd(&hist[1024 - v], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int k = 0; k < 1025; k++) { int c = hist[k]; hist[k] = run; run += c; }
    *counter = 0;
  }
  __syncthreads();
  // stable within a size class: a single thread scatters in index order
  if (threadIdx.x == 0)
    for (int s = 0; s < nsys; s++) {
      int v = n[s];
      v = v < 0 ? 0 : (v > 1024 ? 1024 : v);
      order[hist[1024 - v]++] = s;
    }
}

struct BatchArgs {
  int nsys;
  const int32_t* n;
  const int64_t* rate_off;
  const int64_t* vec_off;
  const double* rates;
  const double* lte;
  double* b;
  double* pop;
  int32_t* iters;
  const int32_t* order;
  int32_t* counter;
  double* slots;       // per workgroup: max_n*max_n + 3*max_n doubles
  size_t slot_stride;
  int precision;
  double tol;
};

__device__ __forceinline__ int next_system(const BatchArgs& A, int* sh)
{
  if (threadIdx.x == 0) *sh = atomicAdd(A.counter, 1);
  __syncthreads();
  int k = *sh;
  __syncthreads();
  return k < A.nsys ? A.order[k] : -1;
}

__global__ void __launch_bounds__(NLTE_THREADS) nlte_batched_f64_kernel(BatchArgs A)
{
  __shared__ ArgMax red[NLTE_THREADS / WAVE];
  __shared__ int perm[1024];
  __shared__ int singular, sh;
  double* M = A.slots + (size_t)blockIdx.x * A.slot_stride;
  while (true) {
    const int s = next_system(A, &sh);
    if (s < 0) break;
    const int n = A.n[s];
    const double* R = A.rates + A.rate_off[s];
    const double* lte = A.lte + A.vec_off[s];
    double* x = M + (size_t)n * n;
    if (threadIdx.x == 0) singular = 0;
    assemble<double>(n, R, lte, M, n);
    lu_solve_f64(n, M, perm, x, red, &singular);
    write_pops(n, x, lte, A.b + A.vec_off[s], A.pop ? A.pop + A.vec_off[s] : nullptr);
    if (A.iters && threadIdx.x == 0) A.iters[s] = 0;
    __syncthreads();
  }
}

// ------------------------------------------------------------------
// FP32 LU in LDS + FP64 refinement
// ------------------------------------------------------------------
// Scaled system: As = Dr M Dc with Dr, Dc powers of two; M b = e becomes
// As y = Dr e, b = Dc y.  LDS holds As (row stride ld = n), the pivot
// permutation lives in a per-workgroup device slot.
struct MixedLds {
  float a[NLTE_LDS_MAX_N * NLTE_LDS_MAX_N];
};

// trailing update A[r0.., c0..] -= L[r0.., j0..j0+nb) * U[j0..j0+nb, c0..]
// on MFMA 32x32x2 f32 tiles; the remainder rows/columns use VALU.
__device__ void trailing_update_mfma(float* a, int ld, int j0, int nb, int r0, int n)
{
  const int m = n - r0;            // rows / cols of the trailing block
  const int tiles = m / 32;        // full 32x32 tiles per dimension
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const int nwaves = blockDim.x / WAVE;
  // MFMA tiles
  for (int t = wave; t < tiles * tiles; t += nwaves) {
    const int ti = t / tiles, tj = t % tiles;
    const int rb = r0 + ti * 32, cb = r0 + tj * 32;
    typedef float f32x16 __attribute__((ext_vector_type(16)));
    f32x16 acc;
    // accumulator layout of v_mfma_f32_32x32x2f32: lane l holds column
    // (l % 32), rows 8*(i/4) + 4*(l/32)... (i = 0..15): row = (i/4)*8 + (l/32)*4 + i%4
#pragma unroll
    for (int i = 0; i < 16; i++) {
      int row = (i / 4) * 8 + (lane / 32) * 4 + (i % 4);
      acc[i] = a[(size_t)(rb + row) * ld + cb + (lane % 32)];
    }
    for (int k = 0; k < nb; k += 2) {
      // A operand: lane l supplies L[rb + l%32][j0 + k + l/32]
      float av = -a[(size_t)(rb + lane % 32) * ld + j0 + k + lane / 32];
      // B operand: lane l supplies U[j0 + k + l/32][cb + l%32]
      float bv = a[(size_t)(j0 + k + lane / 32) * ld + cb + lane % 32];
      if (k + lane / 32 >= nb) { av = 0.f; bv = 0.f; }
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      int row = (i / 4) * 8 + (lane / 32) * 4 + (i % 4);
      a[(size_t)(rb + row) * ld + cb + (lane % 32)] = acc[i];
    }
  }
  // VALU remainder: rows >= r0 + 32*tiles or cols >= r0 + 32*tiles
  const int edge = r0 + tiles * 32;
  for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
    int i = r0 + idx / m, c = r0 + idx % m;
    if (i < edge && c < edge) continue;
    float s = a[(size_t)i * ld + c];
    for (int k = 0; k < nb; ++k) s -= a[(size_t)i * ld + j0 + k] * a[(size_t)(j0 + k) * ld + c];
    a[(size_t)i * ld + c] = s;
  }
}

// blocked right-looking partial-pivot LU in LDS (panel width nb)
__device__ int lu_f32_lds(float* a, int n, int* perm, int nb, ArgMax* red)
{
  const int tid = threadIdx.x, NT = blockDim.x;
  int singular = 0;
  for (int i = tid; i < n; i += NT) perm[i] = i;
  __syncthreads();
  for (int j0 = 0; j0 < n; j0 += nb) {
    const int jb = min(nb, n - j0);
    // panel factorisation (columns j0..j0+jb), updating the full rows
    for (int j = j0; j < j0 + jb; ++j) {
      ArgMax m = {-1.0, n};
      for (int i = j + tid; i < n; i += NT) {
        ArgMax t = {(double)fabsf(a[(size_t)i * n + j]), i};
        m = better(m, t);
      }
      ArgMax p = block_argmax(m, red);
      const int ip = p.i;
      if (ip != j) {
        for (int k = tid; k < n; k += NT) {
          float t = a[(size_t)j * n + k];
          a[(size_t)j * n + k] = a[(size_t)ip * n + k];
          a[(size_t)ip * n + k] = t;
        }
        if (tid == 0) { int t = perm[j]; perm[j] = perm[ip]; perm[ip] = t; }
      }
      __syncthreads();
      const float ajj = a[(size_t)j * n + j];
      if (ajj == 0.f) { singular = 1; continue; }
      const float inv = 1.f / ajj;
      for (int i = j + 1 + tid; i < n; i += NT) a[(size_t)i * n + j] *= inv;
      __syncthreads();
      // update the rest of the panel columns only
      const int mr = n - j - 1, mc = j0 + jb - j - 1;
      if (mc > 0)
        for (int idx = tid; idx < mr * mc; idx += NT) {
          int i = j + 1 + idx / mc, k = j + 1 + idx % mc;
          a[(size_t)i * n + k] -= a[(size_t)i * n + j] * a[(size_t)j * n + k];
        }
      __syncthreads();
    }
    const int r0 = j0 + jb;
    if (r0 >= n) break;
    // U12 = L11^-1 A12 (unit lower triangular solve, columns in parallel)
    for (int c = r0 + tid; c < n; c += NT)
      for (int i = j0 + 1; i < j0 + jb; ++i) {
        float s = a[(size_t)i * n + c];
        for (int k = j0; k < i; ++k) s -= a[(size_t)i * n + k] * a[(size_t)k * n + c];
        a[(size_t)i * n + c] = s;
      }
    __syncthreads();
    trailing_update_mfma(a, n, j0, jb, r0, n);
    __syncthreads();
  }
  return singular;
}

// y = U^-1 L^-1 (P rhs) with the FP32 factors, rhs/y FP64 in device memory
__device__ void lu_f32_apply(const float* a, int n, const int* perm, const double* rhs, double* y, double* tmp)
{
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int i = tid; i < n; i += NT) tmp[i] = rhs[perm[i]];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double v = tmp[k];
    for (int i = k + 1 + tid; i < n; i += NT) tmp[i] -= (double)a[(size_t)i * n + k] * v;
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    if (tid == 0) tmp[k] = tmp[k] / (double)a[(size_t)k * n + k];
    __syncthreads();
    const double v = tmp[k];
    for (int i = tid; i < k; i += NT) tmp[i] -= (double)a[(size_t)i * n + k] * v;
    __syncthreads();
  }
  for (int i = tid; i < n; i += NT) y[i] = tmp[i];
  __syncthreads();
}

__global__ void __launch_bounds__(NLTE_THREADS) nlte_batched_mixed_kernel(BatchArgs A, int nb)
{
  extern __shared__ float lds_a[];
  __shared__ ArgMax red[NLTE_THREADS / WAVE];
  __shared__ double dred[NLTE_THREADS / WAVE];
  __shared__ int sh, sing_f64;
  double* slot = A.slots + (size_t)blockIdx.x * A.slot_stride;
  const int tid = threadIdx.x, NT = blockDim.x;
  while (true) {
    const int s = next_system(A, &sh);
    if (s < 0) break;
    const int n = A.n[s];
    const double* R = A.rates + A.rate_off[s];
    const double* lte = A.lte + A.vec_off[s];
    double* bout = A.b + A.vec_off[s];
    double* pout = A.pop ? A.pop + A.vec_off[s] : nullptr;
    // slot layout: diag[n] | rs[n] | cs[n] | x[n] | r[n] | d[n] | tmp[n] | perm (ints) | M64 (fallback)
    double* diag = slot;
    double* rs = diag + n;
    double* cs = rs + n;
    double* x = cs + n;
    double* r = x + n;
    double* d = r + n;
    double* tmp = d + n;
    int* perm = (int*)(tmp + n);
    double* M64 = (double*)(perm + 2 * ((n + 1) / 2)) ;
    int it = 0;
    bool ok = false;
    if (n <= NLTE_LDS_MAX_N) {
      // diagonal and equilibration: rs_i = 2^-e(max_j |M_ij|)
      for (int i = tid; i < n; i += NT) {
        double Rout = 0.0;
        for (int j = 0; j < n; ++j) Rout += R[(size_t)i * n + j];
        diag[i] = -1 * Rout;
      }
      __syncthreads();
      for (int i = tid; i < n; i += NT) {
        double mx = 0;
        if (i == n - 1) for (int j = 0; j < n; ++j) mx = fmax(mx, fabs(lte[j]));
        else {
          for (int j = 0; j < n; ++j) mx = fmax(mx, j == i ? fabs(diag[i]) : fabs(R[(size_t)j * n + i]));
        }
        int e;
        frexp(mx > 0 ? mx : 1.0, &e);
        rs[i] = ldexp(1.0, -e);
      }
      __syncthreads();
      for (int j = tid; j < n; j += NT) {
        double mx = 0;
        for (int i = 0; i < n; ++i) {
          double v = (i == n - 1) ? lte[j] : (i == j ? diag[i] : R[(size_t)j * n + i]);
          mx = fmax(mx, fabs(v) * rs[i]);
        }
        int e;
        frexp(mx > 0 ? mx : 1.0, &e);
        cs[j] = ldexp(1.0, -e);
      }
      __syncthreads();
      for (int idx = tid; idx < n * n; idx += NT) {
        int i = idx / n, j = idx - i * n;
        double v = (i == n - 1) ? lte[j] : (i == j ? diag[i] : R[(size_t)j * n + i]);
        lds_a[idx] = (float)(v * rs[i] * cs[j]);
      }
      __syncthreads();
      int sing = lu_f32_lds(lds_a, n, perm, nb, red);
      if (!sing) {
        // x = Dc As^-1 Dr e_{n-1}
        for (int i = tid; i < n; i += NT) r[i] = (i == n - 1) ? rs[i] : 0.0;
        __syncthreads();
        lu_f32_apply(lds_a, n, perm, r, d, tmp);
        for (int i = tid; i < n; i += NT) x[i] = d[i] * cs[i];
        __syncthreads();
        for (it = 1; it <= NLTE_MAX_ITERS; ++it) {
          // FP64 residual r = e_{n-1} - M x, scaled by Dr
          for (int i = tid; i < n; i += NT) {
            double acc;
            if (i == n - 1) {
              acc = 0;
              for (int j = 0; j < n; ++j) acc += lte[j] * x[j];
              acc = 1.0 - acc;
            } else {
              acc = diag[i] * x[i];
              for (int j = 0; j < n; ++j) if (j != i) acc += R[(size_t)j * n + i] * x[j];
              acc = -acc;
            }
            r[i] = acc * rs[i];
          }
          __syncthreads();
          lu_f32_apply(lds_a, n, perm, r, d, tmp);
          double dmax = 0, xmax = 0;
          for (int i = tid; i < n; i += NT) {
            double dx = d[i] * cs[i];
            x[i] += dx;
            dmax = fmax(dmax, fabs(dx));
            xmax = fmax(xmax, fabs(x[i]));
          }
          dmax = block_max(dmax, dred);
          xmax = block_max(xmax, dred);
          if (!(dmax == dmax) || !(xmax == xmax)) break;   // NaN: give up on FP32
          if (dmax <= A.tol * xmax) { ok = true; break; }
        }
      }
    }
    if (ok) {
      write_pops(n, x, lte, bout, pout);
      if (A.iters && tid == 0) A.iters[s] = it;
    } else {
      // FP64 fallback in the workgroup's device slot
      if (tid == 0) sing_f64 = 0;
      assemble<double>(n, R, lte, M64, n);
      lu_solve_f64(n, M64, perm, x, red, &sing_f64);
      write_pops(n, x, lte, bout, pout);
      if (A.iters && tid == 0) A.iters[s] = -1;
    }
    __syncthreads();
  }
}

}  // namespace sgpu

using namespace sgpu;

extern "C" int sgpu_nlte_solve(sgpu_ctx* ctx, int32_t n, const double* rates, const double* lte, double* b,
                               double* pop, double* work)
{
  SGPU_REQUIRE(ctx && rates && lte && b && work, "null argument");
  SGPU_REQUIRE(n >= 1 && n <= 1024, "n = %d outside [1, 1024]", n);
  timer_begin(ctx, "nlte");
  hipLaunchKernelGGL(nlte_single_kernel, dim3(1), dim3(NLTE_THREADS), 0, ctx->stream, n, rates, lte, b, pop, work);
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "nlte");
  return SGPU_OK;
}

namespace {

int batched_impl(sgpu_ctx* ctx, int32_t nsys, const int32_t* n, const int64_t* rate_off, const int64_t* vec_off,
                 const double* rates, const double* lte, double* b, double* pop, int32_t precision, double tol,
                 int32_t* iters, int threads, int panel, int wg_per_cu)
{
  SGPU_REQUIRE(ctx && n && rate_off && vec_off && rates && lte && b, "null argument");
  SGPU_REQUIRE(nsys >= 0, "negative system count");
  SGPU_REQUIRE(precision == SGPU_NLTE_FP64 || precision == SGPU_NLTE_FP32_REFINE, "unknown precision %d",
               precision);
  if (nsys == 0) return SGPU_OK;
  int st = 0;
  int32_t* order = (int32_t*)scratch(ctx, "nlte_order", sizeof(int32_t) * ((size_t)nsys + 64), &st);
  if (st) return st;
  int32_t* counter = order + nsys;
  int32_t* mm = order + nsys + 8;
  hipLaunchKernelGGL(size_range_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, nsys, mm);
  SGPU_LAUNCH_CHECK();
  int32_t hmm[2] = {0, 0};
  SGPU_HIP(hipMemcpyAsync(hmm, mm, sizeof(hmm), hipMemcpyDeviceToHost, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  SGPU_REQUIRE(hmm[0] >= 1 && hmm[1] <= 1024, "system sizes must lie in [1, 1024] (got %d..%d)", hmm[0], hmm[1]);
  const int max_n = hmm[1];
  const unsigned nwg = (unsigned)std::min<int64_t>((int64_t)ctx->n_cu * wg_per_cu, nsys);
  // slot: 7 vectors + perm + an FP64 matrix for the fallback / FP64 mode
  const size_t slot = (size_t)max_n * max_n + 8 * (size_t)max_n + 64;
  double* slots = (double*)scratch(ctx, "nlte_slots", sizeof(double) * slot * nwg, &st);
  if (st) return st;
  hipLaunchKernelGGL(order_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, nsys, order, counter);
  SGPU_LAUNCH_CHECK();
  BatchArgs A{nsys, n, rate_off, vec_off, rates, lte, b, pop, iters, order, counter, slots, slot, precision, tol};
  timer_begin(ctx, "nlte");
  if (precision == SGPU_NLTE_FP64) {
    hipLaunchKernelGGL(nlte_batched_f64_kernel, dim3(nwg), dim3(threads), 0, ctx->stream, A);
  } else {
    const size_t lds = sizeof(float) * NLTE_LDS_MAX_N * NLTE_LDS_MAX_N;
    SGPU_HIP(hipFuncSetAttribute((const void*)nlte_batched_mixed_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(nlte_batched_mixed_kernel, dim3(nwg), dim3(threads), lds, ctx->stream, A, panel);
  }
  SGPU_LAUNCH_CHECK();
  timer_end(ctx, "nlte");
  return SGPU_OK;
}

}  // namespace

extern "C" int sgpu_nlte_batched(sgpu_ctx* ctx, int32_t nsys, const int32_t* n, const int64_t* rate_off,
                                 const int64_t* vec_off, const double* rates, const double* lte, double* b,
                                 double* pop, int32_t precision, double tol, int32_t* iters)
{
  SGPU_REQUIRE(ctx, "null context");
  const nlte_tuning& t = ctx->nlte;
  return batched_impl(ctx, nsys, n, rate_off, vec_off, rates, lte, b, pop, precision, tol, iters,
                      t.valid ? t.threads : NLTE_THREADS, t.valid ? t.panel : 16, t.valid ? t.wg_per_cu : 1);
}

extern "C" int sgpu_nlte_autotune(sgpu_ctx* ctx, int32_t nsys, const int32_t* n, const int64_t* rate_off,
                                  const int64_t* vec_off, const double* rates, const double* lte, double* b,
                                  double* pop, int32_t* iters, char* report, size_t report_len)
{
  SGPU_REQUIRE(ctx, "null context");
  const int panels[] = {8, 16, 32};
  const int wgs[] = {1, 2};
  hipEvent_t e0, e1;
  SGPU_HIP(hipEventCreate(&e0));
  SGPU_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bp = 16, bw = 1;
  std::string rep;
  for (int w : wgs)
    for (int p : panels) {
      // one warm-up, then the timed run
      int st = batched_impl(ctx, nsys, n, rate_off, vec_off, rates, lte, b, pop, SGPU_NLTE_FP32_REFINE, 1e-12,
                            iters, NLTE_THREADS, p, w);
      if (st) return st;
      SGPU_HIP(hipEventRecord(e0, ctx->stream));
      st = batched_impl(ctx, nsys, n, rate_off, vec_off, rates, lte, b, pop, SGPU_NLTE_FP32_REFINE, 1e-12, iters,
                        NLTE_THREADS, p, w);
      if (st) return st;
      SGPU_HIP(hipEventRecord(e1, ctx->stream));
      SGPU_HIP(hipEventSynchronize(e1));
      float ms = 0;
      SGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
      char line[128];
      snprintf(line, sizeof(line), "panel=%d wg_per_cu=%d: %.4f ms\n", p, w, ms);
      rep += line;
      if (ms < best) { best = ms; bp = p; bw = w; }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->nlte.valid = 1;
  ctx->nlte.threads = NLTE_THREADS;
  ctx->nlte.panel = bp;
  ctx->nlte.wg_per_cu = bw;
  char line[128];
  snprintf(line, sizeof(line), "chosen: panel=%d wg_per_cu=%d (%.4f ms)\n", bp, bw, best);
  rep += line;
  if (report && report_len) {
    strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  return SGPU_OK;
}
