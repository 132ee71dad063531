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
//                       Blocked right-looking LU in LDS (row stride n | 1):
//                       one wavefront factors each panel and runs the
//                       triangular solves (DPP pivot search, readlane
//                       broadcasts, no workgroup barrier), every wavefront
//                       takes part in the panel's row solve and the trailing
//                       update -- three workgroup barriers per panel.
//     FP64 mode         FP64 LU in LDS for n <= 143 (164 KB), in a
//                       per-workgroup device-memory slot above that.
//     FP32_REFINE mode  systems whose FP64 matrix fits the launch's LDS
//                       (sized for its largest system) are factored directly
//                       in FP64 as in the FP64 mode (iters 0: fewer steps
//                       than FP32 with refinement at these sizes); larger
//                       ones: row/column power-of-two equilibration, FP32 LU
//                       in LDS (n <= 201), trailing updates on MFMA
//                       (v_mfma_f32_32x32x2f32) for the bulk of the matrix,
//                       FP64 residuals r = e - M b straight from the FP64
//                       rates, FP64 correction solves through the FP32
//                       factors, until max|db| <= tol max|b|; systems that do
//                       not converge (or are too large for LDS) are re-solved
//                       in FP64.
//   sgpu_nlte_autotune  times workgroup size x panel width x workgroups per
//                       CU (those whose LDS can share a CU) for both modes.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include "sgpu_internal.h"
#include "sgpu_device.h"

namespace sgpu {

constexpr int NLTE_THREADS = 256;       // default workgroup size
constexpr int NLTE_MAX_THREADS = 512;   // largest the batched kernels accept
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

// Assemble M (row-major, n x n, row stride ld) from rates R (row-major) and
// lte: M_ii = -sum_j R_ij (summed in j order), M_ij = R_ji, last row lte.
// The transpose reads R along its rows (coalesced), eight loads in flight
// per thread, and scatters the writes.
template <class T>
__device__ __forceinline__ void assemble(int n, const double* __restrict__ R, const double* __restrict__ lte,
                                         T* __restrict__ M, int ld)
{
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double Rout = 0.0;
    for (int j = 0; j < n; ++j) Rout += R[(size_t)i * n + j];
    if (i != n - 1) M[(size_t)i * ld + i] = (T)(-1 * Rout);
  }
  const int NT = blockDim.x, nn = n * n;
  int jj = threadIdx.x / n, ii = threadIdx.x - jj * n;   // R[jj][ii] = M[ii][jj]
  const int dj = NT / n, di = NT - dj * n;
  for (int base = threadIdx.x; base < nn; base += 8 * NT) {
    double t[8];
    int ri[8], rj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      ri[u] = ii;
      rj[u] = jj;
      t[u] = base + u * NT < nn ? R[(size_t)jj * n + ii] : 0.0;
      ii += di;
      jj += dj;
      if (ii >= n) { ii -= n; ++jj; }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = ri[u], j = rj[u];
      if (base + u * NT < nn && (i == n - 1 || i != j)) M[(size_t)i * ld + j] = (T)(i == n - 1 ? lte[j] : t[u]);
    }
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

// Blocked form of the same elimination for systems too large for LDS, in
// device memory: the augmented n x (n+1) matrix [M | e_{n-1}] (row stride
// n + 1) is reduced to [U | y] panel by panel, so the right-hand side rides
// along with the row interchanges and eliminations (forward substitution
// done with them) and the multipliers are not kept.  Per panel of BLU_NB
// columns: the panel's rows j0..n-1 are copied to LDS and factored there
// column by column with the first-largest pivot (gsl_linalg_LU_decomp's
// choice); the interchanges are applied to the columns right of the panel
// as one composed gather; the row block U12 = L11^-1 A12 is formed one
// thread per column; the trailing block A22 -= L21 U12 is updated with every
// row read and written once per panel (lanes over consecutive columns,
// wavefronts over rows, the L21 row an LDS broadcast).  Back substitution
// runs in blocks of BLU_NB rows from the bottom: the diagonal block from
// LDS by one thread, the rows above updated with a dot product over the
// block.
// ------------------------------------------------------------------
// wavefront pivot search: DPP within rows of 16 lanes, readlane across them
// ------------------------------------------------------------------
template <class V> struct PivotOf { V v; int i; };

template <class V>
__device__ __forceinline__ PivotOf<V> better_p(PivotOf<V> a, PivotOf<V> b)
{
  // larger magnitude wins; the smaller row index wins ties (first maximum)
  return (b.v > a.v || (b.v == a.v && b.i < a.i)) ? b : a;
}

template <int CTRL>
__device__ __forceinline__ int dpp(int v) { return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false); }

template <int CTRL>
__device__ __forceinline__ PivotOf<float> dpp_better(PivotOf<float> m)
{
  PivotOf<float> o;
  o.v = __int_as_float(dpp<CTRL>(__float_as_int(m.v)));
  o.i = dpp<CTRL>(m.i);
  return better_p(m, o);
}

template <int CTRL>
__device__ __forceinline__ PivotOf<double> dpp_better(PivotOf<double> m)
{
  PivotOf<double> o;
  const long long b = __double_as_longlong(m.v);
  const int lo = dpp<CTRL>((int)(b & 0xffffffff)), hi = dpp<CTRL>((int)(b >> 32));
  o.v = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  o.i = dpp<CTRL>(m.i);
  return better_p(m, o);
}

__device__ __forceinline__ float readlane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double readlane(double v, int l)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// the wavefront's best (largest |value|, then smallest row), uniform result:
// xor 1 and xor 2 inside quads, half-row and row mirrors (every lane then
// holds its 16-lane row's best), the four rows through readlane
template <class V>
__device__ __forceinline__ PivotOf<V> wave_argmax(PivotOf<V> m)
{
  m = dpp_better<0xB1>(m);    // quad_perm [1,0,3,2]
  m = dpp_better<0x4E>(m);    // quad_perm [2,3,0,1]
  m = dpp_better<0x141>(m);   // row_half_mirror
  m = dpp_better<0x140>(m);   // row_mirror
  PivotOf<V> r = {readlane(m.v, 0), __builtin_amdgcn_readlane(m.i, 0)};
  r = better_p(r, PivotOf<V>{readlane(m.v, 16), __builtin_amdgcn_readlane(m.i, 16)});
  r = better_p(r, PivotOf<V>{readlane(m.v, 32), __builtin_amdgcn_readlane(m.i, 32)});
  r = better_p(r, PivotOf<V>{readlane(m.v, 48), __builtin_amdgcn_readlane(m.i, 48)});
  return r;
}


constexpr int BLU_NB = 16;               // panel width
constexpr int BLU_PS = BLU_NB + 1;       // LDS row stride of the panel (doubles)
constexpr int BLU_CW = 64;               // columns per trailing-update chunk
constexpr int BLU_RU = 8;                // rows per thread per trailing-update pass

constexpr int BLU_INTS = 5 * BLU_NB + 1;  // panel pivots, composed interchanges

constexpr int BLU_MAXW = NLTE_MAX_THREADS / WAVE;   // wavefronts of the largest workgroup

// dynamic LDS of lu_solve_f64_blocked: the panel (the back substitution
// reuses its space for y and one diagonal block), the pivot search's two
// buffers of per-wavefront maxima, the ints
__host__ __device__ inline size_t blocked_doubles(int n)
{
  const size_t panel = (size_t)n * BLU_PS, back = (size_t)n + BLU_NB * BLU_NB;
  return panel > back ? panel : back;
}
__host__ __device__ inline size_t blocked_lds_bytes(int n)
{
  return sizeof(double) * blocked_doubles(n) + sizeof(ArgMax) * 2 * BLU_MAXW + sizeof(int) * BLU_INTS;
}

// [M | e_{n-1}] with row stride n + 1 for lu_solve_f64_blocked
__device__ __forceinline__ void assemble_augmented(int n, const double* R, const double* lte, double* M)
{
  for (int i = threadIdx.x; i < n; i += blockDim.x) M[(size_t)i * (n + 1) + n] = (i == n - 1) ? 1.0 : 0.0;
  assemble<double>(n, R, lte, M, n + 1);
}

// One panel step of the blocked LU of the augmented matrix M (n x (n + 1),
// row stride n + 1) at columns j0 .. j0 + jb - 1: the panel loaded into LDS
// and factored with partial pivoting, written back (U11 and L21), its
// interchanges applied to the columns on its right (the rhs included), and
// U12 = L11^-1 A12.  The panel stays in LDS for the trailing update.
// lds: blocked_lds_bytes(n) bytes of dynamic LDS.
//
// With pend (BLU_NB x BLU_NB doubles of LDS, or null), the previous panel's
// trailing update has been applied to this panel's columns only (a delayed
// update): its L columns j0 - BLU_NB .. j0 - 1 take this panel's
// interchanges too, and U12 is formed from A12 - L_prev U12_prev on this
// panel's rows, U12_prev being the previous panel's rows of each column.
__device__ __forceinline__ void blocked_panel(int n, double* __restrict__ M, int j0, double* __restrict__ lds,
                                              ArgMax* red, int* singular, double* pend = nullptr)
{
  const int tid = threadIdx.x, NT = blockDim.x;
  const int ld = n + 1;
  double* P = lds;                                  // panel rows, stride BLU_PS
  ArgMax* wmax = (ArgMax*)(lds + blocked_doubles(n));   // [2][BLU_MAXW]
  int* piv = (int*)((char*)lds + blocked_lds_bytes(n) - sizeof(int) * BLU_INTS);   // BLU_INTS ints
  (void)red;
  {
    const int jb = min(BLU_NB, n - j0);
    const int m = n - j0;
    // the last, narrower panel is padded with zero columns (the trailing
    // update runs over all BLU_NB of them)
    for (int base = 0; base < m * BLU_NB; base += 4 * NT) {
      double v[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int idx = base + e * NT + tid, r = idx / BLU_NB, c = idx % BLU_NB;
        v[e] = (idx < m * BLU_NB && c < jb) ? M[(size_t)(j0 + r) * ld + j0 + c] : 0.0;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int idx = base + e * NT + tid;
        if (idx < m * BLU_NB) P[(idx / BLU_NB) * BLU_PS + idx % BLU_NB] = v[e];
      }
    }
    __syncthreads();
    // Two workgroup barriers per column: the elimination of column t also
    // finds each thread's largest |entry| of column t + 1 among the rows it
    // just updated; wavefront maxima go to one of two LDS buffers (column
    // parity), so a buffer is rewritten only two barriers after its last
    // read.  Ties go to the smaller row, as a sequential search.
    const int nw = NT / WAVE, w = tid / WAVE, lane = tid % WAVE;
    PivotOf<double> am = {-1.0, m};
    for (int r = tid; r < m; r += NT) am = better_p(am, PivotOf<double>{fabs(P[r * BLU_PS]), r});
    for (int t = 0; t < jb; ++t) {
      am = wave_argmax(am);
      ArgMax* buf = wmax + (t & 1) * BLU_MAXW;
      if (lane == 0) buf[w] = ArgMax{am.v, am.i};
      __syncthreads();
      ArgMax p = buf[0];
      for (int k = 1; k < nw; ++k) p = better(p, buf[k]);
      const int pr = p.i < m ? p.i : t;
      if (tid == 0) piv[t] = pr;
      if (pr != t)
        for (int c = tid; c < jb; c += NT) {
          const double a = P[t * BLU_PS + c];
          P[t * BLU_PS + c] = P[pr * BLU_PS + c];
          P[pr * BLU_PS + c] = a;
        }
      __syncthreads();
      const double ajj = P[t * BLU_PS + t];
      const int tn = t + 1;
      am = PivotOf<double>{-1.0, m};
      if (ajj == 0) {
        if (tid == 0) *singular = 1;
        if (tn < jb)
          for (int r = tn + tid; r < m; r += NT) am = better_p(am, PivotOf<double>{fabs(P[r * BLU_PS + tn]), r});
        continue;
      }
      for (int r = tn + tid; r < m; r += NT) {
        const double l = P[r * BLU_PS + t] / ajj;
        P[r * BLU_PS + t] = l;
        for (int c = tn; c < jb; ++c) P[r * BLU_PS + c] -= l * P[t * BLU_PS + c];
        if (tn < jb) am = better_p(am, PivotOf<double>{fabs(P[r * BLU_PS + tn]), r});
      }
    }
    __syncthreads();
    // the factored panel back to device memory (U11 and L21; only U11 is
    // read again, by the back substitution)
    for (int idx = tid; idx < m * jb; idx += NT) {
      const int r = idx / jb, c = idx - r * jb;
      M[(size_t)(j0 + r) * ld + j0 + c] = P[r * BLU_PS + c];
    }
    // the panel's interchanges on the columns to its right (the rhs
    // included), composed into one gather: the rows they touch (at most
    // 2 BLU_NB) and, for each, the row whose value ends there
    int* sw_dst = piv + BLU_NB;
    int* sw_src = sw_dst + 2 * BLU_NB;
    int* sw_n = sw_src + 2 * BLU_NB;
    if (tid < WAVE) {
      // lane q holds touched row dst (the panel's rows 0..jb-1 in lanes
      // 0..jb-1, then each pivot row below them once) and, once the
      // interchanges are applied in order, the row src whose value ends there
      const int lane = tid;
      int dst = lane < jb ? lane : -1;
      int nt = jb;
      for (int t = 0; t < jb; ++t) {
        const int r = piv[t];
        if (r >= jb && __ballot(dst == r) == 0) {
          if (lane == nt) dst = r;
          ++nt;
        }
      }
      int src = dst;
      for (int t = 0; t < jb; ++t) {
        const int r = piv[t];
        if (r != t) {
          const int qb = __ffsll((long long)__ballot(dst == r)) - 1;
          const int sa = __builtin_amdgcn_readlane(src, t), sb = __builtin_amdgcn_readlane(src, qb);
          if (lane == t) src = sb;
          if (lane == qb) src = sa;
        }
      }
      if (lane < nt) {
        sw_dst[lane] = dst;
        sw_src[lane] = src;
      }
      if (lane == 0) *sw_n = nt;
    }
    __syncthreads();
    // one thread per column: the touched rows gathered, then U12 = L11^-1 A12
    // on the panel's rows (lanes 0..jb-1 of the composition hold rows
    // 0..jb-1, so v[0..jb) is the interchanged A12 column), all stored back
    const int nsw = *sw_n;
    if (pend) {
      // the previous panel's L columns take this panel's interchanges, then
      // its rows of this panel (BLU_NB x BLU_NB) are staged in LDS
      for (int k = j0 - BLU_NB + tid; k < j0; k += NT) {
        double v[2 * BLU_NB];
#pragma unroll
        for (int q = 0; q < 2 * BLU_NB; ++q)
          if (q < nsw) v[q] = M[(size_t)(j0 + sw_src[q]) * ld + k];
#pragma unroll
        for (int q = 0; q < 2 * BLU_NB; ++q)
          if (q < nsw) M[(size_t)(j0 + sw_dst[q]) * ld + k] = v[q];
      }
      __syncthreads();
      for (int idx = tid; idx < BLU_NB * BLU_NB; idx += NT) {
        const int t = idx / BLU_NB, q = idx % BLU_NB;
        pend[idx] = t < jb ? M[(size_t)(j0 + t) * ld + j0 - BLU_NB + q] : 0.0;
      }
      __syncthreads();
    }
    for (int k = j0 + jb + tid; k <= n; k += NT) {
      double v[2 * BLU_NB];
#pragma unroll
      for (int q = 0; q < 2 * BLU_NB; ++q)
        if (q < nsw) v[q] = M[(size_t)(j0 + sw_src[q]) * ld + k];
      if (pend) {
        // the delayed update of this panel's rows: v_t -= L_prev[t] U12_prev
        double u[BLU_NB];
#pragma unroll
        for (int q = 0; q < BLU_NB; ++q) u[q] = M[(size_t)(j0 - BLU_NB + q) * ld + k];
#pragma unroll
        for (int t = 0; t < BLU_NB; ++t)
          if (t < jb) {
#pragma unroll
            for (int q = 0; q < BLU_NB; ++q) v[t] -= pend[t * BLU_NB + q] * u[q];
          }
      }
#pragma unroll
      for (int t = 1; t < BLU_NB; ++t)
        if (t < jb) {
#pragma unroll
          for (int q = 0; q < BLU_NB - 1; ++q)
            if (q < t) v[t] -= P[t * BLU_PS + q] * v[q];
        }
#pragma unroll
      for (int q = 0; q < 2 * BLU_NB; ++q)
        if (q < nsw) M[(size_t)(j0 + sw_dst[q]) * ld + k] = v[q];
    }
    __syncthreads();
  }
}

// The trailing update of the panel step at j0 by one workgroup, L21 from the
// panel in LDS: A22 -= L21 U12.  Lanes over BLU_CW consecutive columns
// (coalesced rows), wavefronts over rows; each thread keeps its column of
// U12 in registers and updates BLU_RU rows per pass with their loads issued
// together; the L21 row is an LDS broadcast.
__device__ __forceinline__ void blocked_trailing(int n, double* __restrict__ M, int j0,
                                                 const double* __restrict__ lds)
{
  const int tid = threadIdx.x, NT = blockDim.x;
  const int ld = n + 1;
  const double* P = lds;
  {
    const int jb = min(BLU_NB, n - j0);
    const int rows_per = NT / BLU_CW;
    const int col = tid % BLU_CW, rg = tid / BLU_CW;
    const int i0 = j0 + jb;
    for (int c0 = i0; c0 <= n; c0 += BLU_CW) {
      const int k = c0 + col;
      if (k > n) continue;
      double u[BLU_NB];
#pragma unroll
      for (int t = 0; t < BLU_NB; ++t) u[t] = t < jb ? M[(size_t)(j0 + t) * ld + k] : 0.0;
      int i = i0 + rg;
      for (; i + (BLU_RU - 1) * rows_per < n; i += BLU_RU * rows_per) {
        double v[BLU_RU];
#pragma unroll
        for (int e = 0; e < BLU_RU; ++e) v[e] = M[(size_t)(i + e * rows_per) * ld + k];
#pragma unroll
        for (int e = 0; e < BLU_RU; ++e) {
          const double* pl = P + (size_t)(i + e * rows_per - j0) * BLU_PS;
#pragma unroll
          for (int t = 0; t < BLU_NB; ++t) v[e] -= pl[t] * u[t];
        }
#pragma unroll
        for (int e = 0; e < BLU_RU; ++e) M[(size_t)(i + e * rows_per) * ld + k] = v[e];
      }
      for (; i < n; i += rows_per) {
        const double* pl = P + (size_t)(i - j0) * BLU_PS;
        double v = M[(size_t)i * ld + k];
#pragma unroll
        for (int t = 0; t < BLU_NB; ++t) v -= pl[t] * u[t];
        M[(size_t)i * ld + k] = v;
      }
    }
    __syncthreads();
  }
}

// Back substitution U x = y of the factored augmented matrix (y in column
// n), the running right-hand side in LDS: x is left in lds[0 .. n).
__device__ __forceinline__ void blocked_back_subst(int n, const double* __restrict__ M, double* __restrict__ lds)
{
  const int tid = threadIdx.x, NT = blockDim.x;
  const int ld = n + 1;
  double* Y = lds;                                  // n doubles
  double* D = lds + n;                              // BLU_NB x BLU_NB diagonal block
  for (int i = tid; i < n; i += NT) Y[i] = M[(size_t)i * ld + n];
  __syncthreads();
  const int nblk = (n + BLU_NB - 1) / BLU_NB;
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int b0 = bk * BLU_NB, bs = min(BLU_NB, n - b0);
    for (int idx = tid; idx < bs * bs; idx += NT) {
      const int r = idx / bs, c = idx - r * bs;
      D[r * BLU_NB + c] = M[(size_t)(b0 + r) * ld + b0 + c];
    }
    __syncthreads();
    if (tid == 0) {
      for (int r = bs - 1; r >= 0; --r) {
        double v = Y[b0 + r];
        for (int c = r + 1; c < bs; ++c) v -= D[r * BLU_NB + c] * Y[b0 + c];
        Y[b0 + r] = v / D[r * BLU_NB + r];
      }
    }
    __syncthreads();
    for (int i = tid; i < b0; i += NT) {
      const double* row = M + (size_t)i * ld + b0;
      double v = Y[i];
      for (int c = 0; c < bs; ++c) v -= row[c] * Y[b0 + c];
      Y[i] = v;
    }
    __syncthreads();
  }
}

// Blocked FP64 LU of [M | e_{n-1}] (row stride n + 1, in device memory) by
// one workgroup, panels in LDS, then the back substitution into x.
// lds: blocked_lds_bytes(n) bytes of dynamic LDS.
__device__ __forceinline__ void lu_solve_f64_blocked(int n, double* __restrict__ M, double* __restrict__ x,
                                                     double* __restrict__ lds, ArgMax* red, int* singular)
{
  for (int j0 = 0; j0 < n; j0 += BLU_NB) {
    blocked_panel(n, M, j0, lds, red, singular);
    blocked_trailing(n, M, j0, lds);
  }
  blocked_back_subst(n, M, lds);
  for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = lds[i];
  __syncthreads();
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
// smallest and largest system size and the number of systems above cap
// (single workgroup): mm[0] = min, mm[1] = max, mm[2] = count(n > cap)
__global__ void __launch_bounds__(1024) size_range_kernel(const int32_t* n, int nsys, int cap, int32_t* mm)
{
  __shared__ int lo[1024 / WAVE], hi[1024 / WAVE], cnt[1024 / WAVE];
  int a = 1 << 30, b = -(1 << 30), c = 0;
  for (int s = threadIdx.x; s < nsys; s += blockDim.x) { a = min(a, n[s]); b = max(b, n[s]); c += n[s] > cap; }
  for (int o = 32; o > 0; o >>= 1) {
    a = min(a, __shfl_xor(a, o, WAVE)); b = max(b, __shfl_xor(b, o, WAVE)); c += __shfl_xor(c, o, WAVE);
  }
  if (threadIdx.x % WAVE == 0) { lo[threadIdx.x / WAVE] = a; hi[threadIdx.x / WAVE] = b; cnt[threadIdx.x / WAVE] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x / WAVE); w++) { a = min(a, lo[w]); b = max(b, hi[w]); c += cnt[w]; }
    a = min(a, lo[0]); b = max(b, hi[0]);
    mm[0] = a; mm[1] = b; mm[2] = c;
  }
}

// bucket sort of system indices by size, descending (single workgroup):
// histogram, prefix, then every thread places its systems through an
// atomic cursor (the order within one size is arbitrary; each system's
// result does not depend on it).  The queue cursor starts at the first
// system of at most cap levels (the larger ones take the wide path).
__global__ void __launch_bounds__(1024) order_kernel(const int32_t* n, int nsys, int cap, int32_t* order,
                                                     int32_t* counter)
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
    *counter = hist[1024 - cap];
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
    int v = n[s];
    v = v < 0 ? 0 : (v > 1024 ? 1024 : v);
    order[atomicAdd(&hist[1024 - v], 1)] = s;
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

__global__ void __launch_bounds__(NLTE_MAX_THREADS) nlte_batched_f64_kernel(BatchArgs A, size_t /*lds_bytes*/)
{
  __shared__ ArgMax red[NLTE_MAX_THREADS / WAVE];
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
// Wide systems: the blocked FP64 LU of every system too large for the
// in-LDS paths, spread over the whole GPU.  Each panel step is two
// launches: the panel (one workgroup per system, as blocked_panel) and the
// trailing update A22 -= L21 U12 in 64 x 64 tiles of every system (a
// workgroup per tile), so the update of a large system runs on many CUs
// instead of one.  Wide system w is order[w] (the largest first), its
// augmented matrix at M + w stride.
// ------------------------------------------------------------------
constexpr int WIDE_THREADS = 512;   // back-substitution workgroups
constexpr int WIDE_PANEL_THREADS = 256;   // panel workgroups (measured: 512 took 3 % longer)
// wide systems in interleaved groups, one stream each (the main stream and
// side streams 1 ..), so that groups' panels overlap other groups' updates
// (measured on the per-atom batch: 2 groups 4.3 ms, 3 groups 7.3 ms,
// 4 groups 7.1 ms -- more streams than hardware queues serialise badly)
constexpr int WIDE_GROUPS = 2;
static_assert(WIDE_GROUPS <= sgpu_ctx::N_SIDE, "a side stream per extra group");
constexpr int WIDE_TILE = 64;       // trailing-update tile (rows and columns)
constexpr int WIDE_UPD_THREADS = 256;

struct WideArgs {
  const int32_t* n;
  const int32_t* order;
  const int64_t* rate_off;
  const int64_t* vec_off;
  const double* rates;
  const double* lte;
  double* b;
  double* pop;
  int32_t* iters;
  double* M;
  size_t stride;       // doubles per system
  int iters_value;     // reported iterations (0 FP64 mode, -1 FP64 fallback of the FP32 mode)
  int gstride, goff;   // this launch's group: wide systems goff, goff + gstride, ...
};

// wide system index of the group's g-th system
__device__ __forceinline__ int wide_index(const WideArgs& W, int g) { return g * W.gstride + W.goff; }

// [M | e_{n-1}] of the wide systems as assemble_augmented builds it, grid
// (wide systems, 64-row blocks of R): the workgroup reads R's rows
// i0 .. i0 + 63 in 64 x 64 tiles through LDS (coalesced), sums each row in
// j order (M_ii = -sum_j R_ij, as assemble), and writes the transposed
// tile (M_ji = R_ij) along M's rows; the last row is lte, column n the rhs.
constexpr int WIDE_ASM_THREADS = 256;
__global__ void __launch_bounds__(WIDE_ASM_THREADS) nlte_wide_assemble_kernel(WideArgs W)
{
  __shared__ double T[WIDE_TILE][WIDE_TILE + 1];
  const int w = wide_index(W, blockIdx.x);
  const int s = W.order[w];
  const int n = W.n[s];
  const int i0 = blockIdx.y * WIDE_TILE;
  if (i0 >= n) return;
  const double* R = W.rates + W.rate_off[s];
  const double* lte = W.lte + W.vec_off[s];
  double* M = W.M + (size_t)w * W.stride;
  const int ld = n + 1, tid = threadIdx.x;
  constexpr int PER = WIDE_TILE * WIDE_TILE / WIDE_ASM_THREADS;
  double Rout = 0.0;   // row i0 + tid (threads below WIDE_TILE)
  for (int j0 = 0; j0 < n; j0 += WIDE_TILE) {
    double v[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + e * WIDE_ASM_THREADS, a = idx / WIDE_TILE, b = idx % WIDE_TILE;
      v[e] = (i0 + a < n && j0 + b < n) ? R[(size_t)(i0 + a) * n + j0 + b] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + e * WIDE_ASM_THREADS;
      T[idx / WIDE_TILE][idx % WIDE_TILE] = v[e];
    }
    __syncthreads();
    if (tid < WIDE_TILE) {
      const int jn = min(WIDE_TILE, n - j0);
      for (int b = 0; b < jn; ++b) Rout += T[tid][b];
    }
    // M[j][i] = R[i][j] off the diagonal, rows j < n - 1
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + e * WIDE_ASM_THREADS, b = idx / WIDE_TILE, a = idx % WIDE_TILE;
      const int i = i0 + a, j = j0 + b;
      if (i < n && j < n - 1 && i != j) M[(size_t)j * ld + i] = T[a][b];
    }
    __syncthreads();
  }
  if (tid < WIDE_TILE) {
    const int i = i0 + tid;
    if (i < n) {
      if (i != n - 1) M[(size_t)i * ld + i] = -1 * Rout;
      M[(size_t)(n - 1) * ld + i] = lte[i];
      M[(size_t)i * ld + n] = (i == n - 1) ? 1.0 : 0.0;
    }
  }
}

// pending: the previous panel's trailing update was applied to this panel's
// columns only (blocked_panel's delayed update)
__global__ void __launch_bounds__(WIDE_PANEL_THREADS) nlte_wide_panel_kernel(WideArgs W, int j0, int pending)
{
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ ArgMax red[WIDE_PANEL_THREADS / WAVE];
  __shared__ double pend[BLU_NB * BLU_NB];
  __shared__ int singular;
  const int w = wide_index(W, blockIdx.x);
  const int n = W.n[W.order[w]];
  if (j0 >= n) return;
  blocked_panel(n, W.M + (size_t)w * W.stride, j0, (double*)lds_raw, red, &singular, pending ? pend : nullptr);
}

// grid (tiles of the largest system at this step, wide systems); tiles
// beyond a system's own trailing matrix return at once.  A22 -= L21 U12 for
// the K columns of L from j0 (one panel, or two with a delayed update),
// all columns (c_last < 0) or the next panel's (up to c_last).  Each wavefront updates a 16-row strip of the
// 64 x 64 tile as four 16 x 16 blocks on v_mfma_f64_16x16x4f64 (K / 4
// steps), operands from LDS.
template <int K>
__global__ void __launch_bounds__(WIDE_UPD_THREADS) nlte_wide_update_kernel(WideArgs W, int j0, int c_last)
{
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  __shared__ double Ls[WIDE_TILE][K + 1];
  __shared__ double Us[K][WIDE_TILE + 2];
  const int w = wide_index(W, blockIdx.y);
  const int n = W.n[W.order[w]];
  if (j0 >= n) return;
  const int jb = min(K, n - j0), i0 = j0 + jb;
  // last column updated: every column (the rhs included) for c_last < 0,
  // else the next panel's columns only, up to c_last
  const int c_end = c_last < 0 ? n : min(n - 1, c_last);
  const int mr = n - i0, mc = c_end + 1 - i0;
  if (mr <= 0 || mc <= 0) return;
  const int tc = (mc + WIDE_TILE - 1) / WIDE_TILE, tr = (mr + WIDE_TILE - 1) / WIDE_TILE;
  if ((int)blockIdx.x >= tr * tc) return;
  const int r0 = i0 + ((int)blockIdx.x / tc) * WIDE_TILE, c0 = i0 + ((int)blockIdx.x % tc) * WIDE_TILE;
  double* M = W.M + (size_t)w * W.stride;
  const int ld = n + 1, tid = threadIdx.x;
  // L21 rows r0.. (negated) and U12 columns c0.. of this tile, zero-padded
  // to K; the loads issued together, then stored
  constexpr int PER = WIDE_TILE * K / WIDE_UPD_THREADS;
  double lv[PER], uv[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * WIDE_UPD_THREADS;
    const int r = idx / K, t = idx % K;
    lv[e] = (t < jb && r0 + r < n) ? -M[(size_t)(r0 + r) * ld + j0 + t] : 0.0;
    const int t2 = idx / WIDE_TILE, c = idx % WIDE_TILE;
    uv[e] = (t2 < jb && c0 + c <= c_end) ? M[(size_t)(j0 + t2) * ld + c0 + c] : 0.0;
  }
  // the tile's own entries, four 16 x 16 blocks of this wavefront's strip
  const int wave = tid / WAVE, lane = tid % WAVE, lr = lane % 16, lk = lane / 16;
  const int rs = r0 + wave * 16;
  f64x4 acc[WIDE_TILE / 16];
#pragma unroll
  for (int cb = 0; cb < WIDE_TILE / 16; ++cb) {
    const int col = c0 + cb * 16 + lr;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rs + lk + 4 * v;
      acc[cb][v] = (row < n && col <= c_end) ? M[(size_t)row * ld + col] : 0.0;
    }
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * WIDE_UPD_THREADS;
    Ls[idx / K][idx % K] = lv[e];
    Us[idx / WIDE_TILE][idx % WIDE_TILE] = uv[e];
  }
  __syncthreads();
  double av[K / 4];
#pragma unroll
  for (int k4 = 0; k4 < K / 4; ++k4) av[k4] = Ls[wave * 16 + lr][4 * k4 + lk];
#pragma unroll
  for (int cb = 0; cb < WIDE_TILE / 16; ++cb) {
#pragma unroll
    for (int k4 = 0; k4 < K / 4; ++k4)
      if (4 * k4 < jb)
        acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[k4], Us[4 * k4 + lk][cb * 16 + lr], acc[cb], 0, 0, 0);
  }
#pragma unroll
  for (int cb = 0; cb < WIDE_TILE / 16; ++cb) {
    const int col = c0 + cb * 16 + lr;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rs + lk + 4 * v;
      if (row < n && col <= c_end) M[(size_t)row * ld + col] = acc[cb][v];
    }
  }
}

// Back substitution U x = y (y in column n) with each thread owning rows
// tid and tid + WIDE_THREADS (n <= 2 WIDE_THREADS): a row's 16-column
// segment of the next block is loaded while the current block is solved,
// the diagonal block by the first 16 lanes of wavefront 0 (row r in lane r,
// x broadcast by readlane), the rows above updated from x in LDS.  The
// same operations in the same order as blocked_back_subst; x is left in
// lds[0 .. n).
__device__ __forceinline__ void back_subst_rows(int n, const double* __restrict__ M, double* __restrict__ lds)
{
  constexpr int RPT = 2;
  const int tid = threadIdx.x, lane = tid % WAVE;
  const int ld = n + 1;
  double* X = lds;                     // n doubles
  double* Dl = lds + n;                // BLU_NB rows of stride BLU_PS
  double* Yl = Dl + BLU_NB * BLU_PS;   // BLU_NB
  double y[RPT], seg[RPT][BLU_NB], nxt[RPT][BLU_NB];
  const int nblk = (n + BLU_NB - 1) / BLU_NB;
  {
    const int b0 = (nblk - 1) * BLU_NB, bs = n - b0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int i = tid + q * WIDE_THREADS;
      y[q] = i < n ? M[(size_t)i * ld + n] : 0.0;
#pragma unroll
      for (int c = 0; c < BLU_NB; ++c) seg[q][c] = (i < n && c < bs) ? M[(size_t)i * ld + b0 + c] : 0.0;
    }
  }
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int b0 = bk * BLU_NB, bs = min(BLU_NB, n - b0);
    // the next block's segments of the rows above this block
    if (bk > 0) {
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int i = tid + q * WIDE_THREADS;
#pragma unroll
        for (int c = 0; c < BLU_NB; ++c) nxt[q][c] = i < b0 ? M[(size_t)i * ld + b0 - BLU_NB + c] : 0.0;
      }
    }
    // this block's rows hand their diagonal-block rows and y to LDS
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int i = tid + q * WIDE_THREADS;
      if (i >= b0 && i < b0 + bs) {
#pragma unroll
        for (int c = 0; c < BLU_NB; ++c) Dl[(i - b0) * BLU_PS + c] = seg[q][c];
        Yl[i - b0] = y[q];
      }
    }
    __syncthreads();
    if (tid < WAVE) {
      double Dr[BLU_NB];
      const int rr = lane < BLU_NB ? lane : 0;
#pragma unroll
      for (int c = 0; c < BLU_NB; ++c) Dr[c] = Dl[rr * BLU_PS + c];
      const double yr = Yl[rr];
      double xr[BLU_NB];
#pragma unroll
      for (int r = BLU_NB - 1; r >= 0; --r) {
        xr[r] = 0.0;
        if (r < bs) {
          double v = yr;
#pragma unroll
          for (int c = r + 1; c < BLU_NB; ++c)
            if (c < bs) v -= Dr[c] * xr[c];
          xr[r] = readlane(v / Dr[r], r);
          if (lane == r) X[b0 + r] = xr[r];
        }
      }
    }
    __syncthreads();
    // the rows above: y_i -= U[i][b0 .. b0 + bs) x[b0 .. b0 + bs)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int i = tid + q * WIDE_THREADS;
      if (i < b0) {
        double v = y[q];
#pragma unroll
        for (int c = 0; c < BLU_NB; ++c)
          if (c < bs) v -= seg[q][c] * X[b0 + c];
        y[q] = v;
      }
#pragma unroll
      for (int c = 0; c < BLU_NB; ++c) seg[q][c] = nxt[q][c];
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(WIDE_THREADS) nlte_wide_back_kernel(WideArgs W)
{
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* lds = (double*)lds_raw;
  const int w = wide_index(W, blockIdx.x);
  const int s = W.order[w];
  const int n = W.n[s];
  back_subst_rows(n, W.M + (size_t)w * W.stride, lds);
  write_pops(n, lds, W.lte + W.vec_off[s], W.b + W.vec_off[s], W.pop ? W.pop + W.vec_off[s] : nullptr);
  if (W.iters && threadIdx.x == 0) W.iters[s] = W.iters_value;
}

// ------------------------------------------------------------------
// LU in LDS.  One wavefront factors each panel (DPP pivot search, readlane
// broadcasts, no workgroup barrier inside the panel); every wavefront takes
// part in the panel's row solve and in the trailing update, which runs on
// MFMA for whole and partial tiles.  The row stride is odd (ld = n | 1), so
// that a column (lane-strided by ld) and a row both read LDS without bank
// conflicts.  The loads of a step are issued together, ahead of its stores:
// one LDS round trip per step instead of one per element.
// ------------------------------------------------------------------
template <class T> struct LdsCap;
template <> struct LdsCap<float> { static constexpr int n = 201; };    // 201 * 201 * 4 = 161,604 B
template <> struct LdsCap<double> { static constexpr int n = 143; };   // 143 * 143 * 8 = 163,592 B
constexpr size_t NLTE_LDS_BYTES = 163840;    // LDS of a CU (gfx950)
// the largest matrices and the kernels' static shared memory (< 192 B) fit
static_assert(sizeof(float) * 201 * 201 + 192 <= NLTE_LDS_BYTES, "FP32 LDS capacity");
static_assert(sizeof(double) * 143 * 143 + 192 <= NLTE_LDS_BYTES, "FP64 LDS capacity");
constexpr int NLTE_SLOTS = 4;   // rows per lane in the one-wavefront steps (n <= 256)

__host__ __device__ __forceinline__ int lds_ld(int n) { return n | 1; }

// dynamic LDS of a batched launch whose largest LDS system has cap levels:
// the matrix, and at least 4 cap FP64 partial results for the
// equilibration passes (which run before the matrix is filled)
template <class T>
__host__ __device__ inline size_t lds_bytes_for(int cap)
{
  const size_t m = sizeof(T) * (size_t)cap * lds_ld(cap), st = sizeof(double) * 4 * (size_t)cap;
  return m > st ? m : st;
}

// order LDS accesses of one wavefront (its LDS operations execute in order;
// this keeps the compiler from moving them across)
__device__ __forceinline__ void wave_fence()
{
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <class V>
__device__ __forceinline__ V wave_max(V v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  return v;
}

// The pivot permutation lives in registers of the factoring wavefront:
// lane l keeps perm[l + 64 r] in pv[r].  Interchange of entries j and ip
// (uniform): two readlanes and a select per slot.
__device__ __forceinline__ void perm_swap(int* pv, int j, int ip, int lane)
{
  const int sj = j / WAVE, lj = j % WAVE, sp = ip / WAVE, lp = ip % WAVE;
  const int vj = __builtin_amdgcn_readlane(sj == 0 ? pv[0] : sj == 1 ? pv[1] : sj == 2 ? pv[2] : pv[3], lj);
  const int vp = __builtin_amdgcn_readlane(sp == 0 ? pv[0] : sp == 1 ? pv[1] : sp == 2 ? pv[2] : pv[3], lp);
#pragma unroll
  for (int r = 0; r < NLTE_SLOTS; ++r)
    pv[r] = (r == sj && lane == lj) ? vp : ((r == sp && lane == lp) ? vj : pv[r]);
}

// Panel j0 .. j0+jb of the partial-pivot LU, by one wavefront.  Per column
// j: the pivot (first largest |a_ij|, i >= j) comes from registers -- the
// candidates were formed by the previous column's update; one round trip
// reads rows j and ip (interchanged over all n columns), the pivot and the
// pivot row's panel segment; a second reads each lane's rows below j, which
// are scaled by the pivot and updated in the panel's remaining columns.
// Returns 1 when a pivot is exactly zero (that column is then left as is).
template <class T, int NB>
__device__ __forceinline__ int panel_factor(T* a, int ld, int n, int j0, int jb, int* pv, int lane)
{
  const int ns = (n + WAVE - 1) / WAVE;
  int singular = 0;
  PivotOf<T> m = {(T)-1, n};
#pragma unroll
  for (int r = 0; r < NLTE_SLOTS; ++r) {
    const int i = lane + WAVE * r;
    if (r < ns && i >= j0 && i < n) m = better_p(m, PivotOf<T>{(T)fabs(a[i * ld + j0]), i});
  }
  for (int t = 0; t < jb; ++t) {
    const int j = j0 + t;
    const int mc = jb - t - 1;
    const int ip = wave_argmax(m).i;
    const bool swap = ip != j && ip < n;
    const int pr = (ip < n ? ip : j) * ld;
    T x[NLTE_SLOTS], y[NLTE_SLOTS], u[NB];
    if (swap) {
#pragma unroll
      for (int r = 0; r < NLTE_SLOTS; ++r) {
        const int k = lane + WAVE * r;
        if (r < ns && k < n) { x[r] = a[j * ld + k]; y[r] = a[ip * ld + k]; }
      }
    }
    const T piv = a[pr + j];
#pragma unroll
    for (int c = 0; c < NB - 1; ++c)
      if (c < mc) u[c] = a[pr + j + 1 + c];
    if (swap) {
#pragma unroll
      for (int r = 0; r < NLTE_SLOTS; ++r) {
        const int k = lane + WAVE * r;
        if (r < ns && k < n) { a[j * ld + k] = y[r]; a[ip * ld + k] = x[r]; }
      }
      perm_swap(pv, j, ip, lane);
    }
    wave_fence();
    m = PivotOf<T>{(T)-1, n};
    if (piv == (T)0) {
      singular = 1;
      if (mc > 0) {
#pragma unroll
        for (int r = 0; r < NLTE_SLOTS; ++r) {
          const int i = lane + WAVE * r;
          if (r < ns && i > j && i < n) m = better_p(m, PivotOf<T>{(T)fabs(a[i * ld + j + 1]), i});
        }
      }
      continue;
    }
    const T inv = (T)1 / piv;
#pragma unroll
    for (int r = 0; r < NLTE_SLOTS; ++r) {
      const int i = lane + WAVE * r;
      if (r < ns && i > j && i < n) {
        T sg[NB];
        T l = a[i * ld + j];
#pragma unroll
        for (int c = 0; c < NB - 1; ++c)
          if (c < mc) sg[c] = a[i * ld + j + 1 + c];
        l *= inv;
        a[i * ld + j] = l;
#pragma unroll
        for (int c = 0; c < NB - 1; ++c)
          if (c < mc) {
            sg[c] -= l * u[c];
            a[i * ld + j + 1 + c] = sg[c];
          }
        if (mc > 0) m = better_p(m, PivotOf<T>{(T)fabs(sg[0]), i});
      }
    }
    wave_fence();
  }
  return singular;
}

// The same for a full panel (jb == NB) with every column's shape known at
// compile time (the column index is a template parameter): no per-element
// guards, one wavefront issues about 40 + 5 (NB - t) instructions per row
// slot and column.
template <class T, int NB>
struct LdsPanel {
  T* a;
  int ld, n, j0, lane;
  int rowi[NLTE_SLOTS];
  bool inb[NLTE_SLOTS];
  int singular;
};

template <class T, int NB, int t>
__device__ __forceinline__ void lds_panel_col(LdsPanel<T, NB>& S, PivotOf<T> m, int* pv)
{
  T* a = S.a;
  const int ld = S.ld, n = S.n, lane = S.lane, j = S.j0 + t;
  constexpr int mc = NB - 1 - t;
  const int ip = wave_argmax(m).i;
  const bool swap = ip != j && ip < n;
  const int pr = (ip < n ? ip : j) * ld;
  T x[NLTE_SLOTS], y[NLTE_SLOTS], u[NB];
  if (swap) {
#pragma unroll
    for (int r = 0; r < NLTE_SLOTS; ++r) {
      const int k = lane + WAVE * r;
      if (k < n) { x[r] = a[j * ld + k]; y[r] = a[ip * ld + k]; }
    }
  }
  const T piv = a[pr + j];
#pragma unroll
  for (int c = 0; c < mc; ++c) u[c] = a[pr + j + 1 + c];
  if (swap) {
#pragma unroll
    for (int r = 0; r < NLTE_SLOTS; ++r) {
      const int k = lane + WAVE * r;
      if (k < n) { a[j * ld + k] = y[r]; a[ip * ld + k] = x[r]; }
    }
    perm_swap(pv, j, ip, lane);
  }
  wave_fence();
  PivotOf<T> mn = {(T)-1, n};
  if (piv == (T)0) {
    S.singular = 1;
    if (mc > 0) {
#pragma unroll
      for (int r = 0; r < NLTE_SLOTS; ++r) {
        const int i = lane + WAVE * r;
        if (S.inb[r] && i > j) mn = better_p(mn, PivotOf<T>{(T)fabs(a[S.rowi[r] + j + 1]), i});
      }
    }
  } else {
    const T inv = (T)1 / piv;
#pragma unroll
    for (int r = 0; r < NLTE_SLOTS; ++r) {
      const int i = lane + WAVE * r;
      if (S.inb[r] && i > j) {
        T sg[NB];
        T l = a[S.rowi[r] + j];
#pragma unroll
        for (int c = 0; c < mc; ++c) sg[c] = a[S.rowi[r] + j + 1 + c];
        l *= inv;
        a[S.rowi[r] + j] = l;
#pragma unroll
        for (int c = 0; c < mc; ++c) {
          sg[c] -= l * u[c];
          a[S.rowi[r] + j + 1 + c] = sg[c];
        }
        if (mc > 0) mn = better_p(mn, PivotOf<T>{(T)fabs(sg[0]), i});
      }
    }
  }
  wave_fence();
  if constexpr (t + 1 < NB) lds_panel_col<T, NB, t + 1>(S, mn, pv);
}

template <class T, int NB>
__device__ __forceinline__ int panel_factor_full(T* a, int ld, int n, int j0, int* pv, int lane)
{
  LdsPanel<T, NB> S;
  S.a = a;
  S.ld = ld;
  S.n = n;
  S.j0 = j0;
  S.lane = lane;
  S.singular = 0;
  PivotOf<T> m = {(T)-1, n};
#pragma unroll
  for (int r = 0; r < NLTE_SLOTS; ++r) {
    const int i = lane + WAVE * r;
    S.inb[r] = i < n;
    S.rowi[r] = (i < n ? i : n - 1) * ld;
    if (S.inb[r] && i >= j0) m = better_p(m, PivotOf<T>{(T)fabs(a[S.rowi[r] + j0]), i});
  }
  lds_panel_col<T, NB, 0>(S, m, pv);
  return S.singular;
}

// U12 = L11^-1 A12 (unit lower triangular), one column per thread, the
// column held in registers
template <class T, int NB>
__device__ __forceinline__ void panel_row_solve(T* a, int ld, int n, int j0, int jb, int r0)
{
  for (int c = r0 + (int)threadIdx.x; c < n; c += blockDim.x) {
    T col[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t)
      if (t < jb) col[t] = a[(j0 + t) * ld + c];
#pragma unroll
    for (int i = 1; i < NB; ++i) {
      if (i < jb) {
        T acc = col[i];
#pragma unroll
        for (int k = 0; k < i; ++k) acc -= a[(j0 + i) * ld + j0 + k] * col[k];
        col[i] = acc;
      }
    }
#pragma unroll
    for (int t = 1; t < NB; ++t)
      if (t < jb) a[(j0 + t) * ld + c] = col[t];
  }
}

// trailing update A22 -= L21 U12 in FP32 on v_mfma_f32_32x32x2f32 tiles;
// rows / columns past n are read as zero and not written, so partial tiles
// run on MFMA too.  Lane l supplies A[l%32][k + l/32] and B[k + l/32][l%32];
// it holds accumulator column l%32, rows 8(i/4) + 4(l/32) + i%4, i < 16.
template <int NB>
__device__ __forceinline__ void trailing_update(float* a, int ld, int j0, int jb, int r0, int n)
{
  typedef float f32x16 __attribute__((ext_vector_type(16)));
  const int m = n - r0;
  const int tiles = (m + 31) / 32;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const int nwaves = blockDim.x / WAVE;
  const int lr = lane % 32, lk = lane / 32;
  for (int t = wave; t < tiles * tiles; t += nwaves) {
    const int ti = t / tiles, tj = t - ti * tiles;
    const int rb = r0 + ti * 32, cb = r0 + tj * 32;
    const int col = cb + lr, arow = rb + lr;
    const bool colok = col < n, arowok = arow < n;
    float av[NB / 2], bv[NB / 2];
#pragma unroll
    for (int k2 = 0; k2 < NB / 2; ++k2) {
      const int kk = 2 * k2 + lk;
      const bool kok = kk < jb;
      av[k2] = (kok && arowok) ? -a[arow * ld + j0 + kk] : 0.f;
      bv[k2] = (kok && colok) ? a[(j0 + kk) * ld + col] : 0.f;
    }
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int row = rb + (i / 4) * 8 + lk * 4 + (i % 4);
      acc[i] = (row < n && colok) ? a[row * ld + col] : 0.f;
    }
#pragma unroll
    for (int k2 = 0; k2 < NB / 2; ++k2)
      if (2 * k2 < jb) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[k2], bv[k2], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int row = rb + (i / 4) * 8 + lk * 4 + (i % 4);
      if (row < n && colok) a[row * ld + col] = acc[i];
    }
  }
}

// the same in FP64 on v_mfma_f64_16x16x4f64 tiles: lane l supplies
// A[l%16][k + l/16] and B[k + l/16][l%16]; it holds accumulator column l%16,
// rows l/16 + 4v, v < 4
template <int NB>
__device__ __forceinline__ void trailing_update(double* a, int ld, int j0, int jb, int r0, int n)
{
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  const int m = n - r0;
  const int tiles = (m + 15) / 16;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const int nwaves = blockDim.x / WAVE;
  const int lr = lane % 16, lk = lane / 16;
  for (int t = wave; t < tiles * tiles; t += nwaves) {
    const int ti = t / tiles, tj = t - ti * tiles;
    const int rb = r0 + ti * 16, cb = r0 + tj * 16;
    const int col = cb + lr, arow = rb + lr;
    const bool colok = col < n, arowok = arow < n;
    double av[NB / 4], bv[NB / 4];
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; ++k4) {
      const int kk = 4 * k4 + lk;
      const bool kok = kk < jb;
      av[k4] = (kok && arowok) ? -a[arow * ld + j0 + kk] : 0.0;
      bv[k4] = (kok && colok) ? a[(j0 + kk) * ld + col] : 0.0;
    }
    f64x4 acc;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = rb + lk + 4 * v;
      acc[v] = (row < n && colok) ? a[row * ld + col] : 0.0;
    }
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; ++k4)
      if (4 * k4 < jb) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[k4], bv[k4], acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = rb + lk + 4 * v;
      if (row < n && colok) a[row * ld + col] = acc[v];
    }
  }
}

// blocked right-looking partial-pivot LU in LDS (panel width NB), three
// workgroup barriers per panel; returns 1 when a pivot is exactly zero
template <class T, int NB>
__device__ __forceinline__ int lu_lds(T* a, int ld, int n, int* pv, int* sh_flag)
{
  const int tid = threadIdx.x;
  const int wave = tid / WAVE, lane = tid % WAVE;
#pragma unroll
  for (int r = 0; r < NLTE_SLOTS; ++r) pv[r] = lane + WAVE * r;
  if (tid == 0) *sh_flag = 0;
  __syncthreads();
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int jb = min(NB, n - j0);
    if (wave == 0) {
      const int sing = jb == NB ? panel_factor_full<T, NB>(a, ld, n, j0, pv, lane)
                                : panel_factor<T, NB>(a, ld, n, j0, jb, pv, lane);
      if (lane == 0 && sing) *sh_flag = 1;
    }
    __syncthreads();
    const int r0 = j0 + jb;
    if (r0 >= n) break;
    panel_row_solve<T, NB>(a, ld, n, j0, jb, r0);
    __syncthreads();
    trailing_update<NB>(a, ld, j0, jb, r0, n);
    __syncthreads();
  }
  const int sing = *sh_flag;
  __syncthreads();
  return sing;
}

// y = U^-1 L^-1 (P rhs) with the factors in LDS, by one wavefront, in the
// factors' precision: the right-hand side is scaled by a power of two into
// range first.  Lane l keeps entries l, l + 64, l + 128, l + 192 in
// registers.  The 64-step blocks are unrolled over the slot index, so each
// step is a readlane, one masked update of its own slot and a plain update
// of the later (earlier, going back) slots; the next column's LDS reads are
// issued before the current update.
template <class T>
__device__ __forceinline__ void tri_solve_wave(const T* a, int ld, int n, const int* pv, const double* rhs, double* y,
                                               int lane)
{
  n = __builtin_amdgcn_readfirstlane(n);     // uniform: scalar loop control
  ld = __builtin_amdgcn_readfirstlane(ld);
  const int ns = (n + WAVE - 1) / WAVE;
  double rm = 0;
  for (int i = lane; i < n; i += WAVE) rm = fmax(rm, fabs(rhs[i]));
  rm = wave_max(rm);
  int e = 0;
  if (rm > 0 && rm < HUGE_VAL) frexp(rm, &e);
  const double sc = ldexp(1.0, -e), usc = ldexp(1.0, e);
  constexpr bool recip = sizeof(T) == sizeof(float);
  T v[NLTE_SLOTS], c0[NLTE_SLOTS], c1[NLTE_SLOTS], dg[NLTE_SLOTS];
  int row[NLTE_SLOTS];
#pragma unroll
  for (int r = 0; r < NLTE_SLOTS; ++r) {
    const int i = lane + WAVE * r;
    const int ic = i < n ? i : n - 1;        // rows past n read row n-1; never used
    row[r] = ic * ld;
    v[r] = i < n ? (T)(rhs[pv[r]] * sc) : (T)0;
    dg[r] = a[row[r] + ic];
  }
  if (recip) {
    // FP32: the pivots' reciprocals, once (the refinement absorbs their rounding)
#pragma unroll
    for (int r = 0; r < NLTE_SLOTS; ++r) dg[r] = (T)1 / dg[r];
  }
  // one forward step k = k0 + src with column k in cur[], column kn loaded
  // into nxt[] for the next step
#define SGPU_FWD(src, cur, nxt)                                                   \
  {                                                                               \
    const int kn = min(k0 + (src) + 1, n - 1);                                    \
    _Pragma("unroll") for (int r = rk; r < NLTE_SLOTS; ++r) nxt[r] = a[row[r] + kn]; \
    const T xk = readlane(v[rk], (src));                                          \
    const T u = v[rk] - cur[rk] * xk;                                             \
    v[rk] = lane > (src) ? u : v[rk];                                             \
    _Pragma("unroll") for (int r = rk + 1; r < NLTE_SLOTS; ++r) v[r] -= cur[r] * xk; \
  }
#pragma unroll
  for (int rk = 0; rk < NLTE_SLOTS; ++rk) {   // forward, unit lower
    if (rk < ns) {
      const int k0 = rk * WAVE, cnt = min(WAVE, n - k0);
#pragma unroll
      for (int r = rk; r < NLTE_SLOTS; ++r) c0[r] = a[row[r] + k0];
      int src = 0;
      for (; src + 1 < cnt; src += 2) {
        SGPU_FWD(src, c0, c1)
        SGPU_FWD(src + 1, c1, c0)
      }
      if (src < cnt) SGPU_FWD(src, c0, c1)
    }
  }
#undef SGPU_FWD
  // one backward step k = k0 + src, column k in cur[], column kn into nxt[]
#define SGPU_BWD(src, cur, nxt)                                                   \
  {                                                                               \
    const int kn = max(k0 + (src) - 1, 0);                                        \
    _Pragma("unroll") for (int r = 0; r <= rk; ++r) nxt[r] = a[row[r] + kn];      \
    T xk;                                                                         \
    if (recip) xk = readlane(v[rk] * dg[rk], (src));                              \
    else xk = readlane(v[rk], (src)) / readlane(dg[rk], (src));                   \
    const T u = v[rk] - cur[rk] * xk;                                             \
    v[rk] = lane < (src) ? u : (lane == (src) ? xk : v[rk]);                      \
    _Pragma("unroll") for (int r = 0; r < rk; ++r) v[r] -= cur[r] * xk;           \
  }
#pragma unroll
  for (int rk = NLTE_SLOTS - 1; rk >= 0; --rk) {   // backward
    if (rk < ns) {
      const int k0 = rk * WAVE, cnt = min(WAVE, n - k0);
#pragma unroll
      for (int r = 0; r <= rk; ++r) c0[r] = a[row[r] + k0 + cnt - 1];
      int src = cnt - 1;
      for (; src >= 1; src -= 2) {
        SGPU_BWD(src, c0, c1)
        SGPU_BWD(src - 1, c1, c0)
      }
      if (src == 0) SGPU_BWD(0, c0, c1)
    }
  }
#undef SGPU_BWD
#pragma unroll
  for (int r = 0; r < NLTE_SLOTS; ++r) {
    const int i = lane + WAVE * r;
    if (i < n) y[i] = (double)v[r] * usc;
  }
}

// Global-memory passes over a system's rates: loads issued 16 (row sums: 32)
// at a time so that each thread keeps them in flight, and, where the order
// of the reduction does not matter, q = min(4, NT / n) threads per row of M
// whose partial results meet in part[(q - 1) n].
constexpr int NLTE_UNROLL = 16;

__device__ __forceinline__ int parts_per_row(int n)
{
  const int q = (int)blockDim.x / n;
  return q >= 4 ? 4 : (q >= 2 ? 2 : 1);
}

// diag_i = -sum_j R_ij, summed in j order (one thread per row)
__device__ void row_sums(int n, const double* R, double* diag)
{
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double* row = R + (size_t)i * n;
    double acc = 0.0;
    int j = 0;
    for (; j + 2 * NLTE_UNROLL <= n; j += 2 * NLTE_UNROLL) {
      double t[2 * NLTE_UNROLL];
#pragma unroll
      for (int u = 0; u < 2 * NLTE_UNROLL; ++u) t[u] = row[j + u];
#pragma unroll
      for (int u = 0; u < 2 * NLTE_UNROLL; ++u) acc += t[u];
    }
    for (; j < n; ++j) acc += row[j];
    diag[i] = -1 * acc;
  }
}

__device__ __forceinline__ double pow2_scale(double mx)
{
  int e;
  frexp(mx > 0 ? mx : 1.0, &e);
  return ldexp(1.0, -e);
}

// rs_i = 2^-e(max_j |M_ij|): M_ii = diag_i, M_ij = R_ji, M_{n-1,j} = lte_j
// (threads over the rows of M = columns of R: coalesced reads)
__device__ void row_scales(int n, const double* R, const double* lte, const double* diag, double* rs, double* part)
{
  const int q = parts_per_row(n), len = (n + q - 1) / q;
  for (int t = threadIdx.x; t < q * n; t += blockDim.x) {
    const int p = t / n, i = t - p * n;
    const int j1 = min(n, (p + 1) * len);
    int j = p * len;
    double mx = 0;
    if (i == n - 1) {
      for (; j < j1; ++j) mx = fmax(mx, fabs(lte[j]));
    } else {
      for (; j + NLTE_UNROLL <= j1; j += NLTE_UNROLL) {
        double v[NLTE_UNROLL];
#pragma unroll
        for (int u = 0; u < NLTE_UNROLL; ++u) v[u] = R[(size_t)(j + u) * n + i];
#pragma unroll
        for (int u = 0; u < NLTE_UNROLL; ++u) mx = fmax(mx, j + u == i ? 0.0 : fabs(v[u]));
      }
      for (; j < j1; ++j) mx = fmax(mx, j == i ? 0.0 : fabs(R[(size_t)j * n + i]));
    }
    part[p * n + i] = mx;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double mx = part[i];
    for (int p = 1; p < q; ++p) mx = fmax(mx, part[p * n + i]);
    if (i < n - 1) mx = fmax(mx, fabs(diag[i]));
    rs[i] = pow2_scale(mx);
  }
}

// cs_j = 2^-e(max_i |M_ij| rs_i) (threads over the columns of M = rows of R)
__device__ void col_scales(int n, const double* R, const double* lte, const double* diag, const double* rs,
                           double* cs, double* part)
{
  const int q = parts_per_row(n), len = (n - 1 + q - 1) / q;
  for (int t = threadIdx.x; t < q * n; t += blockDim.x) {
    const int p = t / n, j = t - p * n;
    const double* row = R + (size_t)j * n;
    const int i1 = min(n - 1, (p + 1) * len);
    int i = p * len;
    double mx = 0;
    for (; i + NLTE_UNROLL <= i1; i += NLTE_UNROLL) {
      double v[NLTE_UNROLL], w[NLTE_UNROLL];
#pragma unroll
      for (int u = 0; u < NLTE_UNROLL; ++u) { v[u] = row[i + u]; w[u] = rs[i + u]; }
#pragma unroll
      for (int u = 0; u < NLTE_UNROLL; ++u) mx = fmax(mx, i + u == j ? 0.0 : fabs(v[u]) * w[u]);
    }
    for (; i < i1; ++i) mx = fmax(mx, i == j ? 0.0 : fabs(row[i]) * rs[i]);
    part[p * n + j] = mx;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double mx = fmax(fabs(lte[j]) * rs[n - 1], j < n - 1 ? fabs(diag[j]) * rs[j] : 0.0);
    for (int p = 0; p < q; ++p) mx = fmax(mx, part[p * n + j]);
    cs[j] = pow2_scale(mx);
  }
}

// M (scaled by rs_i cs_j when given) into LDS at row stride ld, T precision,
// transposed from the rows of R: thread k takes the flat indices k, k + NT,
// ... of R (coalesced), 8 loads in flight
template <class T>
__device__ void fill_lds(int n, const double* R, const double* lte, const double* diag, const double* rs,
                         const double* cs, T* a, int ld)
{
  const int NT = blockDim.x, nn = n * n;
  int jj = threadIdx.x / n, ii = threadIdx.x - jj * n;   // R[jj][ii] = M[ii][jj]
  const int dj = NT / n, di = NT - dj * n;
  for (int base = threadIdx.x; base < nn; base += 8 * NT) {
    double t[8];
    int ri[8], rj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      ri[u] = ii;
      rj[u] = jj;
      t[u] = base + u * NT < nn ? R[(size_t)jj * n + ii] : 0.0;
      ii += di;
      jj += dj;
      if (ii >= n) { ii -= n; ++jj; }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (base + u * NT < nn) {
        const int i = ri[u], j = rj[u];
        double v = (i == n - 1) ? lte[j] : (i == j ? diag[i] : t[u]);
        if (rs) v = v * rs[i] * cs[j];
        a[i * ld + j] = (T)v;
      }
    }
  }
}

// r_i = (e_{n-1} - M x)_i rs_i (threads over the rows of M: coalesced reads
// of the columns of R; q threads per row, partial sums in part[(q - 1) n])
__device__ void residual(int n, const double* R, const double* lte, const double* diag, const double* rs,
                         const double* x, double* r, double* part)
{
  const int q = parts_per_row(n), len = (n + q - 1) / q;
  for (int t = threadIdx.x; t < q * n; t += blockDim.x) {
    const int p = t / n, i = t - p * n;
    const int j1 = min(n, (p + 1) * len);
    int j = p * len;
    double acc = 0;
    if (i == n - 1) {
      for (; j + NLTE_UNROLL <= j1; j += NLTE_UNROLL) {
        double v[NLTE_UNROLL], w[NLTE_UNROLL];
#pragma unroll
        for (int u = 0; u < NLTE_UNROLL; ++u) { v[u] = lte[j + u]; w[u] = x[j + u]; }
#pragma unroll
        for (int u = 0; u < NLTE_UNROLL; ++u) acc += v[u] * w[u];
      }
      for (; j < j1; ++j) acc += lte[j] * x[j];
    } else {
      for (; j + NLTE_UNROLL <= j1; j += NLTE_UNROLL) {
        double v[NLTE_UNROLL], w[NLTE_UNROLL];
#pragma unroll
        for (int u = 0; u < NLTE_UNROLL; ++u) { v[u] = R[(size_t)(j + u) * n + i]; w[u] = x[j + u]; }
#pragma unroll
        for (int u = 0; u < NLTE_UNROLL; ++u) acc += j + u == i ? 0.0 : v[u] * w[u];
      }
      for (; j < j1; ++j) if (j != i) acc += R[(size_t)j * n + i] * x[j];
    }
    if (p == 0) r[i] = acc;
    else part[(p - 1) * n + i] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double acc = r[i];
    for (int p = 1; p < q; ++p) acc += part[(p - 1) * n + i];
    acc = (i == n - 1) ? 1.0 - acc : -(diag[i] * x[i] + acc);
    r[i] = acc * rs[i];
  }
}

// M b = e_{n-1} by FP64 LU in LDS (row stride n | 1, the caller checked
// that n (n | 1) doubles fit): row sums, M transposed from the rows of R,
// the blocked LU, one wavefront's triangular solves into x.  diag and r are
// n-vectors of the workgroup's slot.
template <int NB>
__device__ __forceinline__ void solve_f64_lds(int n, const double* R, const double* lte, double* lds_d, double* diag,
                                              double* r, double* x, int* flag)
{
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const int ld = lds_ld(n);
  row_sums(n, R, diag);
  __syncthreads();
  fill_lds<double>(n, R, lte, diag, nullptr, nullptr, lds_d, ld);
  __syncthreads();
  int pv[NLTE_SLOTS];   // pivot permutation (wavefront 0's registers)
  lu_lds<double, NB>(lds_d, ld, n, pv, flag);
  if (wave == 0) {
    for (int i = lane; i < n; i += WAVE) r[i] = (i == n - 1) ? 1.0 : 0.0;
    wave_fence();
    tri_solve_wave(lds_d, ld, n, pv, r, x, lane);
  }
  __syncthreads();
}

// Scaled system: As = Dr M Dc with Dr, Dc powers of two; M b = e becomes
// As y = Dr e, b = Dc y.  The FP32 factors of As sit in LDS (row stride
// n | 1), the pivot permutation in wavefront 0's registers, the FP64
// vectors in the workgroup's slot.
template <int NB>
__global__ void __launch_bounds__(NLTE_MAX_THREADS) nlte_batched_mixed_kernel(BatchArgs A, size_t lds_bytes)
{
  extern __shared__ __align__(16) unsigned char lds_raw[];
  float* lds_a = (float*)lds_raw;
  __shared__ ArgMax red[NLTE_MAX_THREADS / WAVE];
  __shared__ int sh, sing_f64, flag;
  double* slot = A.slots + (size_t)blockIdx.x * A.slot_stride;
  const int tid = threadIdx.x;
  const int wave = tid / WAVE, lane = tid % WAVE;
  while (true) {
    const int s = next_system(A, &sh);
    if (s < 0) break;
    const int n = A.n[s];
    const int ld = lds_ld(n);
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
    double* M64 = (double*)(perm + 2 * ((n + 1) / 2));
    int it = 0;
    bool ok = false;
    if (sizeof(double) * (size_t)n * ld <= lds_bytes) {
      // small enough for an FP64 matrix in this launch's LDS: factored
      // directly in FP64, no refinement (iters 0)
      solve_f64_lds<NB>(n, R, lte, (double*)lds_raw, diag, r, x, &flag);
      write_pops(n, x, lte, bout, pout);
      if (A.iters && tid == 0) A.iters[s] = 0;
      __syncthreads();
      continue;
    }
    if (n <= LdsCap<float>::n) {
      // equilibration: rs_i = 2^-e(max_j |M_ij|), cs_j = 2^-e(max_i |M_ij| rs_i)
      double* part = (double*)lds_raw;   // free until the matrix is filled
      row_sums(n, R, diag);
      __syncthreads();
      row_scales(n, R, lte, diag, rs, part);
      __syncthreads();
      col_scales(n, R, lte, diag, rs, cs, part);
      __syncthreads();
      fill_lds<float>(n, R, lte, diag, rs, cs, lds_a, ld);
      __syncthreads();
      int pv[NLTE_SLOTS];   // pivot permutation (wavefront 0's registers)
      const int sing = lu_lds<float, NB>(lds_a, ld, n, pv, &flag);
      if (!sing) {
        // x = Dc As^-1 Dr e_{n-1}
        if (wave == 0) {
          for (int i = lane; i < n; i += WAVE) r[i] = (i == n - 1) ? rs[i] : 0.0;
          wave_fence();
          tri_solve_wave(lds_a, ld, n, pv, r, d, lane);
          wave_fence();
          for (int i = lane; i < n; i += WAVE) x[i] = d[i] * cs[i];
        }
        __syncthreads();
        double dprev = HUGE_VAL;   // last relative correction (wavefront 0)
        for (it = 1; it <= NLTE_MAX_ITERS; ++it) {
          // FP64 residual r = e_{n-1} - M x, scaled by Dr
          residual(n, R, lte, diag, rs, x, r, M64);
          __syncthreads();
          if (wave == 0) {
            tri_solve_wave(lds_a, ld, n, pv, r, d, lane);
            wave_fence();
            double dmax = 0, xmax = 0;
            for (int i = lane; i < n; i += WAVE) {
              const double dx = d[i] * cs[i];
              const double xi = x[i] + dx;
              x[i] = xi;
              dmax = fmax(dmax, fabs(dx));
              xmax = fmax(xmax, fabs(xi));
            }
            dmax = wave_max(dmax);
            xmax = wave_max(xmax);
            const double rel = dmax / xmax;
            // converged: below tol, or stalled (no longer halving) within 16 tol
            const bool conv = rel <= A.tol || (rel > 0.5 * dprev && rel <= 16 * A.tol);
            dprev = rel;
            // 1: converged, 2: NaN (give up on FP32), 0: continue
            if (lane == 0) flag = (!(dmax == dmax) || !(xmax == xmax)) ? 2 : (conv ? 1 : 0);
          }
          __syncthreads();
          const int f = flag;
          __syncthreads();
          if (f == 2) break;
          if (f == 1) { ok = true; break; }
        }
      }
    }
    if (ok) {
      write_pops(n, x, lte, bout, pout);
      if (A.iters && tid == 0) A.iters[s] = it;
    } else {
      // FP64 fallback in the workgroup's device slot
      if (tid == 0) sing_f64 = 0;
      if (blocked_lds_bytes(n) <= lds_bytes) {
        assemble_augmented(n, R, lte, M64);
        lu_solve_f64_blocked(n, M64, x, (double*)lds_raw, red, &sing_f64);
      } else {
        assemble<double>(n, R, lte, M64, n);
        lu_solve_f64(n, M64, perm, x, red, &sing_f64);
      }
      write_pops(n, x, lte, bout, pout);
      if (A.iters && tid == 0) A.iters[s] = -1;
    }
    __syncthreads();
  }
}

// FP64 LU in LDS (n <= 143) for the FP64 mode; larger systems in the
// workgroup's device slot
template <int NB>
__global__ void __launch_bounds__(NLTE_MAX_THREADS) nlte_batched_f64_lds_kernel(BatchArgs A, size_t lds_bytes)
{
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* lds_d = (double*)lds_raw;
  __shared__ ArgMax red[NLTE_MAX_THREADS / WAVE];
  __shared__ int sh, sing_f64, flag;
  double* slot = A.slots + (size_t)blockIdx.x * A.slot_stride;
  const int tid = threadIdx.x;
  while (true) {
    const int s = next_system(A, &sh);
    if (s < 0) break;
    const int n = A.n[s];
    const double* R = A.rates + A.rate_off[s];
    const double* lte = A.lte + A.vec_off[s];
    // slot layout: diag[n] | x[n] | r[n] | perm (ints) | M64 (larger systems)
    double* diag = slot;
    double* x = diag + n;
    double* r = x + n;
    int* perm = (int*)(r + n);
    double* M64 = (double*)(perm + 2 * ((n + 1) / 2));
    if (n <= LdsCap<double>::n) {
      solve_f64_lds<NB>(n, R, lte, lds_d, diag, r, x, &flag);
    } else {
      if (tid == 0) sing_f64 = 0;
      if (blocked_lds_bytes(n) <= lds_bytes) {
        assemble_augmented(n, R, lte, M64);
        lu_solve_f64_blocked(n, M64, x, lds_d, red, &sing_f64);
      } else {
        assemble<double>(n, R, lte, M64, n);
        lu_solve_f64(n, M64, perm, x, red, &sing_f64);
      }
    }
    write_pops(n, x, lte, A.b + A.vec_off[s], A.pop ? A.pop + A.vec_off[s] : nullptr);
    if (A.iters && tid == 0) A.iters[s] = 0;
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

// the batched kernel for a precision / panel width / largest system, and
// its dynamic LDS (0 for the device-memory FP64 kernel)
const void* batched_kernel(int precision, int panel, int max_n, size_t* lds)
{
  if (precision == SGPU_NLTE_FP64) {
    // systems up to 143 levels in LDS, larger ones in the workgroup's slot
    *lds = lds_bytes_for<double>(std::min(max_n, LdsCap<double>::n));
    return panel == 8 ? (const void*)nlte_batched_f64_lds_kernel<8> : (const void*)nlte_batched_f64_lds_kernel<16>;
  }
  // the FP32 matrix of the largest system, or the FP64 matrix of the
  // largest system that can be factored directly in FP64, if larger
  const int c64 = std::min(max_n, LdsCap<double>::n);
  *lds = std::max(lds_bytes_for<float>(std::min(max_n, LdsCap<float>::n)), lds_bytes_for<double>(c64));
  return panel == 8 ? (const void*)nlte_batched_mixed_kernel<8> : (const void*)nlte_batched_mixed_kernel<16>;
}

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
  // systems above cap take the wide path: above the FP64 LDS capacity in
  // the FP64 mode, above the FP32 one in the FP32 + refinement mode
  const int cap = precision == SGPU_NLTE_FP64 ? LdsCap<double>::n : LdsCap<float>::n;
  hipLaunchKernelGGL(size_range_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, nsys, cap, mm);
  SGPU_LAUNCH_CHECK();
  int32_t hmm[3] = {0, 0, 0};
  SGPU_HIP(hipMemcpyAsync(hmm, mm, sizeof(hmm), hipMemcpyDeviceToHost, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  SGPU_REQUIRE(hmm[0] >= 1 && hmm[1] <= 1024, "system sizes must lie in [1, 1024] (got %d..%d)", hmm[0], hmm[1]);
  const int max_n = hmm[1], n_wide = hmm[2];
  SGPU_REQUIRE(threads >= 2 * WAVE && threads <= NLTE_MAX_THREADS && threads % WAVE == 0,
               "workgroup size %d must be a multiple of %d in [%d, %d]", threads, WAVE, 2 * WAVE, NLTE_MAX_THREADS);
  SGPU_REQUIRE((panel == 8 || panel == 16) && wg_per_cu >= 1,
               "panel width %d (8 or 16) / workgroups per CU %d", panel, wg_per_cu);
  SGPU_REQUIRE(n_wide >= 0 && n_wide <= nsys, "wide system count %d outside [0, %d]", n_wide, nsys);
  hipLaunchKernelGGL(order_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, nsys, cap, order, counter);
  SGPU_LAUNCH_CHECK();
  timer_begin(ctx, "nlte");
  if (n_wide < nsys) {
    // the systems of at most cap levels from the device queue (its cursor
    // starts after the wide systems); beside the wide path, on side stream 0,
    // when there is one
    hipStream_t qs = ctx->stream;
    if (n_wide > 0) {
      st = fork_side(ctx, 0, &qs);
      if (st) return st;
    }
    const int narrow_max = std::min(max_n, cap);
    const unsigned nwg = (unsigned)std::min<int64_t>((int64_t)ctx->n_cu * wg_per_cu, nsys - n_wide);
    // slot: 7 vectors + perm + an FP64 matrix (augmented by its rhs) for the
    // FP64 fallback of the FP32 mode
    const size_t slot = (size_t)narrow_max * (narrow_max + 1) + 10 * (size_t)narrow_max + 64;
    double* slots = (double*)scratch(ctx, "nlte_slots", sizeof(double) * slot * nwg, &st);
    if (st) return st;
    BatchArgs A{nsys, n, rate_off, vec_off, rates, lte, b, pop, iters, order, counter, slots, slot, precision, tol};
    size_t lds = 0;
    const void* fn = batched_kernel(precision, panel, narrow_max, &lds);
    if (lds) SGPU_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    void* args[] = {&A, &lds};
    SGPU_HIP(hipLaunchKernel(fn, dim3(nwg), dim3(threads), args, lds, qs));
    SGPU_LAUNCH_CHECK();
  }
  if (n_wide > 0) {
    // the wide systems are order[0 .. n_wide): one augmented matrix each;
    // WIDE_GROUPS interleaved groups on as many streams, so that one group's
    // panels overlap the others' trailing updates
    const size_t stride = ((size_t)max_n * (max_n + 1) + 31) / 32 * 32;
    double* Mw = (double*)scratch(ctx, "nlte_wide", sizeof(double) * stride * n_wide, &st);
    if (st) return st;
    const size_t plds = blocked_lds_bytes(max_n);
    SGPU_REQUIRE(plds <= NLTE_LDS_BYTES && max_n <= 2 * WIDE_THREADS, "wide system of %d levels exceeds LDS",
                 max_n);
    SGPU_HIP(hipFuncSetAttribute((const void*)nlte_wide_panel_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)plds));
    SGPU_HIP(hipFuncSetAttribute((const void*)nlte_wide_back_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)plds));
    const int groups = std::min(n_wide, WIDE_GROUPS);
    hipStream_t gs[WIDE_GROUPS];
    gs[0] = ctx->stream;
    for (int g = 1; g < groups; ++g) {
      st = fork_side(ctx, g, &gs[g]);
      if (st) return st;
    }
    for (int g = 0; g < groups; ++g) {
      const unsigned ng = (unsigned)((n_wide - g + groups - 1) / groups);
      WideArgs W{n, order, rate_off, vec_off, rates, lte, b, pop, iters, Mw, stride,
                 precision == SGPU_NLTE_FP64 ? 0 : -1, groups, g};
      hipLaunchKernelGGL(nlte_wide_assemble_kernel, dim3(ng, (max_n + WIDE_TILE - 1) / WIDE_TILE),
                         dim3(WIDE_ASM_THREADS), 0, gs[g], W);
      SGPU_LAUNCH_CHECK();
    }
    // panels in pairs (A at j0, B at j0 + BLU_NB) with a delayed update: A
    // factored; A's update applied to B's columns only; B factored, forming
    // U12 of its rows with A's pending update; one rank-2 BLU_NB update of
    // the rest with both panels -- half the trailing-matrix traffic of an
    // update per panel, panels kept narrow
    for (int j0 = 0; j0 < max_n; j0 += 2 * BLU_NB) {
      // rows of the largest system below A, and below A and B
      const int mr1 = max_n - j0 - std::min(BLU_NB, max_n - j0);
      const int mr2 = max_n - j0 - std::min(2 * BLU_NB, max_n - j0);
      const unsigned tiles1 = (unsigned)((mr1 + WIDE_TILE - 1) / WIDE_TILE);
      const unsigned tiles2 = (unsigned)(((mr2 + WIDE_TILE - 1) / WIDE_TILE) * ((mr2 + WIDE_TILE) / WIDE_TILE));
      for (int g = 0; g < groups; ++g) {
        const unsigned ng = (unsigned)((n_wide - g + groups - 1) / groups);
        WideArgs W{n, order, rate_off, vec_off, rates, lte, b, pop, iters, Mw, stride,
                   precision == SGPU_NLTE_FP64 ? 0 : -1, groups, g};
        hipLaunchKernelGGL(nlte_wide_panel_kernel, dim3(ng), dim3(WIDE_PANEL_THREADS), plds, gs[g], W, j0, 0);
        SGPU_LAUNCH_CHECK();
        if (mr1 > 0) {
          hipLaunchKernelGGL(nlte_wide_update_kernel<BLU_NB>, dim3(tiles1, ng), dim3(WIDE_UPD_THREADS), 0, gs[g],
                             W, j0, j0 + 2 * BLU_NB - 1);
          SGPU_LAUNCH_CHECK();
          hipLaunchKernelGGL(nlte_wide_panel_kernel, dim3(ng), dim3(WIDE_PANEL_THREADS), plds, gs[g], W,
                             j0 + BLU_NB, 1);
          SGPU_LAUNCH_CHECK();
        }
        if (mr2 > 0) {
          hipLaunchKernelGGL(nlte_wide_update_kernel<2 * BLU_NB>, dim3(tiles2, ng), dim3(WIDE_UPD_THREADS), 0,
                             gs[g], W, j0, -1);
          SGPU_LAUNCH_CHECK();
        }
      }
    }
    for (int g = 0; g < groups; ++g) {
      const unsigned ng = (unsigned)((n_wide - g + groups - 1) / groups);
      WideArgs W{n, order, rate_off, vec_off, rates, lte, b, pop, iters, Mw, stride,
                 precision == SGPU_NLTE_FP64 ? 0 : -1, groups, g};
      hipLaunchKernelGGL(nlte_wide_back_kernel, dim3(ng), dim3(WIDE_THREADS), plds, gs[g], W);
      SGPU_LAUNCH_CHECK();
    }
    for (int g = 1; g < groups; ++g) {
      st = join_side(ctx, g);
      if (st) return st;
    }
    if (n_wide < nsys) {
      st = join_side(ctx, 0);
      if (st) return st;
    }
  }
  timer_end(ctx, "nlte");
  return SGPU_OK;
}

}  // namespace

extern "C" int sgpu_nlte_batched(sgpu_ctx* ctx, int32_t nsys, const int32_t* n, const int64_t* rate_off,
                                 const int64_t* vec_off, const double* rates, const double* lte, double* b,
                                 double* pop, int32_t precision, double tol, int32_t* iters)
{
  SGPU_REQUIRE(ctx, "null context");
  const nlte_tuning& t = ctx->nlte[precision == SGPU_NLTE_FP32_REFINE ? 1 : 0];
  return batched_impl(ctx, nsys, n, rate_off, vec_off, rates, lte, b, pop, precision, tol, iters,
                      t.valid ? t.threads : NLTE_THREADS, t.valid ? t.panel : 16, t.valid ? t.wg_per_cu : 1);
}

extern "C" int sgpu_nlte_set_tuning(sgpu_ctx* ctx, int32_t precision, int32_t threads, int32_t panel,
                                    int32_t wg_per_cu)
{
  SGPU_REQUIRE(ctx, "null context");
  SGPU_REQUIRE(precision == SGPU_NLTE_FP64 || precision == SGPU_NLTE_FP32_REFINE, "unknown precision %d",
               precision);
  SGPU_REQUIRE(threads >= 2 * WAVE && threads <= NLTE_MAX_THREADS && threads % WAVE == 0,
               "workgroup size %d must be a multiple of %d in [%d, %d]", threads, WAVE, 2 * WAVE, NLTE_MAX_THREADS);
  SGPU_REQUIRE((panel == 8 || panel == 16) && wg_per_cu >= 1 && wg_per_cu <= 16,
               "panel width %d (8 or 16) / workgroups per CU %d (1..16)", panel, wg_per_cu);
  nlte_tuning& t = ctx->nlte[precision == SGPU_NLTE_FP32_REFINE ? 1 : 0];
  t.valid = 1;
  t.threads = threads;
  t.panel = panel;
  t.wg_per_cu = wg_per_cu;
  return SGPU_OK;
}

extern "C" int sgpu_nlte_autotune(sgpu_ctx* ctx, int32_t nsys, const int32_t* n, const int64_t* rate_off,
                                  const int64_t* vec_off, const double* rates, const double* lte, double* b,
                                  double* pop, int32_t* iters, char* report, size_t report_len)
{
  SGPU_REQUIRE(ctx && n, "null argument");
  if (nsys <= 0) return SGPU_OK;
  // the largest system decides the LDS a workgroup takes, and so how many
  // workgroups can share a CU: candidates that cannot co-reside are dropped
  int max_n = 0;
  {
    int st = 0;
    int32_t* mm = (int32_t*)scratch(ctx, "nlte_mm", 64, &st);
    if (st) return st;
    hipLaunchKernelGGL(size_range_kernel, dim3(1), dim3(1024), 0, ctx->stream, n, nsys, 1024, mm);
    SGPU_LAUNCH_CHECK();
    int32_t hmm[2];
    SGPU_HIP(hipMemcpyAsync(hmm, mm, sizeof(hmm), hipMemcpyDeviceToHost, ctx->stream));
    SGPU_HIP(hipStreamSynchronize(ctx->stream));
    max_n = hmm[1];
  }
  const int threads_c[] = {128, 256, 512};
  const int panels[] = {8, 16};
  const int wgs[] = {1, 2, 4};
  const int precs[] = {SGPU_NLTE_FP64, SGPU_NLTE_FP32_REFINE};
  hipEvent_t e0, e1;
  SGPU_HIP(hipEventCreate(&e0));
  SGPU_HIP(hipEventCreate(&e1));
  std::string rep;
  for (int pi = 0; pi < 2; ++pi) {
    const int prec = precs[pi];
    float best = 1e30f;
    int bt = NLTE_THREADS, bp = 16, bw = 1;
    for (int th : threads_c)
      for (int p : panels)
        for (int w : wgs) {
          // only shapes whose workgroups can share a CU (LDS and registers)
          // (the queue kernel takes the systems up to the wide-path cap)
          size_t lds = 0;
          const int cap = prec == SGPU_NLTE_FP64 ? LdsCap<double>::n : LdsCap<float>::n;
          const void* fn = batched_kernel(prec, p, std::min(max_n, cap), &lds);
          if (lds) SGPU_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
          int resident = 0;
          SGPU_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, fn, th, lds));
          if (w > resident) continue;
          // one warm-up, then the timed run
          int st = batched_impl(ctx, nsys, n, rate_off, vec_off, rates, lte, b, pop, prec, 1e-12, iters, th, p, w);
          if (st) return st;
          SGPU_HIP(hipEventRecord(e0, ctx->stream));
          st = batched_impl(ctx, nsys, n, rate_off, vec_off, rates, lte, b, pop, prec, 1e-12, iters, th, p, w);
          if (st) return st;
          SGPU_HIP(hipEventRecord(e1, ctx->stream));
          SGPU_HIP(hipEventSynchronize(e1));
          float ms = 0;
          SGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
          char line[160];
          snprintf(line, sizeof(line), "%s threads=%d panel=%d wg_per_cu=%d: %.4f ms\n",
                   prec == SGPU_NLTE_FP64 ? "fp64" : "fp32_refine", th, p, w, ms);
          rep += line;
          if (ms < best) { best = ms; bt = th; bp = p; bw = w; }
        }
    nlte_tuning& t = ctx->nlte[pi];
    t.valid = 1;
    t.threads = bt;
    t.panel = bp;
    t.wg_per_cu = bw;
    char line[160];
    snprintf(line, sizeof(line), "chosen %s: threads=%d panel=%d wg_per_cu=%d (%.4f ms)\n",
             prec == SGPU_NLTE_FP64 ? "fp64" : "fp32_refine", bt, bp, bw, best);
    rep += line;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (report && report_len) {
    strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  return SGPU_OK;
}
