// sgpu_internal.h -- host-side context and error plumbing shared by the
// library's translation units (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string>
#include <map>
#include "../../include/sedona_gpu.h"

struct sgpu_scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct sgpu_phase_timer {
  hipEvent_t start = nullptr, stop = nullptr;
  bool pending = false;
};

struct nlte_tuning {
  int valid = 0;
  int threads = 256;   // workgroup size of the batched solve
  int panel = 16;      // panel width of the blocked FP32 LU
  int wg_per_cu = 1;   // persistent workgroups per CU
};

struct sgpu_ctx {
  int device = 0;
  int n_cu = 256;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int timing = 0;
  std::map<std::string, sgpu_phase_timer> timers;
  std::map<std::string, sgpu_scratch> scratch;
  nlte_tuning nlte[2];   // per sgpu_nlte_batched precision (FP64, FP32_REFINE)
  // streams for work that runs beside the main stream's (default
  // priority), with the events that fork them from and join them to the
  // main stream; created on first use
  static constexpr int N_SIDE = 4;
  hipStream_t side_stream[N_SIDE] = {};
  hipEvent_t fork_ev = nullptr, join_ev[N_SIDE] = {};
  // HIP graph of a captured step (sgpu_graph_begin / sgpu_graph_end)
  int capturing = 0;
  hipGraphExec_t graph_exec = nullptr;
  size_t graph_nodes = 0;
};

namespace sgpu {
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);
// grow-only named device scratch owned by the context
void* scratch(sgpu_ctx* ctx, const char* name, size_t bytes, int* status);
// side stream i (< sgpu_ctx::N_SIDE) made to wait for the work queued so far on ctx->stream
int fork_side(sgpu_ctx* ctx, int i, hipStream_t* side);
// ctx->stream made to wait for the work queued so far on side stream i
int join_side(sgpu_ctx* ctx, int i);
void timer_begin(sgpu_ctx* ctx, const char* phase);
void timer_end(sgpu_ctx* ctx, const char* phase);
}  // namespace sgpu

#define SGPU_HIP(call)                                                         \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) return sgpu::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define SGPU_LAUNCH_CHECK() SGPU_HIP(hipGetLastError())

#define SGPU_REQUIRE(cond, ...)                                                \
  do {                                                                         \
    if (!(cond)) { sgpu::set_error(__VA_ARGS__); return SGPU_ERR_ARG; }        \
  } while (0)
