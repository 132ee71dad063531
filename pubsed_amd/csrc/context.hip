// context.hip -- sgpu_ctx lifetime, device memory helpers, error text and
// per-phase HIP event timing.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include "sgpu_internal.h"

namespace {
thread_local char g_err[1024] = "";
}

namespace sgpu {

void set_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  return SGPU_ERR_HIP;
}

void* scratch(sgpu_ctx* ctx, const char* name, size_t bytes, int* status)
{
  sgpu_scratch& s = ctx->scratch[name];
  if (s.bytes >= bytes && s.ptr) { *status = SGPU_OK; return s.ptr; }
  if (ctx->capturing) {
    // no allocation inside a captured graph: the same calls must run once
    // uncaptured first, at the same sizes
    set_error("scratch '%s' would grow to %zu bytes during graph capture; run the step once uncaptured first",
              name, bytes);
    *status = SGPU_ERR_ARG;
    return nullptr;
  }
  if (s.ptr) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { *status = hip_fail(e, "sync before scratch regrow", __FILE__, __LINE__); return nullptr; }
    (void)hipFree(s.ptr);
    s.ptr = nullptr; s.bytes = 0;
  }
  size_t b = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&s.ptr, b);
  if (e != hipSuccess) { *status = hip_fail(e, name, __FILE__, __LINE__); s.ptr = nullptr; return nullptr; }
  s.bytes = b;
  *status = SGPU_OK;
  return s.ptr;
}

void timer_begin(sgpu_ctx* ctx, const char* phase)
{
  if (!ctx->timing || ctx->capturing) return;
  sgpu_phase_timer& t = ctx->timers[phase];
  if (!t.start) { (void)hipEventCreate(&t.start); (void)hipEventCreate(&t.stop); }
  (void)hipEventRecord(t.start, ctx->stream);
}

void timer_end(sgpu_ctx* ctx, const char* phase)
{
  if (!ctx->timing || ctx->capturing) return;
  sgpu_phase_timer& t = ctx->timers[phase];
  (void)hipEventRecord(t.stop, ctx->stream);
  t.pending = true;
}

int fork_side(sgpu_ctx* ctx, int i, hipStream_t* side)
{
  if (!ctx->fork_ev) SGPU_HIP(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
  if (!ctx->side_stream[i]) {
    SGPU_HIP(hipStreamCreateWithFlags(&ctx->side_stream[i], hipStreamNonBlocking));
    SGPU_HIP(hipEventCreateWithFlags(&ctx->join_ev[i], hipEventDisableTiming));
  }
  SGPU_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
  SGPU_HIP(hipStreamWaitEvent(ctx->side_stream[i], ctx->fork_ev, 0));
  *side = ctx->side_stream[i];
  return SGPU_OK;
}

int join_side(sgpu_ctx* ctx, int i)
{
  SGPU_HIP(hipEventRecord(ctx->join_ev[i], ctx->side_stream[i]));
  SGPU_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev[i], 0));
  return SGPU_OK;
}

}  // namespace sgpu

extern "C" {

sgpu_ctx* sgpu_create(int device)
{
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { sgpu::hip_fail(e, "hipSetDevice", __FILE__, __LINE__); return nullptr; }
  sgpu_ctx* ctx = new sgpu_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
  e = hipStreamCreate(&ctx->own_stream);   // blocking: ordered with the null stream
  if (e != hipSuccess) { sgpu::hip_fail(e, "hipStreamCreate", __FILE__, __LINE__); delete ctx; return nullptr; }
  ctx->stream = ctx->own_stream;
  return ctx;
}

void sgpu_destroy(sgpu_ctx* ctx)
{
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (hipStream_t q : ctx->side_stream)
    if (q) (void)hipStreamSynchronize(q);
  for (auto& kv : ctx->scratch) if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  for (auto& kv : ctx->timers) {
    if (kv.second.start) (void)hipEventDestroy(kv.second.start);
    if (kv.second.stop) (void)hipEventDestroy(kv.second.stop);
  }
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  for (int i = 0; i < sgpu_ctx::N_SIDE; ++i) {
    if (ctx->join_ev[i]) (void)hipEventDestroy(ctx->join_ev[i]);
    if (ctx->side_stream[i]) (void)hipStreamDestroy(ctx->side_stream[i]);
  }
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
}

int sgpu_graph_begin(sgpu_ctx* ctx)
{
  SGPU_REQUIRE(ctx && !ctx->capturing, "null context, or a capture already open");
  SGPU_REQUIRE(ctx->stream != nullptr, "graph capture needs a created stream, not the null stream");
  SGPU_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = 1;
  return SGPU_OK;
}

int sgpu_graph_end(sgpu_ctx* ctx, int32_t launch)
{
  SGPU_REQUIRE(ctx && ctx->capturing, "no capture open");
  ctx->capturing = 0;
  hipGraph_t g = nullptr;
  SGPU_HIP(hipStreamEndCapture(ctx->stream, &g));
  size_t nodes = 0;
  (void)hipGraphGetNodes(g, nullptr, &nodes);
  ctx->graph_nodes = nodes;
  if (ctx->graph_exec) {
    // the same topology as the last step's graph: update its parameters
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(ctx->graph_exec, g, &err_node, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(ctx->graph_exec);
      ctx->graph_exec = nullptr;
    }
  }
  hipError_t e = hipSuccess;
  if (!ctx->graph_exec) e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
  if (e == hipSuccess && launch) e = hipGraphLaunch(ctx->graph_exec, ctx->stream);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return sgpu::hip_fail(e, "graph instantiate / launch", __FILE__, __LINE__);
  return SGPU_OK;
}

int sgpu_graph_launch(sgpu_ctx* ctx)
{
  SGPU_REQUIRE(ctx && ctx->graph_exec && !ctx->capturing, "no graph to launch");
  SGPU_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
  return SGPU_OK;
}

int64_t sgpu_graph_nodes(sgpu_ctx* ctx) { return ctx ? (int64_t)ctx->graph_nodes : -1; }

int sgpu_use_own_stream(sgpu_ctx* ctx)
{
  SGPU_REQUIRE(ctx, "null context");
  ctx->stream = ctx->own_stream;
  return SGPU_OK;
}

int sgpu_set_stream(sgpu_ctx* ctx, void* s)
{
  SGPU_REQUIRE(ctx, "null context");
  ctx->stream = (hipStream_t)s;   // NULL selects the HIP null (default) stream
  return SGPU_OK;
}

void* sgpu_get_stream(sgpu_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int sgpu_synchronize(sgpu_ctx* ctx)
{
  SGPU_REQUIRE(ctx, "null context");
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  return SGPU_OK;
}

int sgpu_last_error(char* buf, size_t n)
{
  if (!buf || n == 0) return SGPU_ERR_ARG;
  strncpy(buf, g_err, n - 1);
  buf[n - 1] = 0;
  return SGPU_OK;
}

int sgpu_alloc(sgpu_ctx* ctx, void** ptr, size_t bytes)
{
  SGPU_REQUIRE(ctx && ptr, "null argument");
  SGPU_HIP(hipMalloc(ptr, bytes < 1 ? 1 : bytes));
  return SGPU_OK;
}

int sgpu_free(sgpu_ctx* ctx, void* ptr)
{
  SGPU_REQUIRE(ctx, "null context");
  if (ptr) SGPU_HIP(hipFree(ptr));
  return SGPU_OK;
}

int sgpu_memcpy_h2d(sgpu_ctx* ctx, void* dst, const void* src, size_t bytes)
{
  SGPU_REQUIRE(ctx, "null context");
  SGPU_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  return SGPU_OK;
}

int sgpu_memcpy_d2h(sgpu_ctx* ctx, void* dst, const void* src, size_t bytes)
{
  SGPU_REQUIRE(ctx, "null context");
  SGPU_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SGPU_HIP(hipStreamSynchronize(ctx->stream));
  return SGPU_OK;
}

int sgpu_enable_timing(sgpu_ctx* ctx, int on)
{
  SGPU_REQUIRE(ctx, "null context");
  ctx->timing = on;
  return SGPU_OK;
}

double sgpu_last_kernel_ms(sgpu_ctx* ctx, const char* phase)
{
  if (!ctx || !phase) return -1;
  auto it = ctx->timers.find(phase);
  if (it == ctx->timers.end() || !it->second.pending) return -1;
  if (hipEventSynchronize(it->second.stop) != hipSuccess) return -1;
  float ms = -1;
  if (hipEventElapsedTime(&ms, it->second.start, it->second.stop) != hipSuccess) return -1;
  return ms;
}

}  // extern "C"
