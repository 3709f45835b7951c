// Captured rounds: the calls a caller makes on a context between
// pa_gpu_graph_begin and pa_gpu_graph_end are recorded as one HIP graph and
// replayed with pa_gpu_graph_launch, one submission per round instead of a
// launch per kernel and no host round trip inside the round.  Only the
// asynchronous entry points (pa_gpu_seed_generate,
// pa_gpu_tracer_trace_union_round_async, pa_gpu_pool_init_async,
// pa_gpu_pool_run_async, the hand-over kernels) may be captured; a call that
// reads back fails the capture, which pa_gpu_graph_end then reports.  The
// library records no memset nodes: counters are cleared by small kernels.
#include "pa_internal.hpp"

struct pa_gpu_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

using namespace pa_amd;

extern "C" {

pa_status pa_gpu_graph_begin(pa_gpu_ctx* ctx) {
  if (!ctx) return PA_E_INVALID;
  if (ctx->timing) return fail(ctx, PA_E_STATE, "graph capture: turn kernel timing off first");
  if (ctx->capturing) return fail(ctx, PA_E_STATE, "graph capture: already capturing");
  PA_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  ctx->capture_failed = false;
  return PA_OK;
}

pa_status pa_gpu_graph_end(pa_gpu_ctx* ctx, pa_gpu_graph** out) {
  if (!ctx || !out) return PA_E_INVALID;
  *out = nullptr;
  if (!ctx->capturing) return fail(ctx, PA_E_STATE, "graph capture: not capturing");
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->capturing = false;
  if (e != hipSuccess || !g || ctx->capture_failed) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();  // the capture's own error is reported here, not by the next call
    return fail(ctx, PA_E_STATE, "graph capture failed: a captured call reads back or allocates");
  }
  auto* r = new pa_gpu_graph();
  r->graph = g;
  const hipError_t ie = hipGraphInstantiate(&r->exec, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    delete r;
    return check_hip(ctx, ie, "hipGraphInstantiate");
  }
  *out = r;
  return PA_OK;
}

pa_status pa_gpu_graph_launch(pa_gpu_ctx* ctx, const pa_gpu_graph* graph) {
  if (!ctx || !graph || !graph->exec) return PA_E_INVALID;
  PA_HIP(ctx, hipGraphLaunch(graph->exec, ctx->stream));
  return PA_OK;
}

pa_status pa_gpu_graph_destroy(pa_gpu_graph* graph) {
  if (!graph) return PA_E_INVALID;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
  return PA_OK;
}

}  // extern "C"
