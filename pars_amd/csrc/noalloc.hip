// The fused round without allocation: pa_gpu_tracer_trace_union_round_noalloc
// runs the same kernels as pa_gpu_tracer_trace_union_round, with their
// workspace (per-seed index pairs, the finish's offsets, counts and scan
// partials, its group queue) in one fixed device array instead of the
// context's grow-on-demand scratch.  For callers that cannot allocate
// device memory while they run (code under stream capture, tight serving
// loops).  No reference counterpart.
//
// The array is a file-scope __device__ symbol: every device the library
// runs on holds its own copy until the process ends.  Host code never uses
// the symbol's name as an address (on the host it names a stand-in, not
// device memory): static_scratch_setup looks up the device address once per
// device, with that device current, at context creation, and the entry
// point passes only that address to the kernels.  Calls that overlap on one
// device (two contexts or streams at once) share the array and race.
#include <mutex>
#include <string>

#include "pa_internal.hpp"

namespace {

constexpr size_t kStaticScratchBytes = size_t(16) << 20;  // 16 MiB: rounds of up to ~590,000 seeds
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

std::mutex g_lock;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace pa_amd {

pa_status static_scratch_setup(pa_gpu_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, PA_E_DEVICE, "device index above 63");
  std::lock_guard<std::mutex> lock(g_lock);
  if (!g_scratch_ptr[ctx->device]) {
    void* ptr = nullptr;
    PA_HIP(ctx, hipSetDevice(ctx->device));
    // ::scratch, the array (pa_amd::scratch is the context's allocator)
    static_assert(sizeof(::scratch) == kStaticScratchBytes, "the fixed workspace array");
    PA_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch)));
    if (!ptr) return fail(ctx, PA_E_DEVICE, "no device address for the fixed workspace");
    g_scratch_ptr[ctx->device] = ptr;
  }
  ctx->static_scratch = g_scratch_ptr[ctx->device];
  return PA_OK;
}

}  // namespace pa_amd

using namespace pa_amd;

namespace {

size_t max_seeds() {
  size_t lo = 0, hi = size_t(1) << 31;  // largest n whose workspace fits
  while (lo + 1 < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    if (union_round_ws_bytes(int64_t(mid)) <= kStaticScratchBytes) lo = mid; else hi = mid;
  }
  return lo;
}

}  // namespace

extern "C" {

pa_status pa_gpu_tracer_trace_union_round_noalloc(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n,
                                                  size_t region_slots, float* d_slots, int32_t* d_vertex_counts,
                                                  pa_particle* d_continuation, size_t* n_continuation,
                                                  pa_color_mode mode, float parameter, float* d_vertices,
                                                  size_t vertex_capacity, size_t* n_vertices, int32_t* d_indices,
                                                  size_t index_capacity, size_t* n_indices, float* d_colors,
                                                  float* d_endpoints, int32_t* d_seed_steps, uint64_t* total_steps) {
  if (!t || !n_vertices || !n_indices) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (!ctx->static_scratch) return fail(ctx, PA_E_STATE, "fixed workspace not set up on this device");
  if (n > size_t(INT32_MAX) || union_round_ws_bytes(int64_t(n)) > kStaticScratchBytes) {
    *n_vertices = 0;
    *n_indices = 0;
    return fail(ctx, PA_E_CAPACITY,
                "round without allocation: at most " + std::to_string(max_seeds()) +
                    " seeds per call (fixed workspace of " + std::to_string(kStaticScratchBytes >> 20) + " MiB)");
  }
  return union_round(t, d_particles, n, region_slots, d_slots, d_vertex_counts, d_continuation, n_continuation, mode,
                     parameter, d_vertices, vertex_capacity, n_vertices, d_indices, index_capacity, n_indices,
                     d_colors, d_endpoints, d_seed_steps, total_steps, ctx->static_scratch, nullptr);
}

}  // extern "C"
