// Context, memory, timing and error plumbing of the C ABI.
#include <cstring>

#include "pa_internal.hpp"

namespace pa_amd {

pa_status fail(pa_gpu_ctx* ctx, pa_status s, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return s;
}

pa_status check_hip(pa_gpu_ctx* ctx, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(ctx, e == hipErrorOutOfMemory ? PA_E_NOMEM : PA_E_DEVICE, m);
}

pa_status capture_guard(pa_gpu_ctx* ctx) {
  if (!ctx->capturing) return PA_OK;
  ctx->capture_failed = true;
  return fail(ctx, PA_E_STATE, "this call reads back or copies host memory: not allowed while capturing a graph");
}

pa_status scratch(pa_gpu_ctx* ctx, size_t bytes, void** out) {
  if (bytes > ctx->scratch_bytes && ctx->capturing) {
    ctx->capture_failed = true;
    return fail(ctx, PA_E_STATE, "the context's workspace would grow while capturing: run the call once first");
  }
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->scratch) PA_HIP(ctx, hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes + bytes / 4;
    PA_HIP(ctx, hipMalloc(&ctx->scratch, want));
    ctx->scratch_bytes = want;
  }
  *out = ctx->scratch;
  return PA_OK;
}

pa_status round_join(pa_gpu_ctx* ctx) {
  for (auto& r : ctx->round_set)
    if (r.pending) {
      PA_HIP(ctx, hipStreamWaitEvent(ctx->stream, r.finished, 0));
      r.pending = false;
    }
  return PA_OK;
}

static hipEvent_t take_event(pa_gpu_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Fold finished event pairs into the per-kernel totals.
static void drain(pa_gpu_ctx* ctx, bool wait) {
  std::vector<pa_gpu_ctx::timed> keep;
  for (auto& t : ctx->pending) {
    if (!wait && hipEventQuery(t.b) != hipSuccess) {
      keep.push_back(t);
      continue;
    }
    float ms = 0.0f;
    // a pair whose events did not complete is not counted
    if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      auto& slot = ctx->kernel_ms[t.name];
      slot.first += double(ms);
      slot.second += 1;
    }
    ctx->event_pool.push_back(t.a);
    ctx->event_pool.push_back(t.b);
  }
  ctx->pending.swap(keep);
}

void launch_begin(pa_gpu_ctx* ctx, const char* name) {
  if (!ctx->timing) return;
  pa_gpu_ctx::timed t;
  t.a = take_event(ctx);
  t.b = take_event(ctx);
  t.name = name;
  if (!t.a || !t.b) return;
  if (hipEventRecord(t.a, ctx->stream) != hipSuccess) {
    ctx->event_pool.push_back(t.a);
    ctx->event_pool.push_back(t.b);
    return;
  }
  ctx->pending.push_back(t);
}

void launch_end(pa_gpu_ctx* ctx) {
  if (!ctx->timing || ctx->pending.empty()) return;
  if (hipEventRecord(ctx->pending.back().b, ctx->stream) != hipSuccess) {
    ctx->event_pool.push_back(ctx->pending.back().a);
    ctx->event_pool.push_back(ctx->pending.back().b);
    ctx->pending.pop_back();
    return;
  }
  if (ctx->pending.size() > 256) drain(ctx, true);
}

}  // namespace pa_amd

using namespace pa_amd;

extern "C" {

pa_status pa_gpu_ctx_create(int device, pa_gpu_ctx** out) {
  if (!out) return PA_E_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return PA_E_DEVICE;
  if (device < 0 || device >= count) return PA_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return PA_E_DEVICE;
  auto* ctx = new pa_gpu_ctx();
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return PA_E_DEVICE;
  }
  if (hipMalloc(&ctx->counters, 4096) != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PA_E_NOMEM;
  }
  if (hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(ctx->counters);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PA_E_NOMEM;
  }
  if (static_scratch_setup(ctx) != PA_OK) {
    (void)hipHostFree(ctx->pinned);
    (void)hipFree(ctx->counters);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PA_E_DEVICE;
  }
  *out = ctx;
  return PA_OK;
}

pa_status pa_gpu_ctx_destroy(pa_gpu_ctx* ctx) {
  if (!ctx) return PA_E_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->finish_stream) (void)hipStreamSynchronize(ctx->finish_stream);
  drain(ctx, true);
  for (auto& r : ctx->round_set) {
    if (r.traced) (void)hipEventDestroy(r.traced);
    if (r.finished) (void)hipEventDestroy(r.finished);
    if (r.ws) (void)hipFree(r.ws);
    if (r.counters) (void)hipFree(r.counters);
  }
  if (ctx->finish_stream) (void)hipStreamDestroy(ctx->finish_stream);
  // teardown: failures here have nowhere to go
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->counters) (void)hipFree(ctx->counters);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PA_OK;
}

const char* pa_gpu_last_error(const pa_gpu_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

pa_status pa_gpu_synchronize(pa_gpu_ctx* ctx) {
  if (!ctx) return PA_E_INVALID;
  PA_TRY(capture_guard(ctx));
  PA_TRY(round_join(ctx));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PA_OK;
}

pa_status pa_gpu_device_info(pa_gpu_ctx* ctx, char* name, size_t name_len, int* compute_units, size_t* free_bytes,
                             size_t* total_bytes) {
  if (!ctx) return PA_E_INVALID;
  hipDeviceProp_t prop;
  PA_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (compute_units) *compute_units = prop.multiProcessorCount;
  size_t f = 0, t = 0;
  PA_HIP(ctx, hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return PA_OK;
}

pa_status pa_gpu_malloc(pa_gpu_ctx* ctx, size_t bytes, void** d_out) {
  if (!ctx || !d_out) return PA_E_INVALID;
  *d_out = nullptr;
  if (bytes == 0) return PA_OK;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMalloc(d_out, bytes));
  return PA_OK;
}

pa_status pa_gpu_free(pa_gpu_ctx* ctx, void* d_ptr) {
  if (!ctx) return PA_E_INVALID;
  if (!d_ptr) return PA_OK;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  PA_HIP(ctx, hipFree(d_ptr));
  return PA_OK;
}

pa_status pa_gpu_memcpy_h2d(pa_gpu_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
  if (!ctx || (bytes && (!d_dst || !h_src))) return PA_E_INVALID;
  if (!bytes) return PA_OK;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PA_OK;
}

pa_status pa_gpu_memcpy_d2h(pa_gpu_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
  if (!ctx || (bytes && (!h_dst || !d_src))) return PA_E_INVALID;
  if (!bytes) return PA_OK;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PA_OK;
}

pa_status pa_gpu_memcpy_d2d(pa_gpu_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
  if (!ctx || (bytes && (!d_dst || !d_src))) return PA_E_INVALID;
  if (!bytes) return PA_OK;
  PA_HIP(ctx, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return PA_OK;
}

namespace {
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) dst[i] = src[i];
}
__global__ void k_copy_tail(const unsigned char* __restrict__ src, unsigned char* __restrict__ dst, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
}  // namespace

pa_status pa_gpu_copy(pa_gpu_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
  if (!ctx || (bytes && (!d_dst || !d_src))) return PA_E_INVALID;
  if (!bytes) return PA_OK;
  if ((reinterpret_cast<uintptr_t>(d_dst) | reinterpret_cast<uintptr_t>(d_src)) & 15)
    return fail(ctx, PA_E_INVALID, "pa_gpu_copy needs 16-byte aligned pointers");
  const int64_t n4 = int64_t(bytes / 16);
  const int64_t tail = int64_t(bytes % 16);
  launch_begin(ctx, "copy");
  if (n4) {
    int64_t g = (n4 + 255) / 256;
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(k_copy, dim3(unsigned(g)), dim3(256), 0, ctx->stream, static_cast<const float4*>(d_src),
                       static_cast<float4*>(d_dst), n4);
  }
  if (tail)
    hipLaunchKernelGGL(k_copy_tail, dim3(1), dim3(64), 0, ctx->stream,
                       static_cast<const unsigned char*>(d_src) + n4 * 16, static_cast<unsigned char*>(d_dst) + n4 * 16,
                       tail);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

namespace {
// Scattered line reads (calibration): read j (0 <= j < n_reads) takes
// chunk_u4 uint4 from the start of line (mix(j ^ seed) * n_lines) >> 32,
// lines of line_u4 uint4; each lane keeps 8 reads in flight.  The XOR of
// everything read goes to *out (one atomic per wave), so no read is dead.
__device__ __forceinline__ uint32_t mix32(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return uint32_t(z ^ (z >> 31));
}
__global__ __launch_bounds__(256) void k_gather_lines(const uint4* __restrict__ src, uint32_t n_lines, int line_u4,
                                                      int chunk_u4, int64_t n_reads, uint64_t seed,
                                                      unsigned* __restrict__ out) {
  uint32_t acc = 0;
  const int64_t stride = int64_t(gridDim.x) * 256;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n_reads; i += 8 * stride) {
    uint4 v[8][4];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t j = i + k * stride;
      const uint64_t line = (uint64_t(mix32(uint64_t(j) ^ seed)) * n_lines) >> 32;
      const uint4* at = src + line * uint64_t(line_u4);
#pragma unroll
      for (int c = 0; c < 4; ++c) v[k][c] = (j < n_reads && c < chunk_u4) ? at[c] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc ^= v[k][c].x ^ v[k][c].y ^ v[k][c].z ^ v[k][c].w;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc ^= __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) atomicXor(out, acc);
}
}  // namespace

pa_status pa_gpu_gather_lines(pa_gpu_ctx* ctx, const void* d_src, size_t bytes, int32_t line_bytes, int32_t read_bytes,
                              int64_t n_reads, uint64_t seed, uint32_t* d_xor) {
  if (!ctx || !d_src || !d_xor || n_reads < 0) return PA_E_INVALID;
  if (reinterpret_cast<uintptr_t>(d_src) & 15) return fail(ctx, PA_E_INVALID, "gather_lines: 16-byte aligned source");
  if (line_bytes < 16 || line_bytes % 16 || read_bytes < 16 || read_bytes % 16 || read_bytes > 64 ||
      read_bytes > line_bytes)
    return fail(ctx, PA_E_INVALID, "gather_lines: lines and reads in 16-byte units, reads of 16 to 64 bytes");
  const uint64_t n_lines = uint64_t(bytes) / uint64_t(line_bytes);
  if (n_lines < 1 || n_lines > UINT32_MAX) return fail(ctx, PA_E_INVALID, "gather_lines: 1 to 2^32-1 lines");
  if (!n_reads) return PA_OK;
  int cus = 0;
  PA_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  int64_t g = int64_t(cus) * 8;
  const int64_t need = (n_reads + 255) / 256;
  if (g > need) g = need;
  launch_begin(ctx, "gather_lines");
  hipLaunchKernelGGL(k_gather_lines, dim3(unsigned(g)), dim3(256), 0, ctx->stream, static_cast<const uint4*>(d_src),
                     uint32_t(n_lines), line_bytes / 16, read_bytes / 16, n_reads, seed, reinterpret_cast<unsigned*>(d_xor));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_timing_enable(pa_gpu_ctx* ctx, int enable) {
  if (!ctx) return PA_E_INVALID;
  ctx->timing = enable != 0;
  return PA_OK;
}

pa_status pa_gpu_timing_reset(pa_gpu_ctx* ctx) {
  if (!ctx) return PA_E_INVALID;
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  drain(ctx, true);
  ctx->kernel_ms.clear();
  return PA_OK;
}

pa_status pa_gpu_kernel_time(pa_gpu_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches) {
  if (!ctx || !kernel) return PA_E_INVALID;
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  drain(ctx, true);
  auto it = ctx->kernel_ms.find(kernel);
  if (total_ms) *total_ms = it == ctx->kernel_ms.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == ctx->kernel_ms.end() ? 0 : it->second.second;
  return PA_OK;
}

}  // extern "C"
