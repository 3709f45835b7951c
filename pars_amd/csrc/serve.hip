// Device-resident serving loop: a persistent kernel that takes seed requests
// as the host publishes them and runs every particle to its end, with no
// host synchronisation per pass.  No reference counterpart: pars serves one
// request at a time (pars/source/service.cpp:20-69); every particle follows
// the reference round on its own field (particle_tracer.cpp:265-326), like
// the particle pool, so a request's endpoints do not depend on what else is
// being served.
//
// Memory:
//  - host (pinned, coherent): the particle ring the host fills, the
//    endpoint / step ring the device fills, one completion flag per request
//    slot, and the control words (published count, stop flag);
//  - device: the claim cursor and the per-request done counters.
// The host writes a request's particles into the ring, then publishes the
// new end with a release store.  A wave with idle lanes reads the published
// end (system scope, acquire) and claims what it can with a compare-and-swap
// on the cursor, never past the published end, so a wave can leave at any
// time without stranding work.  A finished particle stores its endpoint and
// step count to the host ring, releases them at system scope and counts on
// the device; the lane that completes a request raises its host flag.  Every
// wait is bounded: a wave with nothing to do sleeps between polls and leaves
// after idle_polls empty polls (or when the host asks it to stop); the host
// relaunches the kernel when work is published while none runs.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <vector>

#include "pa_device.hpp"

namespace pa_amd {
namespace {

constexpr int kServeBlock = 256;
constexpr int kServeFields = 16;

struct ServeConfig {
  TraceConfig k;
  FieldDesc f[kServeFields];
  int32_t n_fields;
};

template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kServeBlock) void k_serve(
    const ServeConfig cfg, const pa_particle* __restrict__ ring, float4* __restrict__ out_pos,
    int32_t* __restrict__ out_steps, int32_t* __restrict__ complete, unsigned long long* control,
    unsigned long long* __restrict__ take, int32_t* __restrict__ done, int32_t* req_size,
    unsigned long long* __restrict__ steps_out, unsigned long long capacity, int rule, int idle_polls) {
  const int lane = threadIdx.x & 63;
  bool have = false;
  unsigned long long v = 0;
  float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int it = 0, R = 0, st = 0, slot = 0;
  FieldDesc f = cfg.f[0];
  unsigned long long done_steps = 0;
  int empty = 0;       // consecutive polls with nothing to do
  int since_poll = 0;  // iterations since this wave last read the published end
  for (;;) {
    const uint64_t idle = __ballot(!have);
    const int n_idle = __popcll(idle);
    // poll when the wave is half idle, or every 32 iterations while busy
    if (n_idle && (n_idle >= 32 || ++since_poll >= 32 || n_idle == 64)) {
      since_poll = 0;
      unsigned long long base = 0, got = 0;
      if (lane == 0) {
        const unsigned long long pub = __hip_atomic_load(&control[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        unsigned long long t = __hip_atomic_load(take, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (t < pub) {
          const unsigned long long want = min(static_cast<unsigned long long>(n_idle), pub - t);
          if (__hip_atomic_compare_exchange_strong(take, &t, t + want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            base = t;
            got = want;
            break;
          }
        }
      }
      base = (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(int(base >> 32))) << 32) |
             uint32_t(__builtin_amdgcn_readfirstlane(int(base & 0xffffffffu)));
      got = uint32_t(__builtin_amdgcn_readfirstlane(int(got)));
      if (!have) {
        const unsigned long long r = __popcll(idle & ((uint64_t(1) << lane) - 1));
        if (r < got) {
          v = base + r;
          // the particle record, written by the host before it published
          int32_t* q = reinterpret_cast<int32_t*>(const_cast<pa_particle*>(ring) + v % capacity);
          float pos[4];
#pragma unroll
          for (int c = 0; c < 4; ++c)
            pos[c] = __int_as_float(__hip_atomic_load(q + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
          R = __hip_atomic_load(q + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          const int32_t fi = __hip_atomic_load(q + 5, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          slot = __hip_atomic_load(q + 10, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);  // reserved_[0]
          p = make_float4(pos[0], pos[1], pos[2], pos[3]);
          f = cfg.f[fi >= 0 && fi < cfg.n_fields ? fi : 0];
          it = 1;
          st = 0;
          have = true;
        }
      }
      if (__ballot(have) == 0) {
        const unsigned long long stop = __hip_atomic_load(&control[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (stop || ++empty >= idle_polls) break;
#pragma unroll
        for (int z = 0; z < 8; ++z) __builtin_amdgcn_s_sleep(127);  // ~30 us between polls of an idle wave
        continue;
      }
      empty = 0;
    }
    if (have) {
      bool fin = it >= R - 1;  // the reference breaks at it == R-1 (R < 2: the seed alone)
      if (!fin) {
        float4 next;
        if (advance_rt<S, MODE, INTERP>(cfg.k, f, p, next, rule) != 0) {
          fin = true;
        } else {
          p = next;
          ++it;
          ++st;
          ++done_steps;
        }
      }
      if (fin) {
        const unsigned long long e = v % capacity;
        float* o = reinterpret_cast<float*>(out_pos + e);
        __hip_atomic_store(o + 0, p.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(o + 1, p.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(o + 2, p.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(o + 3, p.w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(out_steps + e, st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        // the outputs before the count; the count before the flag
        const int32_t before = __hip_atomic_fetch_add(done + slot, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        // the size was copied in by the host after an earlier request may have used the slot: read it
        // past the caches
        if (before + 1 == __hip_atomic_load(req_size + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) {
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
          __hip_atomic_store(complete + slot, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        have = false;
      }
    }
  }
  done_steps = wave_sum_u64(done_steps);
  if (lane == 0 && done_steps) atomicAdd(steps_out, done_steps);
}

__global__ void k_serve_zero(unsigned long long* c, int n) {
  if (int(threadIdx.x) < n) c[threadIdx.x] = 0ull;
}
__global__ void k_serve_zero_i32(int32_t* c, int n) {
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) c[i] = 0;
}

}  // namespace
}  // namespace pa_amd

using namespace pa_amd;

struct pa_gpu_server {
  pa_gpu_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;  // the persistent kernel's own stream (default priority)
  ServeConfig cfg;
  int rule = PA_STOP_ISZERO;
  int mode = 0;
  size_t capacity = 0;   // ring entries
  int32_t slots = 0;     // request slots
  // host, pinned and coherent
  pa_particle* ring = nullptr;
  float4* out_pos = nullptr;
  int32_t* out_steps = nullptr;
  int32_t* complete = nullptr;
  unsigned long long* control = nullptr;  // [0] published, [1] stop
  // device
  unsigned long long* d_take = nullptr;   // [0] claim cursor, [1] steps
  int32_t* d_done = nullptr;
  int32_t* d_size = nullptr;
  // host bookkeeping
  unsigned long long published = 0, reclaimed = 0;
  struct Req {
    unsigned long long start = 0;
    size_t n = 0;
    bool used = false, collected = true;
  };
  std::vector<Req> req;
  std::vector<unsigned long long> order;  // start of each live request, oldest first
  int waves_per_cu = 8, idle_polls = 64;
  uint64_t launches = 0;
  bool launched = false;
};

namespace {

pa_status serve_launch(pa_gpu_server* s) {
  pa_gpu_ctx* ctx = s->ctx;
  int cus = 0;
  PA_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int64_t blocks = std::max<int64_t>(1, (int64_t(cus) * s->waves_per_cu + 3) / 4);
  dispatch(s->cfg.f[0].storage, s->mode, s->cfg.k.interp, [&](auto S, auto M, auto I) {
    hipLaunchKernelGGL((k_serve<decltype(S)::value, decltype(M)::value, decltype(I)::value>), dim3(unsigned(blocks)),
                       dim3(kServeBlock), 0, s->stream, s->cfg, s->ring, s->out_pos, s->out_steps, s->complete,
                       s->control, s->d_take, s->d_done, s->d_size, s->d_take + 1,
                       static_cast<unsigned long long>(s->capacity), s->rule, s->idle_polls);
  }, s->cfg.f[0].layout);
  PA_HIP(ctx, hipGetLastError());
  s->launched = true;
  ++s->launches;
  return PA_OK;
}

// Launch the kernel again if work is published and no instance runs.
pa_status serve_kick(pa_gpu_server* s) {
  if (s->launched) {
    const hipError_t q = hipStreamQuery(s->stream);
    if (q == hipErrorNotReady) return PA_OK;
    if (q != hipSuccess) return check_hip(s->ctx, q, "serving kernel");
    s->launched = false;
  }
  bool open = false;
  for (size_t k = 0; k < s->req.size(); ++k)
    if (s->req[k].used && !s->req[k].collected && !__atomic_load_n(&s->complete[k], __ATOMIC_ACQUIRE)) open = true;
  if (!open) return PA_OK;
  return serve_launch(s);
}

}  // namespace

extern "C" {

pa_status pa_gpu_serve_create(pa_gpu_tracer* t, const pa_gpu_field* const* fields, int32_t n_fields,
                              size_t particle_capacity, int32_t request_slots, int32_t waves_per_cu,
                              int32_t idle_polls, pa_gpu_server** out) {
  if (!t || !fields || !out) return PA_E_INVALID;
  *out = nullptr;
  pa_gpu_ctx* ctx = t->ctx;
  if (n_fields < 1 || n_fields > kServeFields) return fail(ctx, PA_E_INVALID, "serve: 1 to 16 fields");
  if (particle_capacity < 64 || request_slots < 1 || request_slots > (1 << 20))
    return fail(ctx, PA_E_INVALID, "serve: at least 64 ring entries and 1 to 2^20 request slots");
  PA_TRY(check_integrator_of(ctx, t->integrator));
  PA_TRY(capture_guard(ctx));
  auto* s = new pa_gpu_server();
  s->ctx = ctx;
  std::memset(&s->cfg, 0, sizeof s->cfg);
  fill_step_constants(t, &s->cfg.k);
  int code = -1;
  for (int k = 0; k < n_fields; ++k) {
    if (!fields[k]) {
      delete s;
      return fail(ctx, PA_E_INVALID, "serve: null field");
    }
    const int c = fields[k]->desc.storage + 3 * fields[k]->desc.layout;
    if (code >= 0 && c != code) {
      delete s;
      return fail(ctx, PA_E_UNSUPPORTED, "serve: all fields must share one storage type and layout");
    }
    code = c;
    s->cfg.f[k] = fields[k]->desc;
  }
  for (int k = n_fields; k < kServeFields; ++k) s->cfg.f[k] = s->cfg.f[0];
  s->cfg.n_fields = n_fields;
  s->rule = t->stop_rule;
  s->mode = mode_of_integrator(t->integrator);
  s->capacity = particle_capacity;
  s->slots = request_slots;
  s->waves_per_cu = waves_per_cu > 0 ? waves_per_cu : 8;
  s->idle_polls = idle_polls > 0 ? idle_polls : 64;
  s->req.resize(size_t(request_slots));
  const unsigned flags = hipHostMallocCoherent | hipHostMallocMapped;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->ring), particle_capacity * sizeof(pa_particle), flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->out_pos), particle_capacity * 16, flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->out_steps), particle_capacity * 4, flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->complete), size_t(request_slots) * 4, flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->control), 64, flags);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_take), 64);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_done), size_t(request_slots) * 4);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_size), size_t(request_slots) * 4);
  if (e != hipSuccess) {
    pa_gpu_serve_destroy(s);
    return check_hip(ctx, e, "serve: allocation");
  }
  std::memset(s->complete, 0, size_t(request_slots) * 4);
  std::memset(s->control, 0, 64);
  hipLaunchKernelGGL(k_serve_zero, dim3(1), dim3(64), 0, s->stream, s->d_take, 8);
  hipLaunchKernelGGL(k_serve_zero_i32, dim3(unsigned((request_slots + 255) / 256)), dim3(256), 0, s->stream, s->d_done,
                     request_slots);
  e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    pa_gpu_serve_destroy(s);
    return check_hip(ctx, e, "serve: set-up");
  }
  *out = s;
  return PA_OK;
}

pa_status pa_gpu_serve_submit(pa_gpu_server* s, const pa_particle* h_particles, size_t n, int32_t* request) {
  if (!s || !request || (n && !h_particles)) return PA_E_INVALID;
  pa_gpu_ctx* ctx = s->ctx;
  *request = -1;
  if (n == 0) return fail(ctx, PA_E_INVALID, "serve: an empty request");
  if (n > s->capacity) return fail(ctx, PA_E_CAPACITY, "serve: request larger than the ring");
  if (s->published + n - s->reclaimed > s->capacity) return fail(ctx, PA_E_CAPACITY, "serve: ring full");
  int32_t slot = -1;
  for (int32_t k = 0; k < s->slots; ++k)
    if (!s->req[size_t(k)].used) {
      slot = k;
      break;
    }
  if (slot < 0) return fail(ctx, PA_E_CAPACITY, "serve: no free request slot");
  for (size_t i = 0; i < n; ++i)
    if (h_particles[i].vector_field_index < 0 || h_particles[i].vector_field_index >= s->cfg.n_fields)
      return fail(ctx, PA_E_INVALID, "serve: a particle names a field index outside [0, n_fields)");
  // the size is on the device before the particles are published
  const int32_t size32 = int32_t(n);
  PA_HIP(ctx, hipMemcpyAsync(s->d_size + slot, &size32, 4, hipMemcpyHostToDevice, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  auto& r = s->req[size_t(slot)];
  r.start = s->published;
  r.n = n;
  r.used = true;
  r.collected = false;
  __atomic_store_n(&s->complete[slot], 0, __ATOMIC_RELAXED);
  for (size_t i = 0; i < n; ++i) {
    pa_particle q = h_particles[i];
    q.reserved_[0] = slot;
    s->ring[(r.start + i) % s->capacity] = q;
  }
  s->order.push_back(r.start);
  s->published += n;
  __atomic_store_n(&s->control[0], s->published, __ATOMIC_RELEASE);
  *request = slot;
  return serve_kick(s);
}

pa_status pa_gpu_serve_poll(pa_gpu_server* s, int32_t request, int32_t* done) {
  if (!s || !done || request < 0 || request >= s->slots || !s->req[size_t(request)].used) return PA_E_INVALID;
  *done = __atomic_load_n(&s->complete[request], __ATOMIC_ACQUIRE);
  if (!*done) return serve_kick(s);
  return PA_OK;
}

pa_status pa_gpu_serve_collect(pa_gpu_server* s, int32_t request, float* h_endpoints, int32_t* h_steps,
                               uint64_t* steps_sum) {
  if (!s || request < 0 || request >= s->slots) return PA_E_INVALID;
  auto& r = s->req[size_t(request)];
  if (!r.used || r.collected) return fail(s->ctx, PA_E_INVALID, "serve: no such request");
  if (!__atomic_load_n(&s->complete[request], __ATOMIC_ACQUIRE)) return fail(s->ctx, PA_E_STATE, "serve: not done");
  uint64_t sum = 0;
  for (size_t i = 0; i < r.n; ++i) {
    const size_t e = (r.start + i) % s->capacity;
    if (h_endpoints) std::memcpy(h_endpoints + 4 * i, &s->out_pos[e], 16);
    if (h_steps) h_steps[i] = s->out_steps[e];
    sum += uint64_t(s->out_steps[e]);
  }
  if (steps_sum) *steps_sum = sum;
  r.collected = true;
  r.used = false;
  // the done counter of the slot starts again from zero for its next request
  const int32_t zero = 0;
  PA_HIP(s->ctx, hipMemcpyAsync(s->d_done + request, &zero, 4, hipMemcpyHostToDevice, s->ctx->stream));
  PA_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  // ring space comes back in publication order
  for (size_t k = 0; k < s->order.size(); ++k)
    if (s->order[k] == r.start) {
      s->order.erase(s->order.begin() + long(k));
      break;
    }
  s->reclaimed = s->order.empty() ? s->published : s->order.front();
  return PA_OK;
}

pa_status pa_gpu_serve_stop(pa_gpu_server* s, uint64_t* total_steps, uint64_t* launches) {
  if (!s) return PA_E_INVALID;
  __atomic_store_n(&s->control[1], 1ull, __ATOMIC_RELEASE);
  PA_HIP(s->ctx, hipStreamSynchronize(s->stream));
  s->launched = false;
  __atomic_store_n(&s->control[1], 0ull, __ATOMIC_RELEASE);
  unsigned long long st = 0;
  PA_HIP(s->ctx, hipMemcpy(&st, s->d_take + 1, 8, hipMemcpyDeviceToHost));
  if (total_steps) *total_steps = st;
  if (launches) *launches = s->launches;
  return PA_OK;
}

pa_status pa_gpu_serve_destroy(pa_gpu_server* s) {
  if (!s) return PA_E_INVALID;
  if (s->stream) {
    if (s->control) __atomic_store_n(&s->control[1], 1ull, __ATOMIC_RELEASE);
    (void)hipStreamSynchronize(s->stream);
  }
  if (s->ring) (void)hipHostFree(s->ring);
  if (s->out_pos) (void)hipHostFree(s->out_pos);
  if (s->out_steps) (void)hipHostFree(s->out_steps);
  if (s->complete) (void)hipHostFree(s->complete);
  if (s->control) (void)hipHostFree(s->control);
  if (s->d_take) (void)hipFree(s->d_take);
  if (s->d_done) (void)hipFree(s->d_done);
  if (s->d_size) (void)hipFree(s->d_size);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return PA_OK;
}

}  // extern "C"
