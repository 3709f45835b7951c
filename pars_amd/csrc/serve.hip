// Device-resident serving loop: a persistent kernel that takes seed requests
// as the host publishes them and runs every particle to its end, with no
// host synchronisation per pass.  No reference counterpart: pars serves one
// request at a time (pars/source/service.cpp:20-69); every particle follows
// the reference round on its own field (particle_tracer.cpp:265-326), like
// the particle pool, so a request's endpoints do not depend on what else is
// being served.
//
// Host memory over the bus is slow for the GPU to touch one word at a time
// (uncached reads and writes of host memory are served a few tens of
// millions per second for the whole chip, and queue behind each other), so
// one wave of the grid, the porter, is the only one that polls host words,
// and three exporter waves the only ones that write results there:
//  - pinned host memory (coherent): the particle ring the host fills (each
//    record carries its request slot and the count the slot's done counter
//    reaches when the request ends), each slot's ring range, the control
//    words (published count, stop flag, requests submitted), and what the
//    exporters write back: the endpoint / step ring and one completion flag
//    per slot;
//  - device memory: the claim cursor, the mirrors of the published count,
//    the submitted count and the stop flag, the per-slot done counters (which only grow: a slot's
//    next request ends at a higher count), the device endpoint / step ring,
//    the queue of ended requests and the step total.
// While the kernel runs the host makes no GPU call (a copy or kernel on
// another stream may share the kernel's hardware queue and wait for it): it
// writes a request's records and range into host memory and then publishes
// the new end.  The porter copies the host's published end, submitted
// count and stop flag into device memory every ~2 us.  A worker wave with
// idle lanes reads the mirror and, when it holds no claim and the cursor is
// behind the end, claims as many indices as it has idle lanes with one
// atomic add (a race may take indices past the end, which the wave keeps
// and hands out once they are published); the whole wave reads the records
// it hands out together (16-byte pieces of host memory, past every cache)
// into LDS.  A finished particle stores its endpoint and step count into
// device memory (past the non-coherent caches); the wave waits for the
// stores and adds one count per request among its finished lanes; the add
// that reaches the request's end count queues the request.  Exporter waves
// take queued requests, copy their endpoints and steps to host memory with
// whole-wave stores, and raise the request's flag after a system-scope
// release.  The kernel stays resident, idle workers polling the device
// mirror every ~16 us, until the host asks it to stop: workers then leave
// once nothing published is left to them, the porter and the exporters
// once every submitted request is exported, and the host restarts the
// cursor at the published end before the next launch.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "pa_device.hpp"

namespace pa_amd {
namespace {

constexpr int kServeBlock = 256;
constexpr int kServeFields = 16;
constexpr size_t kMaxRing = size_t(1) << 24;  // ring byte offsets stay below 2^31 (48 B records)
static_assert(kMaxRing * sizeof(pa_particle) < (size_t(1) << 31), "ring offsets of the buffer loads");

struct ServeConfig {
  TraceConfig k;
  FieldDesc f[kServeFields];
  int32_t n_fields;
};

// Device words of a server (unsigned long long each).
enum {
  kTake = 0,      // claim cursor (virtual ring index)
  kSteps = 1,     // steps served
  kPub = 2,       // the host's published end, as the porter last read it
  kStop = 3,      // the host's stop flag, as the porter last read it
  kQueued = 4,    // requests queued by workers (cumulative)
  kExported = 5,  // requests exported (cumulative)
  kSubmitted = 6, // the host's count of submitted requests, as the porter last read it
  kWordCount = 8
};
// Host control words.
enum { kHostPub = 0, kHostStop = 1, kHostSubmitted = 2 };

struct ServeBuffers {
  const int32_t* ring;                       // host: records (at most 2^24)
  const unsigned long long* range;           // host: per slot {start, n}
  const unsigned long long* control;         // host: control words
  unsigned long long* h_pos;                 // host: endpoints (2 x u64 each)
  int32_t* h_steps;                          // host: step counts
  int32_t* complete;                         // host: per-slot flag
  void* d_pos;                               // device: endpoints (16 B each)
  int32_t* d_steps;                          // device: step counts
  unsigned long long* w;                     // device: words
  uint32_t* done;                            // device: per-slot done counters
  uint32_t* queue;                           // device: ended requests, slot + 1 (0 = not yet written)
  unsigned long long capacity;               // ring entries
  int32_t slots;
};

__device__ __forceinline__ unsigned long long bcast_u64(unsigned long long x) {
  return (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(int(x >> 32))) << 32) |
         uint32_t(__builtin_amdgcn_readfirstlane(int(x & 0xffffffffu)));
}

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kSystemCoherent = 0x11;  // sc0 | sc1: past every cache, as a system-scope atomic access
constexpr int kDeviceCoherent = 0x10;  // sc1: past the non-coherent caches, as an agent-scope atomic access

template <class T>
__device__ __forceinline__ T load_sys(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T>
__device__ __forceinline__ T load_dev(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void store_dev(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The porter (wave 0 of block 0): the grid's only reader of host words.  It mirrors the published end,
// the count of submitted requests and the stop flag into device memory; the stop flag only once it has
// read it twice, so that the submitted count it mirrored before is the host's last.  It leaves once
// stopped and every submitted request is exported (so every particle has ended).
__device__ void serve_porter(const ServeBuffers& b) {
  const int lane = threadIdx.x & 63;
  unsigned long long pub_seen = load_dev(b.w + kPub), sub_seen = ~0ull;
  bool stop_before = false;
  for (;;) {
    unsigned long long host_pub = 0, stop = 0, submitted = 0;
    if (lane == 0) {
      host_pub = load_sys(b.control + kHostPub);
      stop = load_sys(b.control + kHostStop);
      submitted = load_sys(b.control + kHostSubmitted);
    }
    host_pub = bcast_u64(host_pub);
    stop = bcast_u64(stop);
    submitted = bcast_u64(submitted);
    if (host_pub > pub_seen) {
      if (lane == 0) store_dev(b.w + kPub, host_pub);
      pub_seen = host_pub;
    }
    if (submitted != sub_seen) {
      if (lane == 0) store_dev(b.w + kSubmitted, submitted);
      sub_seen = submitted;
    }
    if (stop && stop_before) {
      __builtin_amdgcn_s_waitcnt(0);  // the submitted count is in device memory before the stop flag
      unsigned long long exported = 0;
      if (lane == 0) {
        store_dev(b.w + kStop, 1ull);
        exported = load_dev(b.w + kExported);
      }
      if (bcast_u64(exported) >= submitted) break;
    }
    stop_before = stop != 0;
    __builtin_amdgcn_s_sleep(64);  // ~2 us: host words are read over the bus, which the ring reads share
  }
}

// An exporter (waves 1..kExporters of block 0) takes every kExporters-th ended request in queue order,
// copies its endpoints and steps from the device ring to the host ring with whole-wave plain stores
// (consecutive entries, combined into full lines; atomic stores to host memory go out one word each and
// run ~10x slower), makes them visible with one system-scope release and raises the request's flag.
constexpr int kExporters = 3;
constexpr int kExportDepth = 4;  // entries a lane has in flight

__device__ void serve_exporter(const ServeBuffers& b, int j) {
  const int lane = threadIdx.x & 63;
  const rsrc_t dpos = __builtin_amdgcn_make_buffer_rsrc(b.d_pos, 0, int(b.capacity * 16), 0x00020000);
  const rsrc_t dsteps = __builtin_amdgcn_make_buffer_rsrc(b.d_steps, 0, int(b.capacity * 4), 0x00020000);
  // every request queued before this launch is exported: the queue restarts at the exported count
  unsigned long long k = 0;
  if (lane == 0) k = load_dev(b.w + kExported);
  k = bcast_u64(k) + unsigned(j);
  for (;;) {
    unsigned long long queued = 0;
    uint32_t e = 0;
    if (lane == 0) {
      queued = load_dev(b.w + kQueued);
      if (k < queued) e = load_dev(b.queue + k % unsigned(b.slots));
    }
    e = uint32_t(__builtin_amdgcn_readfirstlane(int(e)));
    if (e != 0) {  // queued and written
      const int32_t slot = int32_t(e) - 1;
      unsigned long long start = 0, n = 0;
      if (lane == 0) {
        store_dev(b.queue + k % unsigned(b.slots), 0u);
        start = load_sys(b.range + 2 * slot);
        n = load_sys(b.range + 2 * slot + 1);
      }
      start = bcast_u64(start);
      n = bcast_u64(n);
      for (unsigned long long i0 = 0; i0 < n; i0 += 64 * kExportDepth) {
        u32x4 d[kExportDepth];
        uint32_t st[kExportDepth];
        unsigned long long r[kExportDepth];
#pragma unroll
        for (int u = 0; u < kExportDepth; ++u) {
          const unsigned long long i = i0 + unsigned(lane + 64 * u);
          r[u] = (start + (i < n ? i : 0)) % b.capacity;
          d[u] = __builtin_amdgcn_raw_buffer_load_b128(dpos, int(r[u] * 16), 0, kDeviceCoherent);
          st[u] = __builtin_amdgcn_raw_buffer_load_b32(dsteps, int(r[u] * 4), 0, kDeviceCoherent);
        }
#pragma unroll
        for (int u = 0; u < kExportDepth; ++u)
          if (i0 + unsigned(lane + 64 * u) < n) {
            reinterpret_cast<u32x4*>(b.h_pos)[r[u]] = d[u];
            b.h_steps[r[u]] = int32_t(st[u]);
          }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) {
        __hip_atomic_store(b.complete + slot, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_fetch_add(b.w + kExported, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      k += kExporters;
      continue;
    }
    unsigned long long stop = 0, done = 0;
    if (lane == 0) {
      stop = load_dev(b.w + kStop);
      // read after the flag: the porter raises it only once the submitted count is final
      if (stop) done = load_dev(b.w + kExported) >= load_dev(b.w + kSubmitted);
    }
    if (bcast_u64(done)) break;
    __builtin_amdgcn_s_sleep(16);  // ~0.5 us
  }
}

template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kServeBlock) void k_serve(const ServeConfig cfg, const ServeBuffers b, int rule) {
  constexpr int kWords = int(sizeof(pa_particle) / 4);  // 12
  if (blockIdx.x == 0) {  // block 0: the porter and the exporters
    static_assert(kServeBlock / 64 == 1 + kExporters, "block 0 holds the porter and the exporters");
    if (threadIdx.x < 64) serve_porter(b); else serve_exporter(b, int(threadIdx.x >> 6) - 1);
    return;
  }
  __shared__ u32x4 staged[kServeBlock / 64][64 * kWords / 4];
  constexpr bool kKeep = keeps_record<S, MODE>();
  __shared__ uint4 keep[kKeep ? keep_words<S>() : 1][kServeBlock];  // the lane's last cell's corners (kKeep)
  uint32_t kept = 0xffffffffu;                         // its cell index; none
  const int lane = threadIdx.x & 63;
  u32x4* mine = staged[threadIdx.x >> 6];
  // the host ring, read 16 bytes a lane past every cache (the host rewrites entries as the ring wraps)
  const rsrc_t ring = __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(b.ring), 0,
                                                        int(b.capacity * sizeof(pa_particle)), 0x00020000);
  const rsrc_t dpos = __builtin_amdgcn_make_buffer_rsrc(b.d_pos, 0, int(b.capacity * 16), 0x00020000);
  const rsrc_t dsteps = __builtin_amdgcn_make_buffer_rsrc(b.d_steps, 0, int(b.capacity * 4), 0x00020000);
  bool have = false;
  unsigned long long v = 0;
  float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int it = 0, R = 0, st = 0, slot = 0;
  uint32_t end_count = 0;
  FieldDesc f = cfg.f[0];
  unsigned long long done_steps = 0;
  // the wave's claimed range [lo, hi) of virtual ring indices not yet handed to lanes (wave-uniform);
  // its part at or past the published end waits for the host to publish it
  unsigned long long lo = 0, hi = 0;
  int since_poll = 0;  // iterations since this wave last looked for work
  int every = 1;       // poll interval of a wave with idle and busy lanes: doubles while nothing arrives
  for (;;) {
    const uint64_t idle = __ballot(!have);
    const int n_idle = __popcll(idle);
    if (n_idle && (n_idle == 64 || ++since_poll >= every)) {
      since_poll = 0;
      unsigned long long pub = 0, stop = 0, base = 0;
      if (lane == 0) {
        pub = load_dev(b.w + kPub);
        if (n_idle == 64) stop = load_dev(b.w + kStop);
        // a new claim only when the wave holds none and the cursor is behind the published end
        if (lo >= hi && load_dev(b.w + kTake) < pub)
          base = __hip_atomic_fetch_add(b.w + kTake, static_cast<unsigned long long>(n_idle), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT) + 1ull;  // + 1: 0 means no claim
      }
      pub = bcast_u64(pub);
      stop = bcast_u64(stop);
      base = bcast_u64(base);
      if (base) {
        lo = base - 1ull;
        hi = lo + unsigned(n_idle);
      }
      // hand idle lanes the published part of the claim
      const unsigned long long avail = lo < pub ? min(hi, pub) - lo : 0ull;
      const unsigned got = unsigned(min(avail, static_cast<unsigned long long>(n_idle)));
      every = got ? 1 : min(every * 2, 64);
      if (got) {
        // the records, read by the whole wave as consecutive 16-byte pieces of the host ring
        for (int u = 0; u < kWords / 4; ++u) {
          const int t = lane + 64 * u;
          if (t < int(got) * (kWords / 4)) {
            const unsigned long long e = (lo + unsigned(t / (kWords / 4))) % b.capacity;
            mine[t] = __builtin_amdgcn_raw_buffer_load_b128(
                ring, int(e * sizeof(pa_particle)) + 16 * (t % (kWords / 4)), 0, kSystemCoherent);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const unsigned r = unsigned(__popcll(idle & ((uint64_t(1) << lane) - 1)));
        if (!have && r < got) {
          const int32_t* q = reinterpret_cast<const int32_t*>(mine + r * (kWords / 4));
          v = lo + r;
          p = make_float4(__int_as_float(q[0]), __int_as_float(q[1]), __int_as_float(q[2]), __int_as_float(q[3]));
          R = q[4];
          const int32_t fi = q[5];
          slot = q[10];                    // reserved_[0]: the request slot
          end_count = uint32_t(q[11]);     // reserved_[1]: the slot's count once the request ends
          f = cfg.f[fi >= 0 && fi < cfg.n_fields ? fi : 0];
          kept = 0xffffffffu;
          it = 1;
          st = 0;
          have = true;
        }
        lo += got;
        // the LDS records are free again only once every lane has read its own
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (__ballot(have) == 0) {
        // Nothing to run.  The wave leaves once the host asks it to stop and nothing published is
        // left to it (the host then resets the cursor to the published end); otherwise it sleeps
        // ~16 us between polls.
        if (stop && (lo >= hi || lo >= pub)) {
          unsigned long long t = 0;
          if (lane == 0) t = load_dev(b.w + kTake);
          if (bcast_u64(t) >= pub) break;
        }
#pragma unroll
        for (int z = 0; z < 4; ++z) __builtin_amdgcn_s_sleep(127);
        continue;
      }
    }
    bool fin = false;
    if (have) {
      fin = it >= R - 1;  // the reference breaks at it == R-1 (R < 2: the seed alone)
      if (!fin) {
        float4 next;
        int out;
        if constexpr (kKeep)
          out = advance_kept<S, MODE, INTERP, kServeBlock>(cfg.k, f, p, next, rule, kept, keep);
        else
          out = advance_rt<S, MODE, INTERP>(cfg.k, f, p, next, rule);
        if (out != 0) {
          fin = true;
        } else {
          p = next;
          ++it;
          ++st;
          ++done_steps;
        }
      }
      if (fin) {
        const unsigned long long e = v % b.capacity;
        const u32x4 d = {__float_as_uint(p.x), __float_as_uint(p.y), __float_as_uint(p.z), __float_as_uint(p.w)};
        __builtin_amdgcn_raw_buffer_store_b128(d, dpos, int(e * 16), 0, kDeviceCoherent);
        __builtin_amdgcn_raw_buffer_store_b32(uint32_t(st), dsteps, int(e * 4), 0, kDeviceCoherent);
      }
    }
    uint64_t ended = __ballot(fin);
    if (ended) {
      // The results went past the non-coherent caches; waiting for them to complete orders them before
      // the counts, which the exporter reads (past the same caches) only after the request is queued.
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the wave's stores are complete
      // One count per request among the lanes that ended (a request's counter is a single address, so
      // a count per particle would queue the whole request's atomics behind each other).
      while (ended) {
        const int leader = __builtin_amdgcn_readfirstlane(int(__builtin_ctzll(ended)));
        const int s0 = __builtin_amdgcn_readlane(slot, leader);
        const uint64_t same = __ballot(fin && slot == s0);
        if (lane == leader) {
          const uint32_t count = uint32_t(__popcll(same));
          const uint32_t before = __hip_atomic_fetch_add(b.done + s0, count, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
          if (uint32_t(end_count - before) - 1u < count) {  // this add reached the request's end count
            const unsigned long long k = __hip_atomic_fetch_add(b.w + kQueued, 1ull, __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT);
            store_dev(b.queue + k % unsigned(b.slots), uint32_t(s0) + 1u);
          }
        }
        ended &= ~same;
      }
      if (fin) have = false;
    }
  }
  done_steps = wave_sum_u64(done_steps);
  if (lane == 0 && done_steps) atomicAdd(b.w + kSteps, done_steps);
}

__global__ void k_serve_set(unsigned long long* w, unsigned long long cursor, unsigned long long published) {
  if (threadIdx.x == 0) {
    w[kTake] = cursor;
    w[kPub] = published;
    w[kStop] = 0;
  }
}
__global__ void k_serve_zero(unsigned long long* c, int n) {
  if (int(threadIdx.x) < n) c[threadIdx.x] = 0ull;
}
__global__ void k_serve_zero_u32(uint32_t* c, int n) {
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
  unsigned long long* range = nullptr;    // per slot {start, n}
  float4* out_pos = nullptr;
  int32_t* out_steps = nullptr;
  int32_t* complete = nullptr;
  unsigned long long* control = nullptr;  // [0] published, [1] stop, [2] requests submitted
  // device
  unsigned long long* d_words = nullptr;
  uint32_t* d_done = nullptr;
  uint32_t* d_queue = nullptr;
  float4* d_pos = nullptr;
  int32_t* d_steps = nullptr;
  std::vector<uint32_t> slot_count;       // each slot's done counter once its latest request ends
  // host bookkeeping
  unsigned long long published = 0, reclaimed = 0, submitted = 0;
  struct Req {
    unsigned long long start = 0;
    size_t n = 0;
    bool used = false, collected = true;
  };
  std::vector<Req> req;
  std::vector<unsigned long long> order;  // start of each live request, oldest first
  int waves_per_cu = 12;
  unsigned long long drained = 0;  // published end when the kernel last left
  uint64_t launches = 0;
  bool launched = false;
  std::chrono::steady_clock::time_point last_query{};
};

namespace {

pa_status serve_launch(pa_gpu_server* s) {
  pa_gpu_ctx* ctx = s->ctx;
  int cus = 0;
  PA_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  // block 0 for the porter and the exporters, then the workers
  const int64_t blocks = 1 + std::max<int64_t>(1, (int64_t(cus) * s->waves_per_cu + 3) / 4);
  // everything before the published end was served by the last instance: its claims past the end are
  // void
  hipLaunchKernelGGL(k_serve_set, dim3(1), dim3(64), 0, s->stream, s->d_words, s->drained, s->published);
  ServeBuffers b;
  b.ring = reinterpret_cast<const int32_t*>(s->ring);
  b.range = s->range;
  b.control = s->control;
  b.h_pos = reinterpret_cast<unsigned long long*>(s->out_pos);
  b.h_steps = s->out_steps;
  b.complete = s->complete;
  b.d_pos = s->d_pos;
  b.d_steps = s->d_steps;
  b.w = s->d_words;
  b.done = s->d_done;
  b.queue = s->d_queue;
  b.capacity = static_cast<unsigned long long>(s->capacity);
  b.slots = s->slots;
  dispatch(s->cfg.f[0].storage, s->mode, s->cfg.k.interp, [&](auto S, auto M, auto I) {
    hipLaunchKernelGGL((k_serve<decltype(S)::value, decltype(M)::value, decltype(I)::value>), dim3(unsigned(blocks)),
                       dim3(kServeBlock), 0, s->stream, s->cfg, b, s->rule);
  }, s->cfg.f[0].layout);
  PA_HIP(ctx, hipGetLastError());
  s->launched = true;
  s->last_query = std::chrono::steady_clock::now();
  ++s->launches;
  return PA_OK;
}

// Launch the kernel if a request is open and no instance runs.  A running instance stays until
// pa_gpu_serve_stop; its stream is queried (for a kernel that failed) at most every 5 ms.
pa_status serve_kick(pa_gpu_server* s) {
  if (s->launched) {
    const auto now = std::chrono::steady_clock::now();
    if (now - s->last_query < std::chrono::milliseconds(5)) return PA_OK;
    s->last_query = now;
    const hipError_t q = hipStreamQuery(s->stream);
    if (q == hipErrorNotReady) return PA_OK;
    if (q != hipSuccess) return check_hip(s->ctx, q, "serving kernel");
    return fail(s->ctx, PA_E_STATE, "serving kernel left before it was asked to stop");
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
  if (particle_capacity < 64 || particle_capacity > kMaxRing || request_slots < 1 || request_slots > (1 << 20))
    return fail(ctx, PA_E_INVALID, "serve: 64 to 2^24 ring entries and 1 to 2^20 request slots");
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
  // 12: on a 1024^3 BF16 cell-record field, saturated serving 24.8 / 30.3 / 34.1 G steps/s at 8 / 12 / 16 waves
  // per CU, and at 1600 requests/s p99 4.1 / 3.7 / 4.4 ms: past 12 the idle waves' polls cost latency
  s->waves_per_cu = waves_per_cu > 0 ? waves_per_cu : 12;
  (void)idle_polls;
  s->req.resize(size_t(request_slots));
  const unsigned flags = hipHostMallocCoherent | hipHostMallocMapped;
  const size_t ns = size_t(request_slots);
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->ring), particle_capacity * sizeof(pa_particle), flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->range), ns * 16, flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->out_pos), particle_capacity * 16, flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->out_steps), particle_capacity * 4, flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->complete), ns * 4, flags);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->control), 64, flags);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_words), kWordCount * 8);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_done), ns * 4);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_queue), ns * 4);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_pos), particle_capacity * 16);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_steps), particle_capacity * 4);
  if (e != hipSuccess) {
    pa_gpu_serve_destroy(s);
    return check_hip(ctx, e, "serve: allocation");
  }
  std::memset(s->complete, 0, ns * 4);
  std::memset(s->control, 0, 64);
  std::memset(s->range, 0, ns * 16);
  // touch every page of the rings now: the first host write of a fresh pinned page faults, which would
  // triple the cost of each request's copy-in and copy-out until the ring has wrapped once
  std::memset(static_cast<void*>(s->ring), 0, particle_capacity * sizeof(pa_particle));
  std::memset(static_cast<void*>(s->out_pos), 0, particle_capacity * 16);
  std::memset(static_cast<void*>(s->out_steps), 0, particle_capacity * 4);
  hipLaunchKernelGGL(k_serve_zero, dim3(1), dim3(64), 0, s->stream, s->d_words, int(kWordCount));
  s->slot_count.assign(ns, 0u);
  hipLaunchKernelGGL(k_serve_zero_u32, dim3(unsigned((request_slots + 255) / 256)), dim3(256), 0, s->stream, s->d_done,
                     request_slots);
  hipLaunchKernelGGL(k_serve_zero_u32, dim3(unsigned((request_slots + 255) / 256)), dim3(256), 0, s->stream,
                     s->d_queue, request_slots);
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
  auto& r = s->req[size_t(slot)];
  r.start = s->published;
  r.n = n;
  r.used = true;
  r.collected = false;
  __atomic_store_n(&s->complete[slot], 0, __ATOMIC_RELAXED);
  // the records and the slot's range, then the new end: the kernel reads the end before the records it
  // covers, and the range once the request has ended
  const uint32_t end_count = s->slot_count[size_t(slot)] + uint32_t(n);  // wraps with the device counter
  s->slot_count[size_t(slot)] = end_count;
  const size_t first = size_t(r.start % s->capacity);
  const size_t run = std::min(n, s->capacity - first);  // records up to the ring's end, then from its start
  for (size_t i = 0; i < n; ++i) {
    pa_particle q = h_particles[i];
    q.reserved_[0] = slot;
    q.reserved_[1] = int32_t(end_count);
    s->ring[i < run ? first + i : i - run] = q;
  }
  s->range[2 * size_t(slot)] = r.start;
  s->range[2 * size_t(slot) + 1] = n;
  s->order.push_back(r.start);
  s->published += n;
  ++s->submitted;
  __atomic_store_n(&s->control[kHostSubmitted], s->submitted, __ATOMIC_RELEASE);
  __atomic_store_n(&s->control[kHostPub], s->published, __ATOMIC_RELEASE);
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
  __atomic_store_n(&s->control[kHostStop], 1ull, __ATOMIC_RELEASE);
  PA_HIP(s->ctx, hipStreamSynchronize(s->stream));
  s->launched = false;
  s->drained = s->published;
  __atomic_store_n(&s->control[kHostStop], 0ull, __ATOMIC_RELEASE);
  unsigned long long st = 0;
  PA_HIP(s->ctx, hipMemcpy(&st, s->d_words + kSteps, 8, hipMemcpyDeviceToHost));
  if (total_steps) *total_steps = st;
  if (launches) *launches = s->launches;
  return PA_OK;
}

pa_status pa_gpu_serve_destroy(pa_gpu_server* s) {
  if (!s) return PA_E_INVALID;
  if (s->stream) {
    if (s->control) __atomic_store_n(&s->control[kHostStop], 1ull, __ATOMIC_RELEASE);
    (void)hipStreamSynchronize(s->stream);
  }
  if (s->ring) (void)hipHostFree(s->ring);
  if (s->range) (void)hipHostFree(s->range);
  if (s->out_pos) (void)hipHostFree(s->out_pos);
  if (s->out_steps) (void)hipHostFree(s->out_steps);
  if (s->complete) (void)hipHostFree(s->complete);
  if (s->control) (void)hipHostFree(s->control);
  if (s->d_words) (void)hipFree(s->d_words);
  if (s->d_done) (void)hipFree(s->d_done);
  if (s->d_queue) (void)hipFree(s->d_queue);
  if (s->d_pos) (void)hipFree(s->d_pos);
  if (s->d_steps) (void)hipFree(s->d_steps);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return PA_OK;
}

}  // extern "C"
