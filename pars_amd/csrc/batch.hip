// Particle pool over up to 16 resident fields: the batched form (many seed
// sets over fields of uneven size in one persistent work-queue kernel) and
// the step engine of the serving loop (bounded passes, requests admitted
// between passes).  No reference counterpart (service.cpp:20-69 serves one
// request at a time); every particle follows the reference round of
// particle_tracer.cpp:265-326 on its own field alone: it = 1 .. R-1, stop on
// leaving the field, on the stop rule or at it == R-1, so its endpoint and
// step count are those of a one-block reference trace.  With vertex output
// (pa_gpu_pool_run_vertices) each entry also writes that trace's slots: the
// seed at its first slot, each stepped vertex, the termination vertex, and
// its vertex count, as the reference round writes them for the particle
// alone (particle_tracer.cpp:243-326).  A particle's result does not depend
// on what else is in the pool.
#include <cstring>
#include <string>

#include "pa_device.hpp"

namespace pa_amd {
namespace {

constexpr int kPoolBlock = 256;
constexpr int kStage = 4;  // vertices a lane stages in LDS before writing them
constexpr int kMaxFields = 16;
constexpr int kChunk = 64;  // pool entries a wave claims at a time
// Request counters a workgroup keeps in LDS (dynamic, n_requests ints) and adds to request_done once, when it
// exits: one add per particle on a request's single counter serialises behind the others and holds up the
// wave's next gather wait (batched workload: 68.8 ms with them, 31 ms without).  More requests than this
// are counted with one add per particle.  4,096 counters (16 KB) leave 16-bit cell records at 5 waves per
// SIMD.
constexpr int kLdsRequests = 4096;

struct PoolConfig {
  TraceConfig k;                // integrator constants, stop rule, interpolation
  FieldDesc f[kMaxFields];      // field i is particle.vector_field_index == i
  int32_t n_fields;
};

// Optional vertex output: entry i owns slots [base[i], base[i] + max(R, 1)) of `slots`.
struct PoolVertices {
  float4* slots = nullptr;       // null: endpoints only
  const int64_t* base = nullptr;
  int32_t* count = nullptr;
  int64_t capacity = 0;          // slots
};

__global__ __launch_bounds__(kPoolBlock) void k_pool_init(const pa_particle* __restrict__ P, int64_t n, int n_fields,
                                                          float4* __restrict__ pos, int32_t* __restrict__ it,
                                                          int32_t* __restrict__ rem, int32_t* __restrict__ steps,
                                                          int32_t* __restrict__ status, int32_t* __restrict__ field,
                                                          int32_t* __restrict__ request, int n_requests,
                                                          unsigned int* __restrict__ invalid) {
  const int64_t i = int64_t(blockIdx.x) * kPoolBlock + threadIdx.x;
  if (i >= n) return;
  const pa_particle p = P[i];
  pos[i] = make_float4(p.position[0], p.position[1], p.position[2], p.position[3]);
  it[i] = 1;
  rem[i] = p.remaining_iterations;
  steps[i] = 0;
  bool ok = p.vector_field_index >= 0 && p.vector_field_index < n_fields;
  if (request) {
    // the request id indexes request_done: an id outside it is never stored
    const bool known = p.original_rank >= 0 && p.original_rank < n_requests;
    request[i] = known ? p.original_rank : 0;
    ok = ok && known;
  }
  status[i] = ok ? PA_POOL_ACTIVE : PA_POOL_INVALID;
  field[i] = ok ? p.vector_field_index : 0;
  if (!ok) atomicAdd(invalid, 1u);
}

// Persistent: each wave takes pool entries in chunks of `chunk` from a work
// counter, the claim of its next chunk always in flight behind the steps; a
// lane that has no entry takes the next of its wave's chunk, whose state
// loads go out with the other lanes' gathers and land at the same wait, and
// starts stepping it in the following iteration.  A lane advances an entry
// for at most `budget` iterations (or to the end), stores the state back and
// takes the next.  Finished entries add one to their request's counter
// (counted in LDS per workgroup up to kLdsRequests requests).
// Entries go to lanes in no fixed order; a particle's result does not depend
// on it.
template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kPoolBlock)
__attribute__((amdgpu_waves_per_eu(keeps_record<S, MODE>() && layout_of(S) == PA_LAYOUT_CELLS &&
                                           storage_of(S) != PA_STORAGE_FP32 ? 5 : 1)))
void k_pool_run(const PoolConfig cfg, float4* __restrict__ pos,
                                                         int32_t* __restrict__ it_a, const int32_t* __restrict__ rem,
                                                         int32_t* __restrict__ steps_a,
                                                         int32_t* __restrict__ status_a,
                                                         const int32_t* __restrict__ field_a,
                                                         const int32_t* __restrict__ request,
                                                         int32_t* __restrict__ request_done, int n_requests,
                                                         int64_t n, int budget, int rule, int chunk,
                                                         unsigned long long* __restrict__ work,
                                                         unsigned long long* __restrict__ steps_out,
                                                         unsigned long long* __restrict__ rejected,
                                                         const PoolVertices V) {
  // The field table in LDS: a lane reads its entry's field by index when it takes the entry.
  static_assert(sizeof(FieldDesc) % 4 == 0, "FieldDesc copies as 32-bit words");
  __shared__ FieldDesc ftab[kMaxFields];
  {
    constexpr int kWords = int(sizeof(FieldDesc)) * kMaxFields / 4;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(cfg.f);
    uint32_t* dst = reinterpret_cast<uint32_t*>(ftab);
    for (int k = threadIdx.x; k < kWords; k += kPoolBlock) dst[k] = src[k];
    __syncthreads();
  }
  // Dynamic LDS: the workgroup's request counters (n_requests ints, when counted here), then, for vertex
  // output, the staged vertices (kStage x kPoolBlock vec4); sized at launch (pool_lds_bytes).
  extern __shared__ uint4 dyn_lds[];
  int32_t* req_count = reinterpret_cast<int32_t*>(dyn_lds);
  const bool lds_count = request && request_done && n_requests <= kLdsRequests;
  auto stage = reinterpret_cast<float4(*)[kPoolBlock]>(dyn_lds + (lds_count ? (n_requests + 3) / 4 : 0));
  if (lds_count) {
    for (int k = threadIdx.x; k < n_requests; k += kPoolBlock) req_count[k] = 0;
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  __shared__ uint8_t order[kPoolBlock];  // lanes whose staged run goes out, in lane order (a wave each 64)
  const uint64_t below = (uint64_t(1) << lane) - 1;
  const unsigned long long un = static_cast<unsigned long long>(n);
  const unsigned long long csize = static_cast<unsigned long long>(chunk);
  // The wave's chunk [cur, end); `ahead` (lane 0) is the start of the chunk claimed next, valid while
  // `claimed` (the claim went out before any chunk reached n).
  unsigned long long cur = 0, end = 0, ahead = 0;
  bool claimed = true;
  if (lane == 0) ahead = atomicAdd(work, csize);
  // Vertex output is staged per lane in LDS (stage, above) and written kStage slots at a time: a global
  // store counts in the same in-order wait as the next gather, so one per step would lengthen every step.
  constexpr bool kKeep = keeps_record<S, MODE>();
  __shared__ uint4 keep[kKeep ? keep_words<S>() : 1][kPoolBlock];  // the lane's last cell's corners (kKeep)
  uint32_t kept = 0xffffffffu;                        // its cell index; none
  int staged = 0;       // vertices waiting in LDS
  int64_t sfirst = 0;   // slot of the first of them
  bool have = false;    // the lane is stepping entry i
  bool pend = false;    // the lane's next entry is loading (the t* values)
  int64_t i = 0, vb = 0;
  float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int it = 0, R = 0, st = 0, left = 0, rq = 0;
  int fi = 0;  // the entry's field: ftab[fi], read from LDS where a step uses it
  int64_t ti = 0, tvb = 0;
  float4 tp = p;
  int32_t tstat = 0, tfi = 0, trem = 0, tit = 0, tst = 0, treq = 0;
  unsigned long long done_steps = 0;
  for (;;) {
    // hand out entries to the lanes that have none
    const uint64_t idle = __ballot(!have && !pend);
    if (idle && (cur < end || claimed)) {
      const uint32_t want = uint32_t(__popcll(idle));
      const uint32_t rank = uint32_t(__popcll(idle & below));
      const unsigned long long avail = end - cur;
      unsigned long long idx = cur + rank, lim = end;
      if (want <= avail) {
        cur += want;
      } else if (claimed) {
        // the current chunk runs out: continue in the one claimed ahead and claim the next
        const unsigned long long nb = uniform(int64_t(ahead));
        const unsigned long long ne = nb < un ? (un - nb > csize ? nb + csize : un) : nb;
        if (rank >= avail) {
          idx = nb + (rank - avail);
          lim = ne;
        }
        cur = nb + (want - avail) < ne ? nb + (want - avail) : ne;
        end = ne;
        claimed = ne < un;
        if (claimed && lane == 0) ahead = atomicAdd(work, csize);
      } else {
        cur = end;
      }
      if (!have && !pend && idx < lim) {
        ti = int64_t(idx);
        tstat = status_a[ti];
        tfi = field_a[ti];
        trem = rem[ti];
        tp = pos[ti];
        tit = it_a[ti];
        tst = steps_a[ti];
        tvb = V.slots ? V.base[ti] : 0;
        treq = request ? request[ti] : 0;
        pend = true;
      }
    }
    if (__ballot(have || pend) == 0) break;  // nothing held and nothing left to take
    if (have) {
      bool fin = it >= R - 1;  // the reference breaks at it == R-1 (R < 2: the seed alone)
      if (!fin) {
        float4 next;
        int out;
        if constexpr (kKeep)
          out = advance_kept<S, MODE, INTERP, kPoolBlock>(cfg.k, ftab[fi], p, next, rule, kept, keep);
        else
          out = advance_rt<S, MODE, INTERP>(cfg.k, ftab[fi], p, next, rule);
        if (out != 0) {
          fin = true;
        } else {
          if (V.slots) stage[staged++][threadIdx.x] = next;  // written below, with the wave's other lanes
          p = next;
          ++it;
          ++st;
          ++done_steps;
        }
      }
      if (fin || --left == 0) {
        pos[i] = p;
        it_a[i] = it;
        steps_a[i] = st;
        if (V.slots) {
          // the termination vertex at the last slot reached (none for R < 2: the seed alone); what is
          // staged goes out at the end of the particle's pass
          if (fin && R >= 2) stage[staged++][threadIdx.x] = make_float4(-2.0f, -2.0f, -2.0f, 0.0f);
          for (int k = 0; k < staged; ++k) V.slots[sfirst + k] = stage[k][threadIdx.x];
          staged = 0;
          if (fin) V.count[i] = R >= 2 ? it + 1 : 1;
        }
        if (fin) {
          status_a[i] = PA_POOL_DONE;
          if (request && request_done && rq >= 0 && rq < n_requests) {
            if (lds_count) atomicAdd(&req_count[rq], 1);
            else atomicAdd(&request_done[rq], 1);
          }
        }
        have = false;
      }
    }
    // Staged vertices go out in runs that end on a 64-B boundary of the vertex buffer (4 slots): a lane
    // writes its run once the run reaches one, so the lines see whole 64-B writes rather than pieces
    // (only a particle's first and last runs are shorter).  The wave writes the runs of all such lanes
    // together, four lanes a run (a store instruction covers 16 runs of up to 64 B, not 64 scattered
    // 16-B pieces).  A lane adds at most one vertex an iteration and a run ends at the latest at the
    // fourth slot, so a lane holds at most kStage.
    if (V.slots) {
      static_assert(kStage == 4, "a run is written by 4 lanes and ends on a 4-slot boundary");
      const bool full = staged > 0 && int(sfirst & 3) + staged >= kStage;
      const uint64_t F = __ballot(full);
      if (F != 0) {
        const int wave_base = threadIdx.x & ~63;
        if (full) order[wave_base + __popcll(F & below)] = static_cast<uint8_t>(lane);
        __builtin_amdgcn_wave_barrier();  // the wave's stage and order writes above come first
        const int nf = __popcll(F);
        const int k = lane & 3;
        for (int g = 0; g < nf; g += 16) {
          const int rr = g + (lane >> 2);
          const int src = rr < nf ? int(order[wave_base + rr]) : lane;
          const int n_src = __shfl(staged, src);
          const long long f_src = __shfl(static_cast<long long>(sfirst), src);
          const float4 v = stage[k][wave_base + src];
          if (rr < nf && k < n_src) V.slots[f_src + k] = v;
        }
        if (full) {
          sfirst += staged;
          staged = 0;
        }
      }
    }
    // a lane whose entry has loaded takes it up: it steps from the next iteration
    if (pend) {
      pend = false;
      if (tstat == PA_POOL_ACTIVE) {
        if (tfi < 0 || tfi >= cfg.n_fields || (V.slots && (tvb < 0 || tvb > V.capacity - (trem > 1 ? trem : 1)))) {
          // a field this launch was not given, or slots outside the vertex buffer: never stepped
          status_a[ti] = PA_POOL_INVALID;
          atomicAdd(rejected, 1ull);
        } else {
          i = ti;
          p = tp;
          it = tit;
          R = trem;
          st = tst;
          rq = treq;
          fi = tfi;
          kept = 0xffffffffu;
          left = budget;
          have = true;
          vb = tvb;
          staged = 0;
          sfirst = vb + it;
          if (V.slots && it == 1) {  // first load: the seed (initialize, :243-252)
            stage[0][threadIdx.x] = p;
            staged = 1;
            sfirst = vb;
          }
        }
      }
    }
  }
  done_steps = wave_sum_u64(done_steps);
  if (lane == 0 && done_steps) atomicAdd(steps_out, done_steps);
  if (lds_count) {
    __syncthreads();
    for (int k = threadIdx.x; k < n_requests; k += kPoolBlock) {
      const int32_t c = req_count[k];
      if (c) atomicAdd(&request_done[k], c);
    }
  }
}

__global__ void k_pool_zero(unsigned long long* c) {
  if (threadIdx.x < 3) c[threadIdx.x] = 0ull;
}

}  // namespace
}  // namespace pa_amd

using namespace pa_amd;

extern "C" {

}  // extern "C"

namespace {
__global__ void k_zero_i32(int32_t* __restrict__ p, int n) {
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_copy_u32(const unsigned* __restrict__ src, unsigned* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = *src;
}

// Enqueues the pool's initialisation; the count of rejected entries is left
// in the context counter 2080.
pa_status pool_init_launch(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n, int32_t n_fields,
                           const pa_pool* pool) {
  if (!d_particles || !pool->position || !pool->iteration || !pool->remaining || !pool->steps || !pool->status ||
      !pool->field)
    return PA_E_INVALID;
  if (n_fields < 1 || n_fields > kMaxFields) return fail(ctx, PA_E_INVALID, "pool: 1 to 16 fields");
  if (pool->request && (!pool->request_done || pool->n_requests < 1))
    return fail(ctx, PA_E_INVALID, "pool: request ids need request_done and n_requests >= 1");
  auto* invalid = reinterpret_cast<unsigned int*>(static_cast<char*>(ctx->counters) + 2080);
  hipLaunchKernelGGL(k_pool_zero, dim3(1), dim3(64), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 2080));
  launch_begin(ctx, "pool_init");
  hipLaunchKernelGGL(k_pool_init, dim3(unsigned((n + kPoolBlock - 1) / kPoolBlock)), dim3(kPoolBlock), 0, ctx->stream,
                     d_particles, int64_t(n), int(n_fields), reinterpret_cast<float4*>(pool->position),
                     pool->iteration, pool->remaining, pool->steps, pool->status, pool->field, pool->request,
                     int(pool->n_requests), invalid);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}
}  // namespace

extern "C" {

pa_status pa_gpu_pool_init(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n, int32_t n_fields,
                           const pa_pool* pool) {
  if (!ctx || !pool) return PA_E_INVALID;
  if (n == 0) return PA_OK;
  PA_TRY(pool_init_launch(ctx, d_particles, n, n_fields, pool));
  auto* invalid = reinterpret_cast<unsigned int*>(static_cast<char*>(ctx->counters) + 2080);
  unsigned bad = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&bad, invalid, sizeof bad, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (bad)
    return fail(ctx, PA_E_INVALID,
                "pool: particles with a field index outside [0, n_fields) or a request id outside [0, n_requests)");
  return PA_OK;
}

}  // extern "C"

namespace {
__global__ void k_pool_result(const unsigned long long* __restrict__ c, unsigned long long* __restrict__ out) {
  if (threadIdx.x < 2) out[threadIdx.x] = c[1 + threadIdx.x];
}

// Enqueues the persistent pool kernel; its step count and rejected entries
// are left in the context counters 2056 and 2064.
pa_status pool_launch(pa_gpu_tracer* t, const pa_gpu_field* const* fields, int32_t n_fields, const pa_pool* pool,
                      size_t n, int32_t max_iterations, int32_t waves_per_cu, const PoolVertices& V = PoolVertices()) {
  pa_gpu_ctx* ctx = t->ctx;
  if (!fields || n_fields < 1 || n_fields > kMaxFields) return fail(ctx, PA_E_INVALID, "pool: 1 to 16 fields");
  if (max_iterations < 0 || waves_per_cu < 0) return fail(ctx, PA_E_INVALID, "pool: negative budget or waves");
  if (pool->request && pool->request_done && pool->n_requests < 1)
    return fail(ctx, PA_E_INVALID, "pool: request ids need n_requests >= 1");
  PA_TRY(check_integrator_of(ctx, t->integrator));
  PoolConfig cfg;
  std::memset(&cfg, 0, sizeof cfg);
  fill_step_constants(t, &cfg.k);
  int code = -1;
  for (int k = 0; k < n_fields; ++k) {
    if (!fields[k]) return fail(ctx, PA_E_INVALID, "pool: null field");
    const int c = fields[k]->desc.storage + 3 * fields[k]->desc.layout;
    if (code >= 0 && c != code)
      return fail(ctx, PA_E_UNSUPPORTED, "pool: all fields must share one storage type and layout");
    code = c;
    cfg.f[k] = fields[k]->desc;
  }
  for (int k = n_fields; k < kMaxFields; ++k) cfg.f[k] = cfg.f[0];
  cfg.n_fields = n_fields;
  auto* counters = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 2048);
  int cus = 0;
  PA_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int64_t need = (int64_t(n) + kPoolBlock - 1) / kPoolBlock;
  hipLaunchKernelGGL(k_pool_zero, dim3(1), dim3(64), 0, ctx->stream, counters);
  const int budget = max_iterations > 0 ? max_iterations : INT32_MAX;
  hipError_t occ = hipSuccess;
  // k_pool_run's dynamic LDS: the request counters (counted there up to kLdsRequests requests), 16-byte
  // aligned, then the vertex stage of vertex output
  const bool lds_count = pool->request && pool->request_done && pool->n_requests <= kLdsRequests;
  const size_t lds_bytes = (lds_count ? (size_t(pool->n_requests) + 3) / 4 * 16 : 0) +
                           (V.slots ? size_t(kStage) * kPoolBlock * sizeof(float4) : 0);
  launch_begin(ctx, "pool_run");
  dispatch(cfg.f[0].storage, mode_of_integrator(t->integrator), cfg.k.interp, [&](auto S, auto M, auto I) {
    const auto kernel = k_pool_run<decltype(S)::value, decltype(M)::value, decltype(I)::value>;
    // waves_per_cu 0: as many workgroups per CU as the kernel keeps resident
    int per_cu_blocks = 0;
    if (waves_per_cu == 0)
      occ = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_blocks, kernel, kPoolBlock, lds_bytes);
    int64_t blocks = waves_per_cu > 0 ? (int64_t(cus) * waves_per_cu + 3) / 4
                                      : int64_t(cus) * (per_cu_blocks > 0 ? per_cu_blocks : 4);
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(kernel, dim3(unsigned(blocks)), dim3(kPoolBlock), lds_bytes, ctx->stream, cfg,
                       reinterpret_cast<float4*>(pool->position), pool->iteration, pool->remaining, pool->steps,
                       pool->status, pool->field, pool->request, pool->request_done, int(pool->n_requests),
                       int64_t(n), budget, t->stop_rule, kChunk, counters, counters + 1, counters + 2, V);
  }, cfg.f[0].layout);
  launch_end(ctx);
  PA_HIP(ctx, occ);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}
}  // namespace

extern "C" {

pa_status pa_gpu_pool_run(pa_gpu_tracer* t, const pa_gpu_field* const* fields, int32_t n_fields, const pa_pool* pool,
                          size_t n, int32_t max_iterations, int32_t waves_per_cu, uint64_t* total_steps) {
  if (!t || !pool) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (total_steps) *total_steps = 0;
  if (n == 0) return PA_OK;
  PA_TRY(pool_launch(t, fields, n_fields, pool, n, max_iterations, waves_per_cu));
  auto* counters = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 2048);
  unsigned long long s[2] = {0, 0};
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(s, counters + 1, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (total_steps) *total_steps = s[0];
  if (s[1])
    return fail(ctx, PA_E_INVALID,
                "pool: " + std::to_string(s[1]) + " entries name a field index >= n_fields; marked invalid");
  return PA_OK;
}

pa_status pa_gpu_pool_run_vertices(pa_gpu_tracer* t, const pa_gpu_field* const* fields, int32_t n_fields,
                                   const pa_pool* pool, size_t n, int32_t max_iterations, int32_t waves_per_cu,
                                   const int64_t* d_slot_base, float* d_slots, size_t slot_capacity,
                                   int32_t* d_vertex_counts, uint64_t* total_steps) {
  if (!t || !pool) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (total_steps) *total_steps = 0;
  if (n == 0) return PA_OK;
  if (!d_slot_base || !d_slots || !d_vertex_counts) return PA_E_INVALID;
  if (slot_capacity > size_t(INT64_MAX)) return fail(ctx, PA_E_INVALID, "pool: slot capacity above 2^63");
  PoolVertices V;
  V.slots = reinterpret_cast<float4*>(d_slots);
  V.base = d_slot_base;
  V.count = d_vertex_counts;
  V.capacity = int64_t(slot_capacity);
  PA_TRY(pool_launch(t, fields, n_fields, pool, n, max_iterations, waves_per_cu, V));
  auto* counters = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 2048);
  unsigned long long s[2] = {0, 0};
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(s, counters + 1, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (total_steps) *total_steps = s[0];
  if (s[1])
    return fail(ctx, PA_E_INVALID,
                "pool: " + std::to_string(s[1]) +
                    " entries name a field index >= n_fields or slots outside the vertex buffer; marked invalid");
  return PA_OK;
}

pa_status pa_gpu_pool_init_async(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n, int32_t n_fields,
                                 const pa_pool* pool, int32_t zero_request_counters, uint32_t* d_rejected) {
  if (!ctx || !pool || !d_rejected) return PA_E_INVALID;
  auto* invalid = reinterpret_cast<unsigned int*>(static_cast<char*>(ctx->counters) + 2080);
  if (zero_request_counters && pool->request_done && pool->n_requests > 0)
    hipLaunchKernelGGL(k_zero_i32, dim3(unsigned((pool->n_requests + 255) / 256)), dim3(256), 0, ctx->stream,
                       pool->request_done, int(pool->n_requests));
  if (n == 0)
    hipLaunchKernelGGL(k_pool_zero, dim3(1), dim3(64), 0, ctx->stream,
                       reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 2080));
  else
    PA_TRY(pool_init_launch(ctx, d_particles, n, n_fields, pool));
  hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(64), 0, ctx->stream, invalid, d_rejected);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_pool_run_async(pa_gpu_tracer* t, const pa_gpu_field* const* fields, int32_t n_fields,
                                const pa_pool* pool, size_t n, int32_t max_iterations, int32_t waves_per_cu,
                                uint64_t* d_result) {
  if (!t || !pool || !d_result) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  auto* counters = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 2048);
  if (n == 0) {
    hipLaunchKernelGGL(k_pool_zero, dim3(1), dim3(64), 0, ctx->stream, counters);
  } else {
    PA_TRY(pool_launch(t, fields, n_fields, pool, n, max_iterations, waves_per_cu));
  }
  hipLaunchKernelGGL(k_pool_result, dim3(1), dim3(64), 0, ctx->stream, counters,
                     reinterpret_cast<unsigned long long*>(d_result));
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

}  // extern "C"
