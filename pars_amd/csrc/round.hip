// The stages after the union trace of one round, fused: prune
// (particle_tracer.cpp:402-409, as the counted prune), index_generator
// (index_generator.cpp:10-24) and color_generator (color_generator.cpp:
// 30-98) on 8-seed grouped slots (slot s of seed p = vec4 at
// ((p / 8) * region + s) * 8 + p % 8): one pass packs the vertices and writes
// the index list and the colours.
//
// Layout of the outputs (as the separate stages produce them):
//   vertices  seed p's slots at O[p] .. O[p] + counts[p] - 1, O = exclusive
//             scan of counts (termination vertices included, as prune keeps
//             them);
//   indices   every i with vertices i and i+1 both not the termination
//             vertex, ascending; seed p owns I[p] .. I[p+1] - 1, where the
//             tracer counted the in-run pairs and the pair that crosses into
//             seed p+1 is added here;
//   colours   colour of segment (i, i+1) at i, and at i+1 when i+1 ends a
//             curve (no index i+1) unless i is the last index; other entries
//             are left as they were, like the colour stage.
#include "pa_color.hpp"
#include "pa_device.hpp"

namespace pa_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;  // 8-slot rows per register block of a packing wave

// icount[p] = in-run pairs + the pair (last of p, first of p+1).
__global__ __launch_bounds__(kThreads) void k_round_counts(const float4* __restrict__ slots, int64_t n, int64_t region,
                                                           const int32_t* __restrict__ counts,
                                                           const int32_t* __restrict__ pairs,
                                                           int32_t* __restrict__ icount, int32_t* __restrict__ cross) {
  const int64_t p = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (p >= n) return;
  int x = 0;
  if (p + 1 < n) {
    const float4 last = slots[group8_slot(p, counts[p] - 1, region)];
    const float4 next = slots[group8_slot(p + 1, 0, region)];
    x = (!is_termination_vertex(last) && !is_termination_vertex(next)) ? 1 : 0;
  }
  cross[p] = x;
  icount[p] = pairs[p] + x;
}

// A DPP row shift; a lane without a source reads 0 (bound control), which
// needs no copy of the source first.  Every use below replaces those lanes
// (j == 7 after row_shl:1, j == 0 after row_shr:1) or reads only lanes that
// have a source (row_shr:7 at j == 7, row_shl:7 at j == 0).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ float4 dpp_f4(float4 v) {
  return make_float4(dpp_f<CTRL>(v.x), dpp_f<CTRL>(v.y), dpp_f<CTRL>(v.z), dpp_f<CTRL>(v.w));
}

// Output stores are raw buffer stores (store16 / store4) whose descriptor
// starts at the group's first output element: a lane with nothing to store
// gives the offset kDrop and the hardware drops the store, so every store is
// issued by the whole wave without branches (exact wait counts) and without
// traffic of its own.  A group's vertex offsets reach (8 * region - 1) * 16
// bytes, which must stay below kDrop = 2^31 - 16 (an offset equal to kDrop
// is dropped like the lanes that store nothing), so region is below 2^24.
constexpr int64_t kMaxRegion = (int64_t(1) << 24) - 1;
static_assert((8 * kMaxRegion - 1) * 16 < kDrop, "largest group-relative vertex offset must stay below kDrop");

// What one lane of a packing wave knows about its seed p (lane 8k + j of the wave
// holds slot s0 + j of seed k of the group).
struct SeedLane {
  int64_t O;         // first packed vertex of the seed
  int vrel, irel;    // first vertex and first index relative to the group's
  int c;             // vertex count
  int j, k, lane;
  int carry;         // index pairs of this seed before the current chunk
  int prev_bit7;     // whether slot s0 - 1 begins a pair
  bool cross;        // the seed's last vertex pairs with the next seed's first
  bool from_prev;    // slot 0 takes the colour of the previous seed's cross pair
  float4 first_next; // first vertex of seed p + 1 (when cross)
  float4 last_prev;  // last vertex of seed p - 1 (when from_prev)
  float4 prev_tail;  // lane 8k: slot s0 - 1 of the seed
};

// The group's vertex and index outputs.
struct GroupOut {
  rsrc_t v, i;
};

// A value select (a select of float4 lvalues would select their addresses
// and keep the operands in scratch memory).
__device__ __forceinline__ float4 pick(bool c, float4 x, float4 y) {
  return make_float4(c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z, c ? x.w : y.w);
}

// Colour work leaves the packing wave as (segment start a, segment end b,
// packed vertex that takes the colour) through single-producer
// single-consumer LDS rings to colour waves of the same workgroup: the colour
// arithmetic of those waves overlaps the memory traffic of the packing wave,
// and only vertices that take a colour cost colour arithmetic.  (A packing
// wave that colours its own work from an LDS queue measured 0.96 ms on the
// share workload, one colour wave per packing wave 0.89 ms, two 0.87 ms.)
constexpr int kRing = 256;    // RingSink entries per producer/consumer pair
constexpr int kSpinLimit = 1 << 22;  // bounded waits (about 0.1 s), then a flag
constexpr unsigned kStalled = 1u << 16;  // flag bit: a finish wave gave up waiting

// One producer/consumer ring.  The producer alone writes head and the
// entries, the consumer alone writes tail; both publish with a store-release
// at workgroup scope and read the other's counter with a load-acquire.
struct Ring {
  float4 a[kRing];
  float4 b[kRing];
  int32_t dst[kRing];  // packed vertex (absolute; totals are below 2^31)
  uint32_t head, tail, done;
};
constexpr int kSplitThreads = 192;  // one packing wave, two colour waves
__shared__ Ring rings[2];

// The producer's sink: lanes 0..31 (seeds 0..3 of the group) feed ring 0,
// lanes 32..63 (seeds 4..7) ring 1, each drained by its own colour wave.
struct RingSink {
  Ring* r;            // two rings
  uint32_t head0, head1;
  unsigned* flags;
  bool ok;
  __device__ __forceinline__ void put(int lane, uint64_t m, bool mine, const float4& a, const float4& b, int,
                                      int64_t dst) {
    if (!ok) return;
    const uint32_t k0 = uint32_t(__popcll(m & 0xFFFFFFFFull)), k1 = uint32_t(__popcll(m >> 32));
    if ((k0 | k1) == 0) return;
    // wait for room (a colour wave frees entries as soon as it has read them)
    for (int spin = 0;; ++spin) {
      const uint32_t t0 = __hip_atomic_load(&r[0].tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
      const uint32_t t1 = __hip_atomic_load(&r[1].tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (head0 + k0 - t0 <= uint32_t(kRing) && head1 + k1 - t1 <= uint32_t(kRing)) break;
      if (spin >= kSpinLimit) {
        if (lane == 0) atomicOr(flags, kStalled);
        ok = false;
        return;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (mine) {
      const bool hi = lane >= 32;
      const uint64_t below = m & ((uint64_t(1) << lane) - 1u) & (hi ? ~0xFFFFFFFFull : 0xFFFFFFFFull);
      const uint32_t e = ((hi ? head1 : head0) + uint32_t(__popcll(below))) & (kRing - 1);
      Ring* q = hi ? r + 1 : r;
      q->a[e] = a;
      q->b[e] = b;
      q->dst[e] = int32_t(dst);
    }
    head0 += k0;
    head1 += k1;
    if (k0) __hip_atomic_store(&r[0].head, head0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k1) __hip_atomic_store(&r[1].head, head1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ __forceinline__ void finish(int) {}
};

// The colour wave of a pair: takes up to 64 entries at a time (64 while the
// producer runs, the rest once it is done), frees them as soon as they are
// read, then colours and stores them.  Ends when the producer is done and
// the ring is empty, or after kSpinLimit idle polls (flagged).
template <int MODE>
__device__ void colour_consumer(Ring* r, int lane, float param, float4* __restrict__ cout, unsigned* flags) {
  uint32_t tail = 0;
  int spin = 0;
  for (;;) {
    const uint32_t done = __hip_atomic_load(&r->done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t head = __hip_atomic_load(&r->head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t avail = head - tail;
    if (avail >= 64u || (done && avail > 0u)) {
      const uint32_t k = min(avail, 64u);
      float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
      int32_t dst = 0;
      if (uint32_t(lane) < k) {
        const uint32_t e = (tail + uint32_t(lane)) & (kRing - 1);
        a = r->a[e];
        b = r->b[e];
        dst = r->dst[e];
      }
      tail += k;
      __hip_atomic_store(&r->tail, tail, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      const float4 col = segment_color(a, b, MODE, param);
      if (uint32_t(lane) < k) cout[dst] = col;
      spin = 0;
    } else if (done) {
      break;
    } else {
      if (++spin >= kSpinLimit) {
        if (lane == 0) atomicOr(flags, kStalled);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
}

// One chunk row (slot s of every seed of the group) of the fused round:
// the vertex, the index pair (s, s + 1) and the colour of slot s.  Colour
// rule of color_generator.cpp:24-38 on the packed arrays: a vertex that
// begins a pair takes that segment's colour; a vertex that does not, but
// follows one that does, takes the previous segment's colour unless that
// pair is the last index of the list.
template <bool ROUND, typename Sink>
__device__ __forceinline__ void finish_row(SeedLane& L, Sink& sink, const float4 v, const float4 next_row, int s,
                                           int64_t NI, int64_t I0, const GroupOut& out) {
  store16(out.v, v, s < L.c ? (L.vrel + s) * 16 : kDrop);
  if constexpr (ROUND) {
    // slot s + 1: the next lane of the seed (DPP row_shl:1); for the last
    // lane, the first slot of the next chunk (row_shr:7 takes lane 8k to
    // lane 8k + 7).  Slot s - 1: row_shr:1, for the first lane the tail kept
    // from the previous chunk.
    float4 nx = dpp_f4<0x101>(v);
    const float4 f = dpp_f4<0x117>(next_row);
    if (L.j == 7) nx = f;
    float4 pv = dpp_f4<0x111>(v);
    if (L.j == 0) pv = L.prev_tail;
    const bool valid = s <= L.c - 2 && !is_termination_vertex(v) && !is_termination_vertex(nx);
    const uint32_t bits = uint32_t(__ballot(valid) >> (8 * L.k)) & 0xFFu;
    const int rel = L.irel + L.carry + __popc(bits & ((1u << L.j) - 1u));
    store4(out.i, int32_t(L.O + s), valid ? rel * 4 : kDrop);
    const bool prev_valid = L.j ? ((bits >> (L.j - 1)) & 1u) != 0u : L.prev_bit7 != 0;
    const bool own_cross = !valid && L.cross && s == L.c - 1;
    const bool after = !valid && !own_cross && s < L.c && prev_valid && I0 + rel != NI;
    const bool first = !valid && !own_cross && s == 0 && L.from_prev;
    const bool colour = valid || own_cross || after || first;
    const uint64_t m = __ballot(colour);
    sink.put(L.lane, m, colour, pick(after, pv, pick(first, L.last_prev, v)),
             pick(valid, nx, pick(own_cross, L.first_next, v)), L.vrel + s, L.O + s);
    L.carry += __popc(bits);
    L.prev_bit7 = int(bits >> 7);
    L.prev_tail = dpp_f4<0x107>(v);
  }
}

// Packing of one group of 8 seeds, whose slots are contiguous, by one wave:
// lane 8k + j holds slot s0 + j of seed k, so every load is one 1-KiB run of
// the group and every seed's 8 vertices (and index pairs) are stored as one
// contiguous run.  Two register blocks alternate so the next block's loads
// are in flight while one is written.  The index position of a pair is the
// seed's offset, the pairs before this chunk, and the prefix inside it; the
// pair that crosses into the next seed is the seed's last index.
template <bool ROUND, typename Sink>
__device__ __forceinline__ void pack_group(int64_t g, Sink& sink, const float4* __restrict__ slots, int64_t n,
                                           int64_t region, const int32_t* __restrict__ counts,
                                           const int32_t* __restrict__ cross, const int32_t* __restrict__ icount,
                                           const int64_t* __restrict__ voff, const int64_t* __restrict__ ioff,
                                           int64_t NI, float4* __restrict__ vout, int32_t* __restrict__ iout,
                                           int64_t& O0) {
  SeedLane L;
  L.lane = threadIdx.x & 63;
  L.k = L.lane >> 3;
  L.j = L.lane & 7;
  const int64_t p = g * 8 + L.k;
  const bool has = p < n;
  L.c = has ? counts[p] : 0;
  int cmax = L.c;
#pragma unroll
  for (int o = 8; o < 64; o <<= 1) cmax = max(cmax, __shfl_xor(cmax, o, 64));
  // the group's first vertex and index (seed 8g, lane 0), wave-uniform
  O0 = uniform(voff[g * 8]);
  const int64_t I0 = ROUND ? uniform(ioff[g * 8]) : 0;
  L.O = has ? voff[p] : O0;
  L.vrel = int(L.O - O0);
  L.irel = (ROUND && has) ? int(ioff[p] - I0) : 0;
  GroupOut out;
  out.v = out_rsrc(vout + O0);
  out.i = out_rsrc(ROUND ? iout + I0 : nullptr);
  L.carry = 0;
  L.prev_bit7 = 0;
  L.cross = ROUND && has && cross[p] != 0;
  L.from_prev = ROUND && has && p >= 1 && cross[p - 1] != 0 && I0 + L.irel != NI;
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  L.first_next = zero;
  L.last_prev = zero;
  L.prev_tail = zero;
  if (L.cross) L.first_next = slots[group8_slot(p + 1, 0, region)];
  if (L.from_prev) L.last_prev = slots[group8_slot(p - 1, counts[p - 1] - 1, region)];
  const float4* base = slots + g * region * 8;
  // slot s of this lane's seed, clamped to the seed's last vertex (rows past
  // a seed's count then fetch no cache line of their own; their values are
  // not stored) and into the region
  const int last = min(max(L.c - 1, 0), int(region - 1));
  auto load = [&](float4 (&b)[kUnroll], int s0) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int s = s0 + 8 * u + L.j;
      b[u] = base[int64_t(s < last ? s : last) * 8 + L.k];
    }
  };
  auto block = [&](const float4 (&b)[kUnroll], const float4& next_first, int s0) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      finish_row<ROUND>(L, sink, b[u], u + 1 < kUnroll ? b[u + 1 < kUnroll ? u + 1 : u] : next_first,
                        s0 + 8 * u + L.j, NI, I0, out);
  };
  float4 A[kUnroll], B[kUnroll];
  load(A, 0);
  for (int s0 = 0;; s0 += 16 * kUnroll) {
    load(B, s0 + 8 * kUnroll);
    block(A, B[0], s0);
    if (s0 + 8 * kUnroll >= cmax) break;
    load(A, s0 + 16 * kUnroll);
    block(B, A[0], s0 + 8 * kUnroll);
    if (s0 + 16 * kUnroll >= cmax) break;
  }
  sink.finish(L.lane);
  if (ROUND && L.cross && L.j == 0) iout[I0 + L.irel + icount[p] - 1] = int32_t(L.O + L.c - 1);
}

struct NoSink {
  __device__ __forceinline__ void put(int, uint64_t, bool, const float4&, const float4&, int, int64_t) {}
  __device__ __forceinline__ void finish(int) {}
};

// Counted prune of grouped slots (vertices only): one wave per group.
__global__ __launch_bounds__(kThreads) void k_group8_pack(const float4* __restrict__ slots, int64_t n, int64_t region,
                                                          const int32_t* __restrict__ counts,
                                                          const int64_t* __restrict__ voff, int64_t vcap,
                                                          float4* __restrict__ vout) {
  const int64_t g = int64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (g * 8 >= n) return;
  if (voff[n] > vcap) return;  // nothing is written past the capacity
  NoSink sink;
  int64_t O0 = 0;
  pack_group<false>(g, sink, slots, n, region, counts, nullptr, nullptr, voff, nullptr, 0, vout, nullptr, O0);
}

// Fused finish with a packing wave and two colour waves per workgroup:
// wave 0 packs group b and hands the colour work of seeds 0..3 and 4..7 to
// waves 1 and 2 through two rings.  Every wait is bounded.
template <int MODE>
__global__ __launch_bounds__(kSplitThreads) void k_group8_split(const float4* __restrict__ slots, int64_t n,
                                                                int64_t region, const int32_t* __restrict__ counts,
                                                                const int32_t* __restrict__ cross,
                                                                const int32_t* __restrict__ icount,
                                                                const int64_t* __restrict__ voff,
                                                                const int64_t* __restrict__ ioff, int64_t vcap,
                                                                int64_t icap, float param,
                                                                float4* __restrict__ vout, int32_t* __restrict__ iout,
                                                                float4* __restrict__ cout, unsigned* flags,
                                                                unsigned* __restrict__ queue) {
  const int64_t NV = voff[n], NI = ioff[n];
  if (NV > vcap || NI > icap || NV > INT32_MAX) return;  // the same for every wave
  const int w = int(threadIdx.x >> 6), lane = int(threadIdx.x & 63);
  if (threadIdx.x < 2) {
    rings[threadIdx.x].head = 0u;
    rings[threadIdx.x].tail = 0u;
    rings[threadIdx.x].done = 0u;
  }
  __syncthreads();
  if (w == 0) {
    // The packing wave's progress bounds the kernel (memory latency per
    // row); the colour waves only need to keep their rings from filling.
    // Issue priority to the packer: 0.89 -> 0.83 ms on the share round.
    __builtin_amdgcn_s_setprio(3);
    // Persistent workgroups: the grid is as many as fit on the device at
    // once; workgroup b packs group b, then the groups that a queue counter
    // (zeroed by k_round_totals) hands out.  A group costs ~4 us of
    // dependent header loads besides 0.8 us per row, and a workgroup per
    // group also left its slots to colour waves draining their rings: the
    // share round's finish 0.82 -> 0.72 ms.  The next group is taken before
    // the current one is packed, so its atomic is in flight meanwhile.
    RingSink sink{rings, 0u, 0u, flags, true};
    const int64_t groups = (n + 7) / 8;
    for (int64_t g = int64_t(blockIdx.x); g < groups;) {
      uint32_t next = 0;
      if (lane == 0) next = atomicAdd(queue, 1u);
      int64_t O0 = 0;
      pack_group<true>(g, sink, slots, n, region, counts, cross, icount, voff, ioff, NI, vout, iout, O0);
      g = int64_t(gridDim.x) + int64_t(__builtin_amdgcn_readfirstlane(next));
    }
    __hip_atomic_store(&rings[0].done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_store(&rings[1].done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    colour_consumer<MODE>(&rings[w - 1], lane, param, cout, flags);
  }
}

}  // namespace

size_t finish_ws_bytes(int64_t n) {
  return size_t(n + 1) * 2 * sizeof(int64_t) + size_t(n) * 2 * sizeof(int32_t) + scan_workspace_bytes(n) + 4096;
}

__global__ void k_round_totals(const int64_t* __restrict__ voff, const int64_t* __restrict__ ioff, int64_t n,
                               int64_t* __restrict__ totals, unsigned* __restrict__ queue) {
  if (threadIdx.x == 0) {
    totals[0] = voff[n];
    totals[1] = ioff[n];
    *queue = 0u;  // the finish's group queue (a kernel store: capturable)
  }
}

pa_status finish_group8(pa_gpu_ctx* ctx, const float* d_slots, int64_t n, int64_t region, const int32_t* d_counts,
                        const int32_t* d_pairs, int mode, float param, float* d_vertices, size_t vertex_capacity,
                        int32_t* d_indices, size_t index_capacity, float* d_colors, void* ws, int64_t* totals,
                        unsigned* flags) {
  if (region > kMaxRegion) return fail(ctx, PA_E_INVALID, "round: region_slots below 2^24");
  char* w = static_cast<char*>(ws);
  auto* voff = reinterpret_cast<int64_t*>(w);
  auto* ioff = voff + (n + 1);
  auto* icount = reinterpret_cast<int32_t*>(ioff + (n + 1));
  int32_t* cross = icount + n;
  char* tail = reinterpret_cast<char*>(cross + n);
  auto* sws = reinterpret_cast<int64_t*>(tail + (256 - (reinterpret_cast<uintptr_t>(tail) & 255)));
  // the group queue sits past the scan workspace, inside the 4 KiB slack
  auto* queue = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(sws) + (scan_workspace_bytes(n) + 255) / 256 * 256);
  const auto* slots = reinterpret_cast<const float4*>(d_slots);
  launch_begin(ctx, "round_counts");
  hipLaunchKernelGGL(k_round_counts, dim3(unsigned((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, ctx->stream,
                     slots, n, region, d_counts, d_pairs, icount, cross);
  PA_TRY(scan_i32(ctx, d_counts, n, voff, sws));
  PA_TRY(scan_i32(ctx, icount, n, ioff, sws));
  hipLaunchKernelGGL(k_round_totals, dim3(1), dim3(64), 0, ctx->stream, voff, ioff, n, totals, queue);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  const int64_t groups = (n + 7) / 8;
  int cus = 0;
  PA_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  launch_begin(ctx, "round_finish");
  if (!with_color_mode(mode, [&](auto M) {
        constexpr int MD = decltype(M)::value;
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_group8_split<MD>, kSplitThreads, 0) != hipSuccess ||
            per_cu < 1)
          per_cu = 1;
        const int64_t grid = std::min<int64_t>(groups, int64_t(cus > 0 ? cus : 1) * per_cu);
        hipLaunchKernelGGL(k_group8_split<MD>, dim3(unsigned(grid)), dim3(kSplitThreads), 0, ctx->stream, slots, n, region,
                           d_counts, cross, icount, voff, ioff, int64_t(vertex_capacity), int64_t(index_capacity),
                           param, reinterpret_cast<float4*>(d_vertices), d_indices,
                           reinterpret_cast<float4*>(d_colors), flags, queue);
      }))
    return fail(ctx, PA_E_INVALID, "unknown color mode");
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

}  // namespace pa_amd

using namespace pa_amd;

extern "C" pa_status pa_gpu_prune_counted_group8(pa_gpu_ctx* ctx, const float* d_slots, size_t region_slots,
                                                 const int32_t* d_vertex_counts, size_t n, float* d_out,
                                                 size_t capacity, size_t* n_out, int64_t* d_offsets) {
  if (!ctx || !n_out || (n && (!d_slots || !d_vertex_counts || !d_out))) return PA_E_INVALID;
  if (n == 0) {
    *n_out = 0;
    return PA_OK;
  }
  if (region_slots > size_t(kMaxRegion)) return fail(ctx, PA_E_INVALID, "prune_counted: region_slots below 2^24");
  void* ws = nullptr;
  const size_t own = d_offsets ? 0 : (n + 1) * sizeof(int64_t);
  PA_TRY(scratch(ctx, own + scan_workspace_bytes(int64_t(n)) + 4096, &ws));
  int64_t* offs = d_offsets ? d_offsets : static_cast<int64_t*>(ws);
  auto* sws = reinterpret_cast<int64_t*>(static_cast<char*>(ws) + (own + 255) / 256 * 256);
  launch_begin(ctx, "prune_scan");
  PA_TRY(scan_i32(ctx, d_vertex_counts, int64_t(n), offs, sws));
  launch_end(ctx);
  int64_t total = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&total, offs + n, sizeof total, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n_out = size_t(total);
  if (size_t(total) > capacity) return fail(ctx, PA_E_CAPACITY, "prune_counted: output capacity too small");
  const int64_t groups = (int64_t(n) + 7) / 8;
  launch_begin(ctx, "prune_counted");
  hipLaunchKernelGGL(k_group8_pack, dim3(unsigned((groups + 3) / 4)), dim3(kThreads), 0, ctx->stream,
                     reinterpret_cast<const float4*>(d_slots), int64_t(n), int64_t(region_slots), d_vertex_counts,
                     offs, int64_t(capacity), reinterpret_cast<float4*>(d_out));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}
