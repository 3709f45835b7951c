// Seed lattice, particle tracer (reference round and union-of-7 form),
// particle advector and flow map.  One lane per particle; the 7 blocks stay
// resident in HBM; every lane keeps its last vertex in registers and writes
// each stepped vertex once.
#include <cmath>
#include <cstring>
#include <limits>

#include "pa_device.hpp"

#pragma clang fp contract(off)

namespace pa_amd {

size_t group_ws_bytes(int64_t n);
pa_status group_handoffs(pa_gpu_ctx* ctx, const int32_t* d_code, const pa_particle* d_tmp, int64_t n,
                         pa_handoff* handoff, void* ws);

namespace {
// Validation flags of the hybrid engine (bit 0: field index outside -1..5,
// bit 1: block not loaded, bit 2: more iterations than region_slots, bit 3:
// negative iterations).
pa_status union_flags(pa_gpu_ctx* ctx, unsigned bad) {
  if (bad & 1u) return fail(ctx, PA_E_INVALID, "vector_field_index outside -1..5");
  if (bad & 8u) return fail(ctx, PA_E_INVALID, "negative remaining_iterations");
  if (bad & 2u) return fail(ctx, PA_E_STATE, "a particle names a block that is not loaded");
  if (bad & 4u) return fail(ctx, PA_E_CAPACITY, "region_slots must be at least the maximum remaining_iterations");
  if (bad & (1u << 16)) return fail(ctx, PA_E_DEVICE, "round: a finish wave stopped waiting (outputs incomplete)");
  return PA_OK;
}
pa_status trace_union_impl(pa_gpu_tracer* t, int engine, const pa_particle* d_particles, size_t n,
                           size_t region_slots, pa_slot_order order, float* d_vertices, int32_t* d_vertex_counts,
                           pa_particle* d_continuation, size_t* n_continuation, float* d_endpoints,
                           int32_t* d_seed_steps, uint64_t* total_steps, int32_t* pairs, bool deferred);
}  // namespace

namespace {

constexpr int kBlock = 256;

// ------------------------------------------------------------------ seeds
// seed_generator.cpp:9-21: index -> (i,j,k) in C order (index.hpp:9-19),
// position = offset + stride * float(ijk), w = 0.
__global__ __launch_bounds__(kBlock) void k_seeds(pa_particle* __restrict__ out, int64_t n, int d1, int d2, float o0,
                                                  float o1, float o2, float s0, float s1, float s2, int32_t R,
                                                  int32_t rank) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int k = int(i % d2);
  const int j = int((i / d2) % d1);
  const int x = int(i / (int64_t(d1) * d2));
  pa_particle p;
  p.position[0] = o0 + s0 * float(x);
  p.position[1] = o1 + s1 * float(j);
  p.position[2] = o2 + s2 * float(k);
  p.position[3] = 0.0f;
  p.remaining_iterations = R;
  p.vector_field_index = -1;
  p.original_rank = rank;
  p.original_voxel[0] = x;
  p.original_voxel[1] = j;
  p.original_voxel[2] = k;
  p.reserved_[0] = 0;
  p.reserved_[1] = 0;
  out[i] = p;
}

__global__ __launch_bounds__(kBlock) void k_max_remaining(const pa_particle* __restrict__ P, int64_t n,
                                                          int* __restrict__ out_max, int* __restrict__ out_min_vfi,
                                                          int* __restrict__ out_max_vfi) {
  int m = INT32_MIN, lo = INT32_MAX, hi = INT32_MIN;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock) {
    m = max(m, P[i].remaining_iterations);
    lo = min(lo, P[i].vector_field_index);
    hi = max(hi, P[i].vector_field_index);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    m = max(m, __shfl_xor(m, o, 64));
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
  }
  // one atomic per workgroup and value (the counters are shared by the grid)
  __shared__ int red[3][kBlock / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = m;
    red[1][w] = lo;
    red[2][w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / 64; ++k) {
      m = max(m, red[0][k]);
      lo = min(lo, red[1][k]);
      hi = max(hi, red[2][k]);
    }
    atomicMax(out_max, m);
    atomicMin(out_min_vfi, lo);
    atomicMax(out_max_vfi, hi);
  }
}

__global__ void k_init_counters(int* c) {
  c[0] = INT32_MIN;
  c[1] = INT32_MAX;
  c[2] = INT32_MIN;
}

// Counter reset as a kernel (no memset node if a caller captures the
// stream in a graph).
__global__ void k_zero_u64(unsigned long long* c, int n) {
  if (int(threadIdx.x) < n) c[threadIdx.x] = 0ull;
}

// ------------------------------------------------------------------ trace
// initialize + trace of one reference round (particle_tracer.cpp:243-337).
// Slot layout: particle p owns slots [p*maxR, p*maxR + maxR) of the round,
// i.e. chunk p / ppic at offset (p % ppic) * maxR.  The loop writes the
// termination vertex at slot `it` after it ends, which is where the
// reference's write-then-overwrite leaves it.
template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kBlock) void k_trace(const TraceConfig cfg, const pa_particle* __restrict__ P, int64_t n,
                                                  float4* const* __restrict__ chunks, int64_t ppic, int64_t maxR,
                                                  int32_t* __restrict__ counts, pa_particle* __restrict__ hand_tmp,
                                                  int32_t* __restrict__ hand_code,
                                                  unsigned long long* __restrict__ steps_out, int stop_rule) {
  const int64_t pi = logical_block(cfg.xcd_remap) * kBlock + threadIdx.x;
  unsigned long long steps = 0;
  if (pi < n) {
    const pa_particle p = P[pi];
    float4* out = chunks[pi / ppic] + (pi % ppic) * maxR;
    float4 last = make_float4(p.position[0], p.position[1], p.position[2], p.position[3]);
    out[0] = last;
    const int b = p.vector_field_index;
    const FieldDesc f = cfg.field[b + 1];  // register copy: no descriptor loads in the step loop
    const int R = p.remaining_iterations;
    int it = 1;
    bool exited = false;
    for (; it < R; ++it) {
      if (it == R - 1) break;
      float4 next;
      const int a = advance_rt<S, MODE, INTERP>(cfg, f, last, next, stop_rule);
      if (a != 0) {
        exited = a == 1;
        break;
      }
      stream_store(&out[it], next);
      last = next;
      ++steps;
    }
    int count = 1;
    if (R >= 2) {
      out[it] = termination_vertex();
      count = it + 1;
    }
    if (counts) counts[pi] = count;
    if (hand_code) {
      int code = -1;
      if (exited) {
        if (b < 0) {
          const int k = face_rule(f, last, cfg.local_mask);  // :288-293
          if (k >= 0) code = k;                               // map 0: out_of_bounds_particles
        } else {
          code = 6 + b;  // map 1: neighbor_out_of_bounds_particles (:299-306)
        }
      }
      if (code >= 0) {
        pa_particle h;  // {last_vertex, R - it, -1} (:282)
        h.position[0] = last.x;
        h.position[1] = last.y;
        h.position[2] = last.z;
        h.position[3] = last.w;
        h.remaining_iterations = R - it;
        h.vector_field_index = -1;
        h.original_rank = 0;
        h.original_voxel[0] = h.original_voxel[1] = h.original_voxel[2] = 0;
        h.reserved_[0] = h.reserved_[1] = 0;
        hand_tmp[pi] = h;
      }
      hand_code[pi] = code;
    }
  }
  steps = wave_sum_u64(steps);
  if ((threadIdx.x & 63) == 0 && steps) atomicAdd(steps_out, steps);
}

// Union-of-7 form: a hand-off to a preloaded neighbour (or back to the own
// block) starts a new segment in the same lane.  Segment start = the
// uncontained vertex, remaining = R - it, exactly what the receiving rank
// would trace in the reference's next round.
template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kBlock) void k_trace_union(const TraceConfig cfg, const pa_particle* __restrict__ P,
                                                        int64_t n, float4* __restrict__ verts, int64_t region,
                                                        int64_t seed_stride, int64_t slot_stride,
                                                        int32_t* __restrict__ counts, pa_particle* __restrict__ cont,
                                                        unsigned int* __restrict__ n_cont,
                                                        unsigned long long* __restrict__ steps_out, int stop_rule,
                                                        float4* __restrict__ endpoints, int32_t* __restrict__ seed_steps) {
  const int64_t pi = logical_block(cfg.xcd_remap) * kBlock + threadIdx.x;
  unsigned long long steps = 0;
  if (pi < n) {
    const pa_particle p = P[pi];
    // Slot s of seed pi lives at verts[pi * seed_stride + s * slot_stride]:
    // seed-major (region, 1) or step-major (1, n).  Step-major makes the
    // lanes of a wave, which step in lockstep, store one contiguous run.
    float4* const base_ptr = verts + pi * seed_stride;
    auto out = [&](int64_t sl) -> float4& { return base_ptr[sl * slot_stride]; };
    float4 last = make_float4(p.position[0], p.position[1], p.position[2], p.position[3]);
    out(0) = last;
    int b = p.vector_field_index;
    int R = p.remaining_iterations;
    int64_t base = 0;
    int64_t used = 1;
    for (;;) {
      const FieldDesc f = cfg.field[b + 1];  // register copy: no descriptor loads in the step loop
      int it = 1;
      bool exited = false;
      for (; it < R; ++it) {
        if (it == R - 1) break;
        float4 next;
        const int a = advance_rt<S, MODE, INTERP>(cfg, f, last, next, stop_rule);
        if (a != 0) {
          exited = a == 1;
          break;
        }
        stream_store(&out(base + it), next);
        last = next;
        ++steps;
      }
        if (R < 2) {
        used = base + 1;
        break;
      }
      out(base + it) = termination_vertex();
      used = base + it + 1;
      if (!exited) break;
      int nb;
      if (b < 0) {
        const int k = face_rule(f, last, cfg.local_mask);
        if (k < 0 || !cfg.has_field[k + 1]) break;
        nb = k;
      } else {
        const int d = face_rule(f, last, cfg.face_mask[b]);
        if (d != (b ^ 1)) break;
        nb = -1;
      }
      const int Rn = R - it;
      if (used + Rn > region) {
        pa_particle c;
        c.position[0] = last.x;
        c.position[1] = last.y;
        c.position[2] = last.z;
        c.position[3] = last.w;
        c.remaining_iterations = Rn;
        c.vector_field_index = nb;
        c.original_rank = p.original_rank;
        c.original_voxel[0] = int32_t(pi);
        c.original_voxel[1] = 0;
        c.original_voxel[2] = 0;
        c.reserved_[0] = c.reserved_[1] = 0;
        cont[atomicAdd(n_cont, 1u)] = c;
        break;
      }
      out(used) = last;
      base = used;
      R = Rn;
      b = nb;
    }
    counts[pi] = int32_t(used);
    if (endpoints) endpoints[pi] = last;
    if (seed_steps) seed_steps[pi] = int32_t(steps);
  }
  steps = wave_sum_u64(steps);
  if ((threadIdx.x & 63) == 0 && steps) atomicAdd(steps_out, steps);
}

// vector_field::contains + interpolate at arbitrary points.
template <int S, int INTERP>
__global__ __launch_bounds__(kBlock) void k_interpolate(const FieldDesc f, const float4* __restrict__ pts, int64_t n,
                                                        float4* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 p = pts[i];
  float d[3], q[3], v[3] = {0.0f, 0.0f, 0.0f};
  const bool inside = locate(f, p.x, p.y, p.z, d, q);
  if (inside) interpolate<S, INTERP>(f, d, q, v);
  out[i] = make_float4(v[0], v[1], v[2], inside ? 1.0f : 0.0f);
}

// ----------------------------------------------------------------- advect
// Whether advect_one keeps each lane's last cell's corner data in LDS
// (advance_kept): for cell records and FP32 nodes.  A lane reads memory only
// when its particle enters another cell (the FTLE flow map on ABC 256^3 FP32
// nodes: 69.6 -> 50.5 ms).  16-bit nodes gather from the lines that their
// neighbouring lanes read as well, and the kept copy costs registers and
// occupancy (56 -> 91 VGPRs): the same flow map on BF16 nodes went
// 57.3 -> 61.7 ms, so they gather every step.
template <int S, int MODE>
constexpr bool lane_keeps() {
  return keeps_record<S, MODE>() && (layout_of(S) == PA_LAYOUT_CELLS || storage_of(S) == PA_STORAGE_FP32);
}

// particle_advector.cpp:61-141 in place.  state 1 inactive, 2 hand-off to
// slot, 3 never stepped (R <= 1).  keep: the block's
// [lane_keeps ? keep_words : 1][kBlock] LDS array.
template <int S, int MODE, int INTERP>
__device__ __forceinline__ int advect_one(const TraceConfig& cfg, const FieldDesc& f, float4& pos, int& R, int& slot,
                                          unsigned long long& steps, int stop_rule, uint4 (*keep)[kBlock]) {
  int state = 3;
  slot = -1;
  uint32_t kept = 0xffffffffu;  // the cell whose corners keep holds: none yet
  for (int it = 1; it < R; ++it) {
    float4 next;
    int a;
    if constexpr (lane_keeps<S, MODE>())
      a = advance_kept<S, MODE, INTERP, kBlock>(cfg, f, pos, next, stop_rule, kept, keep);
    else
      a = stop_rule == PA_STOP_NORM ? advance<S, MODE, INTERP, PA_STOP_NORM>(cfg, f, pos, next)
                                    : advance<S, MODE, INTERP, PA_STOP_ISZERO>(cfg, f, pos, next);
    if (a == 1) {  // :76-100
      R -= it;
      const int k = face_rule(f, pos, cfg.local_mask);
      if (k < 0) {
        R = 0;
        state = 1;
      } else {
        state = 2;
        slot = k;
      }
      return state;
    }
    if (a == 2) {  // :102-109
      R = 0;
      return 1;
    }
    pos = next;
    ++steps;
    if (it + 1 == R) {  // :132-138
      R = 0;
      return 1;
    }
  }
  return state;
}

template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kBlock) void k_advect(const TraceConfig cfg, pa_particle* __restrict__ P, int64_t n,
                                                   int32_t* __restrict__ state, int32_t* __restrict__ slot,
                                                   unsigned long long* __restrict__ steps_out, int stop_rule) {
  // each lane's last cell's corners (the particles are unrelated after a few steps)
  __shared__ uint4 keep[lane_keeps<S, MODE>() ? keep_words<S>() : 1][kBlock];  // each lane's last cell's corners
  const int64_t pi = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long steps = 0;
  if (pi < n) {
    pa_particle p = P[pi];
    float4 pos = make_float4(p.position[0], p.position[1], p.position[2], p.position[3]);
    int R = p.remaining_iterations;
    int sl;
    const int st = advect_one<S, MODE, INTERP>(cfg, cfg.field[0], pos, R, sl, steps, stop_rule, keep);
    p.position[0] = pos.x;
    p.position[1] = pos.y;
    p.position[2] = pos.z;
    p.position[3] = pos.w;
    p.remaining_iterations = R;
    P[pi] = p;
    state[pi] = st;
    slot[pi] = sl;
  }
  steps = wave_sum_u64(steps);
  if ((threadIdx.x & 63) == 0 && steps) atomicAdd(steps_out, steps);
}

// flow_map_generator.cpp:39-52 fused: the seed of lattice point i is made
// in the lane, advected, and its final position scattered to its voxel.
__global__ __launch_bounds__(kBlock) void k_flow_init(float* __restrict__ fm, int64_t n, int d1, int d2, float o0,
                                                      float o1, float o2, float s0, float s1, float s2) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int k = int(i % d2), j = int((i / d2) % d1), x = int(i / (int64_t(d1) * d2));
  fm[3 * i] = o0 + s0 * float(x);
  fm[3 * i + 1] = o1 + s1 * float(j);
  fm[3 * i + 2] = o2 + s2 * float(k);
}

template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kBlock) void k_flow_map(const TraceConfig cfg, int64_t n, int sd1, int sd2, float o0,
                                                     float o1, float o2, float s0, float s1, float s2, int32_t R0,
                                                     float* __restrict__ fm, int fd0, int fd1, int fd2,
                                                     unsigned long long* __restrict__ steps_out, int stop_rule) {
  __shared__ uint4 keep[lane_keeps<S, MODE>() ? keep_words<S>() : 1][kBlock];  // each lane's last cell's corners
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long steps = 0;
  if (i < n) {
    const int k = int(i % sd2), j = int((i / sd2) % sd1), x = int(i / (int64_t(sd1) * sd2));
    float4 pos = make_float4(o0 + s0 * float(x), o1 + s1 * float(j), o2 + s2 * float(k), 0.0f);
    int R = R0, sl;
    const int st = advect_one<S, MODE, INTERP>(cfg, cfg.field[0], pos, R, sl, steps, stop_rule, keep);
    if (st == 1 && x < fd0 && j < fd1 && k < fd2) {
      const int64_t at = (int64_t(x) * fd1 + j) * fd2 + k;
      fm[3 * at] = pos.x;
      fm[3 * at + 1] = pos.y;
      fm[3 * at + 2] = pos.z;
    }
  }
  steps = wave_sum_u64(steps);
  if ((threadIdx.x & 63) == 0 && steps) atomicAdd(steps_out, steps);
}

// Cell-record gathers shared by a wave.  Loaded lane by lane, one wave load
// instruction touches 64 cache lines (one record per lane); here wave load i
// covers chunks 64 i .. 64 i + 63 of the wave's records in order (lane l:
// chunk (64 i + l) % U of record (64 i + l) / U), so it touches about 64 / U
// lines, and goes straight to LDS (no staging registers); every lane then
// reads its own record from LDS.  Only the chunks of records that are wanted
// load (the others' lanes are masked off): a lane whose record is not
// wanted keeps what its LDS slot holds, e.g. the record of the cell it was
// in at the last step.
template <int S>
struct CoopRecords {
  uint64_t addr[64];
  uint4 rec[64][rec_used<S>()];
};

// After the call lane l's record is sh.rec[l]; the lane reads it before the
// wave's next gather rewrites it (a wave executes in order).
template <int S>
__device__ __forceinline__ void coop_gather(CoopRecords<S>& sh, bool want, const uint4* mine, int lane) {
  constexpr int U = rec_used<S>();
  const uint64_t W = __ballot(want);
  if (W == 0) return;
  sh.addr[lane] = reinterpret_cast<uint64_t>(mine);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  auto* dst = (__attribute__((address_space(3))) uint4*)(&sh.rec[0][0]);
#pragma unroll
  for (int i = 0; i < U; ++i) {
    const int g = i * 64 + lane, src = g / U, c = g % U;
    if ((W >> src) & 1)
      __builtin_amdgcn_global_load_lds(
          reinterpret_cast<const __attribute__((address_space(1))) void*>(sh.addr[src] + 16u * c), dst + i * 64, 16,
          0, 0);
  }
  // Loads into LDS count in vmcnt and no fence waits for them: wait here
  // (vmcnt 0; expcnt and lgkmcnt left at their maxima) before any lane
  // reads a record.
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// k_flow_map for cell records with one gather per step: the advection of
// advect_one (particle_advector.cpp:61-141) as a wave-level loop, so the
// wave can gather its records together (coop_gather).  Same arithmetic,
// same results bit for bit.
template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kBlock) void k_flow_map_cells(const TraceConfig cfg, int64_t n, int sd1, int sd2,
                                                           float o0, float o1, float o2, float s0, float s1, float s2,
                                                           int32_t R0, float* __restrict__ fm, int fd0, int fd1,
                                                           int fd2, unsigned long long* __restrict__ steps_out,
                                                           int stop_rule) {
  __shared__ CoopRecords<S> sh[kBlock / 64];
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const FieldDesc& f = cfg.field[0];
  unsigned long long steps = 0;
  int x = 0, j = 0, k = 0;
  float4 pos = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int R = R0, it = 1;
  int state = 3;  // advect_one's states: 1 inactive (written), 2 hand-off, 3 never stepped; 0 running
  // the cell whose record the lane's LDS slot holds (fewer than 2^31 records: no cell is 0xffffffff)
  uint32_t kept = 0xffffffffu;
  if (i < n) {
    k = int(i % sd2);
    j = int((i / sd2) % sd1);
    x = int(i / (int64_t(sd1) * sd2));
    pos = make_float4(o0 + s0 * float(x), o1 + s1 * float(j), o2 + s2 * float(k), 0.0f);
    if (R > 1) state = 0;
  }
  for (;;) {
    const bool run = state == 0;
    if (__ballot(run) == 0) break;
    float d[3], q[3];
    bool inside = false;
    const uint4* mine = nullptr;
    bool need = false;  // inside a cell other than the kept one
    if (run) {
      inside = locate(f, pos.x, pos.y, pos.z, d, q);
      if (inside) {
        int m[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) m[a] = static_cast<int>(floorf(q[a]));
        const uint32_t ci =
            uint32_t(m[0]) * uint32_t(f.cstride0) + uint32_t(m[1]) * uint32_t(f.cstride1) + uint32_t(m[2]);
        need = ci != kept;
        kept = ci;
        mine = cell_record<S>(f, m);
      }
    }
    coop_gather<S>(sh[w], need, mine, lane);
    if (run) {
      float4 next;
      int a = 1;
      if (inside) {
        float kv[24], v[3];
        decode_record<S>(sh[w].rec[lane], kv);
        lerp_planar<INTERP>(f, d, q, kv, v);
        a = finish_step<S, MODE, INTERP, STOP_RUNTIME>(cfg, f, pos, v, next, stop_rule);
      }
      if (a == 1) {  // :76-100
        R -= it;
        state = face_rule(f, pos, cfg.local_mask) < 0 ? 1 : 2;
      } else if (a == 2) {  // :102-109
        state = 1;
      } else {
        pos = next;
        ++steps;
        if (it + 1 == R) state = 1;  // :132-138
        else ++it;
      }
    }
  }
  if (i < n && state == 1 && x < fd0 && j < fd1 && k < fd2) {
    const int64_t at = (int64_t(x) * fd1 + j) * fd2 + k;
    fm[3 * at] = pos.x;
    fm[3 * at + 1] = pos.y;
    fm[3 * at + 2] = pos.z;
  }
  steps = wave_sum_u64(steps);
  if (lane == 0 && steps) atomicAdd(steps_out, steps);
}

// ------------------------------------------------------------ dispatching
int mode_of(int32_t integrator) {
  if (integrator == PA_INTEGRATOR_EULER) return STEP_EULER;
  if (integrator == PA_INTEGRATOR_CLASSIC_RK4) return STEP_CLASSIC;
  return STEP_RK4;
}

unsigned grid_of(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }

// Validates the particle set (vector field indices name loaded blocks,
// remaining iterations non-negative); returns max remaining iterations.
pa_status scan_particles(pa_gpu_ctx* ctx, const TraceConfig& cfg, const pa_particle* d_particles, size_t n,
                         int* max_r) {
  int* c = static_cast<int*>(ctx->counters);
  hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(1), 0, ctx->stream, c);
  int64_t g = (int64_t(n) + kBlock - 1) / kBlock;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_max_remaining, dim3(unsigned(g)), dim3(kBlock), 0, ctx->stream, d_particles, int64_t(n), c,
                     c + 1, c + 2);
  PA_HIP(ctx, hipGetLastError());
  int h[3];
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(h, c, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h[1] < -1 || h[2] > 5) return fail(ctx, PA_E_INVALID, "vector_field_index outside -1..5");
  for (int k = h[1]; k <= h[2]; ++k)
    if (k >= 0 && !cfg.has_field[k + 1])
      return fail(ctx, PA_E_STATE, "a particle names a neighbour block that is not loaded");
  if (h[1] == -1 && !cfg.has_field[0]) return fail(ctx, PA_E_STATE, "no local vector field");
  *max_r = h[0];
  return PA_OK;
}

}  // namespace

pa_status make_config(pa_gpu_tracer* t, TraceConfig* cfg) {
  pa_gpu_ctx* ctx = t->ctx;
  std::memset(cfg, 0, sizeof *cfg);
  int storage = -1;
  auto put = [&](int slot, const pa_gpu_field* f) -> pa_status {
    if (!f) return PA_OK;
    const int code = f->desc.storage + 3 * f->desc.layout;
    if (storage >= 0 && code != storage)
      return fail(ctx, PA_E_UNSUPPORTED, "all blocks of one tracer must share one storage type and layout");
    storage = code;
    cfg->field[slot] = f->desc;
    cfg->has_field[slot] = 1;
    return PA_OK;
  };
  PA_TRY(put(0, t->local));
  for (int k = 0; k < 6; ++k) PA_TRY(put(k + 1, t->neighbors[k]));
  if (storage < 0) return fail(ctx, PA_E_STATE, "no vector field set");
  if (!t->local) cfg->field[0] = cfg->field[1];  // keep a valid descriptor for unused slot 0
  for (int k = 0; k < 6; ++k)
    if (!t->neighbors[k]) cfg->field[k + 1] = t->local ? t->local->desc : cfg->field[0];
  cfg->local_mask = 0;
  for (int k = 0; k < 6; ++k) {
    if (t->neighbor_rank[k] >= 0) cfg->local_mask |= 1 << k;
    cfg->face_mask[k] = t->face_mask[k];
  }
  fill_step_constants(t, cfg);
  return PA_OK;
}

// The sum w + a_0 0 + a_1 0 + ... (left to right, one rounding each) that a
// frozen-derivative step takes of the w component equals w + z for every w:
// each term a_i 0 is +0, -0 (a_i < 0) or NaN (a_i infinite or NaN); adding a
// signed zero leaves every w but -0 unchanged, and -0 stays -0 only while
// every term added is -0.  So z = NaN if a term is NaN, -0 if all are -0,
// else +0.
static float zero_terms(const float* a, int n) {
  bool all_neg = true;
  for (int i = 0; i < n; ++i) {
    const float t = a[i] * 0.0f;
    if (t != t) return std::numeric_limits<float>::quiet_NaN();
    if (!std::signbit(t)) all_neg = false;
  }
  return all_neg ? -0.0f : 0.0f;
}

void fill_step_constants(const pa_gpu_tracer* t, TraceConfig* cfg) {
  const float dt = t->step_size;
  // odeint runge_kutta4 tableau in value_type float: b = {1/6, 1/3, 1/3, 1/6}.
  const float b0 = 1.0f / 6.0f, b1 = 1.0f / 3.0f, b2 = 1.0f / 3.0f, b3 = 1.0f / 6.0f;
  cfg->dt = dt;
  cfg->a0 = b0 * dt;
  cfg->a1 = b1 * dt;
  cfg->a2 = b2 * dt;
  cfg->a3 = b3 * dt;
  cfg->half_dt = (1.0f / 2.0f) * dt;
  cfg->zero_dt = 0.0f * dt;
  cfg->one_dt = 1.0f * dt;
  cfg->stop_eps2 = t->stop_eps * t->stop_eps;
  const float rk4[4] = {cfg->a0, cfg->a1, cfg->a2, cfg->a3};
  cfg->w_rk4 = zero_terms(rk4, 4);
  cfg->w_euler = zero_terms(&cfg->dt, 1);
  cfg->interp = t->interp;
  cfg->xcd_remap = t->xcd_remap;
  cfg->quad_switch = t->quad_switch;
  cfg->pair_live = t->pair_live;
}

int mode_of_integrator(int integrator) { return mode_of(integrator); }

pa_status check_integrator_of(pa_gpu_ctx* ctx, int integ) {
  if (integ == PA_INTEGRATOR_EULER || integ == PA_INTEGRATOR_RUNGE_KUTTA_4 || integ == PA_INTEGRATOR_CLASSIC_RK4)
    return PA_OK;
  if (integ >= 0 && integ <= 7)
    return fail(ctx, PA_E_UNSUPPORTED, "only euler, runge_kutta_4 and classic RK4 run on the GPU");
  return fail(ctx, PA_E_INVALID, "unknown integrator");
}

}  // namespace pa_amd

using namespace pa_amd;

namespace {

pa_status check_integrator(pa_gpu_ctx* ctx, int32_t integ) {
  if (integ == PA_INTEGRATOR_EULER || integ == PA_INTEGRATOR_RUNGE_KUTTA_4 || integ == PA_INTEGRATOR_CLASSIC_RK4)
    return PA_OK;
  if (integ >= 0 && integ <= 7)
    return fail(ctx, PA_E_UNSUPPORTED, "only euler, runge_kutta_4 and classic RK4 run on the GPU");
  return fail(ctx, PA_E_INVALID, "unknown integrator");
}

}  // namespace

extern "C" {

pa_status pa_gpu_seed_count(const float size[3], const float stride[3], int32_t dims_out[3], size_t* n_out) {
  if (!size || !stride || !dims_out || !n_out) return PA_E_INVALID;
  int64_t n = 1;
  for (int i = 0; i < 3; ++i) {
    const float q = size[i] / stride[i];  // seed_generator.cpp:11
    if (!(q >= 0.0f) || q >= 2147483648.0f) return PA_E_INVALID;
    dims_out[i] = int32_t(q);
    n *= dims_out[i];
  }
  *n_out = size_t(n);
  return PA_OK;
}

pa_status pa_gpu_seed_generate(pa_gpu_ctx* ctx, const float offset[3], const float size[3], const float stride[3],
                               int32_t remaining_iterations, int32_t rank, pa_particle* d_out, size_t capacity,
                               size_t* n_out) {
  if (!ctx || !offset || !size || !stride || !n_out) return PA_E_INVALID;
  int32_t d[3];
  size_t n = 0;
  if (pa_gpu_seed_count(size, stride, d, &n) != PA_OK) return fail(ctx, PA_E_INVALID, "bad seed lattice");
  *n_out = n;
  if (n == 0) return PA_OK;
  if (!d_out || capacity < n) return fail(ctx, PA_E_CAPACITY, "seed buffer too small");
  launch_begin(ctx, "seeds");
  hipLaunchKernelGGL(k_seeds, dim3(grid_of(int64_t(n))), dim3(kBlock), 0, ctx->stream, d_out, int64_t(n), d[1], d[2],
                     offset[0], offset[1], offset[2], stride[0], stride[1], stride[2], remaining_iterations, rank);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_tracer_create(pa_gpu_ctx* ctx, pa_gpu_tracer** out) {
  if (!ctx || !out) return PA_E_INVALID;
  *out = new pa_gpu_tracer();
  (*out)->ctx = ctx;
  return PA_OK;
}

pa_status pa_gpu_tracer_destroy(pa_gpu_tracer* t) {
  if (!t) return PA_E_INVALID;
  delete t;
  return PA_OK;
}

pa_status pa_gpu_tracer_set_local_vector_field(pa_gpu_tracer* t, const pa_gpu_field* f) {
  if (!t) return PA_E_INVALID;
  t->local = f;
  return PA_OK;
}

pa_status pa_gpu_tracer_set_neighbor_vector_fields(pa_gpu_tracer* t, const pa_gpu_field* const fields[6],
                                                   const int32_t neighbor_rank[6]) {
  if (!t) return PA_E_INVALID;
  for (int k = 0; k < 6; ++k) {
    t->neighbors[k] = fields ? fields[k] : nullptr;
    t->neighbor_rank[k] = neighbor_rank ? neighbor_rank[k] : -1;
  }
  return PA_OK;
}

pa_status pa_gpu_tracer_set_integrator(pa_gpu_tracer* t, pa_integrator integrator) {
  if (!t) return PA_E_INVALID;
  PA_TRY(check_integrator(t->ctx, integrator));
  t->integrator = integrator;
  return PA_OK;
}

pa_status pa_gpu_tracer_set_step_size(pa_gpu_tracer* t, float step_size) {
  if (!t) return PA_E_INVALID;
  t->step_size = step_size;
  return PA_OK;
}

pa_status pa_gpu_tracer_set_stop_rule(pa_gpu_tracer* t, pa_stop_rule rule, float eps) {
  if (!t) return PA_E_INVALID;
  if (rule != PA_STOP_ISZERO && rule != PA_STOP_NORM) return fail(t->ctx, PA_E_INVALID, "unknown stop rule");
  t->stop_rule = rule;
  t->stop_eps = eps;
  return PA_OK;
}

pa_status pa_gpu_tracer_set_engine(pa_gpu_tracer* t, pa_engine engine, int32_t waves) {
  if (!t) return PA_E_INVALID;
  if (engine != PA_ENGINE_LANE && engine != PA_ENGINE_COOPERATIVE && engine != PA_ENGINE_QUAD &&
      engine != PA_ENGINE_HYBRID)
    return fail(t->ctx, PA_E_INVALID, "unknown engine");
  if (waves < 0) return fail(t->ctx, PA_E_INVALID, "waves must be >= 0");
  t->engine = engine;
  t->waves = waves;
  return PA_OK;
}

pa_status pa_gpu_tracer_set_xcd_remap(pa_gpu_tracer* t, int32_t enable) {
  if (!t) return PA_E_INVALID;
  if (enable < 0 || enable > 64) return fail(t->ctx, PA_E_INVALID, "xcd order: 0, 1 or a run length 2..64");
  t->xcd_remap = enable;
  return PA_OK;
}

pa_status pa_gpu_tracer_set_launch(pa_gpu_tracer* t, int32_t quad_switch, int32_t pair_live) {
  if (!t) return PA_E_INVALID;
  if (quad_switch < 1 || quad_switch > 16 || pair_live < 0 || pair_live > 64)
    return fail(t->ctx, PA_E_INVALID, "launch constants: quad_switch 1..16, pair_live 0..64");
  t->quad_switch = quad_switch;
  t->pair_live = pair_live;
  return PA_OK;
}

pa_status pa_gpu_tracer_set_interpolation(pa_gpu_tracer* t, pa_interp interp) {
  if (!t) return PA_E_INVALID;
  if (interp != PA_INTERP_REFERENCE && interp != PA_INTERP_CONSISTENT)
    return fail(t->ctx, PA_E_INVALID, "unknown interpolation mode");
  t->interp = interp;
  return PA_OK;
}

pa_status pa_gpu_field_interpolate(pa_gpu_ctx* ctx, const pa_gpu_field* field, pa_interp interp,
                                   const float* d_points, size_t n, float* d_out) {
  if (!ctx || !field) return PA_E_INVALID;
  if (n == 0) return PA_OK;
  if (!d_points || !d_out) return PA_E_INVALID;
  if (interp != PA_INTERP_REFERENCE && interp != PA_INTERP_CONSISTENT)
    return fail(ctx, PA_E_INVALID, "unknown interpolation mode");
  launch_begin(ctx, "interpolate");
  dispatch(field->desc.storage, STEP_RK4, interp, [&](auto S, auto M, auto I) {
    if constexpr (decltype(M)::value == STEP_RK4)
      hipLaunchKernelGGL((k_interpolate<decltype(S)::value, decltype(I)::value>), dim3(grid_of(int64_t(n))),
                         dim3(kBlock), 0, ctx->stream, field->desc, reinterpret_cast<const float4*>(d_points),
                         int64_t(n), reinterpret_cast<float4*>(d_out));
  }, field->desc.layout);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_tracer_set_neighbor_face_masks(pa_gpu_tracer* t, const int32_t face_mask[6]) {
  if (!t || !face_mask) return PA_E_INVALID;
  for (int k = 0; k < 6; ++k) t->face_mask[k] = face_mask[k] & 63;
  return PA_OK;
}

pa_status pa_gpu_tracer_compute_round_info(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n,
                                           size_t integral_curve_offset, pa_round_info* out) {
  if (!t || !out) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (n == 0) return fail(ctx, PA_E_INVALID, "compute_round_info of an empty particle set");
  if (!d_particles) return PA_E_INVALID;
  TraceConfig cfg;
  PA_TRY(make_config(t, &cfg));
  int maxr = 0;
  PA_TRY(scan_particles(ctx, cfg, d_particles, n, &maxr));
  if (maxr <= 0) return fail(ctx, PA_E_INVALID, "maximum remaining_iterations must be positive");
  const uint64_t max_vertices = uint64_t(INT32_MAX) / 16u;  // :216, sizeof(vector4)
  const uint64_t ppic = max_vertices / uint64_t(maxr);
  out->maximum_remaining_iterations = uint64_t(maxr);
  out->vertices_per_integral_curve = ppic * uint64_t(maxr);
  out->vertices_per_integral_curve_last = (uint64_t(n) % ppic) * uint64_t(maxr);
  out->integral_curve_offset = integral_curve_offset;
  out->integral_curve_count = 1 + uint64_t(n) / ppic;
  return PA_OK;
}

namespace {
__global__ void k_fill_invalid(float4* v, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    v[i] = make_float4(-1.0f, -1.0f, -1.0f, 0.0f);
}
}  // namespace

pa_status pa_gpu_fill_invalid(pa_gpu_ctx* ctx, float* d_vertices, size_t count) {
  if (!ctx || (count && !d_vertices)) return PA_E_INVALID;
  if (!count) return PA_OK;
  int64_t g = (int64_t(count) + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  launch_begin(ctx, "fill_invalid");
  hipLaunchKernelGGL(k_fill_invalid, dim3(unsigned(g)), dim3(kBlock), 0, ctx->stream,
                     reinterpret_cast<float4*>(d_vertices), int64_t(count));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_tracer_trace(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n, const pa_round_info* info,
                              float* const* d_chunks, int32_t* d_vertex_counts, pa_handoff* handoff,
                              uint64_t* total_steps) {
  if (!t || !info) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (n == 0) {
    if (handoff) std::memset(handoff->counts, 0, sizeof handoff->counts);
    if (total_steps) *total_steps = 0;
    return PA_OK;
  }
  if (!d_particles || !d_chunks) return PA_E_INVALID;
  PA_TRY(check_integrator(ctx, t->integrator));
  TraceConfig cfg;
  PA_TRY(make_config(t, &cfg));
  int maxr = 0;
  PA_TRY(scan_particles(ctx, cfg, d_particles, n, &maxr));
  if (uint64_t(maxr) > info->maximum_remaining_iterations || info->maximum_remaining_iterations == 0)
    return fail(ctx, PA_E_INVALID, "round_info does not cover the particles");
  const uint64_t maxR = info->maximum_remaining_iterations;
  const uint64_t ppic = info->vertices_per_integral_curve / maxR;
  if (ppic == 0 || info->integral_curve_count < 1 + (uint64_t(n) - 1) / ppic)
    return fail(ctx, PA_E_INVALID, "round_info chunk geometry does not cover the particles");
  if (handoff && (!handoff->d_particles || handoff->capacity < n))
    return fail(ctx, PA_E_CAPACITY, "hand-off buffer must hold one entry per particle");
  // Device copy of the chunk pointer table + hand-off scratch.
  const size_t nchunks = size_t(info->integral_curve_count);
  const size_t table = ((nchunks * sizeof(void*) + 255) / 256) * 256;
  const size_t tmp_bytes = handoff ? ((n * (sizeof(pa_particle) + sizeof(int32_t)) + 255) / 256) * 256 : 0;
  const size_t group_bytes = handoff ? group_ws_bytes(int64_t(n)) : 0;
  void* ws = nullptr;
  PA_TRY(scratch(ctx, table + tmp_bytes + group_bytes, &ws));
  char* w = static_cast<char*>(ws);
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(w, d_chunks, nchunks * sizeof(void*), hipMemcpyHostToDevice, ctx->stream));
  pa_particle* tmp = handoff ? reinterpret_cast<pa_particle*>(w + table) : nullptr;
  int32_t* code = handoff ? reinterpret_cast<int32_t*>(w + table + n * sizeof(pa_particle)) : nullptr;
  auto* steps = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 1024);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, steps, 1);
  const int mode = mode_of(t->integrator);
  if (t->engine == PA_ENGINE_COOPERATIVE) {
    CoopLaunch L;
    L.particles = d_particles;
    L.n = int64_t(n);
    L.counts = d_vertex_counts;
    L.stop_rule = t->stop_rule;
    L.waves = t->waves;
    L.chunks = reinterpret_cast<float4* const*>(w);
    L.ppic = int64_t(ppic);
    L.maxR = int64_t(maxR);
    L.hand_tmp = tmp;
    L.hand_code = code;
    if (!d_vertex_counts) {
      // the cooperative engine always records counts; keep them in scratch
      return fail(ctx, PA_E_INVALID, "the cooperative engine needs d_vertex_counts");
    }
    PA_TRY(launch_coop(ctx, cfg, mode, false, L));
  } else {
  launch_begin(ctx, "trace");
  dispatch(cfg.field[0].storage, mode, cfg.interp, [&](auto S, auto M, auto I) {
    hipLaunchKernelGGL((k_trace<decltype(S)::value, decltype(M)::value, decltype(I)::value>), dim3(grid_of(int64_t(n))),
                       dim3(kBlock), 0, ctx->stream, cfg, d_particles, int64_t(n),
                       reinterpret_cast<float4* const*>(w), int64_t(ppic), int64_t(maxR), d_vertex_counts, tmp, code,
                       steps, t->stop_rule);
  }, cfg.field[0].layout);
  launch_end(ctx);
  }
  PA_HIP(ctx, hipGetLastError());
  if (handoff) PA_TRY(group_handoffs(ctx, code, tmp, int64_t(n), handoff, w + table + tmp_bytes));
  unsigned long long s = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&s, steps, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (total_steps) *total_steps = s;
  return PA_OK;
}

pa_status pa_gpu_tracer_trace_union(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n, size_t region_slots,
                                    float* d_vertices, int32_t* d_vertex_counts, pa_particle* d_continuation,
                                    size_t* n_continuation, float* d_endpoints, int32_t* d_seed_steps,
                                    uint64_t* total_steps) {
  return pa_gpu_tracer_trace_union_layout(t, d_particles, n, region_slots, PA_SLOTS_SEED_MAJOR, d_vertices,
                                          d_vertex_counts, d_continuation, n_continuation, d_endpoints, d_seed_steps,
                                          total_steps);
}

pa_status pa_gpu_tracer_trace_union_layout(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n,
                                           size_t region_slots, pa_slot_order order, float* d_vertices,
                                           int32_t* d_vertex_counts, pa_particle* d_continuation,
                                           size_t* n_continuation, float* d_endpoints, int32_t* d_seed_steps,
                                           uint64_t* total_steps) {
  if (!t) return PA_E_INVALID;
  return trace_union_impl(t, t->engine, d_particles, n, region_slots, order, d_vertices, d_vertex_counts,
                          d_continuation, n_continuation, d_endpoints, d_seed_steps, total_steps, nullptr, false);
}

pa_status pa_gpu_tracer_trace_union_round(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n,
                                          size_t region_slots, float* d_slots, int32_t* d_vertex_counts,
                                          pa_particle* d_continuation, size_t* n_continuation, pa_color_mode mode,
                                          float parameter, float* d_vertices, size_t vertex_capacity,
                                          size_t* n_vertices, int32_t* d_indices, size_t index_capacity,
                                          size_t* n_indices, float* d_colors, float* d_endpoints,
                                          int32_t* d_seed_steps, uint64_t* total_steps) {
  if (!t || !n_vertices || !n_indices) return PA_E_INVALID;
  void* ws = nullptr;
  if (n > 0 && n <= size_t(INT32_MAX)) PA_TRY(scratch(t->ctx, union_round_ws_bytes(int64_t(n)), &ws));
  return union_round(t, d_particles, n, region_slots, d_slots, d_vertex_counts, d_continuation, n_continuation, mode,
                     parameter, d_vertices, vertex_capacity, n_vertices, d_indices, index_capacity, n_indices,
                     d_colors, d_endpoints, d_seed_steps, total_steps, ws, nullptr);
}

pa_status pa_gpu_tracer_trace_union_round_async(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n,
                                                size_t region_slots, float* d_slots, int32_t* d_vertex_counts,
                                                pa_particle* d_continuation, pa_color_mode mode, float parameter,
                                                float* d_vertices, size_t vertex_capacity, int32_t* d_indices,
                                                size_t index_capacity, float* d_colors, float* d_endpoints,
                                                int32_t* d_seed_steps, pa_round_result* d_result) {
  if (!t || !d_result) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  void* ws = nullptr;
  if (n > 0 && n <= size_t(INT32_MAX)) {
    const size_t need = union_round_ws_bytes(int64_t(n));
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    PA_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone && need > ctx->scratch_bytes)
      return fail(ctx, PA_E_STATE, "round under capture: run it once uncaptured first (its workspace must exist)");
    PA_TRY(scratch(ctx, need, &ws));
  }
  return union_round(t, d_particles, n, region_slots, d_slots, d_vertex_counts, d_continuation, nullptr, mode,
                     parameter, d_vertices, vertex_capacity, nullptr, d_indices, index_capacity, nullptr, d_colors,
                     d_endpoints, d_seed_steps, nullptr, ws, d_result);
}

pa_status pa_gpu_tracer_trace_union_round_pipelined(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n,
                                                    size_t region_slots, float* d_slots, int32_t* d_vertex_counts,
                                                    pa_particle* d_continuation, pa_color_mode mode, float parameter,
                                                    float* d_vertices, size_t vertex_capacity, int32_t* d_indices,
                                                    size_t index_capacity, float* d_colors, float* d_endpoints,
                                                    int32_t* d_seed_steps, pa_round_result* d_result) {
  if (!t || !d_result) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (ctx->capturing) return fail(ctx, PA_E_STATE, "pipelined rounds run on two streams: not while capturing");
  if (n > size_t(INT32_MAX)) return fail(ctx, PA_E_INVALID, "round: at most 2^31-1 seeds per call");
  if (!ctx->finish_stream) {
    PA_HIP(ctx, hipStreamCreateWithFlags(&ctx->finish_stream, hipStreamNonBlocking));
    for (auto& r : ctx->round_set) {
      PA_HIP(ctx, hipEventCreateWithFlags(&r.traced, hipEventDisableTiming));
      PA_HIP(ctx, hipEventCreateWithFlags(&r.finished, hipEventDisableTiming));
      PA_HIP(ctx, hipMalloc(&r.counters, 4096));
    }
  }
  auto& R = ctx->round_set[ctx->pipelined_rounds & 1];
  // the round before last used this set: its finish must be done before this trace reuses the
  // workspace and counters (and, by the caller's alternation, the buffers)
  if (R.pending) PA_HIP(ctx, hipStreamWaitEvent(ctx->stream, R.finished, 0));
  R.pending = false;
  const size_t need = n ? union_round_ws_bytes(int64_t(n)) : 0;
  if (need > R.ws_bytes) {
    PA_HIP(ctx, hipEventSynchronize(R.finished));
    if (R.ws) PA_HIP(ctx, hipFree(R.ws));
    R.ws = nullptr;
    R.ws_bytes = 0;
    PA_HIP(ctx, hipMalloc(&R.ws, need + need / 4));
    R.ws_bytes = need + need / 4;
  }
  struct CountersRestore {
    pa_gpu_ctx* c;
    void* p;
    ~CountersRestore() { c->counters = p; }
  } restore{ctx, ctx->counters};
  ctx->counters = R.counters;
  PA_TRY(union_round(t, d_particles, n, region_slots, d_slots, d_vertex_counts, d_continuation, nullptr, mode,
                     parameter, d_vertices, vertex_capacity, nullptr, d_indices, index_capacity, nullptr, d_colors,
                     d_endpoints, d_seed_steps, nullptr, n ? R.ws : nullptr, d_result,
                     n ? ctx->finish_stream : nullptr, R.traced));
  if (n) {
    PA_HIP(ctx, hipEventRecord(R.finished, ctx->finish_stream));
    R.pending = true;
  }
  ++ctx->pipelined_rounds;
  return PA_OK;
}

pa_status pa_gpu_round_join(pa_gpu_ctx* ctx) {
  if (!ctx) return PA_E_INVALID;
  return round_join(ctx);
}

pa_status pa_gpu_round_result_check(pa_gpu_ctx* ctx, const pa_round_result* r, size_t vertex_capacity,
                                    size_t index_capacity) {
  if (!ctx || !r) return PA_E_INVALID;
  PA_TRY(union_flags(ctx, r->flags));
  if (size_t(r->n_vertices) > vertex_capacity) return fail(ctx, PA_E_CAPACITY, "round: vertex capacity too small");
  if (size_t(r->n_indices) > index_capacity) return fail(ctx, PA_E_CAPACITY, "round: index capacity too small");
  if (r->n_vertices > INT32_MAX) return fail(ctx, PA_E_CAPACITY, "round: indices are int32 (at most 2^31-1 vertices)");
  return PA_OK;
}

}  // extern "C"

namespace pa_amd {

size_t union_round_ws_bytes(int64_t n) {
  return (size_t(n) * sizeof(int32_t) + 255) / 256 * 256 + finish_ws_bytes(n);
}

namespace {
// The round's result words (counters 1024..1071) into the caller's buffer.
__global__ void k_copy_result(const unsigned long long* __restrict__ src, unsigned long long* __restrict__ dst) {
  if (threadIdx.x < 6) dst[threadIdx.x] = src[threadIdx.x];
}
__global__ void k_zero_result(unsigned long long* __restrict__ dst) {
  if (threadIdx.x < 6) dst[threadIdx.x] = 0ull;
}
}  // namespace

pa_status union_round(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n, size_t region_slots,
                      float* d_slots, int32_t* d_vertex_counts, pa_particle* d_continuation, size_t* n_continuation,
                      pa_color_mode mode, float parameter, float* d_vertices, size_t vertex_capacity,
                      size_t* n_vertices, int32_t* d_indices, size_t index_capacity, size_t* n_indices,
                      float* d_colors, float* d_endpoints, int32_t* d_seed_steps, uint64_t* total_steps, void* ws,
                      pa_round_result* d_result, hipStream_t finish_stream, hipEvent_t traced) {
  if (!t || (!d_result && (!n_vertices || !n_indices))) return PA_E_INVALID;
  if (finish_stream && (!d_result || !traced)) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (n_vertices) *n_vertices = 0;
  if (n_indices) *n_indices = 0;
  if (n == 0) {
    if (n_continuation) *n_continuation = 0;
    if (total_steps) *total_steps = 0;
    if (d_result) {
      hipLaunchKernelGGL(k_zero_result, dim3(1), dim3(64), 0, ctx->stream,
                         reinterpret_cast<unsigned long long*>(d_result));
      PA_HIP(ctx, hipGetLastError());
    }
    return PA_OK;
  }
  if (!d_vertices || !d_indices || !d_colors) return PA_E_INVALID;
  if (mode != PA_COLOR_HSL_CONSTANT_S && mode != PA_COLOR_HSL_CONSTANT_L && mode != PA_COLOR_HSV_CONSTANT_S &&
      mode != PA_COLOR_HSV_CONSTANT_V && mode != PA_COLOR_RGB && mode != PA_COLOR_MAGNITUDE)
    return fail(ctx, PA_E_INVALID, "unknown color mode");
  if (n > size_t(INT32_MAX)) return fail(ctx, PA_E_INVALID, "round: at most 2^31-1 seeds per call");
  if (!ws) return fail(ctx, PA_E_INVALID, "round: no workspace");
  const size_t pair_bytes = (n * sizeof(int32_t) + 255) / 256 * 256;
  auto* pairs = static_cast<int32_t*>(ws);
  // trace and finish are queued back to back; one read-back at the end
  // carries the step count, the continuation count, the validation flags
  // and the two totals
  size_t ncont_ignored = 0;
  PA_TRY(trace_union_impl(t, PA_ENGINE_HYBRID, d_particles, n, region_slots, PA_SLOTS_GROUP8, d_slots,
                          d_vertex_counts, d_continuation, n_continuation ? n_continuation : &ncont_ignored,
                          d_endpoints, d_seed_steps, total_steps, pairs, true));
  char* cbase = static_cast<char*>(ctx->counters);
  auto* totals = reinterpret_cast<int64_t*>(cbase + 1056);
  // A pipelined round finishes on its own stream once the trace is done; the context's stream is free
  // for the next round's trace meanwhile.  Every launch below goes to the finish stream.
  struct StreamRestore {
    pa_gpu_ctx* c;
    hipStream_t s;
    ~StreamRestore() { c->stream = s; }
  } restore{ctx, ctx->stream};
  if (finish_stream) {
    PA_HIP(ctx, hipEventRecord(traced, ctx->stream));
    PA_HIP(ctx, hipStreamWaitEvent(finish_stream, traced, 0));
    ctx->stream = finish_stream;
  }
  PA_TRY(finish_group8(ctx, d_slots, int64_t(n), int64_t(region_slots), d_vertex_counts, pairs, int(mode), parameter,
                       d_vertices, vertex_capacity, d_indices, index_capacity, d_colors,
                       static_cast<char*>(ws) + pair_bytes, totals, reinterpret_cast<unsigned*>(cbase + 1036)));
  if (d_result) {  // no read-back: the caller reads the result words when it needs them
    hipLaunchKernelGGL(k_copy_result, dim3(1), dim3(64), 0, ctx->stream,
                       reinterpret_cast<const unsigned long long*>(cbase + 1024),
                       reinterpret_cast<unsigned long long*>(d_result));
    PA_HIP(ctx, hipGetLastError());
    return PA_OK;
  }
  unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(h, cbase + 1024, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  PA_TRY(union_flags(ctx, unsigned(h[1] >> 32)));
  if (total_steps) *total_steps = h[0];
  *n_continuation = size_t(unsigned(h[1] & 0xffffffffu));
  const int64_t tv = int64_t(h[4]), ti = int64_t(h[5]);
  *n_vertices = size_t(tv);
  *n_indices = size_t(ti);
  if (size_t(tv) > vertex_capacity) return fail(ctx, PA_E_CAPACITY, "round: vertex capacity too small");
  if (size_t(ti) > index_capacity) return fail(ctx, PA_E_CAPACITY, "round: index capacity too small");
  if (tv > INT32_MAX) return fail(ctx, PA_E_CAPACITY, "round: indices are int32 (at most 2^31-1 vertices)");
  return PA_OK;
}

namespace {
pa_status trace_union_impl(pa_gpu_tracer* t, int engine, const pa_particle* d_particles, size_t n,
                           size_t region_slots, pa_slot_order order, float* d_vertices, int32_t* d_vertex_counts,
                           pa_particle* d_continuation, size_t* n_continuation, float* d_endpoints,
                           int32_t* d_seed_steps, uint64_t* total_steps, int32_t* pairs, bool deferred) {
  pa_gpu_ctx* ctx = t->ctx;
  if (n_continuation) *n_continuation = 0;
  if (n == 0) {
    if (total_steps) *total_steps = 0;
    return PA_OK;
  }
  if (!d_particles || !d_vertices || !d_vertex_counts || !d_continuation || !n_continuation) return PA_E_INVALID;
  PA_TRY(check_integrator(ctx, t->integrator));
  TraceConfig cfg;
  PA_TRY(make_config(t, &cfg));
  // The hybrid engine checks the particles itself (field index, loaded
  // block, iterations against region_slots) and reports them with the
  // counters; the others are checked here first.
  if (engine != PA_ENGINE_HYBRID) {
    int maxr = 0;
    PA_TRY(scan_particles(ctx, cfg, d_particles, n, &maxr));
    if (maxr < 0) return fail(ctx, PA_E_INVALID, "negative remaining_iterations");
    if (region_slots < size_t(maxr < 1 ? 1 : maxr))
      return fail(ctx, PA_E_CAPACITY, "region_slots must be at least the maximum remaining_iterations");
  }
  if (region_slots < 1) return fail(ctx, PA_E_CAPACITY, "region_slots must be at least 1");
  if (order != PA_SLOTS_SEED_MAJOR && order != PA_SLOTS_STEP_MAJOR && order != PA_SLOTS_CHUNKED &&
      order != PA_SLOTS_GROUP8)
    return fail(ctx, PA_E_INVALID, "unknown slot order");
  if (order == PA_SLOTS_STEP_MAJOR && engine != PA_ENGINE_LANE)
    return fail(ctx, PA_E_UNSUPPORTED, "step-major slots need the lane engine");
  if ((order == PA_SLOTS_GROUP8) != (engine == PA_ENGINE_HYBRID))
    return fail(ctx, PA_E_UNSUPPORTED, "the hybrid engine writes (only) 8-seed grouped slots");
  if ((order == PA_SLOTS_CHUNKED) != (engine == PA_ENGINE_QUAD))
    return fail(ctx, PA_E_UNSUPPORTED, "the quad engine writes (only) chunked slots");
  const int64_t seed_stride = order == PA_SLOTS_STEP_MAJOR ? 1 : int64_t(region_slots);
  const int64_t slot_stride = order == PA_SLOTS_STEP_MAJOR ? int64_t(n) : 1;
  char* cbase = static_cast<char*>(ctx->counters);
  auto* steps = reinterpret_cast<unsigned long long*>(cbase + 1024);
  auto* ncont = reinterpret_cast<unsigned*>(cbase + 1032);
  const int mode = mode_of(t->integrator);
  if (engine == PA_ENGINE_COOPERATIVE || engine == PA_ENGINE_QUAD || engine == PA_ENGINE_HYBRID) {
    CoopLaunch L;
    L.particles = d_particles;
    L.n = int64_t(n);
    L.counts = d_vertex_counts;
    L.stop_rule = t->stop_rule;
    L.waves = t->waves;
    L.verts = reinterpret_cast<float4*>(d_vertices);
    L.region = int64_t(region_slots);
    L.cont = d_continuation;
    L.endpoints = reinterpret_cast<float4*>(d_endpoints);
    L.seed_steps = d_seed_steps;
    if (engine == PA_ENGINE_HYBRID) {
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, steps, 2);
      launch_begin(ctx, "trace_union");
      const pa_status st = launch_hybrid_union(ctx, cfg, mode, L, steps, ncont, pairs);
      launch_end(ctx);
      if (st != PA_OK) return st;
    } else if (engine == PA_ENGINE_QUAD) {
      auto* work = reinterpret_cast<unsigned long long*>(cbase + 1040);
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, steps, 3);
      launch_begin(ctx, "trace_union");
      const pa_status st = launch_quad_union(ctx, cfg, mode, L, steps, ncont, work);
      launch_end(ctx);
      if (st != PA_OK) return st;
    } else {
      PA_TRY(launch_coop(ctx, cfg, mode, true, L));
    }
  } else {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, steps, 2);
  launch_begin(ctx, "trace_union");
  dispatch(cfg.field[0].storage, mode, cfg.interp, [&](auto S, auto M, auto I) {
    hipLaunchKernelGGL((k_trace_union<decltype(S)::value, decltype(M)::value, decltype(I)::value>),
                       dim3(grid_of(int64_t(n))), dim3(kBlock), 0, ctx->stream, cfg, d_particles, int64_t(n),
                       reinterpret_cast<float4*>(d_vertices), int64_t(region_slots), seed_stride, slot_stride,
                       d_vertex_counts, d_continuation, ncont, steps, t->stop_rule,
                       reinterpret_cast<float4*>(d_endpoints), d_seed_steps);
  }, cfg.field[0].layout);
  launch_end(ctx);
  }
  PA_HIP(ctx, hipGetLastError());
  if (deferred) return PA_OK;
  unsigned long long h[2] = {0, 0};
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(h, steps, 16, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (engine == PA_ENGINE_HYBRID) PA_TRY(union_flags(ctx, unsigned(h[1] >> 32)));
  if (total_steps) *total_steps = h[0];
  *n_continuation = size_t(unsigned(h[1] & 0xffffffffu));
  return PA_OK;
}
}  // namespace
}  // namespace pa_amd

extern "C" {

pa_status pa_gpu_advect(pa_gpu_tracer* t, pa_particle* d_particles, size_t n, int32_t* d_state,
                        int32_t* d_handoff_slot, uint64_t* total_steps) {
  if (!t) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (n == 0) {
    if (total_steps) *total_steps = 0;
    return PA_OK;
  }
  if (!d_particles || !d_state || !d_handoff_slot) return PA_E_INVALID;
  if (!t->local) return fail(ctx, PA_E_STATE, "advect needs the local vector field");
  PA_TRY(check_integrator(ctx, t->integrator));
  TraceConfig cfg;
  PA_TRY(make_config(t, &cfg));
  auto* steps = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 1024);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, steps, 1);
  const int mode = mode_of(t->integrator);
  launch_begin(ctx, "advect");
  dispatch(cfg.field[0].storage, mode, cfg.interp, [&](auto S, auto M, auto I) {
    hipLaunchKernelGGL((k_advect<decltype(S)::value, decltype(M)::value, decltype(I)::value>), dim3(grid_of(int64_t(n))),
                       dim3(kBlock), 0, ctx->stream, cfg, d_particles, int64_t(n), d_state, d_handoff_slot, steps,
                       t->stop_rule);
  }, cfg.field[0].layout);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  unsigned long long s = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&s, steps, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (total_steps) *total_steps = s;
  return PA_OK;
}

pa_status pa_gpu_flow_map_generate(pa_gpu_tracer* t, const int32_t block_size[3], uint64_t iterations,
                                   float resolution_scale, int32_t rank, float* d_flow_map, int32_t dims_out[3],
                                   uint64_t* total_steps) {
  (void)rank;
  if (!t || !block_size || !dims_out) return PA_E_INVALID;
  pa_gpu_ctx* ctx = t->ctx;
  if (!t->local) return fail(ctx, PA_E_STATE, "flow map needs the local vector field");
  for (int k = 0; k < 6; ++k)
    if (t->neighbor_rank[k] >= 0)
      return fail(ctx, PA_E_UNSUPPORTED, "flow map across ranks needs the MPI hand-off (host side)");
  if (!(resolution_scale > 0.0f)) return fail(ctx, PA_E_INVALID, "resolution scale must be positive");
  if (iterations > uint64_t(INT32_MAX)) return fail(ctx, PA_E_INVALID, "iterations exceed int32");
  PA_TRY(check_integrator(ctx, t->integrator));
  TraceConfig cfg;
  PA_TRY(make_config(t, &cfg));
  const FieldDesc& f = t->local->desc;
  float spacing[3];
  int32_t sd[3];
  for (int i = 0; i < 3; ++i) {
    dims_out[i] = int32_t(float(block_size[i]) * resolution_scale);  // flow_map_generator.cpp:247-250
    spacing[i] = f.spacing[i] / resolution_scale;                     // :253-256
  }
  size_t ns = 0;
  if (pa_gpu_seed_count(t->local->size, spacing, sd, &ns) != PA_OK)
    return fail(ctx, PA_E_INVALID, "bad flow-map lattice");
  const int64_t nmap = int64_t(dims_out[0]) * dims_out[1] * dims_out[2];
  if (nmap <= 0) return fail(ctx, PA_E_INVALID, "empty flow map");
  if (!d_flow_map) return PA_E_INVALID;
  auto* steps = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->counters) + 1024);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, steps, 1);
  launch_begin(ctx, "flow_init");
  hipLaunchKernelGGL(k_flow_init, dim3(grid_of(nmap)), dim3(kBlock), 0, ctx->stream, d_flow_map, nmap, dims_out[1],
                     dims_out[2], f.offset[0], f.offset[1], f.offset[2], spacing[0], spacing[1], spacing[2]);
  launch_end(ctx);
  if (ns > 0) {
    const int mode = mode_of(t->integrator);
    launch_begin(ctx, "flow_map");
    dispatch(cfg.field[0].storage, mode, cfg.interp, [&](auto S, auto M, auto I) {
      constexpr int SS = decltype(S)::value, MM = decltype(M)::value, II = decltype(I)::value;
      if constexpr (layout_of(SS) == PA_LAYOUT_CELLS && MM != STEP_CLASSIC)
        hipLaunchKernelGGL((k_flow_map_cells<SS, MM, II>), dim3(grid_of(int64_t(ns))), dim3(kBlock), 0, ctx->stream,
                           cfg, int64_t(ns), sd[1], sd[2], f.offset[0], f.offset[1], f.offset[2], spacing[0],
                           spacing[1], spacing[2], int32_t(iterations), d_flow_map, dims_out[0], dims_out[1],
                           dims_out[2], steps, t->stop_rule);
      else
        hipLaunchKernelGGL((k_flow_map<SS, MM, II>), dim3(grid_of(int64_t(ns))), dim3(kBlock), 0, ctx->stream, cfg,
                           int64_t(ns), sd[1], sd[2], f.offset[0], f.offset[1], f.offset[2], spacing[0], spacing[1],
                           spacing[2], int32_t(iterations), d_flow_map, dims_out[0], dims_out[1], dims_out[2], steps,
                           t->stop_rule);
    }, cfg.field[0].layout);
    launch_end(ctx);
  }
  PA_HIP(ctx, hipGetLastError());
  unsigned long long s = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&s, steps, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (total_steps) *total_steps = s;
  return PA_OK;
}

}  // extern "C"
