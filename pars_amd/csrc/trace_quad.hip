// Quad engine of the union-of-7 tracer: four lanes per particle, lane c of a
// quad owning component c of the particle's vec4 (x, y, z, w).
//
// The step of one particle is a serial chain (gather -> interpolate -> step ->
// locate -> gather), and the longest trajectories of a round set its length.
// Splitting each step over four lanes cuts the arithmetic on that chain by
// about three: every lane divides, floors and lerps along its own axis, and
// the few values that cross axes (subscripts, weights, |v|^2, the face code)
// move between the lanes of the quad with DPP quad permutations.  A cell
// record stores the corners component by component, so the quad reads one
// record as one coalesced line (lane c its quarter).
//
// Waves are persistent: a quad whose particle is done takes the next seed
// from a work counter.  All quads of a wave produce exactly one slot per
// iteration, so the slot index modulo 8 is the same across the wave and a
// chunk of 8 slots is buffered in registers (buf[0..7], statically indexed)
// and written as one 128-B line per quad: component-planar chunks
// x0..x7 y0..y7 z0..z7 w0..w7 (PA_SLOTS_CHUNKED, packed by
// pa_gpu_prune_counted_chunked).  The vertex sequence of every seed is the
// one of the lane engine (particle_tracer.cpp:279-326 per segment, union
// hand-offs as in the lane engine), bit for bit.
#include <hip/hip_runtime.h>

#include "pa_device.hpp"

namespace pa_amd {
namespace {

constexpr int kQuadBlock = 256;  // 4 waves, 64 quads
constexpr int kChunk = 8;        // slots per planar chunk

// DPP quad permutations (quad_perm: lane i of a quad reads lane p_i).
constexpr int kBcast0 = 0x00, kBcast1 = 0x55, kBcast2 = 0xAA;
constexpr int kSwap1 = 0xB1;  // (1,0,3,2)
constexpr int kSwap2 = 0x4E;  // (2,3,0,1)

template <int CTRL>
__device__ __forceinline__ int qmov(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float qmovf(float v) {
  return __int_as_float(qmov<CTRL>(__float_as_int(v)));
}
__device__ __forceinline__ uint32_t quad_sum(uint32_t v) {
  v += static_cast<uint32_t>(qmov<kSwap1>(static_cast<int>(v)));
  v += static_cast<uint32_t>(qmov<kSwap2>(static_cast<int>(v)));
  return v;
}
__device__ __forceinline__ bool quad_all(bool b) {
  int v = b ? 1 : 0;
  v &= qmov<kSwap1>(v);
  v &= qmov<kSwap2>(v);
  return v != 0;
}

// Per-component view of one block (lane c of the quad holds component c;
// the w lane holds neutral values that never fail a test).
struct QEntry {
  float off, h, inv, hi, up;
  uint32_t stride;  // cells: cstride0, cstride1, 1, 0; nodes: stride0, stride1, 1, 0
  uint32_t st0, st1;
};
struct QBlock {
  const void* data;  // cell records or nodes
  int32_t fast;
  int32_t mask;      // face mask: own block -> local_mask, neighbour k -> face_mask[k]
  int32_t has;
  int32_t pad_;
};

struct QLane {
  const void* data;
  float off, h, inv, hi, up;
  uint32_t stride, st0, st1;
  int32_t fast, mask;
};

__device__ __forceinline__ void load_lane(const QEntry (*tab)[4], const QBlock* blk, int b, int c, QLane& L) {
  const QEntry& e = tab[b + 1][c];
  const QBlock& k = blk[b + 1];
  L.data = k.data;
  L.off = e.off;
  L.h = e.h;
  L.inv = e.inv;
  L.hi = e.hi;
  L.up = e.up;
  L.stride = e.stride;
  L.st0 = e.st0;
  L.st1 = e.st1;
  L.fast = k.fast;
  L.mask = k.mask;
}

// contains + interpolate of vector_field.cpp:12-48 at the quad's point
// (lane c: component x); v = component c of the velocity (0 on the w lane).
// Quad-uniform result; false = outside the block.
template <int S, int INTERP>
__device__ __forceinline__ bool quad_eval(const QLane& L, int c, float x, float& v) {
  const bool wl = c == 3;
  const float d = x - L.off;
  const bool fast = quad_all(wl | div_range(d)) & (L.fast != 0);
  float q;
  if (__builtin_expect(fast, 1)) q = div_corrected(d, L.h, L.inv);
  else q = d / L.h;
  const float s = floorf(q);
  if (!quad_all(wl | ((s >= 0.0f) & (s <= L.hi)))) return false;
  const uint32_t m = wl ? 0u : static_cast<uint32_t>(static_cast<int>(s));
  float k0, k1, k2, k3, k4, k5, k6, k7;  // component c of c000 c001 c010 c011 c100 c101 c110 c111
  if constexpr (layout_of(S) == PA_LAYOUT_CELLS) {
    const uint32_t ci = quad_sum(m * L.stride);
    if constexpr (storage_of(S) == PA_STORAGE_FP32) {
      const uint64_t r = uint64_t(ci) * 8 + 2 * c;
      const float4 A = gload<float4>(L.data, r), B = gload<float4>(L.data, r + 1);
      k0 = A.x; k1 = A.y; k2 = A.z; k3 = A.w; k4 = B.x; k5 = B.y; k6 = B.z; k7 = B.w;
    } else {
      constexpr int st = storage_of(S);
      const uint4 U = gload<uint4>(L.data, uint64_t(ci) * 4 + c);
      k0 = dec16(U.x, st); k1 = dec16(U.x >> 16, st); k2 = dec16(U.y, st); k3 = dec16(U.y >> 16, st);
      k4 = dec16(U.z, st); k5 = dec16(U.z >> 16, st); k6 = dec16(U.w, st); k7 = dec16(U.w >> 16, st);
    }
  } else {
    const uint32_t n0 = quad_sum(m * L.stride);
    const uint32_t n2 = n0 + L.st1, n4 = n0 + L.st0, n6 = n4 + L.st1;
    if constexpr (storage_of(S) == PA_STORAGE_FP32) {
      const void* D = L.data;
      k0 = gload<float>(D, uint64_t(n0) * 4 + c); k1 = gload<float>(D, uint64_t(n0 + 1) * 4 + c);
      k2 = gload<float>(D, uint64_t(n2) * 4 + c); k3 = gload<float>(D, uint64_t(n2 + 1) * 4 + c);
      k4 = gload<float>(D, uint64_t(n4) * 4 + c); k5 = gload<float>(D, uint64_t(n4 + 1) * 4 + c);
      k6 = gload<float>(D, uint64_t(n6) * 4 + c); k7 = gload<float>(D, uint64_t(n6 + 1) * 4 + c);
    } else {
      constexpr int st = storage_of(S);
      const void* D = L.data;
      k0 = dec16(gload<uint16_t>(D, uint64_t(n0) * 4 + c), st); k1 = dec16(gload<uint16_t>(D, uint64_t(n0 + 1) * 4 + c), st);
      k2 = dec16(gload<uint16_t>(D, uint64_t(n2) * 4 + c), st); k3 = dec16(gload<uint16_t>(D, uint64_t(n2 + 1) * 4 + c), st);
      k4 = dec16(gload<uint16_t>(D, uint64_t(n4) * 4 + c), st); k5 = dec16(gload<uint16_t>(D, uint64_t(n4 + 1) * 4 + c), st);
      k6 = dec16(gload<uint16_t>(D, uint64_t(n6) * 4 + c), st); k7 = dec16(gload<uint16_t>(D, uint64_t(n6 + 1) * 4 + c), st);
    }
  }
  float w;
  if constexpr (INTERP == PA_INTERP_CONSISTENT) {
    w = q - s;
  } else {
    const float r = exact_fmod(d, L.h, q);
    if (__builtin_expect(quad_all(wl | div_range(r)) & (L.fast != 0), 1)) w = div_corrected(r, L.h, L.inv);
    else w = r / L.h;
  }
  const float w0 = qmovf<kBcast0>(w), w1 = qmovf<kBcast1>(w), w2 = qmovf<kBcast2>(w);
  const float u0 = 1.0f - w0, u1 = 1.0f - w1, u2 = 1.0f - w2;
  const float c00 = lerp1(k0, k1, w2, u2), c01 = lerp1(k2, k3, w2, u2);
  const float c10 = lerp1(k4, k5, w2, u2), c11 = lerp1(k6, k7, w2, u2);
  const float c0 = lerp1(c00, c01, w1, u1), c1 = lerp1(c10, c11, w1, u1);
  v = wl ? 0.0f : lerp1(c0, c1, w0, u0);
  return true;
}

// Stop test of the lane engine (stops<>), quad-uniform.
__device__ __forceinline__ bool quad_stops(float v, int c, int rule, float eps2) {
  if (rule == PA_STOP_NORM) {
    const float sq = v * v;
    const float n2 = (qmovf<kBcast0>(sq) + qmovf<kBcast1>(sq)) + qmovf<kBcast2>(sq);
    return n2 < eps2;
  }
  return quad_all((c == 3) | (fabsf(v) <= 1e-5f));
}

__device__ __forceinline__ float frozen_rk4(const TraceConfig& k, float x, float v) {
  return (((x + k.a0 * v) + k.a1 * v) + k.a2 * v) + k.a3 * v;
}

// One step of advance<> (0 stepped, 1 left the block, 2 stop rule).
template <int S, int MODE, int INTERP>
__device__ __forceinline__ int quad_advance(const TraceConfig& k, const QLane& L, int c, float x, float& next,
                                            int rule) {
  float v;
  if (!quad_eval<S, INTERP>(L, c, x, v)) return 1;
  if (quad_stops(v, c, rule, k.stop_eps2)) return 2;
  if constexpr (MODE == STEP_EULER) {
    next = x + k.dt * v;
  } else if constexpr (MODE == STEP_CLASSIC) {
    float k2 = 0.0f, k3 = 0.0f, k4 = 0.0f;
    bool ok = quad_eval<S, INTERP>(L, c, x + k.half_dt * v, k2);
    if (ok) ok = quad_eval<S, INTERP>(L, c, (x + k.zero_dt * v) + k.half_dt * k2, k3);
    if (ok) ok = quad_eval<S, INTERP>(L, c, ((x + k.zero_dt * v) + k.zero_dt * k2) + k.one_dt * k3, k4);
    next = ok ? (((x + k.a0 * v) + k.a1 * k2) + k.a2 * k3) + k.a3 * k4 : frozen_rk4(k, x, v);
  } else {
    next = frozen_rk4(k, x, v);
  }
  return 0;
}

// face_rule of the lane engine over the quad: first of x<min, x>max, y<min,
// y>max, z<min, z>max whose mask bit is set, or -1.
__device__ __forceinline__ int quad_face(const QLane& L, int c, float x) {
  int code = 0;
  if (c < 3) {
    if (x < L.off && ((L.mask >> (2 * c)) & 1)) code |= 1;
    if (x > L.up && ((L.mask >> (2 * c + 1)) & 1)) code |= 2;
  }
  const int bits = qmov<kBcast0>(code) | (qmov<kBcast1>(code) << 2) | (qmov<kBcast2>(code) << 4);
  return bits ? __builtin_ctz(static_cast<unsigned>(bits)) : -1;
}

enum { Q_IDLE = 0, Q_START = 1, Q_STEP = 2, Q_DONE = 3 };

template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kQuadBlock) void k_quad_union(const TraceConfig cfg, const pa_particle* __restrict__ P,
                                                           int64_t n, float* __restrict__ slots, int64_t region,
                                                           int32_t* __restrict__ counts,
                                                           pa_particle* __restrict__ cont,
                                                           unsigned int* __restrict__ n_cont,
                                                           unsigned long long* __restrict__ steps_out,
                                                           unsigned long long* __restrict__ work, int rule,
                                                           float* __restrict__ endpoints,
                                                           int32_t* __restrict__ seed_steps) {
  __shared__ QEntry tab[7][4];
  __shared__ QBlock blk[7];
  if (threadIdx.x < 28) {
    const int b = threadIdx.x >> 2, c = threadIdx.x & 3;
    const FieldDesc& F = cfg.field[b];
    const bool cells = layout_of(S) == PA_LAYOUT_CELLS;
    QEntry e;
    if (c < 3) {
      e.off = F.offset[c];
      e.h = F.spacing[c];
      e.inv = F.inv_spacing[c];
      e.hi = static_cast<float>(F.dims[c] - 2);
      e.up = F.upper[c];
    } else {
      e.off = 0.0f;
      e.h = 1.0f;
      e.inv = 1.0f;
      e.hi = 3.0e38f;
      e.up = __builtin_inff();
    }
    const int64_t s0 = cells ? F.cstride0 : F.stride0, s1 = cells ? F.cstride1 : F.stride1;
    e.stride = c == 0 ? uint32_t(s0) : c == 1 ? uint32_t(s1) : c == 2 ? 1u : 0u;
    e.st0 = uint32_t(F.stride0);
    e.st1 = uint32_t(F.stride1);
    tab[b][c] = e;
    if (c == 0) {
      QBlock k;
      k.data = cells ? F.cells : F.data;
      k.fast = F.fast_div;
      k.mask = b == 0 ? cfg.local_mask : cfg.face_mask[b - 1];
      k.has = cfg.has_field[b];
      k.pad_ = 0;
      blk[b] = k;
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int c = lane & 3;
  const float term = c == 3 ? 0.0f : -2.0f;  // termination vertex (-2, -2, -2, 0)
  int state = Q_IDLE;
  int64_t pid = 0;
  int b = -1, R = 0, it = 0, slot = 0, nsteps = 0, orank = 0;
  float last = 0.0f;
  QLane L;
  load_lane(tab, blk, -1, c, L);
  bool open = true;
  unsigned long long wave_steps = 0;

  for (;;) {
    // Idle quads take the next seeds (all lanes are converged here).
    const uint64_t leaders = __ballot(state == Q_IDLE && c == 0);
    if (open && leaders) {
      const uint32_t take = uint32_t(__popcll(leaders));
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(work, static_cast<unsigned long long>(take));
      const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(base & 0xffffffffu));
      const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(base >> 32));
      base = (static_cast<unsigned long long>(hi) << 32) | lo;
      if (base + take >= static_cast<unsigned long long>(n)) open = false;
      if (state == Q_IDLE) {
        const uint64_t below = (uint64_t(1) << (lane & ~3)) - 1;
        const unsigned long long p = base + uint32_t(__popcll(leaders & below));
        if (p < static_cast<unsigned long long>(n)) {
          pid = int64_t(p);
          const pa_particle& src = P[pid];
          last = src.position[c];
          R = src.remaining_iterations;
          b = src.vector_field_index;
          orank = src.original_rank;
          load_lane(tab, blk, b, c, L);
          state = Q_START;
          slot = 0;
          nsteps = 0;
        }
      }
    }
    if (__ballot(state != Q_IDLE) == 0) break;

    const bool writes = state != Q_IDLE;
    const int slot0 = slot;  // a multiple of kChunk for every writing quad
    float buf[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      float val = term;
      if (state == Q_START) {
        val = last;
        ++slot;
        if (R < 2) {
          state = Q_DONE;
        } else {
          state = Q_STEP;
          it = 1;
        }
      } else if (state == Q_STEP) {
        ++slot;
        if (it == R - 1) {
          state = Q_DONE;
        } else {
          float next;
          const int a = quad_advance<S, MODE, INTERP>(cfg, L, c, last, next, rule);
          if (a == 0) {
            val = next;
            last = next;
            ++it;
            ++nsteps;
          } else if (a == 2) {
            state = Q_DONE;
          } else {
            // hand-off within the union (trace_union of the lane engine)
            state = Q_DONE;
            const int face = quad_face(L, c, last);
            int nb = -1;
            bool go;
            if (b < 0) {
              go = face >= 0 && blk[face + 1].has != 0;
              nb = face;
            } else {
              go = face == (b ^ 1);
            }
            if (go) {
              const int Rn = R - it;
              if (int64_t(slot) + Rn > region) {
                unsigned idx = 0;
                if (c == 0) idx = atomicAdd(n_cont, 1u);
                idx = static_cast<unsigned>(qmov<kBcast0>(static_cast<int>(idx)));
                pa_particle* dst = cont + idx;
                dst->position[c] = last;
                if (c == 0) {
                  dst->remaining_iterations = Rn;
                  dst->vector_field_index = nb;
                  dst->original_rank = orank;
                  dst->original_voxel[0] = int32_t(pid);
                  dst->original_voxel[1] = 0;
                  dst->original_voxel[2] = 0;
                  dst->reserved_[0] = 0;
                  dst->reserved_[1] = 0;
                }
              } else {
                b = nb;
                R = Rn;
                load_lane(tab, blk, b, c, L);
                state = Q_START;
              }
            }
          }
        }
      }
      buf[j] = val;
    }
    if (writes) {
      float4* dst = reinterpret_cast<float4*>(slots + (pid * region + slot0) * 4 + c * kChunk);
      dst[0] = make_float4(buf[0], buf[1], buf[2], buf[3]);
      dst[1] = make_float4(buf[4], buf[5], buf[6], buf[7]);
    }
    if (state == Q_DONE) {
      if (c == 0) {
        counts[pid] = slot;
        if (seed_steps) seed_steps[pid] = nsteps;
        wave_steps += static_cast<unsigned long long>(nsteps);
      }
      if (endpoints) endpoints[pid * 4 + c] = last;
      state = Q_IDLE;
    }
  }
  wave_steps = wave_sum_u64(wave_steps);
  if (lane == 0 && wave_steps) atomicAdd(steps_out, wave_steps);
}

}  // namespace

pa_status launch_quad_union(pa_gpu_ctx* ctx, const TraceConfig& cfg, int mode, const CoopLaunch& l,
                            unsigned long long* steps, unsigned* ncont, unsigned long long* work) {
  if (l.region % kChunk != 0) return fail(ctx, PA_E_INVALID, "chunked slots: region_slots must be a multiple of 8");
  if (l.n > INT32_MAX) return fail(ctx, PA_E_INVALID, "chunked slots: at most 2^31-1 seeds per call");
  int cus = 0;
  PA_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  int per_cu = 0;
  int64_t blocks = (l.n + 63) / 64;
  dispatch(cfg.field[0].storage, mode, cfg.interp, [&](auto S, auto M, auto I) {
    auto kern = k_quad_union<decltype(S)::value, decltype(M)::value, decltype(I)::value>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kQuadBlock, 0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    const int64_t resident = int64_t(cus) * per_cu;
    if (blocks > resident) blocks = resident;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(kQuadBlock), 0, ctx->stream, cfg, l.particles, l.n,
                       reinterpret_cast<float*>(l.verts), l.region, l.counts, l.cont, ncont, steps, work,
                       l.stop_rule, reinterpret_cast<float*>(l.endpoints), l.seed_steps);
  }, cfg.field[0].layout);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

}  // namespace pa_amd
