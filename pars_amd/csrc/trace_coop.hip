// Wave-cooperative tracer.
//
// A scattered per-lane gather costs the texture-address and -data units one
// cache-line access per lane per load instruction; with 6-8 loads per step
// they, not HBM or the ALUs, bound a lane-per-particle tracer.  Here every
// wave still owns 64 particles (one per lane, same arithmetic as the
// lane-per-particle kernels), but the 8-corner gather is done by groups of
// L lanes: in load j, group g fetches the L 16-byte pieces of the record of
// its member g*L + j, so each load instruction touches 64/L lines instead of
// 64.  The pieces go through LDS back to their owner lane.  The loop is
// wave-uniform and lanes whose particle has finished take the next one from
// a work counter, so long-lived particles do not leave lanes idle.
#include "pa_device.hpp"

#pragma clang fp contract(off)

namespace pa_amd {
namespace {

constexpr int kWave = 64;

template <int S>
struct Gather {
  static constexpr int ST = storage_of(S);
  static constexpr bool CELLS = layout_of(S) == PA_LAYOUT_CELLS;
  static constexpr int L = ST == PA_STORAGE_FP32 ? 8 : 4;                      // lanes (16-B pieces) per record
  static constexpr int USED = CELLS ? (ST == PA_STORAGE_FP32 ? 6 : 3) : L;     // pieces that carry data
  static constexpr int SD = 4 * L + 4;                                         // LDS dwords per owner (padded)
  static constexpr int NODE_BYTES = ST == PA_STORAGE_FP32 ? 16 : 8;
  static constexpr int REC_BYTES = ST == PA_STORAGE_FP32 ? 128 : 64;
};

// 16-byte load through the global address space (a flat load would also
// count against the LDS wait counter and serialise with the LDS traffic).
__device__ __forceinline__ uint4 load_global(uint64_t addr) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(1))) uint4*)(addr);
#else
  return *reinterpret_cast<const uint4*>(addr);
#endif
}

// Address of piece c of the gather described by (base, s0, s1).
template <int S>
__device__ __forceinline__ uint64_t piece_addr(uint64_t base, uint32_t s0, uint32_t s1, int c) {
  using G = Gather<S>;
  if constexpr (G::CELLS) {
    return base + 16u * uint32_t(c);
  } else {
    const int zp = G::L == 8 ? (c >> 1) : c;  // z-pair 0: (0,0) 1: (0,1) 2: (1,0) 3: (1,1) in (x,y)
    return base + ((zp & 2) ? uint64_t(s0) : 0u) + ((zp & 1) ? uint64_t(s1) : 0u) +
           (G::L == 8 ? 16u * uint32_t(c & 1) : 0u);
  }
}

// Owner-side decode of the pieces in LDS into c000..c111 (xyz each).
template <int S>
__device__ __forceinline__ void decode(const uint32_t* d, float* c) {
  using G = Gather<S>;
  if constexpr (G::ST == PA_STORAGE_FP32) {
    if constexpr (G::CELLS) {  // record: x of the 8 corners, then y, then z
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) c[3 * k + j] = __uint_as_float(d[8 * j + k]);
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        c[3 * k] = __uint_as_float(d[4 * k]);
        c[3 * k + 1] = __uint_as_float(d[4 * k + 1]);
        c[3 * k + 2] = __uint_as_float(d[4 * k + 2]);
      }
    }
  } else {
    if constexpr (G::CELLS) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const int h = 8 * j + k;  // 16-bit element
          c[3 * k + j] = dec16(d[h >> 1] >> (16 * (h & 1)), G::ST);
        }
    } else {
      // piece zp = nodes (zp,0) and (zp,1): x|y<<16, z|pad, x|y<<16, z|pad
#pragma unroll
      for (int zp = 0; zp < 4; ++zp) {
        const uint32_t* q = d + 4 * zp;
        c[6 * zp + 0] = dec16(q[0], G::ST);
        c[6 * zp + 1] = dec16(q[0] >> 16, G::ST);
        c[6 * zp + 2] = dec16(q[1], G::ST);
        c[6 * zp + 3] = dec16(q[2], G::ST);
        c[6 * zp + 4] = dec16(q[2] >> 16, G::ST);
        c[6 * zp + 5] = dec16(q[3], G::ST);
      }
    }
  }
}

struct CoopArgs {
  const pa_particle* P;
  int64_t n;
  int32_t* counts;
  unsigned long long* steps_out;
  unsigned long long* next;
  const uint4* dummy;
  int stop_rule;
  // union form
  float4* verts;
  int64_t region;
  pa_particle* cont;
  unsigned* n_cont;
  float4* endpoints;
  int32_t* seed_steps;
  // reference round
  float4* const* chunks;
  int64_t ppic;
  int64_t maxR;
  pa_particle* hand_tmp;
  int32_t* hand_code;
};

template <int S, int MODE, int INTERP, bool UNION>
__global__ __launch_bounds__(kWave) void k_coop(const TraceConfig cfg, const CoopArgs a) {
  using G = Gather<S>;
  __shared__ uint4 info[kWave];
  __shared__ __attribute__((aligned(16))) uint32_t data[kWave * G::SD];
  const int lane = threadIdx.x;

  bool active = false;
  bool drained = false;  // wave-uniform
  int64_t pi = 0;
  float4* out = nullptr;
  float4 last = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int R = 0, b = -1, it = 1, stage = 0, seed_steps_local = 0;
  int64_t seg_base = 0;
  FieldDesc f = cfg.field[0];
  unsigned long long steps_total = 0;
  float w[3] = {0, 0, 0}, u[3] = {1, 1, 1};
  int m[3] = {0, 0, 0};
  float k1[3] = {0, 0, 0}, k2[3] = {0, 0, 0}, k3[3] = {0, 0, 0};

  auto finish_union = [&](int64_t used) {
    a.counts[pi] = int32_t(used);
    if (a.endpoints) a.endpoints[pi] = last;
    if (a.seed_steps) a.seed_steps[pi] = seed_steps_local;
    active = false;
  };
  auto slot = [&](int i) -> float4* { return UNION ? out + seg_base + i : out + i; };
  auto emit = [&](float4 next) {
    *slot(it) = next;
    last = next;
    ++it;
    ++steps_total;
    ++seed_steps_local;
  };
  // The segment ends at iteration it: termination vertex at slot it.
  auto end_segment = [&](bool exited) {
    *slot(it) = termination_vertex();
    if constexpr (UNION) {
      int64_t used = seg_base + it + 1;
      if (!exited) {
        finish_union(used);
        return;
      }
      int nb;
      if (b < 0) {
        const int k = face_rule(f, last, cfg.local_mask);
        if (k < 0 || !cfg.has_field[k + 1]) {
          finish_union(used);
          return;
        }
        nb = k;
      } else {
        const int d = face_rule(f, last, cfg.face_mask[b]);
        if (d != (b ^ 1)) {
          finish_union(used);
          return;
        }
        nb = -1;
      }
      const int Rn = R - it;
      if (used + Rn > a.region) {
        pa_particle c;
        c.position[0] = last.x;
        c.position[1] = last.y;
        c.position[2] = last.z;
        c.position[3] = last.w;
        c.remaining_iterations = Rn;
        c.vector_field_index = nb;
        c.original_rank = 0;
        c.original_voxel[0] = int32_t(pi);
        c.original_voxel[1] = c.original_voxel[2] = 0;
        c.reserved_[0] = c.reserved_[1] = 0;
        a.cont[atomicAdd(a.n_cont, 1u)] = c;
        finish_union(used);
        return;
      }
      out[used] = last;  // the receiving block's seed vertex
      seg_base = used;
      R = Rn;
      b = nb;
      f = cfg.field[b + 1];
      it = 1;
      if (R < 2) finish_union(seg_base + 1);
    } else {
      a.counts[pi] = it + 1;
      if (a.hand_code) {
        int code = -1;
        if (exited) {
          if (b < 0) {
            const int k = face_rule(f, last, cfg.local_mask);
            if (k >= 0) code = k;
          } else {
            code = 6 + b;
          }
        }
        if (code >= 0) {
          pa_particle h;
          h.position[0] = last.x;
          h.position[1] = last.y;
          h.position[2] = last.z;
          h.position[3] = last.w;
          h.remaining_iterations = R - it;
          h.vector_field_index = -1;
          h.original_rank = 0;
          h.original_voxel[0] = h.original_voxel[1] = h.original_voxel[2] = 0;
          h.reserved_[0] = h.reserved_[1] = 0;
          a.hand_tmp[pi] = h;
        }
        a.hand_code[pi] = code;
      }
      active = false;
    }
  };

  for (;;) {
    // ---- refill idle lanes from the work counter (one atomic per wave)
    if (!drained) {
      const uint64_t idle = __ballot(!active);
      if (idle) {
        const int nidle = __popcll(idle);
        unsigned long long base0 = 0;
        if (lane == 0) base0 = atomicAdd(a.next, static_cast<unsigned long long>(nidle));
        base0 = __shfl(base0, 0, kWave);
        if (int64_t(base0) + nidle >= a.n) drained = true;
        if (!active) {
          const int64_t idx = int64_t(base0) + __popcll(idle & ((uint64_t(1) << lane) - 1));
          if (idx < a.n) {
            const pa_particle p = a.P[idx];
            pi = idx;
            last = make_float4(p.position[0], p.position[1], p.position[2], p.position[3]);
            R = p.remaining_iterations;
            b = p.vector_field_index;
            f = cfg.field[b + 1];
            it = 1;
            stage = 0;
            seg_base = 0;
            seed_steps_local = 0;
            if constexpr (UNION) {
              out = a.verts + idx * a.region;
            } else {
              out = a.chunks[idx / a.ppic] + (idx % a.ppic) * a.maxR;
            }
            out[0] = last;  // initialize (particle_tracer.cpp:243-252)
            if (R < 2) {
              if constexpr (UNION) {
                finish_union(1);
              } else {
                a.counts[idx] = 1;
                if (a.hand_code) a.hand_code[idx] = -1;
              }
            } else {
              active = true;
            }
          }
        }
      }
    }
    if (!__ballot(active)) {
      if (drained) break;
      continue;
    }

    // ---- pre-gather: decide which lanes evaluate the field this round
    bool need = false;
    if (active) {
      if (stage == 0) {
        if (it == R - 1) {
          end_segment(false);
        } else {
          float d[3], q[3];
          if (!locate(f, last.x, last.y, last.z, d, q)) {
            end_segment(true);
          } else {
            cell_weights<INTERP>(f, d, q, m, w, u);
            need = true;
          }
        }
      } else {
        need = true;  // classic RK4 stage point, located in the previous round
      }
    }

    // ---- cooperative gather
    uint4 mine = make_uint4(0u, 0u, 0u, 0u);
    if (need) {
      uint64_t base;
      if constexpr (G::CELLS) {
        const int64_t ci = int64_t(m[0]) * f.cstride0 + int64_t(m[1]) * f.cstride1 + m[2];
        base = reinterpret_cast<uint64_t>(f.cells) + uint64_t(ci) * G::REC_BYTES;
      } else {
        const int64_t ni = int64_t(m[0]) * f.stride0 + int64_t(m[1]) * f.stride1 + m[2];
        base = reinterpret_cast<uint64_t>(f.data) + uint64_t(ni) * G::NODE_BYTES;
      }
      mine = make_uint4(uint32_t(base), uint32_t(base >> 32), uint32_t(f.stride0 * G::NODE_BYTES),
                        uint32_t(f.stride1 * G::NODE_BYTES));
    }
    info[lane] = mine;
    __syncthreads();
    {
      // Group g0/L loads, in round j, the pieces of member g0 + j; every lane
      // loads (absent owners read a dummy piece) so the loads stay in flight
      // together, and every lane writes its piece to the member's LDS slot.
      const int g0 = lane & ~(G::L - 1);
      const int c = lane & (G::L - 1);
      const uint64_t dummy = reinterpret_cast<uint64_t>(a.dummy);
      uint4 v[G::L];
#pragma unroll
      for (int j = 0; j < G::L; ++j) {
        const uint4 oi = info[g0 + j];
        const uint64_t base = uint64_t(oi.x) | (uint64_t(oi.y) << 32);
        const uint64_t addr = (base != 0 && c < G::USED) ? piece_addr<S>(base, oi.z, oi.w, c) : dummy;
        v[j] = load_global(addr);
      }
#pragma unroll
      for (int j = 0; j < G::L; ++j) *reinterpret_cast<uint4*>(&data[(g0 + j) * G::SD + 4 * c]) = v[j];
    }
    __syncthreads();

    // ---- post-gather: interpolate and step
    if (need) {
      float corner[24], vel[3];
      decode<S>(&data[lane * G::SD], corner);
      lerp_corners(corner, w, u, vel);
      if (stage == 0) {
        const bool stop = a.stop_rule == PA_STOP_NORM ? stops<PA_STOP_NORM>(vel, cfg.stop_eps2)
                                                      : stops<PA_STOP_ISZERO>(vel, cfg.stop_eps2);
        if (stop) {
          end_segment(false);
        } else if constexpr (MODE == STEP_CLASSIC) {
          k1[0] = vel[0];
          k1[1] = vel[1];
          k1[2] = vel[2];
          const float x0 = last.x + cfg.half_dt * k1[0];
          const float x1 = last.y + cfg.half_dt * k1[1];
          const float x2 = last.z + cfg.half_dt * k1[2];
          float d[3], q[3];
          if (locate(f, x0, x1, x2, d, q)) {
            cell_weights<INTERP>(f, d, q, m, w, u);
            stage = 1;
          } else {
            emit(step_frozen<STEP_RK4>(cfg, last, k1));
          }
        } else {
          emit(step_frozen<MODE>(cfg, last, vel));
        }
      } else if constexpr (MODE == STEP_CLASSIC) {
        float d[3], q[3];
        if (stage == 1) {
          k2[0] = vel[0];
          k2[1] = vel[1];
          k2[2] = vel[2];
          const float x0 = (last.x + cfg.zero_dt * k1[0]) + cfg.half_dt * k2[0];
          const float x1 = (last.y + cfg.zero_dt * k1[1]) + cfg.half_dt * k2[1];
          const float x2 = (last.z + cfg.zero_dt * k1[2]) + cfg.half_dt * k2[2];
          if (locate(f, x0, x1, x2, d, q)) {
            cell_weights<INTERP>(f, d, q, m, w, u);
            stage = 2;
          } else {
            stage = 0;
            emit(step_frozen<STEP_RK4>(cfg, last, k1));
          }
        } else if (stage == 2) {
          k3[0] = vel[0];
          k3[1] = vel[1];
          k3[2] = vel[2];
          const float x0 = ((last.x + cfg.zero_dt * k1[0]) + cfg.zero_dt * k2[0]) + cfg.one_dt * k3[0];
          const float x1 = ((last.y + cfg.zero_dt * k1[1]) + cfg.zero_dt * k2[1]) + cfg.one_dt * k3[1];
          const float x2 = ((last.z + cfg.zero_dt * k1[2]) + cfg.zero_dt * k2[2]) + cfg.one_dt * k3[2];
          if (locate(f, x0, x1, x2, d, q)) {
            cell_weights<INTERP>(f, d, q, m, w, u);
            stage = 3;
          } else {
            stage = 0;
            emit(step_frozen<STEP_RK4>(cfg, last, k1));
          }
        } else {
          float4 o;
          o.x = (((last.x + cfg.a0 * k1[0]) + cfg.a1 * k2[0]) + cfg.a2 * k3[0]) + cfg.a3 * vel[0];
          o.y = (((last.y + cfg.a0 * k1[1]) + cfg.a1 * k2[1]) + cfg.a2 * k3[1]) + cfg.a3 * vel[1];
          o.z = (((last.z + cfg.a0 * k1[2]) + cfg.a1 * k2[2]) + cfg.a2 * k3[2]) + cfg.a3 * vel[2];
          o.w = last.w + cfg.w_rk4;  // (((last.w + a0 0) + a1 0) + a2 0) + a3 0
          stage = 0;
          emit(o);
        }
      }
    }
  }
  steps_total = wave_sum_u64(steps_total);
  if (lane == 0 && steps_total) atomicAdd(a.steps_out, steps_total);
}

__global__ void k_zero2(unsigned long long* c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0ull;
}

}  // namespace

// Number of resident waves for a launch: all particles at once up to the
// chip's wave slots (the refill keeps lanes busy beyond that).
int coop_waves(pa_gpu_ctx* ctx, int64_t n, int requested) {
  if (requested > 0) return requested;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  int64_t w = (n + kWave - 1) / kWave;
  const int64_t cap = int64_t(cus) * 12;
  if (w > cap) w = cap;
  if (w < 1) w = 1;
  return int(w);
}

pa_status launch_coop(pa_gpu_ctx* ctx, const TraceConfig& cfg, int mode, bool union_form, const CoopLaunch& L) {
  CoopArgs a{};
  a.P = L.particles;
  a.n = L.n;
  a.counts = L.counts;
  a.stop_rule = L.stop_rule;
  char* cbase = static_cast<char*>(ctx->counters);
  a.steps_out = reinterpret_cast<unsigned long long*>(cbase + 1024);
  a.next = reinterpret_cast<unsigned long long*>(cbase + 1040);
  a.dummy = reinterpret_cast<const uint4*>(cbase + 3072);
  a.verts = L.verts;
  a.region = L.region;
  a.cont = L.cont;
  a.n_cont = reinterpret_cast<unsigned*>(cbase + 1032);
  a.endpoints = L.endpoints;
  a.seed_steps = L.seed_steps;
  a.chunks = L.chunks;
  a.ppic = L.ppic;
  a.maxR = L.maxR;
  a.hand_tmp = L.hand_tmp;
  a.hand_code = L.hand_code;
  hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, ctx->stream, reinterpret_cast<unsigned long long*>(cbase + 1024));
  hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, ctx->stream, reinterpret_cast<unsigned long long*>(cbase + 1040));
  const int waves = coop_waves(ctx, L.n, L.waves);
  launch_begin(ctx, union_form ? "trace_union" : "trace");
  dispatch(cfg.field[0].storage, mode, cfg.interp, [&](auto Sc, auto Mc, auto Ic) {
    constexpr int S = decltype(Sc)::value, M = decltype(Mc)::value, I = decltype(Ic)::value;
    if (union_form)
      hipLaunchKernelGGL((k_coop<S, M, I, true>), dim3(waves), dim3(kWave), 0, ctx->stream, cfg, a);
    else
      hipLaunchKernelGGL((k_coop<S, M, I, false>), dim3(waves), dim3(kWave), 0, ctx->stream, cfg, a);
  }, cfg.field[0].layout);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

}  // namespace pa_amd
