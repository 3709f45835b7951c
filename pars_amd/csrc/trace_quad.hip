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

#include "pa_color.hpp"
#include "pa_device.hpp"

namespace pa_amd {
namespace {

constexpr int kQuadBlock = 256;  // 4 waves, 64 quads
constexpr int kChunk = 8;        // slots per planar chunk

// DPP quad permutations (quad_perm: lane i of a quad reads lane p_i).
constexpr int kBcast0 = 0x00, kBcast1 = 0x55, kBcast2 = 0xAA;
constexpr int kSwap1 = 0xB1;  // (1,0,3,2)
constexpr int kSwap2 = 0x4E;  // (2,3,0,1)

// Every lane has a source under a quad permutation, so the bound control
// (0 for a lane without one) never applies; it spares a copy of v.
template <int CTRL>
__device__ __forceinline__ int qmov(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
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
  float off, h, inv, inv_lo, hi, up;
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
  float off, h, inv, inv_lo, hi, up;
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
  L.inv_lo = e.inv_lo;
  L.hi = e.hi;
  L.up = e.up;
  L.stride = e.stride;
  L.st0 = e.st0;
  L.st1 = e.st1;
  L.fast = k.fast;
  L.mask = k.mask;
}

// Component c's part of cell record ci: FP32 chunks 2c and 2c + 1 (A, B);
// 16-bit chunk c (A, as bits; B untouched).
template <int S>
__device__ __forceinline__ void load_record(const void* data, int c, uint32_t ci, float4& A, float4& B) {
  if constexpr (storage_of(S) == PA_STORAGE_FP32) {
    const uint64_t r = uint64_t(ci) * 8 + 2 * c;
    A = gload<float4>(data, r);
    B = gload<float4>(data, r + 1);
  } else {
    const uint4 U = gload<uint4>(data, uint64_t(ci) * 4 + c);
    A = make_float4(__uint_as_float(U.x), __uint_as_float(U.y), __uint_as_float(U.z), __uint_as_float(U.w));
  }
}

// Component c of the 8 corners c000 c001 c010 c011 c100 c101 c110 c111
// from the lane's part of a record.
template <int S>
__device__ __forceinline__ void decode_quad(const float4& A, const float4& B, float k[8]) {
  if constexpr (storage_of(S) == PA_STORAGE_FP32) {
    k[0] = A.x; k[1] = A.y; k[2] = A.z; k[3] = A.w; k[4] = B.x; k[5] = B.y; k[6] = B.z; k[7] = B.w;
  } else {
    constexpr int st = storage_of(S);
    const uint32_t u0 = __float_as_uint(A.x), u1 = __float_as_uint(A.y), u2 = __float_as_uint(A.z),
                   u3 = __float_as_uint(A.w);
    k[0] = dec16(u0, st); k[1] = dec16(u0 >> 16, st); k[2] = dec16(u1, st); k[3] = dec16(u1 >> 16, st);
    k[4] = dec16(u2, st); k[5] = dec16(u2 >> 16, st); k[6] = dec16(u3, st); k[7] = dec16(u3 >> 16, st);
  }
}

// Weights (vector_field.cpp:27-31) and the lerp tree (:41-47) over the
// quad: lane c weighs along its axis; the result is component c of the
// velocity (0 on the w lane).  unused: the lane's weight is not used (the
// w lane, or a quad that does not step), so it never takes the slow path.
template <int INTERP>
__device__ __forceinline__ float quad_lerp(const QLane& L, bool wl, float d, float q, float s, bool dfast,
                                          const float k[8], bool unused) {
  float w;
  if constexpr (INTERP == PA_INTERP_CONSISTENT) {
    w = q - s;
  } else {
    // as cell_weights: the remainder from s = floor(q) while d and it are
    // within div_corrected's ranges, else fmodf and division
    const float r = rem_located(d, L.h, s);
    const bool fast = (int(unused) | (int(dfast) & int(rem_range(r)))) != 0;
    w = weight_located(d, L.h, L.inv, L.inv_lo, r);
    if (__builtin_expect(__ballot(!fast) != 0, 0)) {  // as in locate
      if (!fast) w = exact_fmod(d, L.h, q) / L.h;
    }
  }
  const float w0 = qmovf<kBcast0>(w), w1 = qmovf<kBcast1>(w), w2 = qmovf<kBcast2>(w);
  const float u0 = 1.0f - w0, u1 = 1.0f - w1, u2 = 1.0f - w2;
  const float c00 = lerp1(k[0], k[1], w2, u2), c01 = lerp1(k[2], k[3], w2, u2);
  const float c10 = lerp1(k[4], k[5], w2, u2), c11 = lerp1(k[6], k[7], w2, u2);
  const float c0 = lerp1(c00, c01, w1, u1), c1 = lerp1(c10, c11, w1, u1);
  return wl ? 0.0f : lerp1(c0, c1, w0, u0);
}

// The quotient of vector_field::contains (vector_field.cpp:12-21) along
// the lane's axis: d = x - offset, q = d / h, s = floor(q); quad-uniform
// result, false = outside the block.
// fast: d and the spacing were within div_corrected's ranges (or the w lane).
// own: the lane's own axis is inside (0 <= s <= dims - 2; always on the w lane).
__device__ __forceinline__ bool quad_locate(const QLane& L, bool wl, float x, float& d, float& q, float& s,
                                            bool& fast, bool& own) {
  d = x - L.off;
  // per lane: both branches give the correctly rounded quotient
  fast = (int(wl | div_range(d)) & int(L.fast != 0)) != 0;
  // (a zero quotient may carry either sign; the division under a
  // wave-uniform branch, as in locate)
  q = div_unsigned(d, L.h, L.inv, L.inv_lo);
  if (__builtin_expect(__ballot(!fast) != 0, 0)) {
    if (!fast) q = d / L.h;
  }
  s = floorf(q);
  own = wl | (__builtin_amdgcn_fmed3f(s, 0.0f, L.hi) == s);
  return quad_all(own);
}

// contains + interpolate of vector_field.cpp:12-48 at the quad's point
// (lane c: component x); v = component c of the velocity (0 on the w lane).
// Quad-uniform result; false = outside the block.
template <int S, int INTERP>
__device__ __forceinline__ bool quad_eval(const QLane& L, int c, float x, float& v) {
  const bool wl = c == 3;
  float d, q, s;
  bool dfast, own;
  if (!quad_locate(L, wl, x, d, q, s, dfast, own)) return false;
  const uint32_t m = wl ? 0u : static_cast<uint32_t>(static_cast<int>(s));
  float k[8];  // component c of c000 c001 c010 c011 c100 c101 c110 c111
  if constexpr (layout_of(S) == PA_LAYOUT_CELLS) {
    float4 A, B;
    load_record<S>(L.data, c, quad_sum(m * L.stride), A, B);
    decode_quad<S>(A, B, k);
  } else {
    const uint32_t n0 = quad_sum(m * L.stride);
    const uint32_t n2 = n0 + L.st1, n4 = n0 + L.st0, n6 = n4 + L.st1;
    if constexpr (storage_of(S) == PA_STORAGE_FP32) {
      const void* D = L.data;
      k[0] = gload<float>(D, uint64_t(n0) * 4 + c); k[1] = gload<float>(D, uint64_t(n0 + 1) * 4 + c);
      k[2] = gload<float>(D, uint64_t(n2) * 4 + c); k[3] = gload<float>(D, uint64_t(n2 + 1) * 4 + c);
      k[4] = gload<float>(D, uint64_t(n4) * 4 + c); k[5] = gload<float>(D, uint64_t(n4 + 1) * 4 + c);
      k[6] = gload<float>(D, uint64_t(n6) * 4 + c); k[7] = gload<float>(D, uint64_t(n6 + 1) * 4 + c);
    } else {
      constexpr int st = storage_of(S);
      const void* D = L.data;
      k[0] = dec16(gload<uint16_t>(D, uint64_t(n0) * 4 + c), st); k[1] = dec16(gload<uint16_t>(D, uint64_t(n0 + 1) * 4 + c), st);
      k[2] = dec16(gload<uint16_t>(D, uint64_t(n2) * 4 + c), st); k[3] = dec16(gload<uint16_t>(D, uint64_t(n2 + 1) * 4 + c), st);
      k[4] = dec16(gload<uint16_t>(D, uint64_t(n4) * 4 + c), st); k[5] = dec16(gload<uint16_t>(D, uint64_t(n4 + 1) * 4 + c), st);
      k[6] = dec16(gload<uint16_t>(D, uint64_t(n6) * 4 + c), st); k[7] = dec16(gload<uint16_t>(D, uint64_t(n6 + 1) * 4 + c), st);
    }
  }
  v = quad_lerp<INTERP>(L, wl, d, q, s, dfast, k, wl);
  return true;
}

// A quad's record for this iteration, gathered before the quad's state
// machine runs (cell records): in = the quad steps and its point is inside
// the block; d, q, s, fast as quad_locate; A, B as load_record.
struct QGather {
  bool in;
  float d, q, s;
  bool fast;
  float4 A, B;
};

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

// The stop test and the step of advance<> once v = field(x) is known (0
// stepped, 2 stop rule).
template <int S, int MODE, int INTERP>
__device__ __forceinline__ int quad_finish(const TraceConfig& k, const QLane& L, int c, float x, float v, float& next,
                                           int rule) {
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

// One step of advance<> (0 stepped, 1 left the block, 2 stop rule); with
// G the record was gathered before (QGather).
template <int S, int MODE, int INTERP>
__device__ __forceinline__ int quad_advance(const TraceConfig& k, const QLane& L, int c, float x, float& next,
                                            int rule, const QGather* G = nullptr) {
  float v;
  if (G) {
    if (!G->in) return 1;
    float kc[8];
    decode_quad<S>(G->A, G->B, kc);
    v = quad_lerp<INTERP>(L, c == 3, G->d, G->q, G->s, G->fast, kc, c == 3);
  } else if (!quad_eval<S, INTERP>(L, c, x, v)) {
    return 1;
  }
  return quad_finish<S, MODE, INTERP>(k, L, c, x, v, next, rule);
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

// Per-component block table in LDS (threads 0..27 of the workgroup).
template <int S>
__device__ __forceinline__ void fill_tables(const TraceConfig& cfg, QEntry (*tab)[4], QBlock* blk) {
  if (threadIdx.x < 28) {
    const int b = threadIdx.x >> 2, c = threadIdx.x & 3;
    const FieldDesc& F = cfg.field[b];
    const bool cells = layout_of(S) == PA_LAYOUT_CELLS;
    QEntry e;
    if (c < 3) {
      e.off = F.offset[c];
      e.h = F.spacing[c];
      e.inv = F.inv_spacing[c];
      e.inv_lo = F.inv_lo[c];
      e.hi = F.cell_hi[c];
      e.up = F.upper[c];
    } else {
      e.off = 0.0f;
      e.h = 1.0f;
      e.inv = 1.0f;
      e.inv_lo = 0.0f;
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
}

enum { Q_IDLE = 0, Q_START = 1, Q_STEP = 2, Q_DONE = 3 };

// One particle's union-trace state (replicated over the quad in quad mode).
struct QState {
  int state;
  int64_t pid;
  int b, R, it, slot, nsteps, orank;
  int pairs, prevT;  // index-generator pairs so far, last slot was the termination vertex
};

// A continuation record for a seed whose region is full (as the lane engine).
__device__ __forceinline__ void write_continuation(pa_particle* dst, const float4& last, int Rn, int nb, int orank,
                                                   int64_t pid) {
  dst->position[0] = last.x;
  dst->position[1] = last.y;
  dst->position[2] = last.z;
  dst->position[3] = last.w;
  dst->remaining_iterations = Rn;
  dst->vector_field_index = nb;
  dst->original_rank = orank;
  dst->original_voxel[0] = int32_t(pid);
  dst->original_voxel[1] = 0;
  dst->original_voxel[2] = 0;
  dst->reserved_[0] = 0;
  dst->reserved_[1] = 0;
}

// One iteration of an active quad (Q_START or Q_STEP): produces exactly one
// slot, returned as this lane's component; the slot index is st.slot - 1
// afterwards.  Segment logic of the lane engine's trace_union.  G: the
// record of this iteration, gathered before (see quad_advance).
template <int S, int MODE, int INTERP>
__device__ __forceinline__ float quad_iterate(const TraceConfig& cfg, const QEntry (*tab)[4], const QBlock* blk,
                                              QState& st, QLane& L, float& last, int c, int rule, int64_t region,
                                              pa_particle* cont, unsigned* n_cont, const QGather* G = nullptr) {
  const float term = c == 3 ? 0.0f : -2.0f;  // termination vertex (-2, -2, -2, 0)
  float val = term;
  ++st.slot;
  if (st.state == Q_START) {
    val = last;
    if (st.R < 2) {
      st.state = Q_DONE;
    } else {
      st.state = Q_STEP;
      st.it = 1;
    }
  } else if (st.it == st.R - 1) {
    st.state = Q_DONE;
  } else {
    float next;
    const int a = quad_advance<S, MODE, INTERP>(cfg, L, c, last, next, rule, G);
    if (a == 0) {
      val = next;
      last = next;
      ++st.it;
      ++st.nsteps;
    } else {
      st.state = Q_DONE;
      if (a == 1) {
        const int face = quad_face(L, c, last);
        int nb = -1;
        bool go;
        if (st.b < 0) {
          go = face >= 0 && blk[face + 1].has != 0;
          nb = face;
        } else {
          go = face == (st.b ^ 1);
        }
        if (go) {
          const int Rn = st.R - st.it;
          if (int64_t(st.slot) + Rn > region) {
            unsigned idx = 0;
            if (c == 0) idx = atomicAdd(n_cont, 1u);
            idx = static_cast<unsigned>(qmov<kBcast0>(static_cast<int>(idx)));
            const float lx = qmovf<kBcast0>(last), ly = qmovf<kBcast1>(last), lz = qmovf<kBcast2>(last),
                        lw = qmovf<0xFF>(last);
            if (c == 0) write_continuation(cont + idx, make_float4(lx, ly, lz, lw), Rn, nb, st.orank, st.pid);
          } else {
            st.b = nb;
            st.R = Rn;
            load_lane(tab, blk, st.b, c, L);
            st.state = Q_START;
          }
        }
      }
    }
  }
  return val;
}

// quad_iterate for cell records and the frozen-derivative steppers as one
// path of selects for every quad (a quad in the tail runs alone on its
// SIMD, where each branch of the state machine is a stall on the chain);
// only the hand-over of a quad that left its block stays a branch.  The
// slots are those of quad_iterate, bit for bit.  G: this iteration's
// gather (G.in: the quad steps and its point is inside the block).
template <int S, int MODE, int INTERP>
__device__ __forceinline__ float quad_iterate_sel(const TraceConfig& cfg, const QEntry (*tab)[4], const QBlock* blk,
                                                  QState& st, QLane& L, float& last, int c, int rule, int64_t region,
                                                  pa_particle* cont, unsigned* n_cont, const QGather& G) {
  static_assert(MODE != STEP_CLASSIC, "classic RK4 evaluates the field at its stage points: quad_iterate");
  const float term = c == 3 ? 0.0f : -2.0f;  // termination vertex (-2, -2, -2, 0)
  const bool active = st.state != Q_IDLE;
  const bool start = st.state == Q_START;
  float kc[8];
  decode_quad<S>(G.A, G.B, kc);
  const float v = quad_lerp<INTERP>(L, c == 3, G.d, G.q, G.s, G.fast, kc, (c == 3) | !G.in);
  const bool stop = quad_stops(v, c, rule, cfg.stop_eps2);
  const float nx = MODE == STEP_EULER ? last + cfg.dt * v : frozen_rk4(cfg, last, v);
  const bool stepped = G.in & !stop;
  const bool left = active & !start & (st.it != st.R - 1) & !G.in;  // contains failed (quad-uniform)
  const float val = start ? last : stepped ? nx : term;
  int nstate = start ? (st.R < 2 ? Q_DONE : Q_STEP) : stepped ? Q_STEP : Q_DONE;
  st.it = start ? 1 : st.it + int(stepped);
  st.nsteps += int(stepped);
  last = stepped ? nx : last;
  st.slot += int(active);
  if (left) {
    const int face = quad_face(L, c, last);
    int nb = -1;
    bool go;
    if (st.b < 0) {
      go = face >= 0 && blk[face + 1].has != 0;
      nb = face;
    } else {
      go = face == (st.b ^ 1);
    }
    if (go) {
      const int Rn = st.R - st.it;
      if (int64_t(st.slot) + Rn > region) {
        unsigned idx = 0;
        if (c == 0) idx = atomicAdd(n_cont, 1u);
        idx = static_cast<unsigned>(qmov<kBcast0>(static_cast<int>(idx)));
        const float lx = qmovf<kBcast0>(last), ly = qmovf<kBcast1>(last), lz = qmovf<kBcast2>(last),
                    lw = qmovf<0xFF>(last);
        if (c == 0) write_continuation(cont + idx, make_float4(lx, ly, lz, lw), Rn, nb, st.orank, st.pid);
      } else {
        st.b = nb;
        st.R = Rn;
        load_lane(tab, blk, st.b, c, L);
        nstate = Q_START;
      }
    }
  }
  st.state = active ? nstate : Q_IDLE;
  return val;
}

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
  fill_tables<S>(cfg, tab, blk);
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int c = lane & 3;
  QState st;
  st.state = Q_IDLE;
  st.pid = 0;
  st.b = -1;
  st.R = st.it = st.slot = st.nsteps = st.orank = st.pairs = st.prevT = 0;
  float last = 0.0f;
  QLane L;
  load_lane(tab, blk, -1, c, L);
  bool open = true;
  unsigned long long wave_steps = 0;

  for (;;) {
    // Idle quads take the next seeds (all lanes are converged here).
    const uint64_t leaders = __ballot(st.state == Q_IDLE && c == 0);
    if (open && leaders) {
      const uint32_t take = uint32_t(__popcll(leaders));
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(work, static_cast<unsigned long long>(take));
      const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(base & 0xffffffffu));
      const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(base >> 32));
      base = (static_cast<unsigned long long>(hi) << 32) | lo;
      if (base + take >= static_cast<unsigned long long>(n)) open = false;
      if (st.state == Q_IDLE) {
        const uint64_t below = (uint64_t(1) << (lane & ~3)) - 1;
        const unsigned long long p = base + uint32_t(__popcll(leaders & below));
        if (p < static_cast<unsigned long long>(n)) {
          st.pid = int64_t(p);
          const pa_particle& src = P[st.pid];
          last = src.position[c];
          st.R = src.remaining_iterations;
          st.b = src.vector_field_index;
          st.orank = src.original_rank;
          load_lane(tab, blk, st.b, c, L);
          st.state = Q_START;
          st.slot = 0;
          st.nsteps = 0;
        }
      }
    }
    if (__ballot(st.state != Q_IDLE) == 0) break;

    const bool writes = st.state != Q_IDLE;
    const int slot0 = st.slot;  // a multiple of kChunk for every writing quad
    float buf[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      float val = 0.0f;
      if (st.state == Q_START || st.state == Q_STEP)
        val = quad_iterate<S, MODE, INTERP>(cfg, tab, blk, st, L, last, c, rule, region, cont, n_cont);
      buf[j] = val;
    }
    if (writes) {
      float4* dst = reinterpret_cast<float4*>(slots + (st.pid * region + slot0) * 4 + c * kChunk);
      dst[0] = make_float4(buf[0], buf[1], buf[2], buf[3]);
      dst[1] = make_float4(buf[4], buf[5], buf[6], buf[7]);
    }
    if (st.state == Q_DONE) {
      if (c == 0) {
        counts[st.pid] = st.slot;
        if (seed_steps) seed_steps[st.pid] = st.nsteps;
        wave_steps += static_cast<unsigned long long>(st.nsteps);
      }
      if (endpoints) endpoints[st.pid * 4 + c] = last;
      st.state = Q_IDLE;
    }
  }
  wave_steps = wave_sum_u64(wave_steps);
  if (lane == 0 && wave_steps) atomicAdd(steps_out, wave_steps);
}

// FP32 cell records, lane mode.  A lane loads three chunks of record e
// (X) and three of record o (Y).  Alone, both are the lane's own record:
// X its chunks 0-2, Y its chunks 3-5.  While a wave is full (pair), the two
// lanes of a pair (l, l ^ 1) share their loads: e is the even lane's
// record and o the odd lane's, the even lane loads chunks 0-2 of both and
// the odd lane chunks 3-5 of both, so each load instruction touches one
// record line per pair instead of one per lane (the L1 looks up half as
// many tags, which is what limits a full wave's gathers).  Each chunk holds
// one component of the four corners of one x-half of the cell (x c000..c011,
// x c100..c111, y .., y .., z .., z ..), which the lerp tree reduces on its
// own (z, then y) to that half's value; only the x step combines chunks
// 2k and 2k + 1.  So every lane reduces the chunks it loaded with the
// weights of the record they belong to (the partner's weights by a DPP
// swap), and the partner's three half values go back by one more swap:
// three values cross the pair instead of twelve record words, and the
// arithmetic is the lerp tree's own, operation for operation.
struct PairGather {
  uint4 X[3], Y[3];
};

__device__ __forceinline__ float swapf(float v) { return qmovf<kSwap1>(v); }

__device__ __forceinline__ void pair_issue(uint64_t r, bool pair, bool even, PairGather& P) {
  const uint32_t plo = static_cast<uint32_t>(qmov<kSwap1>(static_cast<int>(uint32_t(r))));
  const uint32_t phi = static_cast<uint32_t>(qmov<kSwap1>(static_cast<int>(r >> 32)));
  const uint64_t p = uint64_t(phi) << 32 | plo;
  const uint64_t a = pair && !even ? p + 48 : r;
  const uint64_t b = pair && even ? p : r + 48;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    P.X[k] = gload<uint4>(reinterpret_cast<const uint4*>(a), k);
    P.Y[k] = gload<uint4>(reinterpret_cast<const uint4*>(b), k);
  }
}

// One chunk's x-half value: the z lerps of its two corner pairs, then the
// y lerp (vector_field.cpp:41-46 for one component and one x-half).
__device__ __forceinline__ float chunk_half(const uint4 c, float w1, float u1, float w2, float u2) {
  const float z0 = lerp1(__uint_as_float(c.x), __uint_as_float(c.y), w2, u2);
  const float z1 = lerp1(__uint_as_float(c.z), __uint_as_float(c.w), w2, u2);
  return lerp1(z0, z1, w1, u1);
}

// The lane's velocity from a PairGather (all lanes take part: a lane that
// does not step still reduces its partner's chunks).  w, u: the lane's own
// weights (cell_weights).
template <int INTERP>
__device__ __forceinline__ void pair_lerp(const PairGather& P, bool pair, bool even, const float w[3], const float u[3],
                                          float v[3]) {
  // the partner's y and z weights (u = 1 - w, the same single rounding)
  const float pw1 = swapf(w[1]), pw2 = swapf(w[2]);
  const float pu1 = 1.0f - pw1, pu2 = 1.0f - pw2;
  const bool xe = pair && !even;  // X belongs to the partner
  const bool ye = pair && even;   // Y belongs to the partner
  const float xw1 = xe ? pw1 : w[1], xu1 = xe ? pu1 : u[1], xw2 = xe ? pw2 : w[2], xu2 = xe ? pu2 : u[2];
  const float yw1 = ye ? pw1 : w[1], yu1 = ye ? pu1 : u[1], yw2 = ye ? pw2 : w[2], yu2 = ye ? pu2 : u[2];
  float hx[3], hy[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    hx[k] = chunk_half(P.X[k], xw1, xu1, xw2, xu2);
    hy[k] = chunk_half(P.Y[k], yw1, yu1, yw2, yu2);
  }
  // lo = the own record's half values of chunks 0-2, hi = of chunks 3-5
  float lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float send = even ? hy[k] : hx[k];  // what the partner's record needs
    const float recv = swapf(send);  // used only in pair mode
    lo[k] = pair && !even ? recv : hx[k];
    hi[k] = pair && even ? recv : hy[k];
  }
  v[0] = lerp1(lo[0], lo[1], w[0], u[0]);
  v[1] = lerp1(lo[2], hi[0], w[0], u[0]);
  v[2] = lerp1(hi[1], hi[2], w[0], u[0]);
}

// Pair gathers while more than cfg.pair_live lanes of the wave are live (32 by default).

__device__ __forceinline__ float4 pick4(bool c, float4 x, float4 y) {  // value select (no lvalue select)
  return make_float4(c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z, c ? x.w : y.w);
}

// Hybrid union trace into 8-seed grouped slots (PA_SLOTS_GROUP8: slot s of
// seed p = vec4 at ((p / 8) * region + s) * 8 + p % 8; a wave's stores of one
// slot fill whole 128-B lines, and a group's slots are contiguous).  Every lane starts with one seed
// (lane mode, all four components per lane: the most particles in flight
// while the round is throughput-bound).  Once a wave has at most 16 live
// particles, which is where the longest trajectories leave it latency-bound,
// the survivors move through LDS onto quads (quad mode, one component per
// lane) and finish there.  Both modes run the same per-iteration segment
// machine, so the slots are those of the lane engine, bit for bit.
// The switch happens at cfg.quad_switch live particles (16 by default, at most kQuadSwitch: a wave has
// 16 quads).
constexpr int kQuadSwitch = 16;
// A wave's vertex stores fall in its 8 seed groups: byte offsets up to
// 1024 * region - 4 (the w component of seed 63's last slot), which with
// region < 2^21 stay below kDrop = 2^31 - 16.
constexpr int64_t kHybridMaxRegion = (int64_t(1) << 21) - 1;
// Largest byte offset past the wave's first slot that a vertex store of the
// hybrid engine can reach, plus the 16 bytes it stores: lane 63 (group 7,
// seed 7) at slot region - 1.  Every store must end at or below kDrop, the
// descriptor's range (an offset of kDrop is the dropped-store marker).
__host__ __device__ constexpr int64_t hybrid_store_end(int64_t region) {
  return ((7 * region + (region - 1)) * 8 + 7) * 16 + 16;
}
static_assert(hybrid_store_end(kHybridMaxRegion) <= kDrop, "hybrid vertex stores must stay inside the descriptor");
static_assert(hybrid_store_end(kHybridMaxRegion + 1) > kDrop, "kHybridMaxRegion is the tight bound");

template <int S, int MODE, int INTERP>
__global__ __launch_bounds__(kQuadBlock) void k_hybrid_union(const TraceConfig cfg, const pa_particle* __restrict__ P,
                                                             int64_t n, float4* __restrict__ verts, int64_t region,
                                                             int32_t* __restrict__ counts,
                                                             pa_particle* __restrict__ cont,
                                                             unsigned int* __restrict__ n_cont,
                                                             unsigned long long* __restrict__ steps_out, int rule,
                                                             float4* __restrict__ endpoints,
                                                             int32_t* __restrict__ seed_steps,
                                                             int32_t* __restrict__ pairs_out,
                                                             unsigned int* __restrict__ bad) {
  __shared__ QEntry tab[7][4];
  __shared__ QBlock blk[7];
  __shared__ float4 x_last[kQuadBlock / 64][kQuadSwitch];
  __shared__ int32_t x_int[kQuadBlock / 64][kQuadSwitch][10];
  fill_tables<S>(cfg, tab, blk);
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  unsigned long long wave_steps = 0;
  QState st;
  st.pid = logical_block(cfg.xcd_remap) * kQuadBlock + threadIdx.x;
  st.state = st.pid < n ? Q_START : Q_IDLE;
  st.b = -1;
  st.R = st.it = st.slot = st.nsteps = st.orank = st.pairs = st.prevT = 0;
  float4 last = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  FieldDesc f = cfg.field[0];
  if (st.state == Q_START) {
    const pa_particle p = P[st.pid];
    last = make_float4(p.position[0], p.position[1], p.position[2], p.position[3]);
    st.R = p.remaining_iterations;
    st.b = p.vector_field_index;
    st.orank = p.original_rank;
    // A particle this call cannot trace is kept as its seed alone and
    // flagged: field index outside -1..5 (1), its block not loaded (2),
    // more iterations than slots (4), negative iterations (8).
    const unsigned why = (st.b < -1 || st.b > 5) ? 1u
                         : !cfg.has_field[st.b + 1] ? 2u
                         : int64_t(st.R) > region ? 4u
                         : st.R < 0 ? 8u : 0u;
    if (why) {
      atomicOr(bad, why);
      st.b = -1;
      st.R = 1;
    }
    f = cfg.field[st.b + 1];
  }

  // ---- lane mode
  // Cell records: the vertex of an iteration is stored in the next one,
  // right after that iteration's gather is issued: memory operations retire
  // in order, so the gather's wait then leaves the store in flight instead of
  // waiting for it.  The store is issued by the whole wave (store16 with
  // kDrop for lanes with nothing pending), so the wait counts are the same on
  // every path; its offsets stay below kDrop (kHybridMaxRegion).  Nodes (8
  // scattered loads per gather) keep the gather inside the step and store at
  // once.
  constexpr bool kCells = layout_of(S) == PA_LAYOUT_CELLS;
  const int64_t wave0 = logical_block(cfg.xcd_remap) * kQuadBlock + (threadIdx.x & ~63);
  const int64_t vbase = group8_slot(wave0, 0, region);
  const rsrc_t vrs = out_rsrc(verts + uniform(vbase));
  __builtin_amdgcn_s_waitcnt(0x0F70);  // the set-up loads, once (see the hand-off below)
  int pend = kDrop;
  float4 pval = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  // byte offset of the lane's slot 0 (slot s is 128 B further per slot)
  const int lbase = static_cast<int>((group8_slot(st.pid, 0, region) - vbase) * 16);
  uint64_t live;
  for (;;) {
    live = __ballot(st.state != Q_IDLE);
    if (__popcll(live) <= cfg.quad_switch) break;
    // Cell records: advance_rt of a stepping lane, split around the store
    // (a = 0 stepped, next valid; 1 left the block; 2 stop rule).  Every lane
    // issues the gather, so its registers need no merge of paths (whose
    // copies would wait for it); a lane that does not step gathers the
    // record of the wave's first stepping lane (no extra lines).
    int a = 1;
    float4 next;
    if constexpr (kCells) {
      float d[3], q[3];
      bool dfast;
      const bool located = locate(f, last.x, last.y, last.z, d, q, dfast);
      const bool inside = (int(located) & int(st.state == Q_STEP && st.it != st.R - 1)) != 0;
      int m[3];
#pragma unroll
      for (int i = 0; i < 3; ++i)  // floor(q), within the block, for a located lane; cell 0 for the others
        m[i] = located ? static_cast<int>(floorf(q[i])) : 0;
      uint64_t r = reinterpret_cast<uint64_t>(cell_record<S>(f, m));
      const uint64_t W = __ballot(inside);
      if (W != 0 && !inside) {
        const int first = __builtin_ctzll(W);
        // readlane returns int: both halves go through uint32_t so that the
        // low half is not sign-extended into the high one
        const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r >> 32), first));
        const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(uint32_t(r)), first));
        r = uint64_t(hi) << 32 | lo;
      }
      [[maybe_unused]] Gather<S> g;
      [[maybe_unused]] PairGather P;
      const bool pair = __popcll(live) > cfg.pair_live;  // wave-uniform
      const bool even = (lane & 1) == 0;
      if constexpr (storage_of(S) == PA_STORAGE_FP32) pair_issue(r, pair, even, P);
      else gather_record<S>(reinterpret_cast<const uint4*>(r), g);
      store16<kStream>(vrs, pval, pend);
      pend = kDrop;
      // the weights while the loads are in flight (every lane: a pair's
      // partner needs them)
      int mm[3];
      float wt[3], ut[3], v[3];
      cell_weights<INTERP>(f, d, q, dfast, mm, wt, ut, !inside);
      // keep the lerp (which waits for the loads) after the store
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (storage_of(S) == PA_STORAGE_FP32) {
        pair_lerp<INTERP>(P, pair, even, wt, ut, v);
      } else {
        float k[24];
        gather_corners<S>(g, k);
        lerp_weighted(k, wt, ut, v);
      }
      if (inside) a = finish_step<S, MODE, INTERP, STOP_RUNTIME>(cfg, f, last, v, next, rule);
    }
    // The segment machine of one iteration, as selects where it can be (a
    // merge of many paths costs register copies in every iteration): START
    // emits the seed vertex, STEP emits the stepped vertex, or the
    // termination vertex when the segment ends (last iteration, stop rule,
    // left the block); a particle that leaves the block may continue in a
    // neighbour (START again) or as a continuation record.
    {
      const bool active = st.state != Q_IDLE;
      const bool start = st.state == Q_START;
      const bool stepping = active && !start && st.it != st.R - 1;
      if constexpr (!kCells) {
        if (stepping) a = advance_rt<S, MODE, INTERP>(cfg, f, last, next, rule);
      }
      const bool stepped = stepping && a == 0;
      const float4 val = pick4(start, last, pick4(stepped, next, termination_vertex()));
      int nstate = start ? (st.R < 2 ? Q_DONE : Q_STEP) : (stepped ? Q_STEP : Q_DONE);
      st.it = start ? 1 : st.it + int(stepped);
      st.nsteps += int(stepped);
      last = pick4(stepped, next, last);
      st.slot += int(active);
      if (stepping && a == 1) {  // left the block (last is the point that did)
        int nb = -1;
        bool go;
        if (st.b < 0) {
          const int k = face_rule(f, last, cfg.local_mask);
          go = k >= 0 && cfg.has_field[k + 1];
          nb = k;
        } else {
          go = face_rule(f, last, cfg.face_mask[st.b]) == (st.b ^ 1);
        }
        if (go) {
          const int Rn = st.R - st.it;
          if (int64_t(st.slot) + Rn > region) {
            write_continuation(cont + atomicAdd(n_cont, 1u), last, Rn, nb, st.orank, st.pid);
          } else {
            st.b = nb;
            st.R = Rn;
            f = cfg.field[st.b + 1];
            // wait for the descriptor here (a rare path): otherwise its
            // registers count as pending at the loop head of every
            // iteration, whose wait would then take the vertex store too
            __builtin_amdgcn_s_waitcnt(0x0F70);
            nstate = Q_START;
          }
        }
      }
      st.state = active ? nstate : Q_IDLE;
      if constexpr (kCells) {
        pend = active ? lbase + (st.slot - 1) * 128 : kDrop;
        pval = val;
      } else {
        if (active) stream_store(&verts[group8_slot(st.pid, st.slot - 1, region)], val);
      }
      const bool isT = is_termination_vertex(val);
      st.pairs += int(active && st.slot > 1 && !st.prevT && !isT);
      st.prevT = active ? isT : st.prevT;
      if (st.state == Q_DONE) {
        counts[st.pid] = st.slot;
        if (seed_steps) seed_steps[st.pid] = st.nsteps;
        if (pairs_out) pairs_out[st.pid] = st.pairs;
        if (endpoints) endpoints[st.pid] = last;
        wave_steps += static_cast<unsigned long long>(st.nsteps);
        st.state = Q_IDLE;
      }
    }
  }

  if constexpr (kCells) store16<kStream>(vrs, pval, pend);

  // ---- quad mode: live particle k of the wave goes to quad k
  const int nlive = __popcll(live);
  if (nlive > 0) {
    if (st.state != Q_IDLE) {
      const int k = __popcll(live & ((uint64_t(1) << lane) - 1));
      x_last[w][k] = last;
      int32_t* xi = x_int[w][k];
      xi[0] = st.state;
      xi[1] = int32_t(st.pid);
      xi[2] = st.b;
      xi[3] = st.R;
      xi[4] = st.it;
      xi[5] = st.slot;
      xi[6] = st.nsteps;
      xi[7] = st.orank;
      xi[8] = st.pairs;
      xi[9] = st.prevT;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int q = lane >> 2, c = lane & 3;
    QState qs;
    float ql = 0.0f;
    QLane L;
    load_lane(tab, blk, -1, c, L);  // every quad gathers (below): a valid block for the idle ones
    qs.state = Q_IDLE;
    qs.pid = wave0;
    if (q < nlive) {
      const float4 xl = x_last[w][q];
      ql = c == 0 ? xl.x : c == 1 ? xl.y : c == 2 ? xl.z : xl.w;
      const int32_t* xi = x_int[w][q];
      qs.state = xi[0];
      qs.pid = xi[1];
      qs.b = xi[2];
      qs.R = xi[3];
      qs.it = xi[4];
      qs.slot = xi[5];
      qs.nsteps = xi[6];
      qs.orank = xi[7];
      qs.pairs = xi[8];
      qs.prevT = xi[9];
      load_lane(tab, blk, qs.b, c, L);
    }
    // As in the lane mode, a quad's vertex component is stored in the next
    // iteration, right after that iteration's gather; with cell records the
    // gather runs before the state machine, every quad issuing it (a quad
    // that does not step gathers its own cell clamped into the block).
    __builtin_amdgcn_s_waitcnt(0x0F70);  // the lane mode's memory operations, once
    int qpend = kDrop;
    float qval = 0.0f;
    // byte offset of component c of the quad's slot 0 (128 B per slot)
    const int qbase = static_cast<int>((group8_slot(qs.pid, 0, region) - vbase) * 16 + 4 * c);
    for (;;) {
      if (__ballot(qs.state != Q_IDLE) == 0) break;
      QGather G;
      if constexpr (kCells) {
        const bool wl = c == 3;
        bool own;
        G.in = quad_locate(L, wl, ql, G.d, G.q, G.s, G.fast, own) & (qs.state == Q_STEP) & (qs.it != qs.R - 1);
        // the lane's subscript where its axis is inside, else 0: a valid cell for every quad
        const uint32_t ci = quad_sum(wl || !own ? 0u : static_cast<uint32_t>(static_cast<int>(G.s)) * L.stride);
        load_record<S>(L.data, c, ci, G.A, G.B);
      }
      store4<kStream>(vrs, __float_as_int(qval), qpend);
      qpend = kDrop;
      if constexpr (kCells && MODE != STEP_CLASSIC) {
        const bool act = qs.state != Q_IDLE;
        const float val = quad_iterate_sel<S, MODE, INTERP>(cfg, tab, blk, qs, L, ql, c, rule, region, cont, n_cont,
                                                            G);
        qpend = act ? qbase + (qs.slot - 1) * 128 : kDrop;
        qval = val;
        const bool isT = quad_all(val == (c == 3 ? 0.0f : -2.0f));
        qs.pairs += int(act & (qs.slot > 1) & !qs.prevT & !isT);
        qs.prevT = act ? int(isT) : qs.prevT;
      } else if (qs.state != Q_IDLE) {
        const float val = quad_iterate<S, MODE, INTERP>(cfg, tab, blk, qs, L, ql, c, rule, region, cont, n_cont,
                                                        kCells ? &G : nullptr);
        qpend = qbase + (qs.slot - 1) * 128;
        qval = val;
        const bool isT = quad_all(val == (c == 3 ? 0.0f : -2.0f));
        if (qs.slot > 1 && !qs.prevT && !isT) ++qs.pairs;
        qs.prevT = isT;
      }
      {
        if (qs.state == Q_DONE) {
          if (c == 0) {
            counts[qs.pid] = qs.slot;
            if (seed_steps) seed_steps[qs.pid] = qs.nsteps;
            if (pairs_out) pairs_out[qs.pid] = qs.pairs;
            wave_steps += static_cast<unsigned long long>(qs.nsteps);
          }
          if (endpoints) reinterpret_cast<float*>(endpoints)[qs.pid * 4 + c] = ql;
          qs.state = Q_IDLE;
        }
      }
    }
    store4<kStream>(vrs, __float_as_int(qval), qpend);
  }
  wave_steps = wave_sum_u64(wave_steps);
  if (lane == 0 && wave_steps) atomicAdd(steps_out, wave_steps);
}

}  // namespace

pa_status launch_hybrid_union(pa_gpu_ctx* ctx, const TraceConfig& cfg, int mode, const CoopLaunch& l,
                              unsigned long long* steps, unsigned* ncont, int32_t* pairs) {
  if (l.n > INT32_MAX) return fail(ctx, PA_E_INVALID, "hybrid engine: at most 2^31-1 seeds per call");
  if (l.region > kHybridMaxRegion || l.region < 1 || hybrid_store_end(l.region) > kDrop)
    return fail(ctx, PA_E_INVALID, "hybrid engine: region_slots below 2^21");
  const int64_t blocks = (l.n + kQuadBlock - 1) / kQuadBlock;
  dispatch(cfg.field[0].storage, mode, cfg.interp, [&](auto S, auto M, auto I) {
    hipLaunchKernelGGL((k_hybrid_union<decltype(S)::value, decltype(M)::value, decltype(I)::value>),
                       dim3(unsigned(blocks)), dim3(kQuadBlock), 0, ctx->stream, cfg, l.particles, l.n, l.verts,
                       l.region, l.counts, l.cont, ncont, steps, l.stop_rule, l.endpoints, l.seed_steps, pairs,
                       ncont + 1);
  }, cfg.field[0].layout);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

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
