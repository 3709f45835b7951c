// Device-side arithmetic of the advection path.  Every floating-point
// operation here is a single IEEE-rounded operation in the order of the
// reference (contraction is off for the whole library; division and square
// root are the correctly rounded forms), so the FP32 results are
// bit-identical to the reference's SSE arithmetic on the same inputs.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "pa_internal.hpp"

#pragma clang fp contract(off)

namespace pa_amd {

enum { STEP_RK4 = 0, STEP_EULER = 1, STEP_CLASSIC = 2 };
enum { END_NORMAL = 0, END_EXIT = 1 };

__device__ __forceinline__ float bf16_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }
__device__ __forceinline__ float f16_lo(uint32_t u) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(u & 0xffffu)));
}
__device__ __forceinline__ float f16_hi(uint32_t u) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(u >> 16)));
}

// Element i of a global-memory array whose pointer the compiler cannot
// place (e.g. read from LDS): a global load rather than a flat one.
template <class T>
__device__ __forceinline__ T gload(const void* base, uint64_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(1))) T*)(base))[i];
#else
  return static_cast<const T*>(base)[i];
#endif
}

// Template storage code S = storage + 3 * layout.
constexpr int storage_of(int S) { return S % 3; }
constexpr int layout_of(int S) { return S / 3; }

__device__ __forceinline__ float dec16(uint32_t h, int storage) {
  if (storage == PA_STORAGE_BF16) return __uint_as_float(h << 16);
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(h & 0xffffu)));
}

// Correctly rounded a / h from 1/h as an unevaluated sum y + yl (y =
// RN(1/h), yl ~ RN(1/h - y)) without the division sequence.  q0 = RN(a y +
// RN(a yl)) is within a hair of half an ulp of a/h, so within one ulp, and
// one FMA correction is then RN(a/h) by Markstein's theorem (the residual a
// - q0 h of a quotient within one ulp is exact, and RN(q0 + r y) with y the
// correctly rounded reciprocal is the correctly rounded quotient; P.
// Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, 2nd ed., sec. 4.7).  It needs no intermediate
// underflow or overflow: the callers take it only for a in {0} or
// [2^-40, 2^40] with h in [2^-20, 2^20].  A zero a may give a zero of either
// sign: div_corrected restores the sign of a; div_unsigned leaves it to
// callers that give the result its sign themselves or only floor and
// compare it (where -0 and +0 act alike).
__device__ __forceinline__ float div_unsigned(float a, float h, float y, float yl) {
  const float q0 = __builtin_fmaf(a, y, a * yl);
  return __builtin_fmaf(__builtin_fmaf(-q0, h, a), y, q0);
}
__device__ __forceinline__ float div_corrected(float a, float h, float y, float yl) {
  return __builtin_copysignf(div_unsigned(a, h, y, yl), a);
}

__device__ __forceinline__ bool div_range(float a) {
  const float x = fabsf(a);
  return (int(x == 0.0f) | (int(x >= 0x1p-40f) & int(x <= 0x1p40f))) != 0;
}

// fmodf(d, h) of a located point from its cell subscript m = floor(q),
// q = RN(d / h), for d >= 0 (or -0) with d / h not underflowing (the
// callers' range check): q rounds the true quotient t either to within
// [m, m + 1), where d - m h is the remainder, or up to the integer m from
// just below, where d - m h lies in (-h, 0) and the remainder is
// d - m h + h.  Both are exact in FP32 (a remainder is always
// representable, and d - m h then is a small multiple of ulp(d)), so the
// FMA and the add round nothing.  A zero remainder may come out as -0 or
// +0: the caller gives the weight the sign of d (fmodf's sign rule).
__device__ __forceinline__ float rem_located(float d, float h, float m) {
  const float r0 = __builtin_fmaf(-m, h, d);
  return r0 < 0.0f ? r0 + h : r0;
}

// A remainder r in [0, h) that div_corrected may divide is 0 or >= 2^-40,
// i.e. rem_key(r) >= kRemKeyMin (unsigned): a zero's key wraps to the top,
// one of (0, 2^-40) falls below.  The keys of three remainders take one
// unsigned minimum.
__device__ __forceinline__ uint32_t rem_key(float r) { return __float_as_uint(r) - 1u; }
constexpr uint32_t kRemKeyMin = 0x2B800000u - 1u;  // bits of 2^-40, minus 1
__device__ __forceinline__ bool rem_range(float r) { return rem_key(r) >= kRemKeyMin; }

// fmod(d, h) / h of the reference's weight (vector_field.cpp:28-30) from
// the remainder r = rem_located(d, h, floor(q)): correctly rounded, with
// the sign of d, for d and r within div_corrected's ranges.
__device__ __forceinline__ float weight_located(float d, float h, float y, float yl, float r) {
  return __builtin_copysignf(div_unsigned(r, h, y, yl), d);
}

// Exact fmodf(d, h) for h > 0 and a located point (0 <= q = d/h rounded <
// 2^22, so |d| < 2^22 h): n = floor(q) is the true integer quotient or off
// by one, the FMA residual d - n h is exact, and its sign (or its reaching
// h) says which neighbour of n is the true one.  Branch-free.
__device__ __forceinline__ float exact_fmod(float d, float h, float q) {
  const float ax = fabsf(d);
  float n = floorf(fabsf(q));
  const float r = __builtin_fmaf(-n, h, ax);
  n += static_cast<float>(static_cast<int>(r >= h) - static_cast<int>(r < 0.0f));
  return __builtin_copysignf(__builtin_fmaf(-n, h, ax), d);
}

// vector_field::contains (vector_field.cpp:12-21) sharing the quotients
// with interpolate: q_i = (p_i - o_i) / h_i; inside iff 0 <= floor(q_i) <=
// dims_i - 2 (a negative or NaN subscript fails through the reference's
// size_t cast).  A zero quotient may carry either sign: every use floors
// it, compares it or subtracts its floor, where -0 and +0 act alike.
// fast: d and the spacing were within div_corrected's ranges (cell_weights
// takes it).
__device__ __forceinline__ bool locate(const FieldDesc& f, float px, float py, float pz, float d[3], float q[3],
                                       bool& fast) {
  d[0] = px - f.offset[0];
  d[1] = py - f.offset[1];
  d[2] = pz - f.offset[2];
  // d within [2^-40, 2^40] on every axis (a zero d takes the division, which
  // gives the same quotient; a NaN d gives a NaN quotient either way);
  // bitwise on purpose: no short-circuit branches
  const float amin = fminf(fminf(fabsf(d[0]), fabsf(d[1])), fabsf(d[2]));
  const float amax = fmaxf(fmaxf(fabsf(d[0]), fabsf(d[1])), fabsf(d[2]));
  fast = (int(f.fast_div != 0) & int(amin >= 0x1p-40f) & int(amax <= 0x1p40f)) != 0;
  // the corrected quotient everywhere; the division only where some lane of
  // the wave needs it (a wave-uniform branch, not an if/else per lane)
#pragma unroll
  for (int i = 0; i < 3; ++i) q[i] = div_unsigned(d[i], f.spacing[i], f.inv_spacing[i], f.inv_lo[i]);
  if (__builtin_expect(__ballot(!fast) != 0, 0)) {
    if (!fast) {
#pragma unroll
      for (int i = 0; i < 3; ++i) q[i] = d[i] / f.spacing[i];
    }
  }
  // 0 <= s <= dims - 2 as s == median(s, 0, dims - 2) (false for a NaN s)
  int inside = 1;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float s = floorf(q[i]);
    inside &= int(__builtin_amdgcn_fmed3f(s, 0.0f, f.cell_hi[i]) == s);
  }
  return inside != 0;
}
__device__ __forceinline__ bool locate(const FieldDesc& f, float px, float py, float pz, float d[3], float q[3]) {
  bool fast;
  return locate(f, px, py, pz, d, q, fast);
}

__device__ __forceinline__ float lerp1(float a, float b, float w, float u) { return u * a + w * b; }

// Cell index and weights of vector_field::interpolate (vector_field.cpp:
// 27-31) at a located point: m = floor(q); w = fmod(d, h) / h (reference)
// or q - m (consistent); u = 1 - w.  The reference weight takes the
// remainder from m (rem_located) and one corrected quotient while d and the
// remainder are within div_corrected's ranges, else fmodf and division.
// dfast: locate's flag for d; unused: the caller will not use these weights
// (a lane that does not step), so they never take the slow path.
template <int INTERP>
__device__ __forceinline__ void cell_weights(const FieldDesc& f, const float d[3], const float q[3], bool dfast,
                                             int m[3], float w[3], float u[3], bool unused = false) {
  float fl[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    fl[i] = floorf(q[i]);
    m[i] = static_cast<int>(fl[i]);
  }
  if constexpr (INTERP == PA_INTERP_CONSISTENT) {
#pragma unroll
    for (int i = 0; i < 3; ++i) w[i] = q[i] - fl[i];
  } else {
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = rem_located(d[i], f.spacing[i], fl[i]);
    const uint32_t key = min(min(rem_key(r[0]), rem_key(r[1])), rem_key(r[2]));
    const bool fast = (int(unused) | (int(dfast) & int(key >= kRemKeyMin))) != 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) w[i] = weight_located(d[i], f.spacing[i], f.inv_spacing[i], f.inv_lo[i], r[i]);
    if (__builtin_expect(__ballot(!fast) != 0, 0)) {  // as in locate
      if (!fast) {
#pragma unroll
        for (int i = 0; i < 3; ++i) w[i] = exact_fmod(d[i], f.spacing[i], q[i]) / f.spacing[i];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) u[i] = 1.0f - w[i];
}
template <int INTERP>
__device__ __forceinline__ void cell_weights(const FieldDesc& f, const float d[3], const float q[3], int m[3],
                                             float w[3], float u[3]) {
  const bool dfast =
      (int(f.fast_div != 0) & int(div_range(d[0])) & int(div_range(d[1])) & int(div_range(d[2]))) != 0;
  cell_weights<INTERP>(f, d, q, dfast, m, w, u);
}

// The lerp tree of vector_field.cpp:41-47 over the 8 corners
// c[0..23] = c000 c001 c010 c011 c100 c101 c110 c111 (xyz each).
__device__ __forceinline__ void lerp_corners(const float* c, const float w[3], const float u[3], float v[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float c00 = lerp1(c[0 + k], c[3 + k], w[2], u[2]);
    const float c01 = lerp1(c[6 + k], c[9 + k], w[2], u[2]);
    const float c10 = lerp1(c[12 + k], c[15 + k], w[2], u[2]);
    const float c11 = lerp1(c[18 + k], c[21 + k], w[2], u[2]);
    const float c0 = lerp1(c00, c01, w[1], u[1]);
    const float c1 = lerp1(c10, c11, w[1], u[1]);
    v[k] = lerp1(c0, c1, w[0], u[0]);
  }
}

// Cell records (PA_LAYOUT_CELLS): one record per cell holds x of c000 c001
// c010 c011 c100 c101 c110 c111, then y, then z; FP32 records are 8 uint4
// (6 used), 16-bit records 4 uint4 (3 used).
template <int S>
constexpr int rec_used() { return storage_of(S) == PA_STORAGE_FP32 ? 6 : 3; }
template <int S>
constexpr int rec_size() { return storage_of(S) == PA_STORAGE_FP32 ? 8 : 4; }

// The record of the cell with subscripts m (fewer than 2^31 records,
// checked when the layout is built).
template <int S>
__device__ __forceinline__ const uint4* cell_record(const FieldDesc& f, const int m[3]) {
  const uint32_t ci = uint32_t(m[0]) * uint32_t(f.cstride0) + uint32_t(m[1]) * uint32_t(f.cstride1) + uint32_t(m[2]);
  return static_cast<const uint4*>(f.cells) + uint64_t(ci) * rec_size<S>();
}

// The 24 corner values of a record (its rec_used<S>() uint4 as loaded),
// component-planar: k[comp * 8 + corner].
template <int S>
__device__ __forceinline__ void decode_record(const uint4* r, float k[24]) {
  if constexpr (storage_of(S) == PA_STORAGE_FP32) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      k[4 * i + 0] = __uint_as_float(r[i].x);
      k[4 * i + 1] = __uint_as_float(r[i].y);
      k[4 * i + 2] = __uint_as_float(r[i].z);
      k[4 * i + 3] = __uint_as_float(r[i].w);
    }
  } else {
    constexpr int st = storage_of(S);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      k[8 * i + 0] = dec16(r[i].x, st); k[8 * i + 1] = dec16(r[i].x >> 16, st);
      k[8 * i + 2] = dec16(r[i].y, st); k[8 * i + 3] = dec16(r[i].y >> 16, st);
      k[8 * i + 4] = dec16(r[i].z, st); k[8 * i + 5] = dec16(r[i].z >> 16, st);
      k[8 * i + 6] = dec16(r[i].w, st); k[8 * i + 7] = dec16(r[i].w >> 16, st);
    }
  }
}

// The lerp tree of vector_field.cpp:41-47 (z, then y, then x) over
// component-planar corners k[comp * 8 + corner], corner order c000 c001
// c010 c011 c100 c101 c110 c111.
__device__ __forceinline__ void lerp_weighted(const float k[24], const float w[3], const float u[3], float v[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float* c = k + 8 * i;
    const float c00 = lerp1(c[0], c[1], w[2], u[2]), c01 = lerp1(c[2], c[3], w[2], u[2]);
    const float c10 = lerp1(c[4], c[5], w[2], u[2]), c11 = lerp1(c[6], c[7], w[2], u[2]);
    const float c0 = lerp1(c00, c01, w[1], u[1]), c1 = lerp1(c10, c11, w[1], u[1]);
    v[i] = lerp1(c0, c1, w[0], u[0]);
  }
}
template <int INTERP>
__device__ __forceinline__ void lerp_planar(const FieldDesc& f, const float d[3], const float q[3], const float k[24],
                                            float v[3]) {
  int m[3];
  float w[3], u[3];
  cell_weights<INTERP>(f, d, q, m, w, u);
  lerp_weighted(k, w, u, v);
}

// The loads of one cell's 8 corners as issued (gather_issue), decoded later
// (gather_corners), so that a caller can place other memory operations
// between the issue and the first use.  Cell records: the rec_used<S>()
// uint4 of the record; nodes: the 8 nodes c000 c001 c010 c011 c100 c101
// c110 c111 (FP32 as uint4, 16-bit in .x .y).
template <int S>
struct Gather {
  uint4 r[8];
};

template <int S>
__device__ __forceinline__ void gather_record(const uint4* r, Gather<S>& g) {
#pragma unroll
  for (int i = 0; i < rec_used<S>(); ++i) g.r[i] = gload<uint4>(r, i);
}

template <int S>
__device__ __forceinline__ void gather_issue(const FieldDesc& f, const int m[3], Gather<S>& g) {
  if constexpr (layout_of(S) == PA_LAYOUT_CELLS) {
    gather_record<S>(cell_record<S>(f, m), g);
  } else {
    // fewer than 2^32 nodes (checked when the field is made)
    const uint32_t s0 = uint32_t(f.stride0), s1 = uint32_t(f.stride1);
    const uint32_t i000 = uint32_t(m[0]) * s0 + uint32_t(m[1]) * s1 + uint32_t(m[2]);
    const uint32_t at[8] = {i000, i000 + 1, i000 + s1, i000 + s1 + 1, i000 + s0, i000 + s0 + 1, i000 + s0 + s1,
                            i000 + s0 + s1 + 1};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if constexpr (storage_of(S) == PA_STORAGE_FP32) {
        g.r[j] = static_cast<const uint4*>(f.data)[at[j]];
      } else {
        const uint2 u = static_cast<const uint2*>(f.data)[at[j]];
        g.r[j].x = u.x;
        g.r[j].y = u.y;
      }
    }
  }
}

// The 24 corner values, component-planar: k[comp * 8 + corner].
template <int S>
__device__ __forceinline__ void gather_corners(const Gather<S>& g, float k[24]) {
  if constexpr (layout_of(S) == PA_LAYOUT_CELLS) {
    decode_record<S>(g.r, k);
  } else {
    constexpr int st = storage_of(S);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if constexpr (st == PA_STORAGE_FP32) {
        k[j] = __uint_as_float(g.r[j].x);
        k[8 + j] = __uint_as_float(g.r[j].y);
        k[16 + j] = __uint_as_float(g.r[j].z);
      } else {
        k[j] = dec16(g.r[j].x, st);
        k[8 + j] = dec16(g.r[j].x >> 16, st);
        k[16 + j] = dec16(g.r[j].y, st);
      }
    }
  }
}

// vector_field::interpolate (vector_field.cpp:22-48): m = floor(q), w =
// fmod(d, h) / h, lerp (linear_interpolate.hpp:6-10) along z, y, x.  With
// INTERP == PA_INTERP_CONSISTENT the weight is q - floor(q) (exact in FP32
// for 0 <= q < 2^23).  The gathers are issued first so that they are in
// flight while the weights are computed.
template <int S, int INTERP>
__device__ __forceinline__ void interpolate(const FieldDesc& f, const float d[3], const float q[3], float v[3]) {
  int m[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) m[i] = static_cast<int>(floorf(q[i]));
  Gather<S> g;
  gather_issue<S>(f, m, g);
  float k[24];
  gather_corners<S>(g, k);
  lerp_planar<INTERP>(f, d, q, k, v);
}

// Stop test (particle_tracer.cpp:312): Eigen isZero() with the float
// dummy precision 1e-5f, or the |v| < eps extension.
enum { STOP_RUNTIME = -1 };  // the rule is a runtime argument, selected without a branch

template <int STOP>
__device__ __forceinline__ bool stops(const float v[3], float eps2, int rule = STOP) {
  if constexpr (STOP == STOP_RUNTIME) {
    const float n2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    const bool zero = (int(fabsf(v[0]) <= 1e-5f) & int(fabsf(v[1]) <= 1e-5f) & int(fabsf(v[2]) <= 1e-5f)) != 0;
    return rule == PA_STOP_NORM ? n2 < eps2 : zero;
  } else if constexpr (STOP == PA_STOP_NORM) {
    const float n2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    return n2 < eps2;
  } else {
    return fabsf(v[0]) <= 1e-5f && fabsf(v[1]) <= 1e-5f && fabsf(v[2]) <= 1e-5f;
  }
}

// Frozen-derivative step (particle_tracer.cpp:315-324 with odeint's
// runge_kutta4 final stage: 1*x + b0dt*k + b1dt*k + b2dt*k + b3dt*k, left to
// right; euler: 1*x + dt*k).  w is stepped with derivative 0 like the
// reference's vector4 state.
template <int MODE>
__device__ __forceinline__ float4 step_frozen(const TraceConfig& c, float4 p, const float v[3]) {
  float4 o;
  if constexpr (MODE == STEP_EULER) {
    o.x = p.x + c.dt * v[0];
    o.y = p.y + c.dt * v[1];
    o.z = p.z + c.dt * v[2];
    o.w = p.w + c.w_euler;  // p.w + dt 0
  } else {
    o.x = (((p.x + c.a0 * v[0]) + c.a1 * v[0]) + c.a2 * v[0]) + c.a3 * v[0];
    o.y = (((p.y + c.a0 * v[1]) + c.a1 * v[1]) + c.a2 * v[1]) + c.a3 * v[1];
    o.z = (((p.z + c.a0 * v[2]) + c.a1 * v[2]) + c.a2 * v[2]) + c.a3 * v[2];
    o.w = p.w + c.w_rk4;  // (((p.w + a0 0) + a1 0) + a2 0) + a3 0
  }
  return o;
}

// Textbook RK4 (extension): odeint's generic-RK stage sums with the field
// evaluated at every stage point.  False if a stage point leaves the block.
template <int S, int INTERP>
__device__ __forceinline__ bool step_classic(const TraceConfig& c, const FieldDesc& f, float4 p, const float k1[3],
                                             float4& out) {
  float d[3], q[3], k2[3], k3[3], k4[3];
  float x0 = p.x + c.half_dt * k1[0];
  float x1 = p.y + c.half_dt * k1[1];
  float x2 = p.z + c.half_dt * k1[2];
  if (!locate(f, x0, x1, x2, d, q)) return false;
  interpolate<S, INTERP>(f, d, q, k2);
  x0 = (p.x + c.zero_dt * k1[0]) + c.half_dt * k2[0];
  x1 = (p.y + c.zero_dt * k1[1]) + c.half_dt * k2[1];
  x2 = (p.z + c.zero_dt * k1[2]) + c.half_dt * k2[2];
  if (!locate(f, x0, x1, x2, d, q)) return false;
  interpolate<S, INTERP>(f, d, q, k3);
  x0 = ((p.x + c.zero_dt * k1[0]) + c.zero_dt * k2[0]) + c.one_dt * k3[0];
  x1 = ((p.y + c.zero_dt * k1[1]) + c.zero_dt * k2[1]) + c.one_dt * k3[1];
  x2 = ((p.z + c.zero_dt * k1[2]) + c.zero_dt * k2[2]) + c.one_dt * k3[2];
  if (!locate(f, x0, x1, x2, d, q)) return false;
  interpolate<S, INTERP>(f, d, q, k4);
  out.x = (((p.x + c.a0 * k1[0]) + c.a1 * k2[0]) + c.a2 * k3[0]) + c.a3 * k4[0];
  out.y = (((p.y + c.a0 * k1[1]) + c.a1 * k2[1]) + c.a2 * k3[1]) + c.a3 * k4[1];
  out.z = (((p.z + c.a0 * k1[2]) + c.a1 * k2[2]) + c.a2 * k3[2]) + c.a3 * k4[2];
  out.w = p.w + c.w_rk4;  // (((p.w + a0 0) + a1 0) + a2 0) + a3 0
  return true;
}

// Neighbour choice of particle_tracer.cpp:288-293: first of x<min, x>max,
// y<min, y>max, z<min, z>max whose neighbour exists (mask bit).
__device__ __forceinline__ int face_rule(const FieldDesc& f, float4 p, int mask) {
  if (p.x < f.offset[0] && (mask & 1)) return 0;
  if (p.x > f.upper[0] && (mask & 2)) return 1;
  if (p.y < f.offset[1] && (mask & 4)) return 2;
  if (p.y > f.upper[1] && (mask & 8)) return 3;
  if (p.z < f.offset[2] && (mask & 16)) return 4;
  if (p.z > f.upper[2] && (mask & 32)) return 5;
  return -1;
}

__device__ __forceinline__ float4 termination_vertex() { return make_float4(-2.0f, -2.0f, -2.0f, 0.0f); }

// The stop test and the step of advance() once v = field(last) is known.
template <int S, int MODE, int INTERP, int STOP>
__device__ __forceinline__ int finish_step(const TraceConfig& c, const FieldDesc& f, float4 last, const float v[3],
                                           float4& next, int rule = STOP) {
  if (stops<STOP>(v, c.stop_eps2, rule)) return 2;
  if constexpr (MODE == STEP_CLASSIC) {
    // A stage point outside the block: take the frozen-derivative step so
    // the particle reaches the face and is handed off by the next contains.
    if (!step_classic<S, INTERP>(c, f, last, v, next)) next = step_frozen<STEP_RK4>(c, last, v);
  } else {
    next = step_frozen<MODE>(c, last, v);
  }
  return 0;
}

// One field evaluation plus one step from `last`.  Returns 0 = stepped
// (next valid), 1 = left the block (contains failed), 2 = stop rule.
template <int S, int MODE, int INTERP, int STOP>
__device__ __forceinline__ int advance(const TraceConfig& c, const FieldDesc& f, float4 last, float4& next) {
  float d[3], q[3], v[3];
  if (!locate(f, last.x, last.y, last.z, d, q)) return 1;
  interpolate<S, INTERP>(f, d, q, v);
  return finish_step<S, MODE, INTERP, STOP>(c, f, last, v, next);
}

// advance() with the stop rule as a runtime argument, so a step loop that
// serves both rules has a single body.
template <int S, int MODE, int INTERP>
__device__ __forceinline__ int advance_rt(const TraceConfig& c, const FieldDesc& f, float4 last, float4& next,
                                          int rule) {
  float d[3], q[3], v[3];
  if (!locate(f, last.x, last.y, last.z, d, q)) return 1;
  interpolate<S, INTERP>(f, d, q, v);
  return finish_step<S, MODE, INTERP, STOP_RUNTIME>(c, f, last, v, next, rule);
}

// Whether a step loop over unrelated particles (the pool, the serving
// kernel) keeps each lane's last cell's corner data in LDS
// (keep_words<S>() uint4 a lane), for the steppers with one field
// evaluation per step.
template <int S, int MODE>
constexpr bool keeps_record() {
  return MODE != STEP_CLASSIC;
}
// uint4 a lane keeps: a cell record's used part; FP32 nodes, the 8 corner
// nodes; 16-bit nodes, the 8 corners as uint2, two to a uint4.
template <int S>
constexpr int keep_words() {
  return layout_of(S) == PA_LAYOUT_CELLS ? rec_used<S>() : storage_of(S) == PA_STORAGE_FP32 ? 8 : 4;
}

// advance_rt() for a lane that keeps its last cell's corner data in LDS
// (keep[0..keep_words<S>()-1][threadIdx.x]; `kept` names the cell: its
// record's index, or its corner-000 node's; 0xffffffff for none, which no
// cell has: fewer than 2^31 records and 2^32 nodes): a step that stays in
// the kept cell reads the corners from LDS, any other step gathers them and
// keeps them.  Lanes that follow unrelated particles make vector-L1 misses
// of their own with every gather; most steps stay in their cell and now
// make none.  The values are those of the gather.
template <int S, int MODE, int INTERP, int BLOCK>
__device__ __forceinline__ int advance_kept(const TraceConfig& c, const FieldDesc& f, float4 last, float4& next,
                                            int rule, uint32_t& kept, uint4 (*keep)[BLOCK]) {
  float d[3], q[3], v[3];
  if (!locate(f, last.x, last.y, last.z, d, q)) return 1;
  int m[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) m[i] = static_cast<int>(floorf(q[i]));
  constexpr bool kCells = layout_of(S) == PA_LAYOUT_CELLS;
  const uint32_t ci = kCells ? uint32_t(m[0]) * uint32_t(f.cstride0) + uint32_t(m[1]) * uint32_t(f.cstride1) +
                                   uint32_t(m[2])
                             : uint32_t(m[0]) * uint32_t(f.stride0) + uint32_t(m[1]) * uint32_t(f.stride1) +
                                   uint32_t(m[2]);
  const bool hit = ci == kept;
  constexpr int W = keep_words<S>();
  constexpr bool kPairs = !kCells && storage_of(S) != PA_STORAGE_FP32;  // two uint2 nodes to a uint4
  Gather<S> g;
  if (hit) {
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint4 w = keep[k][threadIdx.x];
      if constexpr (kPairs) {
        g.r[2 * k] = make_uint4(w.x, w.y, 0u, 0u);
        g.r[2 * k + 1] = make_uint4(w.z, w.w, 0u, 0u);
      } else {
        g.r[k] = w;
      }
    }
  } else if constexpr (kCells) {
    const uint4* r = static_cast<const uint4*>(f.cells) + uint64_t(ci) * rec_size<S>();
#pragma unroll
    for (int k = 0; k < W; ++k) g.r[k] = gload<uint4>(r, k);
  } else {
    gather_issue<S>(f, m, g);
  }
  float k[24];
  gather_corners<S>(g, k);
  lerp_planar<INTERP>(f, d, q, k, v);
  if (!hit) {
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if constexpr (kPairs)
        keep[j][threadIdx.x] = make_uint4(g.r[2 * j].x, g.r[2 * j].y, g.r[2 * j + 1].x, g.r[2 * j + 1].y);
      else
        keep[j][threadIdx.x] = g.r[j];
    }
    kept = ci;
  }
  return finish_step<S, MODE, INTERP, STOP_RUNTIME>(c, f, last, v, next, rule);
}

// Raw buffer stores whose descriptor (out_rsrc) covers byte offsets below
// kDrop: a lane with nothing to store passes kDrop and the hardware drops
// the store, so the whole wave issues it without a branch (the same wait
// counts on every path) and without traffic for the dropped lanes.
constexpr int kDrop = 0x7FFFFFF0;
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t out_rsrc(void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, kDrop, 0x00020000);
}
//
// Cache policy of a store: kCached, or kStream, the non-temporal hint (the
// L2 evicts such lines first).  The tracers stream their vertices: each
// vertex line is written once and read by a later kernel, and streaming it
// leaves the L2 to the field records that the next steps read again (share
// round, R = 1002: trace 1.455 -> 1.408 ms FP32, 1.36 -> 1.31 ms BF16).
// The finish kernels keep cached stores (streamed: 0.83 -> 1.27 ms).
enum : int { kCached = 0, kStream = 2 };
template <int POLICY = kCached>
__device__ __forceinline__ void store16(rsrc_t r, float4 v, int off) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const u4 d = {__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w)};
  __builtin_amdgcn_raw_buffer_store_b128(d, r, off, 0, POLICY);
}
template <int POLICY = kCached>
__device__ __forceinline__ void store4(rsrc_t r, int32_t v, int off) {
  __builtin_amdgcn_raw_buffer_store_b32(uint32_t(v), r, off, 0, POLICY);
}
// A streaming store through a plain pointer (same hint as kStream).
__device__ __forceinline__ void stream_store(float4* p, float4 v) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 d = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(d, reinterpret_cast<v4*>(p));
}

// A value the whole wave shares, held in scalar registers (a buffer
// descriptor built from it then needs no per-lane loop).
__device__ __forceinline__ int64_t uniform(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(uint64_t(x)));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(uint64_t(x) >> 32));
  return int64_t((uint64_t(hi) << 32) | lo);
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup i runs on
// XCD i % 8).  With remap on, XCD x processes the contiguous x-th eighth of
// the logical workgroups, so particles that start close together share an
// XCD's L2.  A bijection for any grid size.
//
// remap >= 2 keeps the work spread over the XCDs instead: XCD x takes runs of
// `remap` consecutive logical workgroups, x, x + 8, ... (run-wise round robin),
// so neighbouring particles share an L2 within a run while every XCD gets a
// share of every part of the list (seed lists whose long-lived particles sit
// in one part of the list would otherwise load a few XCDs only).  Workgroups
// past the last whole round of 8 runs keep their own index (a bijection).
__device__ __forceinline__ int64_t logical_block(int remap) {
  const int64_t b = blockIdx.x;
  if (!remap) return b;
  const int64_t nb = gridDim.x;
  if (remap >= 2) {
    const int64_t c = remap, whole = nb / (8 * c) * (8 * c);
    if (b >= whole) return b;
    const int64_t x = b % 8, k = b / 8;
    return ((k / c) * 8 + x) * c + k % c;
  }
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t x = b % 8, k = b / 8;
  return x * q + (x < r ? x : r) + k;
}

template <int V>
using IC = std::integral_constant<int, V>;

// Calls f(IC<storage>, IC<mode>, IC<interp>) with the run-time choices as
// compile-time constants (one kernel instantiation per combination).
template <class F>
inline void dispatch(int storage, int mode, int interp, F&& f, int layout = PA_LAYOUT_NODES) {
  auto with_i = [&](auto s, auto m) {
    if (interp == PA_INTERP_CONSISTENT) f(s, m, IC<PA_INTERP_CONSISTENT>{});
    else f(s, m, IC<PA_INTERP_REFERENCE>{});
  };
  auto with_m = [&](auto s) {
    if (mode == STEP_EULER) with_i(s, IC<STEP_EULER>{});
    else if (mode == STEP_CLASSIC) with_i(s, IC<STEP_CLASSIC>{});
    else with_i(s, IC<STEP_RK4>{});
  };
  if (layout == PA_LAYOUT_CELLS) {
    if (storage == PA_STORAGE_BF16) with_m(IC<PA_STORAGE_BF16 + 3>{});
    else if (storage == PA_STORAGE_FP16) with_m(IC<PA_STORAGE_FP16 + 3>{});
    else with_m(IC<PA_STORAGE_FP32 + 3>{});
  } else {
    if (storage == PA_STORAGE_BF16) with_m(IC<PA_STORAGE_BF16>{});
    else if (storage == PA_STORAGE_FP16) with_m(IC<PA_STORAGE_FP16>{});
    else with_m(IC<PA_STORAGE_FP32>{});
  }
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

}  // namespace pa_amd
