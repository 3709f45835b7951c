// Colour mapping of pa::color_generator (color_generator.cpp:30-98) for
// one segment (vertex a -> vertex b), shared by the colour stage and the
// fused round.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "pa_internal.hpp"

#pragma clang fp contract(off)

namespace pa_amd {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

// color_generator.cpp:55-64.
__device__ __forceinline__ float4 hue_to_rgba(float h) {
  return make_float4(clamp01(fabsf(6.0f * h - 3.0f) - 1.0f), clamp01(2.0f - fabsf(6.0f * h - 2.0f)),
                     clamp01(2.0f - fabsf(6.0f * h - 4.0f)), 1.0f);
}

// color_generator.cpp:65-72.
__device__ __forceinline__ float4 hsl_to_rgba(float h, float s, float l) {
  float4 c = hue_to_rgba(h);
  const float pre = (1.0f - fabsf(2.0f * l - 1.0f)) * s;
  c.x = (c.x - 0.5f) * pre + l;
  c.y = (c.y - 0.5f) * pre + l;
  c.z = (c.z - 0.5f) * pre + l;
  c.w = 1.0f;
  return c;
}

// color_generator.cpp:73-79.
__device__ __forceinline__ float4 hsv_to_rgba(float h, float s, float v) {
  float4 c = hue_to_rgba(h);
  c.x = ((c.x - 1.0f) * s + 1.0f) * v;
  c.y = ((c.y - 1.0f) * s + 1.0f) * v;
  c.z = ((c.z - 1.0f) * s + 1.0f) * v;
  c.w = 1.0f;
  return c;
}

__device__ __forceinline__ float4 map_color(float4 t, float4 d, int mode, float param) {
  if (mode == PA_COLOR_RGB) return make_float4(fabsf(t.x), fabsf(t.z), fabsf(t.y), 1.0f);  // :95-98
  if (mode == PA_COLOR_MAGNITUDE) {
    const float len = sqrtf((d.x * d.x + d.y * d.y) + d.z * d.z);
    return hue_to_rgba(clamp01(len / param) * (2.0f / 3.0f));
  }
  // to_spherical (coordinates.hpp:8-16): std::pow(float,int) promotes to
  // double, so the radius is a double sum and square root stored to float.
  const double r = sqrt((double(t.x) * double(t.x) + double(t.y) * double(t.y)) + double(t.z) * double(t.z));
  const float s0 = float(r);
  float s1 = atan2f(t.y, t.x);
  float s2 = acosf(t.z / s0);
  // correct_range (color_generator.cpp:41-54).
  if (s1 < 0.0f) s1 += float(kPi);
  if (s1 >= float(kPi)) s1 -= float(kPi);
  s1 = float(kPi) - s1;
  if (s2 < 0.0f) s2 = fabsf(s2);
  if (s2 >= float(kPi / 2.0)) s2 = float(kPi) - s2;
  s1 = float(double(s1) / kPi);
  s2 = float(double(s2) / (kPi / 2.0f));
  switch (mode) {
    case PA_COLOR_HSL_CONSTANT_S: return hsl_to_rgba(s1, param, s2);
    case PA_COLOR_HSL_CONSTANT_L: return hsl_to_rgba(s1, s2, param);
    case PA_COLOR_HSV_CONSTANT_S: return hsv_to_rgba(s1, param, s2);
    default: return hsv_to_rgba(s1, s2, param);
  }
}

// Eigen normalized() of the Vector4f tangent b - a (Dot.h:122-133) with the
// SSE packet reduction (a0 + a2) + (a1 + a3) (arch/SSE/PacketMath.h:540-551),
// then map_color.
__device__ __forceinline__ float4 segment_color(float4 a, float4 b, int mode, float param) {
  const float4 d = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, b.w - a.w);
  const float z = (d.x * d.x + d.z * d.z) + (d.y * d.y + d.w * d.w);
  float4 t = d;
  if (z > 0.0f) {
    const float s = sqrtf(z);
    t = make_float4(d.x / s, d.y / s, d.z / s, d.w / s);
  }
  return map_color(t, d, mode, param);
}

// Calls f(std::integral_constant<int, MODE>) for a runtime colour mode, so
// kernels specialise on it (the other modes' arithmetic folds away).
template <typename F>
inline bool with_color_mode(int mode, F&& f) {
  switch (mode) {
    case PA_COLOR_HSL_CONSTANT_S: f(std::integral_constant<int, PA_COLOR_HSL_CONSTANT_S>{}); return true;
    case PA_COLOR_HSL_CONSTANT_L: f(std::integral_constant<int, PA_COLOR_HSL_CONSTANT_L>{}); return true;
    case PA_COLOR_HSV_CONSTANT_S: f(std::integral_constant<int, PA_COLOR_HSV_CONSTANT_S>{}); return true;
    case PA_COLOR_HSV_CONSTANT_V: f(std::integral_constant<int, PA_COLOR_HSV_CONSTANT_V>{}); return true;
    case PA_COLOR_RGB: f(std::integral_constant<int, PA_COLOR_RGB>{}); return true;
    case PA_COLOR_MAGNITUDE: f(std::integral_constant<int, PA_COLOR_MAGNITUDE>{}); return true;
    default: return false;
  }
}

__device__ __forceinline__ bool is_termination_vertex(float4 a) {
  return a.x == -2.0f && a.y == -2.0f && a.z == -2.0f && a.w == 0.0f;
}

}  // namespace pa_amd
