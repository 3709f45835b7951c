// pa::color_generator::generate on the device: one lane per index, tangent
// of the segment (vertex idx -> idx+1) mapped to RGBA.
#include "pa_internal.hpp"

#pragma clang fp contract(off)

namespace pa_amd {
namespace {

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

__global__ __launch_bounds__(256) void k_color(const float4* __restrict__ v, const int32_t* __restrict__ idx,
                                               int64_t n_idx, int mode, float param, float4* __restrict__ colors) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_idx) return;
  const int64_t k = idx[i];
  const float4 a = v[k], b = v[k + 1];
  const float4 d = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, b.w - a.w);
  // Eigen normalized() on Vector4f (Dot.h:122-133) with the SSE packet
  // reduction (a0 + a2) + (a1 + a3) (arch/SSE/PacketMath.h:540-551).
  const float z = (d.x * d.x + d.z * d.z) + (d.y * d.y + d.w * d.w);
  float4 t = d;
  if (z > 0.0f) {
    const float s = sqrtf(z);
    t = make_float4(d.x / s, d.y / s, d.z / s, d.w / s);
  }
  const float4 c = map_color(t, d, mode, param);
  colors[k] = c;
  // The last vertex of a curve gets its segment's colour
  // (color_generator.cpp:36-37); the final index never does.
  if (i != n_idx - 1 && int64_t(idx[i + 1]) != k + 1) colors[k + 1] = c;
}

}  // namespace
}  // namespace pa_amd

using namespace pa_amd;

extern "C" pa_status pa_gpu_color_generate(pa_gpu_ctx* ctx, const float* d_vertices, size_t n_vertices,
                                           const int32_t* d_indices, size_t n_indices, pa_color_mode mode,
                                           float parameter, float* d_colors) {
  if (!ctx) return PA_E_INVALID;
  if (n_indices == 0) return PA_OK;
  if (!d_vertices || !d_indices || !d_colors || n_vertices < 2) return PA_E_INVALID;
  if (mode != PA_COLOR_HSL_CONSTANT_S && mode != PA_COLOR_HSL_CONSTANT_L && mode != PA_COLOR_HSV_CONSTANT_S &&
      mode != PA_COLOR_HSV_CONSTANT_V && mode != PA_COLOR_RGB && mode != PA_COLOR_MAGNITUDE)
    return fail(ctx, PA_E_INVALID, "unknown color mode");
  const int64_t blocks = (int64_t(n_indices) + 255) / 256;
  launch_begin(ctx, "color");
  hipLaunchKernelGGL(k_color, dim3(unsigned(blocks)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const float4*>(d_vertices), d_indices, int64_t(n_indices), int(mode), parameter,
                     reinterpret_cast<float4*>(d_colors));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}
