// pa::color_generator::generate on the device: one lane per index, tangent
// of the segment (vertex idx -> idx+1) mapped to RGBA.
#include "pa_color.hpp"

#pragma clang fp contract(off)

namespace pa_amd {
namespace {

template <int MODE>
__global__ __launch_bounds__(256) void k_color(const float4* __restrict__ v, const int32_t* __restrict__ idx,
                                               int64_t n_idx, float param, float4* __restrict__ colors) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_idx) return;
  const int64_t k = idx[i];
  const float4 c = segment_color(v[k], v[k + 1], MODE, param);
  colors[k] = c;
  // The last vertex of a curve gets its segment's colour
  // (color_generator.cpp:36-37); the final index never does.
  if (i != n_idx - 1 && int64_t(idx[i + 1]) != k + 1) colors[k + 1] = c;
}

}  // namespace

pa_status color_launch(pa_gpu_ctx* ctx, const float4* d_vertices, const int32_t* d_indices, int64_t n_indices,
                       int mode, float param, float4* d_colors) {
  if (n_indices == 0) return PA_OK;
  const int64_t blocks = (n_indices + 255) / 256;
  launch_begin(ctx, "color");
  const bool known = with_color_mode(mode, [&](auto M) {
    hipLaunchKernelGGL(k_color<decltype(M)::value>, dim3(unsigned(blocks)), dim3(256), 0, ctx->stream, d_vertices,
                       d_indices, n_indices, param, d_colors);
  });
  launch_end(ctx);
  if (!known) return fail(ctx, PA_E_INVALID, "unknown color mode");
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

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
  return color_launch(ctx, reinterpret_cast<const float4*>(d_vertices), d_indices, int64_t(n_indices), int(mode),
                      parameter, reinterpret_cast<float4*>(d_colors));
}
