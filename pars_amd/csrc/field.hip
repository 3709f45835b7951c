// Device-resident blocks: upload (data_io block load replacement),
// synthesis, read-back.  Nodes are stored as padded 4-component records in
// C order (z fastest), FP32 (16 B) or BF16/FP16 (8 B).
#include <cmath>
#include <cstring>

#include "pa_internal.hpp"

#pragma clang fp contract(off)

namespace pa_amd {
namespace {

__device__ __forceinline__ uint32_t to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;  // quiet NaN
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;                    // round to nearest even
}
__device__ __forceinline__ uint32_t to_f16(float f) {
  const _Float16 h = static_cast<_Float16>(f);
  return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h));
}

__device__ __forceinline__ void store_node(void* base, int64_t i, int storage, float x, float y, float z) {
  if (storage == PA_STORAGE_FP32) {
    static_cast<float4*>(base)[i] = make_float4(x, y, z, 0.0f);
  } else if (storage == PA_STORAGE_BF16) {
    static_cast<uint2*>(base)[i] = make_uint2(to_bf16(x) | (to_bf16(y) << 16), to_bf16(z));
  } else {
    static_cast<uint2*>(base)[i] = make_uint2(to_f16(x) | (to_f16(y) << 16), to_f16(z));
  }
}

__device__ __forceinline__ float dec(uint32_t half, int storage) {
  if (storage == PA_STORAGE_BF16) return __uint_as_float(half << 16);
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(half)));
}

__global__ void k_encode(const float* __restrict__ xyz, void* __restrict__ nodes, int64_t n, int storage) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    store_node(nodes, i, storage, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

__global__ void k_decode(const void* __restrict__ nodes, float* __restrict__ xyz, int64_t n, int storage) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    float x, y, z;
    if (storage == PA_STORAGE_FP32) {
      const float4 v = static_cast<const float4*>(nodes)[i];
      x = v.x; y = v.y; z = v.z;
    } else {
      const uint2 u = static_cast<const uint2*>(nodes)[i];
      x = dec(u.x & 0xffffu, storage);
      y = dec(u.x >> 16, storage);
      z = dec(u.y & 0xffffu, storage);
    }
    xyz[3 * i] = x;
    xyz[3 * i + 1] = y;
    xyz[3 * i + 2] = z;
  }
}

// ABC flow and random-Fourier curl fields evaluated in FP64 at node
// (origin + ijk) * h, then rounded to the storage type.
__global__ void k_synth(void* __restrict__ nodes, int64_t n, int d1, int d2, int o0, int o1, int o2, double h,
                        int kind, const double* __restrict__ params, int n_modes, int storage) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t z = i % d2;
    const int64_t y = (i / d2) % d1;
    const int64_t x = i / (int64_t(d1) * d2);
    const double px = double(o0 + x) * h, py = double(o1 + y) * h, pz = double(o2 + z) * h;
    double u0, u1, u2;
    if (kind == 0) {
      const double A = params[0], B = params[1], C = params[2];
      u0 = A * sin(pz) + C * cos(py);
      u1 = B * sin(px) + A * cos(pz);
      u2 = C * sin(py) + B * cos(px);
    } else {
      u0 = u1 = u2 = 0.0;
      for (int m = 0; m < n_modes; ++m) {
        const double* md = params + 7 * m;
        const double arg = md[0] * px + md[1] * py + md[2] * pz + md[6];
        const double cs = cos(arg);
        u0 += cs * md[3];
        u1 += cs * md[4];
        u2 += cs * md[5];
      }
    }
    store_node(nodes, i, storage, float(u0), float(u1), float(u2));
  }
}

// Cell records: corners c000, c001, c010, c011, c100, c101, c110, c111 (the
// order vector_field::interpolate lerps them), xyz each.
__global__ void k_cells(const void* __restrict__ nodes, void* __restrict__ cells, int64_t ncells, int d1, int d2,
                        int storage) {
  const int64_t cs1 = d2 - 1, cs0 = int64_t(d1 - 1) * (d2 - 1);
  const int64_t ns1 = d2, ns0 = int64_t(d1) * d2;
  for (int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < ncells; c += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = c / cs0, j = (c / cs1) % (d1 - 1), k = c % cs1;
    const int64_t n0 = i * ns0 + j * ns1 + k;
    const int64_t corner[8] = {n0, n0 + 1, n0 + ns1, n0 + ns1 + 1, n0 + ns0, n0 + ns0 + 1, n0 + ns0 + ns1,
                               n0 + ns0 + ns1 + 1};
    // Record = x of the 8 corners, then y, then z, then padding (one
    // component of all corners is one contiguous quarter of the record).
    if (storage == PA_STORAGE_FP32) {
      float r[32];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float4 v = static_cast<const float4*>(nodes)[corner[q]];
        r[q] = v.x;
        r[8 + q] = v.y;
        r[16 + q] = v.z;
      }
#pragma unroll
      for (int q = 24; q < 32; ++q) r[q] = 0.0f;
      float4* out = static_cast<float4*>(cells) + c * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) out[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
    } else {
      uint32_t hw[32];  // 16-bit halves
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint2 u = static_cast<const uint2*>(nodes)[corner[q]];
        hw[q] = u.x & 0xffffu;
        hw[8 + q] = u.x >> 16;
        hw[16 + q] = u.y & 0xffffu;
      }
#pragma unroll
      for (int q = 24; q < 32; ++q) hw[q] = 0u;
      uint4* out = static_cast<uint4*>(cells) + c * 4;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        out[q] = make_uint4(hw[8 * q] | (hw[8 * q + 1] << 16), hw[8 * q + 2] | (hw[8 * q + 3] << 16),
                            hw[8 * q + 4] | (hw[8 * q + 5] << 16), hw[8 * q + 6] | (hw[8 * q + 7] << 16));
    }
  }
}

int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return int(g);
}

pa_status make_field(pa_gpu_ctx* ctx, const int32_t dims[3], const float offset[3], const float size[3],
                     const float spacing[3], pa_storage storage, pa_gpu_field** out) {
  if (!ctx || !dims || !offset || !size || !spacing || !out) return PA_E_INVALID;
  if (storage != PA_STORAGE_FP32 && storage != PA_STORAGE_BF16 && storage != PA_STORAGE_FP16)
    return fail(ctx, PA_E_INVALID, "unknown storage");
  for (int i = 0; i < 3; ++i) {
    if (dims[i] < 2) return fail(ctx, PA_E_INVALID, "a block needs at least 2 nodes per axis");
    if (!(spacing[i] > 0.0f) || !std::isfinite(spacing[i]))
      return fail(ctx, PA_E_INVALID, "spacing must be positive and finite");
  }
  for (int i = 0; i < 3; ++i)
    if (dims[i] > (1 << 22)) return fail(ctx, PA_E_INVALID, "a block axis holds at most 2^22 nodes");
  if (int64_t(dims[0]) * dims[1] * dims[2] > int64_t(UINT32_MAX))
    return fail(ctx, PA_E_INVALID, "a block holds fewer than 2^32 nodes");
  auto* f = new pa_gpu_field();
  const int64_t n = int64_t(dims[0]) * dims[1] * dims[2];
  const size_t node_bytes = storage == PA_STORAGE_FP32 ? 16 : 8;
  f->bytes = size_t(n) * node_bytes;
  hipError_t e = hipMalloc(&f->d_nodes, f->bytes);
  if (e != hipSuccess) {
    delete f;
    return check_hip(ctx, e, "hipMalloc(field)");
  }
  FieldDesc& d = f->desc;
  d.data = f->d_nodes;
  d.storage = storage;
  for (int i = 0; i < 3; ++i) {
    d.dims[i] = dims[i];
    d.offset[i] = offset[i];
    d.spacing[i] = spacing[i];
    f->size[i] = size[i];
    d.cell_hi[i] = static_cast<float>(dims[i] - 2);
    d.upper[i] = offset[i] + size[i];  // particle_tracer.cpp:263, one FP32 add
    d.inv_spacing[i] = 1.0f / spacing[i];
    // 1/h as an unevaluated sum hi + lo: the FP64 reciprocal is within 2^-53
    // of 1/h, so lo is RN(1/h - hi) up to a part in 2^29 of itself
    d.inv_lo[i] = static_cast<float>(1.0 / double(spacing[i]) - double(d.inv_spacing[i]));
  }
  d.fast_div = 1;
  for (int i = 0; i < 3; ++i)
    if (!(spacing[i] >= 0x1p-20f && spacing[i] <= 0x1p20f)) d.fast_div = 0;
  d.stride1 = dims[2];
  d.stride0 = int64_t(dims[1]) * dims[2];
  d.cells = nullptr;
  d.cstride0 = d.cstride1 = 0;
  d.layout = PA_LAYOUT_NODES;
  *out = f;
  return PA_OK;
}

}  // namespace
}  // namespace pa_amd

using namespace pa_amd;

extern "C" {

pa_status pa_gpu_field_create_device(pa_gpu_ctx* ctx, const float* d_xyz, const int32_t dims[3], const float offset[3],
                                     const float size[3], const float spacing[3], pa_storage storage,
                                     pa_gpu_field** out) {
  if (!d_xyz) return PA_E_INVALID;
  PA_TRY(make_field(ctx, dims, offset, size, spacing, storage, out));
  const int64_t n = int64_t(dims[0]) * dims[1] * dims[2];
  launch_begin(ctx, "field_encode");
  hipLaunchKernelGGL(k_encode, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d_xyz, (*out)->d_nodes, n,
                     int(storage));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PA_OK;
}

pa_status pa_gpu_field_create(pa_gpu_ctx* ctx, const float* h_xyz, const int32_t dims[3], const float offset[3],
                              const float size[3], const float spacing[3], pa_storage storage, pa_gpu_field** out) {
  if (!ctx || !h_xyz || !dims) return PA_E_INVALID;
  PA_TRY(capture_guard(ctx));
  for (int i = 0; i < 3; ++i)
    if (dims[i] < 2) return fail(ctx, PA_E_INVALID, "a block needs at least 2 nodes per axis");
  const size_t n = size_t(dims[0]) * size_t(dims[1]) * size_t(dims[2]);
  float* staging = nullptr;
  PA_HIP(ctx, hipMalloc(&staging, n * 12));
  hipError_t e = hipMemcpyAsync(staging, h_xyz, n * 12, hipMemcpyHostToDevice, ctx->stream);
  pa_status s = e == hipSuccess ? pa_gpu_field_create_device(ctx, staging, dims, offset, size, spacing, storage, out)
                                : check_hip(ctx, e, "hipMemcpy(field)");
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(staging);
  return s;
}

pa_status pa_gpu_field_synthesize(pa_gpu_ctx* ctx, int kind, const double* params, int32_t n_params,
                                  const int32_t dims[3], const int32_t origin[3], double h, const float offset[3],
                                  const float size[3], const float spacing[3], pa_storage storage,
                                  pa_gpu_field** out) {
  if (!ctx || !params || !origin) return PA_E_INVALID;
  if (kind == 0 && n_params != 3) return fail(ctx, PA_E_INVALID, "ABC flow takes 3 parameters");
  if (kind == 1 && (n_params <= 0 || n_params % 7 != 0))
    return fail(ctx, PA_E_INVALID, "Fourier modes take 7 parameters each");
  if (kind != 0 && kind != 1) return fail(ctx, PA_E_INVALID, "unknown field kind");
  PA_TRY(make_field(ctx, dims, offset, size, spacing, storage, out));
  double* d_params = nullptr;
  PA_HIP(ctx, hipMalloc(&d_params, sizeof(double) * size_t(n_params)));
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(d_params, params, sizeof(double) * size_t(n_params), hipMemcpyHostToDevice, ctx->stream));
  const int64_t n = int64_t(dims[0]) * dims[1] * dims[2];
  launch_begin(ctx, "field_synthesize");
  hipLaunchKernelGGL(k_synth, dim3(grid_for(n)), dim3(256), 0, ctx->stream, (*out)->d_nodes, n, dims[1], dims[2],
                     origin[0], origin[1], origin[2], h, kind, d_params, n_params / 7, int(storage));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  PA_HIP(ctx, hipFree(d_params));
  return PA_OK;
}

pa_status pa_gpu_field_read(pa_gpu_ctx* ctx, const pa_gpu_field* field, float* h_xyz) {
  if (!ctx || !field || !h_xyz) return PA_E_INVALID;
  PA_TRY(capture_guard(ctx));
  const FieldDesc& d = field->desc;
  const int64_t n = int64_t(d.dims[0]) * d.dims[1] * d.dims[2];
  float* staging = nullptr;
  PA_HIP(ctx, hipMalloc(&staging, size_t(n) * 12));
  hipLaunchKernelGGL(k_decode, dim3(grid_for(n)), dim3(256), 0, ctx->stream, field->d_nodes, staging, n,
                     d.storage);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h_xyz, staging, size_t(n) * 12, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(staging);
  if (e != hipSuccess) return check_hip(ctx, e, "field_read");
  return PA_OK;
}

pa_status pa_gpu_field_set_layout(pa_gpu_ctx* ctx, pa_gpu_field* field, pa_layout layout) {
  if (!ctx || !field) return PA_E_INVALID;
  FieldDesc& d = field->desc;
  if (layout == PA_LAYOUT_NODES) {
    d.layout = PA_LAYOUT_NODES;
    return PA_OK;
  }
  if (layout != PA_LAYOUT_CELLS) return fail(ctx, PA_E_INVALID, "unknown layout");
  if (!field->d_cells) {
    const int64_t ncells = int64_t(d.dims[0] - 1) * (d.dims[1] - 1) * (d.dims[2] - 1);
    if (ncells >= (int64_t(1) << 31)) return fail(ctx, PA_E_INVALID, "cell records: a block holds fewer than 2^31 cells");
    field->cell_bytes = size_t(ncells) * (d.storage == PA_STORAGE_FP32 ? 128 : 64);
    PA_HIP(ctx, hipMalloc(&field->d_cells, field->cell_bytes));
    launch_begin(ctx, "field_cells");
    hipLaunchKernelGGL(k_cells, dim3(grid_for(ncells)), dim3(256), 0, ctx->stream, field->d_nodes, field->d_cells,
                       ncells, d.dims[1], d.dims[2], d.storage);
    launch_end(ctx);
    PA_HIP(ctx, hipGetLastError());
    PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  d.cells = field->d_cells;
  d.cstride1 = d.dims[2] - 1;
  d.cstride0 = int64_t(d.dims[1] - 1) * (d.dims[2] - 1);
  d.layout = PA_LAYOUT_CELLS;
  return PA_OK;
}

pa_status pa_gpu_field_info(const pa_gpu_field* field, int32_t dims[3], float offset[3], float size[3],
                            float spacing[3], pa_storage* storage) {
  if (!field) return PA_E_INVALID;
  for (int i = 0; i < 3; ++i) {
    if (dims) dims[i] = field->desc.dims[i];
    if (offset) offset[i] = field->desc.offset[i];
    if (size) size[i] = field->size[i];
    if (spacing) spacing[i] = field->desc.spacing[i];
  }
  if (storage) *storage = pa_storage(field->desc.storage);
  return PA_OK;
}

pa_status pa_gpu_field_bytes(const pa_gpu_field* field, size_t* bytes) {
  if (!field || !bytes) return PA_E_INVALID;
  *bytes = field->bytes + field->cell_bytes;
  return PA_OK;
}

pa_status pa_gpu_field_destroy(pa_gpu_ctx* ctx, pa_gpu_field* field) {
  if (!ctx || !field) return PA_E_INVALID;
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  PA_HIP(ctx, hipFree(field->d_nodes));
  if (field->d_cells) PA_HIP(ctx, hipFree(field->d_cells));
  delete field;
  return PA_OK;
}

}  // extern "C"
