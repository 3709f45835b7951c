// What the path hands to its neighbours in the pipeline, straight from HBM:
//  - the renderer: a round's pruned vertices, colours and indices split into
//    chunks of at most 134,217,727 vertices (int32 indices relative to the
//    chunk), the form pars gives OSPRay as shared FLOAT3A / FLOAT4 / INT
//    buffers (pars/source/stages/ray_tracer.cpp:156-212; chunk size of
//    particle_tracer.cpp:216-222);
//  - host MPI: particles as the 40-byte records of particle::serialize
//    (pa/include/pa/math/particle.hpp:16-28), and the surplus lists of
//    load_balance_distribute (particle_tracer.cpp:168-184) taken off the end
//    of the device particle list.
#include <algorithm>
#include <cmath>
#include <string>

#include "pa_internal.hpp"

namespace pa_amd {
namespace {

constexpr int kThreads = 256;
// std::numeric_limits<int32_t>::max() / sizeof(vector4): the largest chunk
// (particle_tracer.cpp:216).
constexpr int64_t kMaxChunkVertices = 2147483647ll / 16;

// Exclusive offsets of the chunk boundaries: vstart[c] = first vertex of
// chunk c (the seed c * ppic), then the first index of the chunk (the first
// index >= vstart[c]) and one past its last (the first index >= vstart[c+1]
// - 1: the pair that joins chunk c's last vertex to chunk c+1's first is no
// index of either chunk, as index_generator runs per chunk).
__global__ void k_chunk_bounds(const int64_t* __restrict__ voff, int64_t n, int64_t ppic, int64_t chunks,
                               const int32_t* __restrict__ idx, int64_t ni, int64_t* __restrict__ out) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= chunks) return;
  const int64_t v0 = voff[std::min(c * ppic, n)];
  const int64_t v1 = voff[std::min((c + 1) * ppic, n)];
  auto lower = [&](int64_t key) {
    int64_t lo = 0, hi = ni;
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if (int64_t(idx[mid]) < key) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  out[4 * c + 0] = v0;
  out[4 * c + 1] = v1;
  out[4 * c + 2] = lower(v0);
  out[4 * c + 3] = c + 1 == chunks ? ni : lower(v1 - 1);
}

// Indices of chunk c (c >= 1) relative to its first vertex.
__global__ __launch_bounds__(kThreads) void k_rebase(int32_t* __restrict__ idx, const int64_t* __restrict__ bounds,
                                                     int64_t chunks, int64_t first, int64_t last) {
  const int64_t i = first + int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= last) return;
  int64_t lo = 1, hi = chunks;  // the chunk c whose [begin, end) holds i: last c with begin <= i
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (bounds[4 * mid + 2] <= i) lo = mid; else hi = mid;
  }
  if (i < bounds[4 * lo + 3]) idx[i] = int32_t(int64_t(idx[i]) - bounds[4 * lo]);
}

// particle -> the field sequence of particle::serialize, 10 x 4 bytes.
__global__ __launch_bounds__(kThreads) void k_to_wire(const pa_particle* __restrict__ p, int64_t n,
                                                      int32_t* __restrict__ w) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const pa_particle q = p[i];
  int32_t* o = w + 10 * i;
  for (int k = 0; k < 4; ++k) o[k] = __float_as_int(q.position[k]);
  o[4] = q.remaining_iterations;
  o[5] = q.vector_field_index;
  o[6] = q.original_rank;
  for (int k = 0; k < 3; ++k) o[7 + k] = q.original_voxel[k];
}

__global__ __launch_bounds__(kThreads) void k_from_wire(const int32_t* __restrict__ w, int64_t n,
                                                        pa_particle* __restrict__ p) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const int32_t* o = w + 10 * i;
  pa_particle q;
  for (int k = 0; k < 4; ++k) q.position[k] = __int_as_float(o[k]);
  q.remaining_iterations = o[4];
  q.vector_field_index = o[5];
  q.original_rank = o[6];
  for (int k = 0; k < 3; ++k) q.original_voxel[k] = o[7 + k];
  q.reserved_[0] = q.reserved_[1] = 0;
  p[i] = q;
}

// The surplus for neighbour k: particles [src, src + m) in order, their
// vector field index set to the slot the receiver knows this process by
// (its -x neighbour is our +x and so on: k ^ 1), as serialize records.
__global__ __launch_bounds__(kThreads) void k_take(const pa_particle* __restrict__ p, int64_t src, int64_t m,
                                                   int32_t slot, int32_t* __restrict__ w) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= m) return;
  const pa_particle q = p[src + i];
  int32_t* o = w + 10 * i;
  for (int k = 0; k < 4; ++k) o[k] = __float_as_int(q.position[k]);
  o[4] = q.remaining_iterations;
  o[5] = slot;
  o[6] = q.original_rank;
  for (int k = 0; k < 3; ++k) o[7 + k] = q.original_voxel[k];
}

unsigned blocks_of(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

// Point-splat stand-in for the streamline ray tracer: a pinhole camera as
// ray_tracer::set_camera / set_image_size configure OSPRay's (pos, dir, up,
// fovy 68 degrees, aspect = width / height; ray_tracer.cpp:63, 215-227), one
// pixel per vertex, nearest vertex wins (depth and colour packed into one
// 64-bit key, atomicMin), colours encoded as sRGB 8-bit like OSPRay's
// OSP_FB_SRGBA framebuffer; row 0 is the bottom row.
struct Camera {
  float pos[3], right[3], up[3], fwd[3];  // orthonormal basis
  float tan_half;                         // tan(fovy / 2)
  float aspect;
  int w, h;
};

__global__ __launch_bounds__(kThreads) void k_fill_u64(unsigned long long* __restrict__ p, int64_t n,
                                                       unsigned long long v) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ uint32_t srgb8(float c) {
  c = fminf(fmaxf(c, 0.0f), 1.0f);
  const float e = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
  return uint32_t(fminf(fmaxf(e * 255.0f + 0.5f, 0.0f), 255.0f));
}

__global__ __launch_bounds__(kThreads) void k_splat(const float4* __restrict__ v, const float4* __restrict__ col,
                                                    int64_t n, const Camera cam,
                                                    unsigned long long* __restrict__ zbuf) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const float4 p = v[i];
  // the sentinels of integral_curves.hpp:18-19 are not vertices
  if (p.w == 0.0f && p.x == p.y && p.y == p.z && (p.x == -1.0f || p.x == -2.0f)) return;
  const float d[3] = {p.x - cam.pos[0], p.y - cam.pos[1], p.z - cam.pos[2]};
  const float z = d[0] * cam.fwd[0] + d[1] * cam.fwd[1] + d[2] * cam.fwd[2];
  if (!(z > 1e-6f)) return;
  const float x = (d[0] * cam.right[0] + d[1] * cam.right[1] + d[2] * cam.right[2]) / (z * cam.tan_half * cam.aspect);
  const float y = (d[0] * cam.up[0] + d[1] * cam.up[1] + d[2] * cam.up[2]) / (z * cam.tan_half);
  const float px = (x * 0.5f + 0.5f) * float(cam.w), py = (y * 0.5f + 0.5f) * float(cam.h);
  if (!(px >= 0.0f && px < float(cam.w) && py >= 0.0f && py < float(cam.h))) return;
  const float4 c = col ? col[i] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  const uint32_t rgba = srgb8(c.x) | srgb8(c.y) << 8 | srgb8(c.z) << 16 | uint32_t(fminf(fmaxf(c.w, 0.0f), 1.0f) * 255.0f + 0.5f) << 24;
  const unsigned long long key = (static_cast<unsigned long long>(__float_as_uint(z)) << 32) | rgba;
  atomicMin(&zbuf[int64_t(int(py)) * cam.w + int(px)], key);
}

__global__ __launch_bounds__(kThreads) void k_resolve(const unsigned long long* __restrict__ zbuf, int64_t n,
                                                      uint32_t* __restrict__ rgba) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = zbuf[i];
  rgba[i] = k == ~0ull ? 0u : uint32_t(k & 0xFFFFFFFFull);
}

}  // namespace
}  // namespace pa_amd

using namespace pa_amd;

extern "C" {

pa_status pa_gpu_round_chunks(pa_gpu_ctx* ctx, const int32_t* d_vertex_counts, size_t n, size_t region_slots,
                              size_t max_vertices_per_chunk, float* d_vertices, size_t n_vertices, int32_t* d_indices,
                              size_t n_indices, float* d_colors, pa_curve_chunk* chunks, size_t max_chunks,
                              size_t* n_chunks) {
  if (!ctx || !n_chunks || (n && (!d_vertex_counts || !chunks))) return PA_E_INVALID;
  *n_chunks = 0;
  if (n == 0) return PA_OK;
  if (region_slots < 1) return fail(ctx, PA_E_INVALID, "round_chunks: region_slots must be at least 1");
  const int64_t cap = max_vertices_per_chunk ? std::min<int64_t>(int64_t(max_vertices_per_chunk), kMaxChunkVertices)
                                             : kMaxChunkVertices;
  // particles per chunk as compute_round_info has it, with the per-seed
  // region in place of the largest remaining_iterations: a seed's segments
  // never hold more than region_slots vertices, so no chunk exceeds cap
  const int64_t ppic = cap / int64_t(region_slots);
  if (ppic < 1) return fail(ctx, PA_E_INVALID, "round_chunks: region_slots above the chunk size");
  const int64_t count = (int64_t(n) + ppic - 1) / ppic;
  *n_chunks = size_t(count);
  if (size_t(count) > max_chunks) return fail(ctx, PA_E_CAPACITY, "round_chunks: more chunks than max_chunks");
  if (n_indices && !d_indices) return PA_E_INVALID;
  const size_t scan_bytes = (n + 1) * sizeof(int64_t);
  const size_t bound_bytes = size_t(count) * 4 * sizeof(int64_t);
  void* ws = nullptr;
  PA_TRY(scratch(ctx, scan_bytes + bound_bytes + scan_workspace_bytes(int64_t(n)) + 1024, &ws));
  auto* voff = static_cast<int64_t*>(ws);
  auto* bounds = voff + (n + 1);
  auto* sws = reinterpret_cast<int64_t*>(reinterpret_cast<char*>(bounds + 4 * count) + 256 -
                                         (reinterpret_cast<uintptr_t>(bounds + 4 * count) & 255));
  launch_begin(ctx, "round_chunks");
  PA_TRY(scan_i32(ctx, d_vertex_counts, int64_t(n), voff, sws));
  hipLaunchKernelGGL(k_chunk_bounds, dim3(unsigned((count + 63) / 64)), dim3(64), 0, ctx->stream, voff, int64_t(n),
                     ppic, count, d_indices, int64_t(n_indices), bounds);
  std::vector<int64_t> h(size_t(count) * 4);
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(h.data(), bounds, h.size() * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h[size_t(count) * 4 - 3] != int64_t(n_vertices)) {
    launch_end(ctx);
    return fail(ctx, PA_E_INVALID, "round_chunks: the vertex counts do not add up to n_vertices");
  }
  if (count > 1 && h[4 + 2] < int64_t(n_indices))
    hipLaunchKernelGGL(k_rebase, dim3(blocks_of(int64_t(n_indices) - h[6])), dim3(kThreads), 0, ctx->stream,
                       d_indices, bounds, count, h[6], int64_t(n_indices));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  for (int64_t c = 0; c < count; ++c) {
    pa_curve_chunk& k = chunks[c];
    const int64_t v0 = h[4 * c], v1 = h[4 * c + 1], i0 = h[4 * c + 2], i1 = h[4 * c + 3];
    k.vertices = d_vertices ? d_vertices + 4 * v0 : nullptr;
    k.colors = d_colors ? d_colors + 4 * v0 : nullptr;
    k.indices = d_indices ? d_indices + i0 : nullptr;
    k.n_vertices = size_t(v1 - v0);
    k.n_indices = size_t(i1 - i0);
    k.first_particle = size_t(c * ppic);
    k.n_particles = size_t(std::min<int64_t>(ppic, int64_t(n) - c * ppic));
  }
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PA_OK;
}

pa_status pa_gpu_particles_to_wire(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n, void* d_wire) {
  if (!ctx || (n && (!d_particles || !d_wire))) return PA_E_INVALID;
  if (n == 0) return PA_OK;
  launch_begin(ctx, "wire");
  hipLaunchKernelGGL(k_to_wire, dim3(blocks_of(int64_t(n))), dim3(kThreads), 0, ctx->stream, d_particles, int64_t(n),
                     static_cast<int32_t*>(d_wire));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_particles_from_wire(pa_gpu_ctx* ctx, const void* d_wire, size_t n, pa_particle* d_particles) {
  if (!ctx || (n && (!d_particles || !d_wire))) return PA_E_INVALID;
  if (n == 0) return PA_OK;
  launch_begin(ctx, "wire");
  hipLaunchKernelGGL(k_from_wire, dim3(blocks_of(int64_t(n))), dim3(kThreads), 0, ctx->stream,
                     static_cast<const int32_t*>(d_wire), int64_t(n), d_particles);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_load_balance_take(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n,
                                   const size_t surplus_counts[6], void* d_surplus_wire, size_t surplus_offsets[7],
                                   size_t* n_remaining) {
  if (!ctx || !surplus_counts || !surplus_offsets || !n_remaining) return PA_E_INVALID;
  // particle_tracer.cpp:168-184: neighbour k in order 0..5 takes the last
  // surplus_counts[k] particles of the list as it stands, unless that is more
  // than the list holds
  size_t end = n, total = 0;
  size_t take[6], from[6];
  for (int k = 0; k < 6; ++k) {
    take[k] = surplus_counts[k] <= end ? surplus_counts[k] : 0;
    from[k] = end - take[k];
    end -= take[k];
    surplus_offsets[k] = total;
    total += take[k];
  }
  surplus_offsets[6] = total;
  *n_remaining = end;
  if (total && (!d_particles || !d_surplus_wire)) return PA_E_INVALID;
  launch_begin(ctx, "load_balance_take");
  for (int k = 0; k < 6; ++k)
    if (take[k])
      hipLaunchKernelGGL(k_take, dim3(blocks_of(int64_t(take[k]))), dim3(kThreads), 0, ctx->stream, d_particles,
                         int64_t(from[k]), int64_t(take[k]), int32_t(k ^ 1),
                         static_cast<int32_t*>(d_surplus_wire) + 10 * surplus_offsets[k]);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_render_points(pa_gpu_ctx* ctx, const float* d_vertices, const float* d_colors, size_t n,
                               const float position[3], const float forward[3], const float up[3], int32_t width,
                               int32_t height, uint8_t* d_rgba) {
  if (!ctx || !position || !forward || !up || !d_rgba || (n && !d_vertices)) return PA_E_INVALID;
  if (width < 1 || height < 1 || int64_t(width) * height > (int64_t(1) << 28))
    return fail(ctx, PA_E_INVALID, "render_points: image of 1 to 2^28 pixels");
  Camera cam;
  // basis: forward, then right = forward x up, then the true up = right x forward
  float f[3] = {forward[0], forward[1], forward[2]};
  double fl = std::sqrt(double(f[0]) * f[0] + double(f[1]) * f[1] + double(f[2]) * f[2]);
  float r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double rl = std::sqrt(double(r[0]) * r[0] + double(r[1]) * r[1] + double(r[2]) * r[2]);
  if (!(fl > 0) || !(rl > 0)) return fail(ctx, PA_E_INVALID, "render_points: forward and up must not be parallel");
  for (int i = 0; i < 3; ++i) {
    cam.pos[i] = position[i];
    cam.fwd[i] = float(f[i] / fl);
    cam.right[i] = float(r[i] / rl);
  }
  cam.up[0] = cam.right[1] * cam.fwd[2] - cam.right[2] * cam.fwd[1];
  cam.up[1] = cam.right[2] * cam.fwd[0] - cam.right[0] * cam.fwd[2];
  cam.up[2] = cam.right[0] * cam.fwd[1] - cam.right[1] * cam.fwd[0];
  cam.tan_half = float(std::tan(68.0 * 3.14159265358979323846 / 360.0));
  cam.aspect = float(width) / float(height);
  cam.w = width;
  cam.h = height;
  const int64_t px = int64_t(width) * height;
  void* ws = nullptr;
  PA_TRY(scratch(ctx, size_t(px) * 8, &ws));
  auto* zbuf = static_cast<unsigned long long*>(ws);
  launch_begin(ctx, "render");
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_of(px)), dim3(kThreads), 0, ctx->stream, zbuf, px, ~0ull);
  if (n)
    hipLaunchKernelGGL(k_splat, dim3(blocks_of(int64_t(n))), dim3(kThreads), 0, ctx->stream,
                       reinterpret_cast<const float4*>(d_vertices), reinterpret_cast<const float4*>(d_colors),
                       int64_t(n), cam, zbuf);
  hipLaunchKernelGGL(k_resolve, dim3(blocks_of(px)), dim3(kThreads), 0, ctx->stream, zbuf, px,
                     reinterpret_cast<uint32_t*>(d_rgba));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

}  // extern "C"
