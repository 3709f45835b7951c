// The stages after the union trace of one round, fused: prune
// (particle_tracer.cpp:402-409, as the counted prune), index_generator
// (index_generator.cpp:10-24) and color_generator (color_generator.cpp:
// 30-98) on 8-seed grouped slots (slot s of seed p = vec4 at
// ((p / 8) * region + s) * 8 + p % 8): one pass packs the vertices and writes
// the index list, a dense pass over the indices writes the colours.
//
// Layout of the outputs (as the separate stages produce them):
//   vertices  seed p's slots at O[p] .. O[p] + counts[p] - 1, O = exclusive
//             scan of counts (termination vertices included, as prune keeps
//             them);
//   indices   every i with vertices i and i+1 both not the termination
//             vertex, ascending; seed p owns I[p] .. I[p+1] - 1, where the
//             tracer counted the in-run pairs and the pair that crosses into
//             seed p+1 is added here;
//   colours   colour of segment (i, i+1) at i, and at i+1 when i+1 ends a
//             curve (no index i+1) unless i is the last index; other entries
//             are left as they were, like the colour stage.
#include "pa_color.hpp"

namespace pa_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;  // 8-slot chunks per pass of k_group8

// icount[p] = in-run pairs + the pair (last of p, first of p+1).
__global__ __launch_bounds__(kThreads) void k_round_counts(const float4* __restrict__ slots, int64_t n, int64_t region,
                                                           const int32_t* __restrict__ counts,
                                                           const int32_t* __restrict__ pairs,
                                                           int32_t* __restrict__ icount, int32_t* __restrict__ cross) {
  const int64_t p = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (p >= n) return;
  int x = 0;
  if (p + 1 < n) {
    const float4 last = slots[group8_slot(p, counts[p] - 1, region)];
    const float4 next = slots[group8_slot(p + 1, 0, region)];
    x = (!is_termination_vertex(last) && !is_termination_vertex(next)) ? 1 : 0;
  }
  cross[p] = x;
  icount[p] = pairs[p] + x;
}

// One wave per group of 8 seeds, whose slots are contiguous: lane 8k + j
// holds slot s0 + j of seed k, so every load is one 1-KiB run of the group
// and every seed's 8 vertices (and its valid index pairs) are stored as one
// contiguous run.  The index position of a pair is the seed's offset, the
// pairs before this chunk, and the prefix inside the chunk.
template <bool INDEX>
__global__ __launch_bounds__(kThreads) void k_group8(const float4* __restrict__ slots, int64_t n, int64_t region,
                                                     const int32_t* __restrict__ counts,
                                                     const int32_t* __restrict__ cross,
                                                     const int32_t* __restrict__ icount,
                                                     const int64_t* __restrict__ voff,
                                                     const int64_t* __restrict__ ioff, float4* __restrict__ vout,
                                                     int32_t* __restrict__ iout) {
  const int64_t g = int64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (g * 8 >= n) return;
  const int lane = threadIdx.x & 63, k = lane >> 3, j = lane & 7;
  const int64_t p = g * 8 + k;
  const bool has = p < n;
  const int c = has ? counts[p] : 0;
  int cmax = c;
#pragma unroll
  for (int o = 8; o < 64; o <<= 1) cmax = max(cmax, __shfl_xor(cmax, o, 64));
  const int64_t O = has ? voff[p] : 0;
  const int64_t I = (INDEX && has) ? ioff[p] : 0;
  const float4* base = slots + g * region * 8;
  int carry = 0;
  // kUnroll chunks of 8 slots per pass: their loads are in flight together
  for (int s0 = 0; s0 < cmax; s0 += 8 * kUnroll) {
    float4 v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int s = s0 + 8 * u + j;
      v[u] = s < c ? base[int64_t(s) * 8 + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    float4 after = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // slot s0 + 8 kUnroll of this lane's seed
    if constexpr (INDEX) {
      const int s = s0 + 8 * kUnroll;
      if (j == 7 && s < c) after = base[int64_t(s) * 8 + k];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int s = s0 + 8 * u + j;
      if (s < c) vout[O + s] = v[u];
      if constexpr (INDEX) {
        // slot s + 1: the next lane of this seed, or the first row of the next chunk
        float4 nx;
        nx.x = __shfl_down(v[u].x, 1, 64);
        nx.y = __shfl_down(v[u].y, 1, 64);
        nx.z = __shfl_down(v[u].z, 1, 64);
        nx.w = __shfl_down(v[u].w, 1, 64);
        if (u + 1 < kUnroll) {
          const int src = 8 * k;
          const float4 f = make_float4(__shfl(v[u + 1 < kUnroll ? u + 1 : u].x, src, 64),
                                       __shfl(v[u + 1 < kUnroll ? u + 1 : u].y, src, 64),
                                       __shfl(v[u + 1 < kUnroll ? u + 1 : u].z, src, 64),
                                       __shfl(v[u + 1 < kUnroll ? u + 1 : u].w, src, 64));
          if (j == 7) nx = f;
        } else if (j == 7) {
          nx = after;
        }
        const bool valid = s <= c - 2 && !is_termination_vertex(v[u]) && !is_termination_vertex(nx);
        const uint32_t bits = uint32_t(__ballot(valid) >> (8 * k)) & 0xFFu;
        if (valid) iout[I + carry + __popc(bits & ((1u << j) - 1u))] = int32_t(O + s);
        carry += __popc(bits);
      }
    }
  }
  if (INDEX && has && j == 0 && cross[p]) iout[I + icount[p] - 1] = int32_t(O + c - 1);
}

}  // namespace

size_t finish_ws_bytes(int64_t n) {
  return size_t(n + 1) * 2 * sizeof(int64_t) + size_t(n) * 2 * sizeof(int32_t) + scan_workspace_bytes(n) + 4096;
}

pa_status finish_group8(pa_gpu_ctx* ctx, const float* d_slots, int64_t n, int64_t region, const int32_t* d_counts,
                        const int32_t* d_pairs, int mode, float param, float* d_vertices, size_t vertex_capacity,
                        size_t* n_vertices, int32_t* d_indices, size_t index_capacity, size_t* n_indices,
                        float* d_colors, void* ws) {
  char* w = static_cast<char*>(ws);
  auto* voff = reinterpret_cast<int64_t*>(w);
  auto* ioff = voff + (n + 1);
  auto* icount = reinterpret_cast<int32_t*>(ioff + (n + 1));
  int32_t* cross = icount + n;
  char* tail = reinterpret_cast<char*>(cross + n);
  auto* sws = reinterpret_cast<int64_t*>(tail + (256 - (reinterpret_cast<uintptr_t>(tail) & 255)));
  const auto* slots = reinterpret_cast<const float4*>(d_slots);
  launch_begin(ctx, "round_counts");
  hipLaunchKernelGGL(k_round_counts, dim3(unsigned((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, ctx->stream,
                     slots, n, region, d_counts, d_pairs, icount, cross);
  PA_TRY(scan_i32(ctx, d_counts, n, voff, sws));
  PA_TRY(scan_i32(ctx, icount, n, ioff, sws));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  int64_t tot[2] = {0, 0};
  PA_HIP(ctx, hipMemcpyAsync(&tot[0], voff + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipMemcpyAsync(&tot[1], ioff + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n_vertices = size_t(tot[0]);
  *n_indices = size_t(tot[1]);
  if (size_t(tot[0]) > vertex_capacity) return fail(ctx, PA_E_CAPACITY, "round: vertex capacity too small");
  if (size_t(tot[1]) > index_capacity) return fail(ctx, PA_E_CAPACITY, "round: index capacity too small");
  if (tot[0] > INT32_MAX) return fail(ctx, PA_E_CAPACITY, "round: indices are int32 (at most 2^31-1 vertices)");
  const int64_t groups = (n + 7) / 8;
  launch_begin(ctx, "round_finish");
  hipLaunchKernelGGL(k_group8<true>, dim3(unsigned((groups + 3) / 4)), dim3(kThreads), 0, ctx->stream, slots, n,
                     region, d_counts, cross, icount, voff, ioff, reinterpret_cast<float4*>(d_vertices), d_indices);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  // colours: one lane per index over the packed vertices (dense)
  return color_launch(ctx, reinterpret_cast<const float4*>(d_vertices), d_indices, tot[1], mode, param,
                      reinterpret_cast<float4*>(d_colors));
}

}  // namespace pa_amd

using namespace pa_amd;

extern "C" pa_status pa_gpu_prune_counted_group8(pa_gpu_ctx* ctx, const float* d_slots, size_t region_slots,
                                                 const int32_t* d_vertex_counts, size_t n, float* d_out,
                                                 size_t capacity, size_t* n_out, int64_t* d_offsets) {
  if (!ctx || !n_out || (n && (!d_slots || !d_vertex_counts || !d_out))) return PA_E_INVALID;
  if (n == 0) {
    *n_out = 0;
    return PA_OK;
  }
  void* ws = nullptr;
  const size_t own = d_offsets ? 0 : (n + 1) * sizeof(int64_t);
  PA_TRY(scratch(ctx, own + scan_workspace_bytes(int64_t(n)) + 4096, &ws));
  int64_t* offs = d_offsets ? d_offsets : static_cast<int64_t*>(ws);
  auto* sws = reinterpret_cast<int64_t*>(static_cast<char*>(ws) + (own + 255) / 256 * 256);
  launch_begin(ctx, "prune_scan");
  PA_TRY(scan_i32(ctx, d_vertex_counts, int64_t(n), offs, sws));
  launch_end(ctx);
  int64_t total = 0;
  PA_HIP(ctx, hipMemcpyAsync(&total, offs + n, sizeof total, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n_out = size_t(total);
  if (size_t(total) > capacity) return fail(ctx, PA_E_CAPACITY, "prune_counted: output capacity too small");
  const int64_t groups = (int64_t(n) + 7) / 8;
  launch_begin(ctx, "prune_counted");
  hipLaunchKernelGGL(k_group8<false>, dim3(unsigned((groups + 3) / 4)), dim3(kThreads), 0, ctx->stream,
                     reinterpret_cast<const float4*>(d_slots), int64_t(n), int64_t(region_slots), d_vertex_counts,
                     nullptr, nullptr, offs, nullptr, reinterpret_cast<float4*>(d_out), nullptr);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}
