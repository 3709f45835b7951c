// Order-preserving compaction on the device: exclusive scans, the prune and
// index stages (value-predicate stream compaction with wave ballots and
// LDS prefix sums), the counted prune of a traced round, and the grouping
// of hand-off particles by (map, neighbour slot).
#include "pa_internal.hpp"
#include "pa_color.hpp"

#pragma clang fp contract(off)

namespace pa_amd {
namespace {

constexpr int kBlock = 256;        // threads per block (4 waves of 64)
constexpr int kScanItems = 4;      // items per thread in the scan
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kSelItems = 16;      // items per thread in select
constexpr int kSelTile = kBlock * kSelItems;

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

__device__ __forceinline__ int64_t wave_incl_scan(int64_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// Exclusive prefix of x over the 256-thread block; *total = block sum.
__device__ __forceinline__ int64_t block_excl_scan(int64_t x, int64_t* total) {
  __shared__ int64_t ws[kBlock / 64];
  const int64_t inc = wave_incl_scan(x);
  if (lane_id() == 63) ws[wave_id()] = inc;
  __syncthreads();
  int64_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; ++i) {
    if (i < wave_id()) off += ws[i];
    tot += ws[i];
  }
  __syncthreads();
  *total = tot;
  return off + inc - x;
}

// Exclusive scan over m = n+1 virtual items (item n = 0), so out[n] is the
// total.  Pass 1: per-tile scan + tile sums.
template <class T>
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const T* __restrict__ in, int64_t n, int64_t* __restrict__ out,
                                                       int64_t* __restrict__ sums) {
  const int64_t m = n + 1;
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  int64_t v[kScanItems];
  int64_t local = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int64_t i = base + j;
    v[j] = i < n ? int64_t(in[i]) : 0;
    local += v[j];
  }
  int64_t total;
  int64_t run = block_excl_scan(local, &total);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int64_t i = base + j;
    if (i < m) out[i] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_add(int64_t* __restrict__ out, int64_t m,
                                                     const int64_t* __restrict__ offs) {
  const int64_t add = offs[blockIdx.x];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int64_t i = base + j;
    if (i < m) out[i] += add;
  }
}

int64_t tiles(int64_t m, int64_t tile) { return (m + tile - 1) / tile; }

// Workspace of a scan over n items: tile sums of every level.
size_t scan_ws_bytes(int64_t n) {
  size_t b = 0;
  int64_t m = n + 1;
  while (true) {
    const int64_t t = tiles(m, kScanTile);
    b += size_t(t + 1) * sizeof(int64_t);
    if (t <= 1) break;
    m = t + 1;
  }
  return b + 256;
}

template <class T>
pa_status scan_rec(pa_gpu_ctx* ctx, const T* d_in, int64_t n, int64_t* d_out, int64_t* ws) {
  const int64_t m = n + 1;
  const int64_t t = tiles(m, kScanTile);
  int64_t* sums = ws;
  hipLaunchKernelGGL(k_scan_tiles<T>, dim3(unsigned(t)), dim3(kBlock), 0, ctx->stream, d_in, n, d_out, sums);
  PA_HIP(ctx, hipGetLastError());
  if (t > 1) {
    int64_t* offs = ws + (t + 1);
    // exclusive scan of the t tile sums (t+1 outputs) into offs
    PA_TRY(scan_rec<int64_t>(ctx, sums, t, offs, offs + (t + 1)));
    hipLaunchKernelGGL(k_scan_add, dim3(unsigned(t)), dim3(kBlock), 0, ctx->stream, d_out, m, offs);
    PA_HIP(ctx, hipGetLastError());
  }
  return PA_OK;
}

size_t scan_ws_total(int64_t n) {
  // Each level needs its sums and the scanned offsets of the level above.
  return 2 * scan_ws_bytes(n) + 1024;
}

// ----------------------------------------------------------------- select
struct PruneOp {
  const float4* v;
  float4* out;
  __device__ __forceinline__ bool keep(int64_t i) const {
    const float4 a = v[i];
    return !(a.x == -1.0f && a.y == -1.0f && a.z == -1.0f && a.w == 0.0f);
  }
  __device__ __forceinline__ void emit(int64_t i, int64_t pos) const { out[pos] = v[i]; }
};

__device__ __forceinline__ bool is_termination(float4 a) {
  return a.x == -2.0f && a.y == -2.0f && a.z == -2.0f && a.w == 0.0f;
}

struct IndexOp {
  const float4* v;
  int32_t* out;
  __device__ __forceinline__ bool keep(int64_t i) const { return !is_termination(v[i]) && !is_termination(v[i + 1]); }
  __device__ __forceinline__ void emit(int64_t i, int64_t pos) const { out[pos] = int32_t(i); }
};

template <class Op>
__global__ __launch_bounds__(kBlock) void k_select_count(Op op, int64_t n, int64_t* __restrict__ counts) {
  const int64_t base = int64_t(blockIdx.x) * kSelTile;
  int64_t c = 0;
#pragma unroll 4
  for (int j = 0; j < kSelItems; ++j) {
    const int64_t i = base + int64_t(j) * kBlock + threadIdx.x;
    if (i < n && op.keep(i)) ++c;
  }
  int64_t total;
  block_excl_scan(c, &total);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

template <class Op>
__global__ __launch_bounds__(kBlock) void k_select_scatter(Op op, int64_t n, const int64_t* __restrict__ offs) {
  __shared__ int32_t wc[kBlock / 64];
  const int64_t base = int64_t(blockIdx.x) * kSelTile;
  int64_t run = offs[blockIdx.x];
  const int lane = lane_id(), w = wave_id();
  for (int j = 0; j < kSelItems; ++j) {
    const int64_t i = base + int64_t(j) * kBlock + threadIdx.x;
    const bool k = i < n && op.keep(i);
    const uint64_t mask = __ballot(k);
    const int below = __popcll(mask & ((uint64_t(1) << lane) - 1));
    if (lane == 0) wc[w] = __popcll(mask);
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) {
      if (q < w) pre += wc[q];
      tot += wc[q];
    }
    if (k) op.emit(i, run + pre + below);
    run += tot;
    __syncthreads();
  }
}

template <class Op>
pa_status select(pa_gpu_ctx* ctx, Op op, int64_t n, size_t* n_out, const char* name) {
  const int64_t t = tiles(n, kSelTile);
  const size_t need = size_t(t + 1) * 2 * sizeof(int64_t) + scan_ws_total(t) + 4096;
  void* ws = nullptr;
  PA_TRY(scratch(ctx, need, &ws));
  int64_t* counts = static_cast<int64_t*>(ws);
  int64_t* offs = counts + (t + 1);
  int64_t* sws = offs + (t + 1);
  launch_begin(ctx, name);
  hipLaunchKernelGGL(k_select_count<Op>, dim3(unsigned(t)), dim3(kBlock), 0, ctx->stream, op, n, counts);
  PA_HIP(ctx, hipGetLastError());
  PA_TRY(scan_rec<int64_t>(ctx, counts, t, offs, sws));
  hipLaunchKernelGGL(k_select_scatter<Op>, dim3(unsigned(t)), dim3(kBlock), 0, ctx->stream, op, n, offs);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  int64_t total = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&total, offs + t, sizeof total, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n_out = size_t(total);
  return PA_OK;
}

// ---------------------------------------------------------- counted prune
// One wave per particle copies its count[p] slots (coalesced 16-B lanes).
__global__ __launch_bounds__(kBlock) void k_copy_counted(const float4* __restrict__ slots, int64_t stride,
                                                         const int32_t* __restrict__ counts,
                                                         const int64_t* __restrict__ offs, int64_t n,
                                                         float4* __restrict__ out) {
  const int64_t waves = int64_t(gridDim.x) * (kBlock / 64);
  for (int64_t p = int64_t(blockIdx.x) * (kBlock / 64) + wave_id(); p < n; p += waves) {
    const int32_t c = counts[p];
    const float4* src = slots + p * stride;
    float4* dst = out + offs[p];
    for (int j = lane_id(); j < c; j += 64) dst[j] = src[j];
  }
}

// Counted prune, index generation and colouring of seed-major slots in one
// pass (pipeline.cpp:186-260 runs prune, index_generator.cpp:10-24 and
// color_generator.cpp:16-98 in turn over each chunk).  A wave per seed p
// copies its c = count[p] slots to O = offs[p], 64 at a time, and checks
// that the curve has the shape the tracers write: c >= 2, no termination
// vertex before the last slot, a termination vertex last.  For such curves
// the index generator's pairs are the c - 2 pairs of the non-termination
// vertices, at index positions O - 2p onward, and the chunk has m - 2n
// indices; the colours are k_color's: colors[k] = segment(v[k], v[k+1]) for
// each index k, and the curve's last vertex takes its last segment's colour
// unless that segment is the chunk's final index (then every later seed has
// c == 2).  Indices and colours are written from the vertices in registers
// (v[k+1] from the next lane) as the copy goes, before the whole curve has
// been checked; a curve of any other shape sets *irregular, and the host
// then redoes indices and colours with the general kernels, which overwrite
// every position that they write.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_finish_counted(const float4* __restrict__ slots, int64_t stride,
                                                           const int32_t* __restrict__ counts,
                                                           const int64_t* __restrict__ offs, int64_t n,
                                                           float4* __restrict__ out, int32_t* __restrict__ idx,
                                                           float4* __restrict__ colors, float param,
                                                           int32_t* __restrict__ irregular) {
  const int64_t waves = int64_t(gridDim.x) * (kBlock / 64);
  const int64_t n_idx = offs[n] - 2 * n;
  const int lane = lane_id();
  for (int64_t p = int64_t(blockIdx.x) * (kBlock / 64) + wave_id(); p < n; p += waves) {
    const int32_t c = counts[p];
    const int64_t o = offs[p];
    const float4* src = slots + p * stride;
    float4* dst = out + o;
    const int64_t first = o - 2 * p;  // index position of the curve's first pair
    const bool last_is_final = first + (c - 2) == n_idx;
    bool bad = c < 2;
    for (int j0 = 0; j0 < c; j0 += 64) {
      const int j = j0 + lane;
      float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (j < c) {
        v = src[j];
        dst[j] = v;
        bad |= is_termination(v) != (j == c - 1);
      }
      float4 w;  // v[j + 1]
      w.x = __shfl_down(v.x, 1);
      w.y = __shfl_down(v.y, 1);
      w.z = __shfl_down(v.z, 1);
      w.w = __shfl_down(v.w, 1);
      if (lane == 63 && j + 1 < c) w = src[j + 1];
      if (j < c - 2) {
        const float4 col = segment_color(v, w, MODE, param);
        idx[first + j] = int32_t(o + j);
        colors[o + j] = col;
        if (j == c - 3 && !last_is_final) colors[o + j + 1] = col;
      }
    }
    if (__ballot(bad) != 0 && lane == 0) atomicOr(irregular, 1);
  }
}

__global__ void k_zero_flag(int32_t* f) { *f = 0; }

// Step-major slots (slot s of seed p at s * n + p) into the pruned layout:
// a workgroup takes 64 consecutive seeds and moves 64-slot tiles through LDS,
// reading whole rows (1 KiB per wave) and writing each seed's run contiguously.
constexpr int kTile = 64;

__global__ __launch_bounds__(kBlock) void k_copy_counted_step_major(const float4* __restrict__ slots, int64_t n,
                                                                    const int32_t* __restrict__ counts,
                                                                    const int64_t* __restrict__ offs,
                                                                    float4* __restrict__ out) {
  __shared__ float4 tile[kTile][kTile + 1];
  __shared__ int32_t cnt[kTile];
  __shared__ int64_t dst[kTile];
  __shared__ int32_t wave_max[kBlock / 64];
  const int64_t p0 = int64_t(blockIdx.x) * kTile;
  const int t = threadIdx.x, lane = lane_id(), w = wave_id();
  if (t < kTile) {
    const int64_t p = p0 + t;
    cnt[t] = p < n ? counts[p] : 0;
    dst[t] = p < n ? offs[p] : 0;
  }
  __syncthreads();
  int32_t m = cnt[lane];
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
  if (lane == 0) wave_max[w] = m;
  __syncthreads();
  const int32_t cmax = wave_max[0];
  for (int32_t s0 = 0; s0 < cmax; s0 += kTile) {
    // read: wave w takes rows w, w + 4, ...; lane = seed
    const int32_t c = cnt[lane];
    for (int r = w; r < kTile; r += kBlock / 64) {
      const int32_t s = s0 + r;
      if (s < c) tile[r][lane] = slots[int64_t(s) * n + p0 + lane];
    }
    __syncthreads();
    // write: element e -> seed e / kTile, slot e % kTile
    for (int e = t; e < kTile * kTile; e += kBlock) {
      const int q = e / kTile, j = e % kTile;
      if (s0 + j < cnt[q]) out[dst[q] + s0 + j] = tile[j][q];
    }
    __syncthreads();
  }
}

// Chunked slots (8-slot component-planar chunks per seed) into the pruned
// layout: one wave per seed, lane = vertex; the four component loads of a
// wave hit the same few lines and every store is a contiguous 1 KiB run.
__global__ __launch_bounds__(kBlock) void k_copy_counted_chunked(const float* __restrict__ slots, int64_t region,
                                                                 const int32_t* __restrict__ counts,
                                                                 const int64_t* __restrict__ offs, int64_t n,
                                                                 float4* __restrict__ out) {
  const int64_t waves = int64_t(gridDim.x) * (kBlock / 64);
  for (int64_t p = int64_t(blockIdx.x) * (kBlock / 64) + wave_id(); p < n; p += waves) {
    const int32_t c = counts[p];
    const float* src = slots + p * region * 4;
    float4* dst = out + offs[p];
    for (int j = lane_id(); j < c; j += 64) {
      const float* ch = src + int64_t(j >> 3) * 32 + (j & 7);
      dst[j] = make_float4(ch[0], ch[8], ch[16], ch[24]);
    }
  }
}

// --------------------------------------------------------- hand-off groups
constexpr int kBuckets = 12;  // (map 0/1) x (neighbour slot 0..5)

__global__ __launch_bounds__(kBlock) void k_bucket_count(const int32_t* __restrict__ code, int64_t n,
                                                         int64_t* __restrict__ counts, int64_t nb) {
  __shared__ int32_t c[kBuckets];
  if (threadIdx.x < kBuckets) c[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) {
    const int32_t k = code[i];
    if (k >= 0) atomicAdd(&c[k], 1);
  }
  __syncthreads();
  if (threadIdx.x < kBuckets) counts[int64_t(threadIdx.x) * nb + blockIdx.x] = c[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_bucket_scatter(const int32_t* __restrict__ code,
                                                           const pa_particle* __restrict__ tmp, int64_t n,
                                                           const int64_t* __restrict__ offs, int64_t nb,
                                                           pa_particle* __restrict__ out, int64_t* __restrict__ src) {
  __shared__ int32_t wc[kBlock / 64];
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int32_t k = i < n ? code[i] : -1;
  const int lane = lane_id(), w = wave_id();
  for (int b = 0; b < kBuckets; ++b) {
    const bool mine = k == b;
    const uint64_t mask = __ballot(mine);
    if (lane == 0) wc[w] = __popcll(mask);
    __syncthreads();
    int pre = 0;
    for (int q = 0; q < w; ++q) pre += wc[q];
    if (mine) {
      const int64_t pos = offs[int64_t(b) * nb + blockIdx.x] + pre + __popcll(mask & ((uint64_t(1) << lane) - 1));
      out[pos] = tmp[i];
      if (src) src[pos] = i;
    }
    __syncthreads();
  }
}

__global__ void k_bucket_totals(const int64_t* __restrict__ offs, int64_t nb, int64_t* __restrict__ totals) {
  const int b = threadIdx.x;
  if (b < kBuckets) totals[b] = offs[int64_t(b + 1) * nb] - offs[int64_t(b) * nb];
}

}  // namespace

size_t scan_workspace_bytes(int64_t n) { return scan_ws_total(n); }

pa_status scan_i32(pa_gpu_ctx* ctx, const int32_t* d_in, int64_t n, int64_t* d_out, int64_t* ws) {
  return scan_rec<int32_t>(ctx, d_in, n, d_out, ws);
}

pa_status exclusive_scan_i64(pa_gpu_ctx* ctx, const int64_t* d_in, int64_t* d_out, size_t n) {
  void* ws = nullptr;
  PA_TRY(scratch(ctx, scan_ws_total(int64_t(n)), &ws));
  return scan_rec<int64_t>(ctx, d_in, int64_t(n), d_out, static_cast<int64_t*>(ws));
}

size_t group_ws_bytes(int64_t n) {
  const int64_t nb = (n + kBlock - 1) / kBlock;
  const int64_t m = nb * kBuckets;
  return size_t(m + 1) * 2 * sizeof(int64_t) + scan_ws_total(m) + 4096;
}

// Grouping of the trace kernel's per-particle hand-off codes (-1 none,
// map*6 + slot) into the caller's buffer; writes counts into handoff.  ws
// holds group_ws_bytes(n).
pa_status group_handoffs(pa_gpu_ctx* ctx, const int32_t* d_code, const pa_particle* d_tmp, int64_t n,
                         pa_handoff* handoff, void* ws) {
  const int64_t nb = (n + kBlock - 1) / kBlock;
  const int64_t m = nb * kBuckets;
  int64_t* counts = static_cast<int64_t*>(ws);
  int64_t* offs = counts + (m + 1);
  int64_t* sws = offs + (m + 1);
  launch_begin(ctx, "handoff_group");
  hipLaunchKernelGGL(k_bucket_count, dim3(unsigned(nb)), dim3(kBlock), 0, ctx->stream, d_code, n, counts, nb);
  PA_HIP(ctx, hipGetLastError());
  PA_TRY(scan_rec<int64_t>(ctx, counts, m, offs, sws));
  hipLaunchKernelGGL(k_bucket_scatter, dim3(unsigned(nb)), dim3(kBlock), 0, ctx->stream, d_code, d_tmp, n, offs,
                     nb, handoff->d_particles, handoff->d_sources);
  PA_HIP(ctx, hipGetLastError());
  int64_t* totals = static_cast<int64_t*>(ctx->counters) + 64;
  hipLaunchKernelGGL(k_bucket_totals, dim3(1), dim3(64), 0, ctx->stream, offs, nb, totals);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  int64_t h[kBuckets];
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(h, totals, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < kBuckets; ++b) handoff->counts[b / 6][b % 6] = size_t(h[b]);
  return PA_OK;
}

}  // namespace pa_amd

using namespace pa_amd;

namespace {
// The seed-major counted prunes' output offsets: offs[0..n] (d_offsets, or
// scratch) = the exclusive scan of counts, *total = offs[n] (read back; more
// than capacity is PA_E_CAPACITY).  `extra` more bytes of scratch, after the
// scan's workspace, come back in *extra_out for the caller's own use.
pa_status counted_offsets(pa_gpu_ctx* ctx, const int32_t* d_counts, size_t n, size_t capacity, int64_t* d_offsets,
                          size_t extra, int64_t** offs_out, int32_t** extra_out, size_t* total_out) {
  int64_t* offs = d_offsets;
  void* ws = nullptr;
  const size_t own = d_offsets ? 0 : (n + 1) * sizeof(int64_t);
  PA_TRY(scratch(ctx, own + scan_ws_total(int64_t(n)) + 4096 + extra, &ws));
  if (!offs) offs = static_cast<int64_t*>(ws);
  int64_t* sws = reinterpret_cast<int64_t*>(static_cast<char*>(ws) + own);
  if (extra_out)
    *extra_out = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + own + scan_ws_total(int64_t(n)) + 4096);
  launch_begin(ctx, "prune_scan");
  PA_TRY(scan_rec<int32_t>(ctx, d_counts, int64_t(n), offs, sws));
  launch_end(ctx);
  int64_t total = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&total, offs + n, sizeof total, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *total_out = size_t(total);
  *offs_out = offs;
  if (size_t(total) > capacity) return fail(ctx, PA_E_CAPACITY, "prune_counted: output capacity too small");
  return PA_OK;
}

}  // namespace

extern "C" {

pa_status pa_gpu_prune(pa_gpu_ctx* ctx, const float* d_vertices, size_t n, float* d_out, size_t* n_out) {
  if (!ctx || !n_out || (n && (!d_vertices || !d_out))) return PA_E_INVALID;
  if (n == 0) {
    *n_out = 0;
    return PA_OK;
  }
  PruneOp op{reinterpret_cast<const float4*>(d_vertices), reinterpret_cast<float4*>(d_out)};
  return select(ctx, op, int64_t(n), n_out, "prune");
}

pa_status pa_gpu_index_generate(pa_gpu_ctx* ctx, const float* d_vertices, size_t n, int32_t* d_indices,
                                size_t* n_out) {
  if (!ctx || !n_out) return PA_E_INVALID;
  if (n == 0) return fail(ctx, PA_E_INVALID, "index generation of an empty vertex array");
  if (!d_vertices || !d_indices) return PA_E_INVALID;
  if (n > size_t(INT32_MAX)) return fail(ctx, PA_E_INVALID, "int32 indices cannot address the vertex array");
  if (n == 1) {
    *n_out = 0;
    return PA_OK;
  }
  IndexOp op{reinterpret_cast<const float4*>(d_vertices), d_indices};
  return select(ctx, op, int64_t(n) - 1, n_out, "index");
}

pa_status pa_gpu_prune_counted(pa_gpu_ctx* ctx, const float* d_slots, size_t stride_slots,
                               const int32_t* d_vertex_counts, size_t n, float* d_out, size_t capacity, size_t* n_out,
                               int64_t* d_offsets) {
  if (!ctx || !n_out || (n && (!d_slots || !d_vertex_counts || !d_out))) return PA_E_INVALID;
  if (n == 0) {
    *n_out = 0;
    return PA_OK;
  }
  int64_t* offs = nullptr;
  PA_TRY(counted_offsets(ctx, d_vertex_counts, n, capacity, d_offsets, 0, &offs, nullptr, n_out));
  int64_t blocks = (int64_t(n) + 3) / 4;
  if (blocks > 65536) blocks = 65536;
  launch_begin(ctx, "prune_counted");
  hipLaunchKernelGGL(k_copy_counted, dim3(unsigned(blocks)), dim3(kBlock), 0, ctx->stream,
                     reinterpret_cast<const float4*>(d_slots), int64_t(stride_slots), d_vertex_counts, offs,
                     int64_t(n), reinterpret_cast<float4*>(d_out));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_finish_counted(pa_gpu_ctx* ctx, const float* d_slots, size_t stride_slots,
                                const int32_t* d_vertex_counts, size_t n, float* d_vertices, size_t capacity,
                                size_t* n_vertices, int64_t* d_offsets, int32_t* d_indices, size_t* n_indices,
                                pa_color_mode mode, float parameter, float* d_colors) {
  if (!ctx || !n_vertices || !n_indices) return PA_E_INVALID;
  if (n && (!d_slots || !d_vertex_counts || !d_vertices || !d_indices || !d_colors)) return PA_E_INVALID;
  if (!with_color_mode(int(mode), [](auto) {})) return fail(ctx, PA_E_INVALID, "unknown color mode");
  *n_vertices = 0;
  *n_indices = 0;
  if (n == 0) return PA_OK;
  if (capacity > size_t(INT32_MAX) + 1) capacity = size_t(INT32_MAX) + 1;  // int32 indices address the chunk
  int64_t* offs = nullptr;
  int32_t* flag = nullptr;
  size_t m = 0;
  PA_TRY(counted_offsets(ctx, d_vertex_counts, n, capacity, d_offsets, 256, &offs, &flag, &m));
  *n_vertices = m;
  hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, ctx->stream, flag);
  int64_t blocks = (int64_t(n) + 3) / 4;
  if (blocks > 65536) blocks = 65536;
  launch_begin(ctx, "finish_counted");
  with_color_mode(int(mode), [&](auto M) {
    hipLaunchKernelGGL(k_finish_counted<decltype(M)::value>, dim3(unsigned(blocks)), dim3(kBlock), 0, ctx->stream,
                       reinterpret_cast<const float4*>(d_slots), int64_t(stride_slots), d_vertex_counts, offs,
                       int64_t(n), reinterpret_cast<float4*>(d_vertices), d_indices,
                       reinterpret_cast<float4*>(d_colors), parameter, flag);
  });
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  int32_t irregular = 0;
  PA_HIP(ctx, hipMemcpyAsync(&irregular, flag, sizeof irregular, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (!irregular) {
    *n_indices = m - 2 * n;
    return PA_OK;
  }
  // a curve of another shape: the general index generator and colouring over the pruned array
  if (m < 2) return PA_OK;
  size_t ni = 0;
  PA_TRY(pa_gpu_index_generate(ctx, d_vertices, m, d_indices, &ni));
  *n_indices = ni;
  return color_launch(ctx, reinterpret_cast<const float4*>(d_vertices), d_indices, int64_t(ni), int(mode),
                      parameter, reinterpret_cast<float4*>(d_colors));
}

pa_status pa_gpu_prune_counted_chunked(pa_gpu_ctx* ctx, const float* d_slots, size_t region_slots,
                                       const int32_t* d_vertex_counts, size_t n, float* d_out, size_t capacity,
                                       size_t* n_out, int64_t* d_offsets) {
  if (!ctx || !n_out || (n && (!d_slots || !d_vertex_counts || !d_out))) return PA_E_INVALID;
  if (region_slots % 8 != 0) return fail(ctx, PA_E_INVALID, "chunked slots: region_slots must be a multiple of 8");
  if (n == 0) {
    *n_out = 0;
    return PA_OK;
  }
  int64_t* offs = d_offsets;
  void* ws = nullptr;
  const size_t own = d_offsets ? 0 : (n + 1) * sizeof(int64_t);
  PA_TRY(scratch(ctx, own + scan_ws_total(int64_t(n)) + 4096, &ws));
  if (!offs) offs = static_cast<int64_t*>(ws);
  int64_t* sws = reinterpret_cast<int64_t*>(static_cast<char*>(ws) + own);
  launch_begin(ctx, "prune_scan");
  PA_TRY(scan_rec<int32_t>(ctx, d_vertex_counts, int64_t(n), offs, sws));
  launch_end(ctx);
  int64_t total = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&total, offs + n, sizeof total, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n_out = size_t(total);
  if (size_t(total) > capacity) return fail(ctx, PA_E_CAPACITY, "prune_counted: output capacity too small");
  int64_t blocks = (int64_t(n) + 3) / 4;
  if (blocks > 65536) blocks = 65536;
  launch_begin(ctx, "prune_counted");
  hipLaunchKernelGGL(k_copy_counted_chunked, dim3(unsigned(blocks)), dim3(kBlock), 0, ctx->stream, d_slots,
                     int64_t(region_slots), d_vertex_counts, offs, int64_t(n), reinterpret_cast<float4*>(d_out));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

pa_status pa_gpu_prune_counted_step_major(pa_gpu_ctx* ctx, const float* d_slots, size_t n,
                                          const int32_t* d_vertex_counts, float* d_out, size_t capacity,
                                          size_t* n_out, int64_t* d_offsets) {
  if (!ctx || !n_out || (n && (!d_slots || !d_vertex_counts || !d_out))) return PA_E_INVALID;
  if (n == 0) {
    *n_out = 0;
    return PA_OK;
  }
  int64_t* offs = d_offsets;
  void* ws = nullptr;
  const size_t own = d_offsets ? 0 : (n + 1) * sizeof(int64_t);
  PA_TRY(scratch(ctx, own + scan_ws_total(int64_t(n)) + 4096, &ws));
  if (!offs) offs = static_cast<int64_t*>(ws);
  int64_t* sws = reinterpret_cast<int64_t*>(static_cast<char*>(ws) + own);
  launch_begin(ctx, "prune_scan");
  PA_TRY(scan_rec<int32_t>(ctx, d_vertex_counts, int64_t(n), offs, sws));
  launch_end(ctx);
  int64_t total = 0;
  PA_TRY(capture_guard(ctx));
  PA_HIP(ctx, hipMemcpyAsync(&total, offs + n, sizeof total, hipMemcpyDeviceToHost, ctx->stream));
  PA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n_out = size_t(total);
  if (size_t(total) > capacity) return fail(ctx, PA_E_CAPACITY, "prune_counted: output capacity too small");
  const int64_t blocks = (int64_t(n) + kTile - 1) / kTile;
  if (blocks > INT32_MAX) return fail(ctx, PA_E_INVALID, "prune_counted: too many seeds");
  launch_begin(ctx, "prune_counted");
  hipLaunchKernelGGL(k_copy_counted_step_major, dim3(unsigned(blocks)), dim3(kBlock), 0, ctx->stream,
                     reinterpret_cast<const float4*>(d_slots), int64_t(n), d_vertex_counts, offs,
                     reinterpret_cast<float4*>(d_out));
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

}  // extern "C"
