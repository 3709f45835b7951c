// Internal declarations shared by the HIP translation units of libpa_gpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pa_gpu.h"

static_assert(sizeof(pa_particle) == 48, "pa_particle must match pa::particle (48 B)");
static_assert(sizeof(pa_round_info) == 40, "pa_round_info layout");
static_assert(sizeof(pa_round_result) == 48, "pa_round_result: the six result words of the round counters");

namespace pa_amd {

// Device-side description of one block.  upper = offset + size is computed
// once on the host in FP32, the same single rounding the reference performs
// at particle_tracer.cpp:263.
struct FieldDesc {
  const void* data;   // nodes, storage-dependent (4 components each)
  int32_t dims[3];    // ghosted node counts
  int32_t storage;    // pa_storage
  float offset[3];
  float spacing[3];
  float upper[3];
  float cell_hi[3];   // dims - 2 as FP32: the largest cell subscript
  int64_t stride0;    // dims[1] * dims[2]
  int64_t stride1;    // dims[2]
  const void* cells;  // cell records (PA_LAYOUT_CELLS) or null
  int64_t cstride0;   // (dims[1]-1) * (dims[2]-1)
  int64_t cstride1;   // dims[2]-1
  int32_t layout;     // pa_layout
  float inv_spacing[3];  // RN(1 / spacing), for the corrected quotient
  int32_t fast_div;      // spacing within [2^-20, 2^20] on every axis
  float inv_lo[3];       // RN(1 / spacing - inv_spacing): the reciprocal's low part
};

// Vec4 index of slot s of seed p in 8-seed grouped slots (PA_SLOTS_GROUP8).
__host__ __device__ __forceinline__ int64_t group8_slot(int64_t p, int64_t s, int64_t region) {
  return ((p >> 3) * region + s) * 8 + (p & 7);
}

// Per-launch tracer configuration.  Slot 0 is the own block (vector field
// index -1), slots 1..6 the neighbours 0..5.
struct TraceConfig {
  FieldDesc field[7];
  int32_t has_field[7];
  int32_t local_mask;        // bit k: own block has a neighbour across face k
  int32_t face_mask[6];      // bit d: neighbour block k has a neighbour across face d
  float dt;
  float a0, a1, a2, a3;      // frozen RK4 coefficients b_k * dt
  float half_dt;             // (1/2) * dt for classic RK4 stage points
  float zero_dt;             // 0 * dt (odeint's zero tableau entries)
  float one_dt;              // 1 * dt
  float stop_eps2;
  // The w component of a frozen-derivative step, w + a_0 0 + a_1 0 + ..., is
  // w + w_rk4 (four terms) or w + w_euler (dt 0): +0, -0 if every term is -0,
  // NaN if one is NaN (zero_terms in trace.hip).
  float w_rk4, w_euler;
  int32_t interp;
  int32_t xcd_remap;
  int32_t quad_switch;       // hybrid engine: a wave moves to quads at this many live particles (1..16)
  int32_t pair_live;         // hybrid engine: FP32 pair gathers while more lanes than this are live
};

}  // namespace pa_amd

struct pa_gpu_field {
  pa_amd::FieldDesc desc;
  float size[3] = {0.0f, 0.0f, 0.0f};  // the block's extent as the caller gave it
  void* d_nodes = nullptr;
  size_t bytes = 0;
  void* d_cells = nullptr;
  size_t cell_bytes = 0;
};

struct pa_gpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  // Kernel timing
  bool timing = false;
  struct timed {
    hipEvent_t a, b;
    std::string name;
  };
  std::vector<timed> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, uint64_t>> kernel_ms;
  // Scratch that grows on demand (block sums, counters, flags).
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  void* counters = nullptr;  // 4 KiB of device counters
  void* pinned = nullptr;    // 4 KiB of pinned host memory for counter readback
  void* static_scratch = nullptr;  // the device's fixed no-allocation workspace (never freed)
  bool capturing = false;          // between pa_gpu_graph_begin and pa_gpu_graph_end
  bool capture_failed = false;     // a call that reads back was made while capturing
  // Pipelined rounds: a round's finish runs on finish_stream while the next round traces on stream;
  // two sets of workspace and counters alternate, each free again once its finish event has passed.
  hipStream_t finish_stream = nullptr;
  struct RoundSet {
    void* ws = nullptr;
    size_t ws_bytes = 0;
    void* counters = nullptr;  // 4 KiB, used in place of the context's counters
    hipEvent_t traced = nullptr, finished = nullptr;
    bool pending = false;      // a finish was queued and the context's stream has not waited for it
  };
  RoundSet round_set[2];
  uint64_t pipelined_rounds = 0;
};

struct pa_gpu_tracer {
  pa_gpu_ctx* ctx = nullptr;
  const pa_gpu_field* local = nullptr;
  const pa_gpu_field* neighbors[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int32_t neighbor_rank[6] = {-1, -1, -1, -1, -1, -1};
  int32_t face_mask[6] = {0, 0, 0, 0, 0, 0};
  int32_t integrator = PA_INTEGRATOR_EULER;  // the reference default (particle_tracer.hpp:68)
  float step_size = 1.0f;                    // particle_tracer.hpp:69
  int32_t stop_rule = PA_STOP_ISZERO;
  float stop_eps = 0.0f;
  int32_t interp = PA_INTERP_REFERENCE;
  int32_t engine = PA_ENGINE_LANE;
  int32_t waves = 0;  // resident waves of the cooperative engine (0 = automatic)
  int32_t xcd_remap = 1;
  int32_t quad_switch = 16;  // launch constants of the hybrid engine (pa_gpu_tracer_set_launch)
  int32_t pair_live = 16;
};

namespace pa_amd {

pa_status fail(pa_gpu_ctx* ctx, pa_status s, const std::string& msg);
pa_status check_hip(pa_gpu_ctx* ctx, hipError_t e, const char* what);
// Grow-only scratch on the context (PA_E_STATE if it would grow while the
// stream is being captured).
pa_status scratch(pa_gpu_ctx* ctx, size_t bytes, void** out);
// Calls that read back or copy from pageable host memory cannot be part of
// a captured graph: they fail with PA_E_STATE while the context captures,
// and the capture is then discarded by pa_gpu_graph_end.
pa_status capture_guard(pa_gpu_ctx* ctx);
// Timing brackets around a launch.
void launch_begin(pa_gpu_ctx* ctx, const char* name);
void launch_end(pa_gpu_ctx* ctx);
// Build the per-launch configuration from a tracer.
pa_status make_config(pa_gpu_tracer* t, TraceConfig* cfg);
// Integrator constants, stop rule and interpolation of a tracer (no fields).
void fill_step_constants(const pa_gpu_tracer* t, TraceConfig* cfg);
int mode_of_integrator(int integrator);
pa_status check_integrator_of(pa_gpu_ctx* ctx, int integrator);

// Launch of the wave-cooperative tracer (trace_coop.hip).
struct CoopLaunch {
  const pa_particle* particles = nullptr;
  int64_t n = 0;
  int32_t* counts = nullptr;
  int stop_rule = PA_STOP_ISZERO;
  int waves = 0;
  float4* verts = nullptr;
  int64_t region = 0;
  pa_particle* cont = nullptr;
  float4* endpoints = nullptr;
  int32_t* seed_steps = nullptr;
  float4* const* chunks = nullptr;
  int64_t ppic = 0;
  int64_t maxR = 0;
  pa_particle* hand_tmp = nullptr;
  int32_t* hand_code = nullptr;
};
pa_status launch_coop(pa_gpu_ctx* ctx, const TraceConfig& cfg, int mode, bool union_form, const CoopLaunch& l);
// Union trace on the quad engine into chunked slots; steps, ncont and work
// are zeroed device counters.
pa_status launch_quad_union(pa_gpu_ctx* ctx, const TraceConfig& cfg, int mode, const CoopLaunch& l,
                            unsigned long long* steps, unsigned* ncont, unsigned long long* work);
// Union trace on the hybrid engine (lane mode, then quad mode) into
// step-major slots.
// pairs (optional): per-seed index-generator pairs inside the seed's run.
pa_status launch_hybrid_union(pa_gpu_ctx* ctx, const TraceConfig& cfg, int mode, const CoopLaunch& l,
                              unsigned long long* steps, unsigned* ncont, int32_t* pairs = nullptr);
// Exclusive scan of n int32 values into n+1 int64 outputs with the caller's
// workspace of scan_workspace_bytes(n) bytes.
size_t scan_workspace_bytes(int64_t n);
pa_status scan_i32(pa_gpu_ctx* ctx, const int32_t* d_in, int64_t n, int64_t* d_out, int64_t* ws);
// Prune + index + colour of grouped union slots (the rest of a fused
// round); pairs = per-seed in-run index pairs from the hybrid tracer.
// Asynchronous: the totals land in totals[0] (vertices) and totals[1]
// (indices) on the device; nothing is written when they exceed the
// capacities (or 2^31-1 vertices), which the caller checks after its one
// synchronisation; a finish wave that gives up a bounded wait sets bit 16
// of *flags.
pa_status finish_group8(pa_gpu_ctx* ctx, const float* d_slots, int64_t n, int64_t region, const int32_t* d_counts,
                        const int32_t* d_pairs, int mode, float param, float* d_vertices, size_t vertex_capacity,
                        int32_t* d_indices, size_t index_capacity, float* d_colors, void* ws, int64_t* totals,
                        unsigned* flags);
size_t finish_ws_bytes(int64_t n);
// The fused round (trace on the hybrid engine + prune + index + colour)
// with the caller's device workspace of union_round_ws_bytes(n) bytes; the
// rest is pa_gpu_tracer_trace_union_round.  With finish_stream (and
// d_result), the finish waits for the trace through the event traced and
// runs on finish_stream; the trace stays on the context's stream.
size_t union_round_ws_bytes(int64_t n);
pa_status union_round(pa_gpu_tracer* t, const pa_particle* d_particles, size_t n, size_t region_slots,
                      float* d_slots, int32_t* d_vertex_counts, pa_particle* d_continuation, size_t* n_continuation,
                      pa_color_mode mode, float parameter, float* d_vertices, size_t vertex_capacity,
                      size_t* n_vertices, int32_t* d_indices, size_t index_capacity, size_t* n_indices,
                      float* d_colors, float* d_endpoints, int32_t* d_seed_steps, uint64_t* total_steps, void* ws,
                      pa_round_result* d_result, hipStream_t finish_stream = nullptr, hipEvent_t traced = nullptr);
// Make the context's stream wait for every pipelined round's finish.
pa_status round_join(pa_gpu_ctx* ctx);
// Fixed workspace of the no-allocation fused round (noalloc.hip): its
// device address on the context's device, looked up once per device at
// context creation.
pa_status static_scratch_setup(pa_gpu_ctx* ctx);
// The colour stage's kernel (one lane per index).
pa_status color_launch(pa_gpu_ctx* ctx, const float4* d_vertices, const int32_t* d_indices, int64_t n_indices,
                       int mode, float param, float4* d_colors);

// Exclusive scan of n int64 values in place on the device (n+1 outputs:
// out[n] = total).  Uses ctx scratch.
pa_status exclusive_scan_i64(pa_gpu_ctx* ctx, const int64_t* d_in, int64_t* d_out, size_t n);

}  // namespace pa_amd

#define PA_HIP(ctx, expr)                                           \
  do {                                                              \
    hipError_t e__ = (expr);                                        \
    if (e__ != hipSuccess) return pa_amd::check_hip(ctx, e__, #expr); \
  } while (0)

#define PA_TRY(expr)                 \
  do {                               \
    pa_status s__ = (expr);          \
    if (s__ != PA_OK) return s__;    \
  } while (0)
