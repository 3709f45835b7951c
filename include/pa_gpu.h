/*
 * pa_gpu.h — C ABI of the MI355X (gfx950) implementation of pars'
 * per-process advection path.
 *
 * Every entry point stands in for one stage of the reference `pa` library
 * (acdemiralp/pars); the comment above each names the reference interface it
 * replaces (file:line).  The C++ classes in pars_amd/include/pa_amd/ wrap
 * these functions with the reference's class and method names.
 *
 * Conventions
 *   - Every function returns pa_status (0 = PA_OK).  pa_gpu_last_error(ctx)
 *     returns the message of the last failure on that context.  No C++
 *     exception crosses this boundary.
 *   - Pointers named d_* are device (HBM) pointers, e.g. from pa_gpu_malloc;
 *     pointers named h_* or without prefix are host pointers.
 *   - Work is ordered on the context's HIP stream.  Calls that return host
 *     values (counts, sizes) synchronise that stream before returning; the
 *     others return once the work is enqueued.  pa_gpu_synchronize() waits.
 *   - One host thread drives a context at a time (the reference runs its
 *     stages under MPI_THREAD_SERIALIZED, pars/source/pipeline.cpp:14).
 */
#ifndef PA_GPU_H
#define PA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PA_GPU_ABI_VERSION 2

typedef enum pa_status {
  PA_OK = 0,
  PA_E_INVALID = 1,      /* invalid argument */
  PA_E_UNSUPPORTED = 2,  /* valid in the reference but not implemented on the GPU (e.g. odeint dopri5) */
  PA_E_DEVICE = 3,       /* HIP runtime error */
  PA_E_NOMEM = 4,        /* device allocation failed */
  PA_E_CAPACITY = 5,     /* a caller-provided buffer is too small */
  PA_E_STATE = 6         /* call order violated (e.g. trace without a field) */
} pa_status;

typedef struct pa_gpu_ctx pa_gpu_ctx;
typedef struct pa_gpu_field pa_gpu_field;
typedef struct pa_gpu_tracer pa_gpu_tracer;
typedef struct pa_gpu_graph pa_gpu_graph;
typedef struct pa_gpu_server pa_gpu_server;

/* pa::particle (pa/include/pa/math/particle.hpp:30-35).  48 bytes: the
 * Eigen::Vector4f member gives the reference struct 16-byte alignment, so
 * 40 bytes of members are padded to 48. */
typedef struct pa_particle {
  float   position[4];
  int32_t remaining_iterations;
  int32_t vector_field_index; /* -1 = own block, 0..5 = neighbour -x,+x,-y,+y,-z,+z */
  int32_t original_rank;
  int32_t original_voxel[3];
  int32_t reserved_[2];
} pa_particle;

/* Field storage on the device: padded 4-component nodes. */
typedef enum pa_storage { PA_STORAGE_FP32 = 0, PA_STORAGE_BF16 = 1, PA_STORAGE_FP16 = 2 } pa_storage;

/* pa::variant_integrator alternatives (pa/include/pa/math/integrators.hpp:16-33),
 * same order.  The reference's system() freezes the derivative at the last
 * vertex (pa/source/stages/particle_tracer.cpp:315-318), which the GPU
 * reproduces for EULER and RUNGE_KUTTA_4.  PA_INTEGRATOR_CLASSIC_RK4 is the
 * textbook four-evaluation RK4, a build extension. */
typedef enum pa_integrator {
  PA_INTEGRATOR_EULER = 0,
  PA_INTEGRATOR_MODIFIED_MIDPOINT = 1,
  PA_INTEGRATOR_RUNGE_KUTTA_4 = 2,
  PA_INTEGRATOR_RUNGE_KUTTA_CASH_KARP_54 = 3,
  PA_INTEGRATOR_RUNGE_KUTTA_DORMAND_PRINCE_5 = 4,
  PA_INTEGRATOR_RUNGE_KUTTA_FEHLBERG_78 = 5,
  PA_INTEGRATOR_ADAMS_BASHFORTH_2 = 6,
  PA_INTEGRATOR_ADAMS_BASHFORTH_MOULTON_2 = 7,
  PA_INTEGRATOR_CLASSIC_RK4 = 100
} pa_integrator;

/* Stop test before each step: the reference's vector.isZero()
 * (particle_tracer.cpp:312; every |v_i| <= 1e-5f) or the extension
 * |v| < eps. */
typedef enum pa_stop_rule { PA_STOP_ISZERO = 0, PA_STOP_NORM = 1 } pa_stop_rule;

/* Interpolation weights.  PA_INTERP_REFERENCE reproduces
 * vector_field::interpolate (pa/source/math/vector_field.cpp:22-48): cell
 * from floor((p-o)/h), weight from fmod(p-o, h)/h, which reads one cell too
 * far when the rounded quotient lands on an integer the true quotient is just
 * below.  PA_INTERP_CONSISTENT (extension) uses w = q - floor(q) with the
 * same cell: the intended trilinear interpolant, and three divisions and
 * three remainders cheaper per evaluation. */
typedef enum pa_interp { PA_INTERP_REFERENCE = 0, PA_INTERP_CONSISTENT = 1 } pa_interp;

/* pa::color_generator::mode (pa/include/pa/stages/color_generator.hpp:12-19),
 * same order; PA_COLOR_MAGNITUDE is a build extension (hue from segment
 * length / parameter). */
typedef enum pa_color_mode {
  PA_COLOR_HSL_CONSTANT_S = 0,
  PA_COLOR_HSL_CONSTANT_L = 1,
  PA_COLOR_HSV_CONSTANT_S = 2,
  PA_COLOR_HSV_CONSTANT_V = 3,
  PA_COLOR_RGB = 4,
  PA_COLOR_MAGNITUDE = 100
} pa_color_mode;

/* pa::ftle_map_generator::mode (pa/include/pa/stages/ftle_map_generator.hpp:16-20). */
typedef enum pa_ftle_mode { PA_FTLE_REGULAR = 0, PA_FTLE_FRACTIONAL_ANISOTROPY = 1 } pa_ftle_mode;

/* pa::particle_tracer::round_info without the two hand-off maps
 * (pa/include/pa/stages/particle_tracer.hpp:26-37). */
typedef struct pa_round_info {
  uint64_t maximum_remaining_iterations;
  uint64_t vertices_per_integral_curve;
  uint64_t vertices_per_integral_curve_last;
  uint64_t integral_curve_offset;
  uint64_t integral_curve_count;
} pa_round_info;

/* The two hand-off maps of round_info (particle_tracer.hpp:34-35), compacted
 * on the device into one caller buffer: first out_of_bounds_particles for
 * neighbour slots 0..5, then neighbor_out_of_bounds_particles for slots
 * 0..5, each group in source-particle order.  counts[kind][slot] is set on
 * return.  The reference keys the maps by neighbour rank; the slot is the
 * index into the neighbour array. */
typedef struct pa_handoff {
  pa_particle* d_particles; /* device buffer, capacity >= particle count */
  int64_t*     d_sources;   /* optional: source particle index per entry */
  size_t       capacity;
  size_t       counts[2][6];
} pa_handoff;

/* ------------------------------------------------------------------ context */
pa_status   pa_gpu_ctx_create(int device, pa_gpu_ctx** out);
pa_status   pa_gpu_ctx_destroy(pa_gpu_ctx* ctx);
const char* pa_gpu_last_error(const pa_gpu_ctx* ctx);
pa_status   pa_gpu_synchronize(pa_gpu_ctx* ctx);
/* Device name and compute-unit count (for reports). */
pa_status   pa_gpu_device_info(pa_gpu_ctx* ctx, char* name, size_t name_len, int* compute_units,
                               size_t* free_bytes, size_t* total_bytes);

/* Device memory and copies (stream-ordered on the context stream). */
pa_status   pa_gpu_malloc(pa_gpu_ctx* ctx, size_t bytes, void** d_out);
pa_status   pa_gpu_free(pa_gpu_ctx* ctx, void* d_ptr);
pa_status   pa_gpu_memcpy_h2d(pa_gpu_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
pa_status   pa_gpu_memcpy_d2h(pa_gpu_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
pa_status   pa_gpu_memcpy_d2d(pa_gpu_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* Streaming device copy by a plain kernel (float4 loads and stores), e.g.
 * to calibrate hardware byte counters against a known transfer. */
pa_status   pa_gpu_copy(pa_gpu_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* Scattered line reads by a plain kernel (calibration, no counterpart in the
 * reference): n_reads reads of read_bytes (16 to 64, 16-byte units) at the
 * start of pseudo-random line_bytes lines of [d_src, d_src + bytes), 8 in
 * flight per lane; the XOR of every 32-bit word read is XORed into *d_xor.
 * Its rate bounds kernels whose gathers land on unrelated lines (the
 * particle pool's). */
pa_status   pa_gpu_gather_lines(pa_gpu_ctx* ctx, const void* d_src, size_t bytes, int32_t line_bytes,
                                int32_t read_bytes, int64_t n_reads, uint64_t seed, uint32_t* d_xor);

/* Kernel timing: when enabled, every kernel launch of the library is
 * bracketed with HIP events on the context stream; per-kernel totals are
 * read with pa_gpu_kernel_time (name as in the kernel table, e.g.
 * "trace", "prune_counted", "index", "color", "flow_map", "ftle"). */
pa_status   pa_gpu_timing_enable(pa_gpu_ctx* ctx, int enable);
pa_status   pa_gpu_timing_reset(pa_gpu_ctx* ctx);
pa_status   pa_gpu_kernel_time(pa_gpu_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches);

/* ------------------------------------------------------------------- fields */
/* Block upload.  Replaces the data_io block load into pa::vector_field
 * (pa/source/stages/data_io.cpp:91-115; vector_field.hpp:14-24): h_xyz is the
 * reference's boost::multi_array<vector3,3> payload, [d0][d1][d2][3] FP32,
 * C order (z fastest), dims are the ghosted node counts.  offset, size and
 * spacing are the vector_field members.  The device copy is stored as padded
 * 4-component nodes in the requested precision (BF16/FP16 rounded to
 * nearest even). */
pa_status   pa_gpu_field_create(pa_gpu_ctx* ctx, const float* h_xyz, const int32_t dims[3], const float offset[3],
                                const float size[3], const float spacing[3], pa_storage storage,
                                pa_gpu_field** out);
/* Same, from an FP32 xyz array already on the device. */
pa_status   pa_gpu_field_create_device(pa_gpu_ctx* ctx, const float* d_xyz, const int32_t dims[3],
                                       const float offset[3], const float size[3], const float spacing[3],
                                       pa_storage storage, pa_gpu_field** out);
/* Synthesised on the device (no host array): kind 0 = ABC flow
 * u = (A sin z + C cos y, B sin x + A cos z, C sin y + B cos x) with
 * params = {A, B, C}; kind 1 = sum of Fourier modes u = sum_m cos(k_m.x +
 * phi_m) c_m with params = n_modes x {k0,k1,k2,c0,c1,c2,phi}.  Node (i,j,k)
 * sits at (origin + (i,j,k)) * h, evaluated in FP64, rounded to storage. */
pa_status   pa_gpu_field_synthesize(pa_gpu_ctx* ctx, int kind, const double* params, int32_t n_params,
                                    const int32_t dims[3], const int32_t origin[3], double h, const float offset[3],
                                    const float size[3], const float spacing[3], pa_storage storage,
                                    pa_gpu_field** out);
/* Extension: device layout of a block.  PA_LAYOUT_CELLS stores, for every
 * cell, its 8 corner nodes in one record (FP32: 96 B used of a 128-B record,
 * BF16/FP16: 48 B of 64 B), so a trilinear gather reads one cache line
 * instead of ~4.5; it costs 8x the node storage.  Node values are copied bit
 * for bit, so results do not change.  The node array is kept for read-back. */
typedef enum pa_layout { PA_LAYOUT_NODES = 0, PA_LAYOUT_CELLS = 1 } pa_layout;
pa_status   pa_gpu_field_set_layout(pa_gpu_ctx* ctx, pa_gpu_field* field, pa_layout layout);

/* data_io::load_vector_field (pa/source/stages/data_io.cpp:91-115): read
 * one rank's ghosted block from an HDF5 file (dataset "vectors"
 * [X][Y][Z][3] float; file attribute "spacing" float[3]) into host memory.
 * h_xyz receives ghosted[0] * ghosted[1] * ghosted[2] * 3 floats, C order,
 * from the hyperslab at offset (the rank's integer offset); spacing_out is
 * the attribute divided by its largest component, offset_out =
 * float(offset) * spacing and size_out = float(block) * spacing, FP32 as
 * data_io.cpp:108-114 computes them.  libhdf5 1.10+ is opened at run time
 * (PA_E_UNSUPPORTED without it).  Needs no GPU.  error (optional)
 * receives a message. */
pa_status   pa_gpu_read_vector_block(const char* path, const int64_t offset[3], const int32_t ghosted[3],
                                     const int32_t block[3], float* h_xyz, float offset_out[3], float size_out[3],
                                     float spacing_out[3], char* error, size_t error_len);
/* data_io::load_dimensions (pa/source/stages/data_io.cpp:36-40): the node
 * counts of the dataset "vectors" [X][Y][Z][3].  Needs no GPU. */
pa_status   pa_gpu_read_vector_dims(const char* path, int64_t dims[3], char* error, size_t error_len);
/* The same block straight into a device field of the given storage (read
 * into pinned memory, uploaded and encoded on the device). */
pa_status   pa_gpu_field_load(pa_gpu_ctx* ctx, const char* path, const int64_t offset[3], const int32_t ghosted[3],
                              const int32_t block[3], pa_storage storage, pa_gpu_field** out);

/* Decoded FP32 node values back to the host, [d0][d1][d2][3]. */
pa_status   pa_gpu_field_read(pa_gpu_ctx* ctx, const pa_gpu_field* field, float* h_xyz);
pa_status   pa_gpu_field_bytes(const pa_gpu_field* field, size_t* bytes);
/* vector_field's public members (vector_field.hpp:14-24): ghosted node
 * counts, offset, size, spacing, and the storage type (any may be NULL). */
pa_status   pa_gpu_field_info(const pa_gpu_field* field, int32_t dims[3], float offset[3], float size[3],
                              float spacing[3], pa_storage* storage);
pa_status   pa_gpu_field_destroy(pa_gpu_ctx* ctx, pa_gpu_field* field);

/* vector_field::contains + interpolate (pa/source/math/vector_field.cpp:12-48)
 * at n points: d_points and d_out are n vec4; d_out = (u, v, w, inside) with
 * inside 1.0 or 0.0 (u, v, w are 0 when outside). */
pa_status   pa_gpu_field_interpolate(pa_gpu_ctx* ctx, const pa_gpu_field* field, pa_interp interp,
                                     const float* d_points, size_t n, float* d_out);

/* -------------------------------------------------------------------- seeds */
/* pa::seed_generator::generate (pa/source/stages/seed_generator.cpp:9-21;
 * pa/include/pa/stages/seed_generator.hpp:13): a strided lattice of
 * int(size/stride) particles per axis, unravelled in C order.  Query the
 * count with pa_gpu_seed_count first. */
pa_status   pa_gpu_seed_count(const float size[3], const float stride[3], int32_t dims_out[3], size_t* n_out);
pa_status   pa_gpu_seed_generate(pa_gpu_ctx* ctx, const float offset[3], const float size[3], const float stride[3],
                                 int32_t remaining_iterations, int32_t rank, pa_particle* d_out, size_t capacity,
                                 size_t* n_out);

/* ------------------------------------------------------------------- tracer */
/* pa::particle_tracer (pa/include/pa/stages/particle_tracer.hpp:39-62). */
pa_status   pa_gpu_tracer_create(pa_gpu_ctx* ctx, pa_gpu_tracer** out);
pa_status   pa_gpu_tracer_destroy(pa_gpu_tracer* tracer);
/* set_local_vector_field (particle_tracer.hpp:46). */
pa_status   pa_gpu_tracer_set_local_vector_field(pa_gpu_tracer* tracer, const pa_gpu_field* field);
/* set_neighbor_vector_fields (particle_tracer.hpp:47) plus the neighbour
 * ranks of partitioner::neighbor_rank_info (partitioner.hpp:39); a rank of
 * -1 means no neighbour across that face; a NULL field means the block is
 * not preloaded (the reference loads neighbours only with load balancing,
 * pars/source/pipeline.cpp:96-98). */
pa_status   pa_gpu_tracer_set_neighbor_vector_fields(pa_gpu_tracer* tracer, const pa_gpu_field* const fields[6],
                                                     const int32_t neighbor_rank[6]);
/* set_integrator (particle_tracer.hpp:48): EULER, RUNGE_KUTTA_4 and
 * CLASSIC_RK4 are supported; the other odeint steppers return
 * PA_E_UNSUPPORTED. */
pa_status   pa_gpu_tracer_set_integrator(pa_gpu_tracer* tracer, pa_integrator integrator);
/* set_step_size (particle_tracer.hpp:49). */
pa_status   pa_gpu_tracer_set_step_size(pa_gpu_tracer* tracer, float step_size);
/* Extension: stop rule (default PA_STOP_ISZERO, the reference's). */
pa_status   pa_gpu_tracer_set_stop_rule(pa_gpu_tracer* tracer, pa_stop_rule rule, float eps);
/* Extension: interpolation weights (default PA_INTERP_REFERENCE). */
pa_status   pa_gpu_tracer_set_interpolation(pa_gpu_tracer* tracer, pa_interp interp);
/* Extension: tracer engine.  PA_ENGINE_LANE runs one lane per particle
 * with per-lane gathers; PA_ENGINE_COOPERATIVE keeps one lane per
 * particle for the arithmetic but gathers the 8 corners cooperatively (groups
 * of lanes load one particle's cell record or z-pairs as coalesced 16-B
 * pieces, redistributed through LDS) and refills finished lanes from a work
 * counter.  PA_ENGINE_QUAD (union trace only, chunked slots) splits every
 * particle over the four lanes of a quad, one vec4 component each, which
 * shortens the serial step chain that the longest trajectories of a round
 * sit on; persistent waves refill quads from a work counter.
 * PA_ENGINE_HYBRID (union trace only, grouped slots) runs lane mode while
 * a wave holds more than 16 live particles and moves the survivors onto
 * quads after that.  All engines produce identical results;
 * PA_ENGINE_LANE is the default.  waves = resident waves of the cooperative
 * engine (0 = automatic). */
typedef enum pa_engine {
  PA_ENGINE_LANE = 0,
  PA_ENGINE_COOPERATIVE = 1,
  PA_ENGINE_QUAD = 2,
  PA_ENGINE_HYBRID = 3
} pa_engine;
pa_status   pa_gpu_tracer_set_engine(pa_gpu_tracer* tracer, pa_engine engine, int32_t waves);
/* Extension: XCD-aware workgroup order (default 1).  The 8 XCDs take
 * workgroups round-robin; with 1 each processes a contiguous eighth of the
 * particle list, so particles seeded close together share an L2; with 0
 * the hardware order stays (every XCD gets every part of the list); with
 * 2..64 each XCD takes runs of that many consecutive workgroups in turn
 * (locality within a run, balance across the list).  Results are
 * unchanged. */
pa_status   pa_gpu_tracer_set_xcd_remap(pa_gpu_tracer* tracer, int32_t enable);
/* Extension: launch constants of the hybrid engine, for an autotuner: a
 * wave moves its live particles onto quads once at most quad_switch (1..16,
 * default 16) are left; FP32 cell records are gathered by lane pairs while
 * more than pair_live (0..64, default 16) lanes are live.  Results are
 * unchanged; only the speed is. */
pa_status   pa_gpu_tracer_set_launch(pa_gpu_tracer* tracer, int32_t quad_switch, int32_t pair_live);
/* Union-of-7 topology: bit d of face_mask[k] set when neighbour block k has
 * its own neighbour across face d (partitioner.cpp:47-52 order). */
pa_status   pa_gpu_tracer_set_neighbor_face_masks(pa_gpu_tracer* tracer, const int32_t face_mask[6]);

/* compute_round_info (pa/source/stages/particle_tracer.cpp:210-234).
 * Rejects an empty set or a non-positive maximum (undefined in the
 * reference). */
pa_status   pa_gpu_tracer_compute_round_info(pa_gpu_tracer* tracer, const pa_particle* d_particles, size_t n,
                                             size_t integral_curve_offset, pa_round_info* out);
/* allocate (particle_tracer.cpp:235-242): fill count vertices with
 * invalid_vertex (-1,-1,-1,0). */
pa_status   pa_gpu_fill_invalid(pa_gpu_ctx* ctx, float* d_vertices, size_t count);
/* initialize + trace (particle_tracer.cpp:243-337): one trace round.
 * d_chunks[c] (host array of integral_curve_count device pointers) holds the
 * round's chunk c with the size given by info; slots the reference leaves as
 * invalid_vertex are not written, so callers that need the reference's
 * pre-prune contents fill them first with pa_gpu_fill_invalid.
 * d_vertex_counts[p] (optional) receives the slots written for particle p
 * (seed through termination vertex); handoff (optional) receives the two
 * hand-off maps; total_steps (optional) the number of integrator steps. */
pa_status   pa_gpu_tracer_trace(pa_gpu_tracer* tracer, const pa_particle* d_particles, size_t n,
                                const pa_round_info* info, float* const* d_chunks, int32_t* d_vertex_counts,
                                pa_handoff* handoff, uint64_t* total_steps);

/* Union-of-7 trace (no reference counterpart as one call; equal to the
 * reference's round loop restricted to the own block and its six preloaded
 * neighbours, pars/source/pipeline.cpp:153-209): a particle handed to a
 * preloaded neighbour continues there at once, and one handed back from a
 * neighbour continues in the own block.  Each seed owns region_slots
 * consecutive slots of d_vertices (n * region_slots vec4); its segments are
 * written one after another, each [start, stepped..., termination].
 * d_vertex_counts[p] = slots used by seed p.  A seed whose segments would
 * overflow its region is stopped at the segment boundary and written to
 * d_continuation (capacity n; vector_field_index = block, original_voxel[0]
 * = seed index), *n_continuation counts them; tracing them again with this
 * function continues them exactly.  region_slots must be >= the maximum
 * remaining_iterations.  d_endpoints (optional, n vec4) receives each seed's
 * last vertex before termination in this call, d_seed_steps (optional) its
 * integrator steps in this call. */
pa_status   pa_gpu_tracer_trace_union(pa_gpu_tracer* tracer, const pa_particle* d_particles, size_t n,
                                      size_t region_slots, float* d_vertices, int32_t* d_vertex_counts,
                                      pa_particle* d_continuation, size_t* n_continuation, float* d_endpoints,
                                      int32_t* d_seed_steps, uint64_t* total_steps);

/* Extension: order of the per-seed slots of the union trace.
 * PA_SLOTS_SEED_MAJOR: slot s of seed p at d_vertices[p * region_slots + s]
 * (the layout above).  PA_SLOTS_STEP_MAJOR: at d_vertices[s * n + p]; the
 * lanes of a wave step in lockstep, so every vertex store is one contiguous
 * run, which is what the memory system rewards.  Pack either into the
 * reference's pruned layout with pa_gpu_prune_counted /
 * pa_gpu_prune_counted_step_major / pa_gpu_prune_counted_chunked /
 * pa_gpu_prune_counted_group8.  Step-major needs PA_ENGINE_LANE, chunked
 * PA_ENGINE_QUAD, grouped PA_ENGINE_HYBRID. */
typedef enum pa_slot_order {
  PA_SLOTS_SEED_MAJOR = 0,
  PA_SLOTS_STEP_MAJOR = 1,
  /* seed-major 8-slot chunks, component-planar: slot s of seed p is float
   * (p * region_slots + s - s % 8) * 4 + 8 * c + s % 8 for component c;
   * region_slots a multiple of 8; written by PA_ENGINE_QUAD only. */
  PA_SLOTS_CHUNKED = 2,
  /* 8-seed groups, step-major inside a group: slot s of seed p is vec4
   * ((p / 8) * region_slots + s) * 8 + p % 8 (allocate ceil(n / 8) * 8 *
   * region_slots vec4); written by PA_ENGINE_HYBRID only, with
   * region_slots below 2^21 (PA_E_INVALID otherwise). */
  PA_SLOTS_GROUP8 = 3
} pa_slot_order;
pa_status   pa_gpu_tracer_trace_union_layout(pa_gpu_tracer* tracer, const pa_particle* d_particles, size_t n,
                                             size_t region_slots, pa_slot_order order, float* d_vertices,
                                             int32_t* d_vertex_counts, pa_particle* d_continuation,
                                             size_t* n_continuation, float* d_endpoints, int32_t* d_seed_steps,
                                             uint64_t* total_steps);
/* Extension: one whole round of the union trace, fused: trace_union on the
 * hybrid engine into grouped d_slots (PA_SLOTS_GROUP8, ceil(n / 8) * 8 *
 * region_slots vec4, scratch for the caller), then prune + index_generator +
 * color_generator of those slots.  d_vertices / d_indices / d_colors receive
 * exactly what pa_gpu_prune_counted_group8, pa_gpu_index_generate and
 * pa_gpu_color_generate would produce on them (colours of vertices that no
 * segment colours are left untouched).  Replaces the stage sequence
 * particle_tracer::trace (pars/source/pipeline.cpp:186-189) -> prune
 * (:213-220) -> index_generator (:224-234) -> color_generator (:238-260)
 * for one rank's share traced over its 7 preloaded blocks.  The particles
 * are checked on the device: a field index outside -1..5 (PA_E_INVALID), a
 * block that is not loaded (PA_E_STATE), remaining_iterations above
 * region_slots (PA_E_CAPACITY) or negative (PA_E_INVALID) fail the call;
 * region_slots is below 2^21.  The call reads back once, at its end. */
pa_status   pa_gpu_tracer_trace_union_round(pa_gpu_tracer* tracer, const pa_particle* d_particles, size_t n,
                                            size_t region_slots, float* d_slots, int32_t* d_vertex_counts,
                                            pa_particle* d_continuation, size_t* n_continuation,
                                            pa_color_mode mode, float parameter, float* d_vertices,
                                            size_t vertex_capacity, size_t* n_vertices, int32_t* d_indices,
                                            size_t index_capacity, size_t* n_indices, float* d_colors,
                                            float* d_endpoints, int32_t* d_seed_steps, uint64_t* total_steps);
/* No counterpart in the reference: the fused round above for callers that
 * cannot allocate device memory while they run (code under stream capture,
 * tight serving loops).  Same arguments and results as
 * pa_gpu_tracer_trace_union_round; its workspace (index pairs, the
 * finish's offsets and scan partials) lives in a fixed 16 MiB device array
 * that each device holds for the life of the process, so a call allocates
 * nothing.  At most ~590,000 seeds per call (PA_E_CAPACITY beyond, with the
 * limit in the message).  Calls that overlap on one device share that array
 * and race: run them on one stream.  Kernel timing (pa_gpu_timing_enable)
 * creates its events on demand; leave it off in such callers. */
pa_status   pa_gpu_tracer_trace_union_round_noalloc(pa_gpu_tracer* tracer, const pa_particle* d_particles, size_t n,
                                                    size_t region_slots, float* d_slots, int32_t* d_vertex_counts,
                                                    pa_particle* d_continuation, size_t* n_continuation,
                                                    pa_color_mode mode, float parameter, float* d_vertices,
                                                    size_t vertex_capacity, size_t* n_vertices, int32_t* d_indices,
                                                    size_t index_capacity, size_t* n_indices, float* d_colors,
                                                    float* d_endpoints, int32_t* d_seed_steps,
                                                    uint64_t* total_steps);
/* The fused round's result words, written on the device by the
 * asynchronous form below (48 bytes). */
typedef struct pa_round_result {
  uint64_t total_steps;
  uint32_t n_continuation;
  uint32_t flags;         /* particle checks and finish waits; see pa_gpu_round_result_check */
  uint64_t reserved_[2];
  int64_t  n_vertices;    /* nothing was written when above vertex_capacity */
  int64_t  n_indices;     /* nothing was written when above index_capacity */
} pa_round_result;
/* Extension: pa_gpu_tracer_trace_union_round without its read-back, for
 * callers that keep several rounds in flight or capture them in a graph
 * (pa_gpu_graph_begin): the counts land in d_result on the device and the
 * call returns once the work is enqueued.  pa_gpu_round_result_check turns
 * the result words (copied to the host) into the status the synchronous call
 * returns.  Under capture the context's workspace must already be large
 * enough (run the round once uncaptured first): PA_E_STATE otherwise. */
pa_status   pa_gpu_tracer_trace_union_round_async(pa_gpu_tracer* tracer, const pa_particle* d_particles, size_t n,
                                                  size_t region_slots, float* d_slots, int32_t* d_vertex_counts,
                                                  pa_particle* d_continuation, pa_color_mode mode, float parameter,
                                                  float* d_vertices, size_t vertex_capacity, int32_t* d_indices,
                                                  size_t index_capacity, float* d_colors, float* d_endpoints,
                                                  int32_t* d_seed_steps, pa_round_result* d_result);
pa_status   pa_gpu_round_result_check(pa_gpu_ctx* ctx, const pa_round_result* h_result, size_t vertex_capacity,
                                      size_t index_capacity);
/* Extension: the asynchronous round for a loop of rounds
 * (pars/source/pipeline.cpp:153-209 traces round after round and uses the
 * curves once the loop ends).  The trace runs on the context's stream; the
 * finish (pack, index, colour, result words) runs on a second stream of
 * the context once the trace is done, so the next round's trace, queued
 * right after, runs while this round finishes.  Successive calls alternate
 * two internal workspaces: a call waits (on the device) for the finish of
 * the call before last.  The caller alternates two sets of round buffers
 * likewise (slots, counts, continuation, outputs, endpoints, result): the
 * next call must not touch this call's.  Work queued on the context later
 * is ordered after this round's trace but not its finish: pa_gpu_round_join
 * orders the context's stream after every queued finish, and
 * pa_gpu_synchronize waits for them.  Not while capturing (PA_E_STATE). */
pa_status   pa_gpu_tracer_trace_union_round_pipelined(pa_gpu_tracer* tracer, const pa_particle* d_particles,
                                                      size_t n, size_t region_slots, float* d_slots,
                                                      int32_t* d_vertex_counts, pa_particle* d_continuation,
                                                      pa_color_mode mode, float parameter, float* d_vertices,
                                                      size_t vertex_capacity, int32_t* d_indices,
                                                      size_t index_capacity, float* d_colors, float* d_endpoints,
                                                      int32_t* d_seed_steps, pa_round_result* d_result);
pa_status   pa_gpu_round_join(pa_gpu_ctx* ctx);
/* Counted prune of step-major slots (row stride n seeds) into the pruned
 * particle-major layout; same outputs as pa_gpu_prune_counted. */
pa_status   pa_gpu_prune_counted_step_major(pa_gpu_ctx* ctx, const float* d_slots, size_t n,
                                            const int32_t* d_vertex_counts, float* d_out, size_t capacity,
                                            size_t* n_out, int64_t* d_offsets);
/* Counted prune of 8-seed grouped slots (PA_SLOTS_GROUP8) into the pruned
 * layout; region_slots below 2^24 (PA_E_INVALID otherwise). */
pa_status   pa_gpu_prune_counted_group8(pa_gpu_ctx* ctx, const float* d_slots, size_t region_slots,
                                        const int32_t* d_vertex_counts, size_t n, float* d_out, size_t capacity,
                                        size_t* n_out, int64_t* d_offsets);
/* pa_gpu_prune_counted, pa_gpu_index_generate and pa_gpu_color_generate of
 * one chunk in one pass over the slots (pipeline.cpp:186-260 runs
 * particle_tracer's prune, index_generator.cpp:10-24 and
 * color_generator.cpp:16-98 in turn on each chunk): the same vertices,
 * indices and colours at the positions color_generate writes.  Vertices at
 * most capacity and below 2^31 (int32 indices); d_offsets optional (n + 1
 * int64).  Curves of the shape the tracers write (count >= 2, one
 * termination vertex, last) are indexed and coloured in the same pass, and
 * the other colour positions keep what d_colors held; a chunk with any
 * other curve is indexed and coloured by the general kernels afterwards,
 * and there the positions that color_generate does not write are
 * undefined. */
pa_status   pa_gpu_finish_counted(pa_gpu_ctx* ctx, const float* d_slots, size_t stride_slots,
                                  const int32_t* d_vertex_counts, size_t n, float* d_vertices, size_t capacity,
                                  size_t* n_vertices, int64_t* d_offsets, int32_t* d_indices, size_t* n_indices,
                                  pa_color_mode mode, float parameter, float* d_colors);
/* Counted prune of chunked slots (PA_SLOTS_CHUNKED) into the pruned layout. */
pa_status   pa_gpu_prune_counted_chunked(pa_gpu_ctx* ctx, const float* d_slots, size_t region_slots,
                                         const int32_t* d_vertex_counts, size_t n, float* d_out, size_t capacity,
                                         size_t* n_out, int64_t* d_offsets);

/* Extension: particle pool over up to 16 resident fields.  It serves the
 * batched form (many seed sets over fields of uneven size in one persistent
 * work-queue kernel) and is the step engine of the serving loop (bounded
 * passes, requests admitted between passes).  No reference counterpart:
 * pars serves one request at a time (pars/source/service.cpp:20-69).  Each
 * particle follows the reference round (particle_tracer.cpp:265-326) on
 * its own field alone (field index = its vector_field_index; no
 * neighbours, stop on leaving the field): the endpoint and step count equal
 * a one-block trace's last vertex and step count; vertices are not kept.
 * A particle's result does not depend on the rest of the pool. */
enum { PA_POOL_ACTIVE = 0, PA_POOL_DONE = 1, PA_POOL_INVALID = 2 };
typedef struct pa_pool {
  float*   position;     /* n vec4: seed position in, endpoint out */
  int32_t* iteration;    /* the reference loop counter it (starts at 1) */
  int32_t* remaining;    /* remaining_iterations R */
  int32_t* steps;        /* integrator steps taken so far */
  int32_t* status;       /* PA_POOL_* */
  int32_t* field;        /* field index */
  int32_t* request;      /* optional: request id (the particle's original_rank) */
  int32_t* request_done; /* optional: finished particles per request id (zeroed by the caller; a
                          * run adds its counts as its workgroups end, all of them by its end) */
  int32_t  n_requests;   /* entries of request_done; request ids are in [0, n_requests) */
} pa_pool;
/* Fill pool entries [0, n) from particles; entries whose field index is not
 * in [0, n_fields), or whose request id (original_rank) is not in
 * [0, n_requests) when request is given, become PA_POOL_INVALID and the call
 * returns PA_E_INVALID after initialising the rest.  request needs
 * request_done and n_requests >= 1. */
pa_status   pa_gpu_pool_init(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n, int32_t n_fields,
                             const pa_pool* pool);
/* Advance every active entry of [0, n) by at most max_iterations reference
 * iterations (0 = to the end) with the tracer's integrator, step size, stop
 * rule and interpolation; waves_per_cu = waves of the persistent kernel
 * per CU (0 = as many as one CU keeps resident, 12 to 20 by variant).  All fields share one storage and layout.  An
 * active entry whose field index is n_fields or more becomes
 * PA_POOL_INVALID without being stepped and the call returns PA_E_INVALID
 * once the others have run. */
pa_status   pa_gpu_pool_run(pa_gpu_tracer* tracer, const pa_gpu_field* const* fields, int32_t n_fields,
                            const pa_pool* pool, size_t n, int32_t max_iterations, int32_t waves_per_cu,
                            uint64_t* total_steps);
/* pa_gpu_pool_run that also keeps each entry's curve: entry i writes the
 * slots of its one-block reference round (particle_tracer.cpp:243-326) to
 * d_slots (vec4) from slot d_slot_base[i] on: the seed when it is first
 * stepped, each stepped vertex, and the termination vertex (none for
 * remaining_iterations < 2), and on its end its vertex count (seed + steps
 * + termination) to d_vertex_counts[i].  Slots past the count are not
 * written.  With a seed set's entries consecutive and its bases i * R, the
 * set's slots are the reference's chunk layout for that set, ready for
 * pa_gpu_prune_counted / index / colour once the set has ended (its request
 * counter).  An entry whose slots [base, base + max(R, 1)) do not lie in
 * [0, slot_capacity) becomes PA_POOL_INVALID like a bad field index.
 * Bounded passes (max_iterations) keep writing where they left off. */
pa_status   pa_gpu_pool_run_vertices(pa_gpu_tracer* tracer, const pa_gpu_field* const* fields, int32_t n_fields,
                                     const pa_pool* pool, size_t n, int32_t max_iterations, int32_t waves_per_cu,
                                     const int64_t* d_slot_base, float* d_slots, size_t slot_capacity,
                                     int32_t* d_vertex_counts, uint64_t* total_steps);

/* Asynchronous forms of pa_gpu_pool_init and pa_gpu_pool_run (capturable):
 * the rejected-entry count (uint32) and the (steps, rejected) pair (2 x
 * uint64) land on the device; with zero_request_counters the request
 * counters are cleared first (a kernel, not a memset). */
pa_status   pa_gpu_pool_init_async(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n, int32_t n_fields,
                                   const pa_pool* pool, int32_t zero_request_counters, uint32_t* d_rejected);
pa_status   pa_gpu_pool_run_async(pa_gpu_tracer* tracer, const pa_gpu_field* const* fields, int32_t n_fields,
                                  const pa_pool* pool, size_t n, int32_t max_iterations, int32_t waves_per_cu,
                                  uint64_t* d_result);

/* Extension: device-resident serving loop.  No reference counterpart: pars
 * serves one request at a time (pars/source/service.cpp:20-69).  A
 * persistent kernel on the server's own stream takes the particles of every
 * submitted request as soon as they are published and runs each to its end
 * on its field (vector_field_index; the pool's one-block reference round),
 * so requests of any size overlap on the device with no host round trip per
 * pass.  One wave of the grid reads the host's control words and mirrors
 * them into device memory; workers claim ring entries from the device
 * cursor, read the records from the pinned ring and leave endpoints and
 * step counts in device memory; the request's last particle queues it, and
 * exporter waves copy its results to pinned host memory and raise its
 * completion flag.  Ring space and request slots are admitted by capacity
 * (PA_E_CAPACITY when full); a request's results equal its trace alone.
 * While the kernel runs the host makes no GPU call (submit and collect
 * only touch pinned memory); the kernel stays resident, idle waves polling
 * device memory every ~16 us, until pa_gpu_serve_stop, and submit or poll
 * launch it when needed.  Other work of the process on the same device may
 * share its hardware queue: stop the server before it.  particle_capacity
 * 64 .. 2^24 (68 B of pinned host memory and 20 B of device memory each);
 * waves_per_cu 0 = 12; idle_polls is reserved (0). */
pa_status   pa_gpu_serve_create(pa_gpu_tracer* tracer, const pa_gpu_field* const* fields, int32_t n_fields,
                                size_t particle_capacity, int32_t request_slots, int32_t waves_per_cu,
                                int32_t idle_polls, pa_gpu_server** out);
/* Publish a request (host particles, copied): *request = its slot. */
pa_status   pa_gpu_serve_submit(pa_gpu_server* server, const pa_particle* h_particles, size_t n, int32_t* request);
/* Non-blocking: *done = 1 once every particle of the request has ended. */
pa_status   pa_gpu_serve_poll(pa_gpu_server* server, int32_t request, int32_t* done);
/* A done request's endpoints (n vec4) and step counts (either optional),
 * then its slot and ring space are free again. */
pa_status   pa_gpu_serve_collect(pa_gpu_server* server, int32_t request, float* h_endpoints, int32_t* h_steps,
                                 uint64_t* steps_sum);
/* Drain what is published, wait for the kernel to leave; total steps served
 * and kernel launches so far. */
pa_status   pa_gpu_serve_stop(pa_gpu_server* server, uint64_t* total_steps, uint64_t* launches);
pa_status   pa_gpu_serve_destroy(pa_gpu_server* server);

/* Extension: graph-captured rounds.  The asynchronous calls made on ctx
 * between pa_gpu_graph_begin and pa_gpu_graph_end (seeds, the asynchronous
 * round and pool calls, hand-over kernels) are recorded once and replayed by
 * pa_gpu_graph_launch on ctx's stream: one submission per round, the same
 * buffers and arguments every replay.  A call that reads back during
 * capture makes pa_gpu_graph_end fail with PA_E_STATE.  Kernel timing must
 * be off while capturing.  The round loop it serves is pars' per-request
 * pipeline (pars/source/pipeline.cpp:153-260). */
pa_status   pa_gpu_graph_begin(pa_gpu_ctx* ctx);
pa_status   pa_gpu_graph_end(pa_gpu_ctx* ctx, pa_gpu_graph** out);
pa_status   pa_gpu_graph_launch(pa_gpu_ctx* ctx, const pa_gpu_graph* graph);
pa_status   pa_gpu_graph_destroy(pa_gpu_graph* graph);

/* prune (particle_tracer.cpp:402-409): order-preserving removal of
 * invalid_vertex from one chunk.  *n_out receives the kept count. */
pa_status   pa_gpu_prune(pa_gpu_ctx* ctx, const float* d_vertices, size_t n, float* d_out, size_t* n_out);
/* prune of a traced round without the invalid fill: particle p's
 * d_vertex_counts[p] slots, starting at slot p*stride_slots of the
 * concatenated d_slots, are packed in particle order into d_out.
 * d_offsets (optional, n+1) receives the exclusive scan. */
pa_status   pa_gpu_prune_counted(pa_gpu_ctx* ctx, const float* d_slots, size_t stride_slots,
                                 const int32_t* d_vertex_counts, size_t n, float* d_out, size_t capacity,
                                 size_t* n_out, int64_t* d_offsets);

/* pa::index_generator::generate (pa/source/stages/index_generator.cpp:10-24):
 * index k is emitted iff neither vertex k nor k+1 is the termination vertex;
 * d_indices needs capacity n.  Rejects n == 0 (size()-1 underflows in the
 * reference). */
pa_status   pa_gpu_index_generate(pa_gpu_ctx* ctx, const float* d_vertices, size_t n, int32_t* d_indices,
                                  size_t* n_out);

/* pa::color_generator::generate (pa/source/stages/color_generator.cpp:16-40).
 * d_colors has n_vertices vec4; colors the reference leaves uninitialised are
 * left untouched here (callers zero them for comparisons). */
pa_status   pa_gpu_color_generate(pa_gpu_ctx* ctx, const float* d_vertices, size_t n_vertices,
                                  const int32_t* d_indices, size_t n_indices, pa_color_mode mode, float parameter,
                                  float* d_colors);

/* ------------------------------------------------------------- hand-overs */
/* One chunk of a round's output, as pars hands it to the renderer:
 * ray_tracer::set_integral_curves (pars/source/stages/ray_tracer.cpp:156-212)
 * shares each integral_curves chunk with OSPRay as vertex (OSP_FLOAT3A:
 * vec4 FP32, w = 0), vertex.color (OSP_FLOAT4) and index (OSP_INT) buffers.
 * The pointers are device (HBM) pointers into the round's outputs. */
typedef struct pa_curve_chunk {
  float*   vertices;       /* n_vertices vec4 */
  float*   colors;         /* n_vertices vec4 */
  int32_t* indices;        /* n_indices, relative to vertices */
  size_t   n_vertices;
  size_t   n_indices;
  size_t   first_particle; /* seeds [first_particle, first_particle + n_particles) */
  size_t   n_particles;
} pa_curve_chunk;
/* Split a fused round's outputs (pa_gpu_tracer_trace_union_round) into the
 * reference's chunks: particle_tracer::compute_round_info
 * (pa/source/stages/particle_tracer.cpp:210-222) gives each chunk
 * int32_max / sizeof(vector4) / R particles so that no chunk exceeds
 * 134,217,727 vertices and int32 indices stay valid; here R is region_slots
 * (a seed's segments never exceed it) and max_vertices_per_chunk (0 = the
 * reference's 134,217,727) may lower the cap.  Indices are rewritten in
 * place relative to their chunk, and each chunk's index list is what
 * index_generator::generate (index_generator.cpp:10-24) gives on that chunk
 * alone: the pair joining one chunk's last vertex to the next chunk's first
 * belongs to neither.  Colours at the positions color_generator writes per
 * chunk are unchanged.  Call once per round; *n_chunks is set even when
 * max_chunks is too small (PA_E_CAPACITY). */
pa_status   pa_gpu_round_chunks(pa_gpu_ctx* ctx, const int32_t* d_vertex_counts, size_t n, size_t region_slots,
                                size_t max_vertices_per_chunk, float* d_vertices, size_t n_vertices,
                                int32_t* d_indices, size_t n_indices, float* d_colors, pa_curve_chunk* chunks,
                                size_t max_chunks, size_t* n_chunks);
/* Stand-in for the streamline ray tracer (OSPRay, out of scope): splats
 * n vertices with their colours into a width x height RGBA8 image
 * (d_rgba, 4 bytes per pixel, row 0 at the bottom, sRGB-encoded like
 * OSPRay's OSP_FB_SRGBA framebuffer that ray_tracer::serialize sends,
 * pars/source/stages/ray_tracer.cpp:240-253) through a pinhole camera at
 * position looking along forward with up, fovy 68 degrees and aspect
 * width / height (ray_tracer.cpp:63, 215-227); one pixel per vertex, the
 * nearest wins; sentinel vertices are skipped; d_colors may be NULL
 * (white).  Asynchronous. */
pa_status   pa_gpu_render_points(pa_gpu_ctx* ctx, const float* d_vertices, const float* d_colors, size_t n,
                                 const float position[3], const float forward[3], const float up[3], int32_t width,
                                 int32_t height, uint8_t* d_rgba);
/* Particles as the records boost::mpi sends: the field sequence of
 * particle::serialize (pa/include/pa/math/particle.hpp:16-28), position[0..3],
 * remaining_iterations, vector_field_index, original_rank,
 * original_voxel[0..2], 4 bytes each, 40 bytes per particle, packed.  The
 * hand-off lists of pa_gpu_tracer_trace and the surplus lists below are
 * turned into them on the device so the host's MPI sends straight from a
 * device (or pinned) buffer. */
pa_status   pa_gpu_particles_to_wire(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n, void* d_wire);
pa_status   pa_gpu_particles_from_wire(pa_gpu_ctx* ctx, const void* d_wire, size_t n, pa_particle* d_particles);
/* The particle side of particle_tracer::load_balance_distribute
 * (pa/source/stages/particle_tracer.cpp:168-184); the host keeps the count
 * exchange and the averages (:63-166).  For neighbour slot k = 0..5 in
 * order, the last surplus_counts[k] particles of the list as it stands
 * (none if that is more than it holds) are written, in order, as serialize
 * records to d_surplus_wire + surplus_offsets[k] records with
 * vector_field_index = k ^ 1 (the receiver's slot for this process); the
 * list keeps its first *n_remaining particles. */
pa_status   pa_gpu_load_balance_take(pa_gpu_ctx* ctx, const pa_particle* d_particles, size_t n,
                                     const size_t surplus_counts[6], void* d_surplus_wire,
                                     size_t surplus_offsets[7], size_t* n_remaining);

/* ----------------------------------------------------------- advector / FTLE */
/* pa::particle_advector::advect, one round in place over one block
 * (pa/source/stages/particle_advector.cpp:61-141).  d_state[p]: 1 =
 * inactive (returned to original_rank), 2 = handed to neighbour slot
 * d_handoff_slot[p], 3 = never stepped (remaining_iterations <= 1; the
 * reference drops it).  Uses the tracer's local field, neighbour ranks,
 * integrator, step size and stop rule. */
pa_status   pa_gpu_advect(pa_gpu_tracer* tracer, pa_particle* d_particles, size_t n, int32_t* d_state,
                          int32_t* d_handoff_slot, uint64_t* total_steps);

/* pa::flow_map_generator::generate (pa/source/stages/flow_map_generator.cpp:
 * 12-52) for a block without neighbours: dims = block_size * scale,
 * spacing / scale, one seed per voxel, advected in place, final xyz written
 * to d_flow_map[(x*d1+y)*d2+z][3] (voxels of lost particles keep the seed
 * position).  Seeds are generated inside the kernel; no particle array is
 * materialised. */
pa_status   pa_gpu_flow_map_generate(pa_gpu_tracer* tracer, const int32_t block_size[3], uint64_t iterations,
                                     float resolution_scale, int32_t rank, float* d_flow_map, int32_t dims_out[3],
                                     uint64_t* total_steps);

/* pa::ftle_map_generator::generate (pa/source/stages/ftle_map_generator.cpp:
 * 22-60) in textbook form, all in FP64 on the device: flow-map Jacobian by
 * central differences (one-sided at the lattice border), C = J^T J, symmetric
 * 3x3 eigenvalues, then ln(sqrt(lambda_max))/|T| or fractional anisotropy,
 * stored FP32.  d_lambda_max (optional) receives lambda_max in FP64. */
pa_status   pa_gpu_ftle_generate(pa_gpu_ctx* ctx, const float* d_flow_map, const int32_t dims[3],
                                 const float spacing[3], double time, pa_ftle_mode mode, float* d_ftle,
                                 double* d_lambda_max);

/* The FP64 symmetric 3x3 eigen solve of pa_gpu_ftle_generate on its own
 * (the role of Eigen's SelfAdjointEigenSolver at ftle_map_generator.cpp:
 * 40-48): d_sym holds n matrices as (a00, a01, a02, a11, a12, a22), d_eigenvalues
 * receives 3 eigenvalues per matrix, ascending.  Smith's trigonometric
 * closed form; absolute error a few 1e-9 of the largest |eigenvalue| near
 * repeated eigenvalues, ~1e-15 elsewhere. */
pa_status   pa_gpu_sym3_eigenvalues(pa_gpu_ctx* ctx, const double* d_sym, size_t n, double* d_eigenvalues);

#ifdef __cplusplus
}
#endif

#endif /* PA_GPU_H */
