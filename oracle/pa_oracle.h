/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of pars' per-process advection path (acdemiralp/pars, `pa`
 * library), used as the checker for the HIP product and as the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load liboracle.so.  The product (pars_amd/libpa_gpu.so) never links,
 * loads or calls anything here.
 *
 * Every function cites the reference file:line it restates.  The FP32 entry
 * points reproduce the reference's single-precision operation order exactly
 * (built with -ffp-contract=off, IEEE division, no fast-math); the *_f64
 * entry points are the same algorithms in double precision and serve as the
 * FP64 error oracle.
 *
 * Interpolation weights.  The reference takes the cell from the rounded
 * quotient floor((p-o)/h) but the weight from the exact fmod(p-o, h)/h
 * (vector_field.cpp:29-30).  When the quotient rounds up to an integer k
 * while the true quotient is just below k (common for seeds placed on node
 * planes), it reads cell k with weight ~1, i.e. one cell too far.  The FP32
 * path reproduces this by default (interp = ORC_INTERP_REFERENCE);
 * ORC_INTERP_CONSISTENT uses w = q - floor(q) (a build extension).  The FP64
 * twin always uses the consistent weight: it is the error oracle for the
 * intended trilinear scheme.
 *
 * Pinning status (what checks each function; see tests/test_oracle.py):
 *   seed lattice            brute force on small lattices, closed form
 *   contains / interpolate  constant field, node values, affine fields
 *   trace (stepper A)       constant-field 5-term sum, affine-field Euler
 *                           closed form, brute force round bookkeeping
 *   odeint final-sum order  parity unpinned (Boost.Odeint 1.69 is not in
 *                           this container; no reference test pins it)
 *   prune / index           brute force, uniform-lifetime closed form
 *   color                   axis-aligned closed forms, HSV/HSL identities
 *   advect / flow map       constant / affine fields, exit freeze
 *   eigen / FTLE (FP64)     numpy.linalg.eigh fixtures, saddle closed form,
 *                           rotation gives zero
 */
#ifndef PA_ORACLE_H
#define PA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pa::particle (pa/include/pa/math/particle.hpp:30-35): 48 B, the Eigen
 * Vector4f member gives the struct 16-B alignment. */
typedef struct orc_particle {
  float   position[4];
  int32_t remaining_iterations;
  int32_t vector_field_index;
  int32_t original_rank;
  int32_t original_voxel[3];
  int32_t reserved_[2];
} orc_particle;

/* pa::vector_field (pa/include/pa/math/vector_field.hpp:14-24) with the
 * boost::multi_array<vector3,3> flattened: data[(x*d1 + y)*d2 + z][3]. */
typedef struct orc_field {
  const float* data;
  int32_t      dims[3];
  float        offset[3];
  float        size[3];
  float        spacing[3];
} orc_field;

typedef struct orc_field64 {
  const double* data;
  int32_t       dims[3];
  double        offset[3];
  double        size[3];
  double        spacing[3];
} orc_field64;

/* Integrators, in the order of pa::variant_integrator
 * (pa/include/pa/math/integrators.hpp:25-33).  Only euler and
 * runge_kutta_4 are restated; ORC_CLASSIC_RK4 is the textbook four-gather
 * RK4 (a build extension without a reference counterpart).
 * ORC_RK4_DT_OVER_6 is not a reference integrator: it is the frozen step
 * with the coefficients Boost.Odeint's runge_kutta4_classic would use
 * (dt/6, dt/3, dt/3, dt/6 rounded once each) instead of runge_kutta4's
 * (1/6)*dt etc., used only by the tests that measure how much the unpinned
 * rounding of the final stage could move a trajectory. */
enum { ORC_EULER = 0, ORC_RK4 = 2, ORC_CLASSIC_RK4 = 100, ORC_RK4_DT_OVER_6 = 102 };

/* Stop rules: the reference's Eigen isZero() test (every |v_i| <= 1e-5f,
 * particle_tracer.cpp:312) or the extension |v| < eps. */
enum { ORC_STOP_ISZERO = 0, ORC_STOP_NORM = 1 };

/* Interpolation weights (see the note at the top). */
enum { ORC_INTERP_REFERENCE = 0, ORC_INTERP_CONSISTENT = 1 };

typedef struct orc_trace_config {
  const orc_field* local;          /* own block (vector_field_index == -1) */
  const orc_field* neighbors[6];   /* -x,+x,-y,+y,-z,+z; NULL = no neighbour */
  int32_t          neighbor_rank[6];
  /* Union-of-7 mode only: bit d of neighbor_face_mask[k] is set when block k
   * has a neighbour across face d (partitioner.cpp:47-52 order). */
  int32_t          neighbor_face_mask[6];
  float            step_size;
  int32_t          integrator;
  int32_t          stop_rule;
  float            stop_eps;
  int32_t          interp;
} orc_trace_config;

typedef struct orc_trace_config64 {
  const orc_field64* local;
  const orc_field64* neighbors[6];
  int32_t            neighbor_rank[6];
  int32_t            neighbor_face_mask[6];
  double             step_size;
  int32_t            integrator;
  int32_t            stop_rule;
  double             stop_eps;
  int32_t            interp;          /* ignored: the FP64 twin is consistent */
} orc_trace_config64;

/* One hand-off record: particle_tracer.cpp:282-306.  kind 0 =
 * out_of_bounds_particles[neighbour slot], kind 1 =
 * neighbor_out_of_bounds_particles[vector_field_index]. */
typedef struct orc_handoff {
  orc_particle particle;
  int32_t      kind;
  int32_t      slot;
  int64_t      source;
} orc_handoff;

void     orc_set_threads(int n);
int      orc_get_threads(void);
/* std::thread::hardware_concurrency() of this host (for reports). */
int      orc_hardware_concurrency(void);

/* seed_generator.cpp:9-21 */
int64_t  orc_seed_count(const float size[3], const float stride[3], int32_t dims_out[3]);
void     orc_seed_generate(const float offset[3], const float size[3], const float stride[3],
                           int32_t remaining_iterations, int32_t rank, orc_particle* out);

/* vector_field.cpp:12-48 */
int      orc_contains(const orc_field* f, const float position[4]);
void     orc_interpolate(const orc_field* f, const float position[4], float out[3]);
void     orc_interpolate_consistent(const orc_field* f, const float position[4], float out[3]);
int      orc_contains_f64(const orc_field64* f, const double position[4]);
void     orc_interpolate_f64(const orc_field64* f, const double position[4], double out[3]);

/* One step of the frozen-derivative stepper (particle_tracer.cpp:311-324). */
void     orc_step(int32_t integrator, const float position[4], const float vec[3], float dt, float out[4]);

/* particle_tracer.cpp:210-234.  out = {maximum_remaining_iterations,
 * vertices_per_integral_curve, vertices_per_integral_curve_last,
 * integral_curve_offset, integral_curve_count}.  Returns 0, or -1 for an
 * empty particle set (the reference dereferences max_element of an empty
 * range there). */
int      orc_round_info(const orc_particle* particles, size_t n, uint64_t curve_offset, uint64_t out[5]);

/* particle_tracer.cpp:235-242: fill with invalid_vertex. */
void     orc_fill_invalid(float* vertices, size_t n);

/* particle_tracer.cpp:243-252 (initialize) + 253-337 (trace).  chunks[c] is
 * chunk integral_curve_offset + c of this round, already filled with
 * invalid_vertex.  vertex_counts[p] (optional) = slots written for particle p
 * (seed .. termination).  handoffs (capacity n) receive the hand-off lists
 * sorted by (kind, slot, source).  Returns 0 or a negative error. */
int      orc_trace(const orc_trace_config* cfg, const orc_particle* particles, size_t n,
                   const uint64_t info[5], float* const* chunks, int32_t* vertex_counts,
                   orc_handoff* handoffs, size_t* n_handoffs, uint64_t* total_steps);

/* Union-of-7 emulation: the reference's round loop restricted to the own
 * block and its six face neighbours, each particle continued in the block it
 * is handed to.  Segments are emitted per seed in segment order; each
 * segment is [start, stepped..., termination].  Returns 0, or -2 when a
 * capacity is exceeded. */
int      orc_trace_union(const orc_trace_config* cfg, const orc_particle* particles, size_t n,
                         float* out_vertices, size_t vertex_capacity, size_t* n_vertices,
                         int64_t* segment_seed, int32_t* segment_block, int64_t* segment_begin,
                         int32_t* segment_length, size_t segment_capacity, size_t* n_segments,
                         uint64_t* total_steps);

/* Per-seed trajectory endpoints in FP32 and FP64 (last stepped or start
 * vertex of the final segment), for error reporting.  union_mode selects
 * the union-of-7 continuation. */
int      orc_endpoints(const orc_trace_config* cfg, const orc_particle* particles, size_t n,
                       int union_mode, float* endpoints, int32_t* steps);
int      orc_endpoints_f64(const orc_trace_config64* cfg, const double* positions, size_t n,
                           int32_t remaining_iterations, int union_mode, double* endpoints,
                           int32_t* steps);

/* particle_tracer.cpp:402-409 */
size_t   orc_prune(const float* vertices, size_t n, float* out);
/* index_generator.cpp:10-24.  Returns the index count, or -1 for an empty
 * vertex array (size()-1 underflows in the reference). */
int64_t  orc_index_generate(const float* vertices, size_t n, int32_t* indices);
/* color_generator.cpp:16-98.  colors must be pre-zeroed by the caller (the
 * reference leaves unwritten colors uninitialised).  mode: 0 hsl_constant_s,
 * 1 hsl_constant_l, 2 hsv_constant_s, 3 hsv_constant_v, 4 rgb, 100
 * magnitude (extension). */
void     orc_color_generate(const float* vertices, size_t n_vertices, const int32_t* indices,
                            size_t n_indices, int32_t mode, float parameter, float* colors);

/* particle_advector.cpp:61-141 (one round, in place).  state[p]: 0 = still
 * active (never returned by a single-block round), 1 = inactive (returned
 * to original_rank), 2 = handed to a neighbour slot handoff_slot[p], 3 =
 * lost (remaining_iterations <= 1). */
int      orc_advect(const orc_trace_config* cfg, orc_particle* particles, size_t n,
                    int32_t* state, int32_t* handoff_slot, uint64_t* total_steps);

/* flow_map_generator.cpp:12-52 for one block without neighbours.
 * flow_map[(x*d1+y)*d2+z][3]; dims_out = block_size * scale. Voxels whose
 * particle is lost keep the seed position. */
int      orc_flow_map(const orc_field* f, const int32_t block_size[3], uint64_t iterations,
                      float resolution_scale, int32_t rank, int32_t integrator, float step_size,
                      float* flow_map, int32_t dims_out[3], uint64_t* total_steps);

/* Textbook FTLE in FP64 (the reference's gradient at vector_field.cpp:49-93
 * is defective; the structure follows ftle_map_generator.cpp:22-60).
 * mode 0 = ln(sqrt(lambda_max))/|T|, 1 = fractional anisotropy. */
void     orc_ftle(const float* flow_map, const int32_t dims[3], const float spacing[3],
                  double time, int32_t mode, float* ftle, double* lambda_max_out);

/* Symmetric 3x3 eigenvalues in FP64 (cyclic Jacobi), ascending.
 * a = {a00, a01, a02, a11, a12, a22}. */
void     orc_eigh3(const double a[6], double eigenvalues[3]);

/* Synthetic fields: ABC (u = (A sin z + C cos y, B sin x + A cos z,
 * C sin y + B cos x)) sampled at global node g*h; random Fourier curl
 * field from explicit modes.  Output [d0][d1][d2][3]. */
void     orc_abc_field_f64(const int32_t dims[3], const int32_t origin[3], double h,
                           double A, double B, double C, double* out);
void     orc_fourier_field_f64(const int32_t dims[3], const int32_t origin[3], double h,
                               const double* modes /* [m][7]: k0 k1 k2 c0 c1 c2 phase */,
                               int32_t n_modes, double* out);

#ifdef __cplusplus
}
#endif

#endif
