// ORACLE — TEST INFRASTRUCTURE ONLY.  CPU restatement of pars' per-process
// advection path; the checker for the HIP product and the timed CPU
// baseline.  Never linked into or loaded by the product.  Build flags
// (oracle/Makefile): -O2 -ffp-contract=off, no fast-math, x86-64 baseline
// ISA, so every float operation is a single IEEE-rounded SSE operation in
// the order written here.
//
// Citations are to acdemiralp/pars.

#include "pa_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Threading: dynamic chunked scheduling, the role tbb::parallel_for with
// grain 1 plays in the reference (particle_tracer.cpp:258).
// ---------------------------------------------------------------------------
int g_threads = 0;

int thread_count() {
  if (g_threads > 0) return g_threads;
  unsigned h = std::thread::hardware_concurrency();
  return h ? int(h) : 1;
}

template <class Fn>
void parallel_for(size_t n, size_t grain, Fn&& fn) {
  const int t = thread_count();
  if (t <= 1 || n <= grain) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const size_t b = next.fetch_add(grain);
      if (b >= n) break;
      const size_t e = std::min(n, b + grain);
      for (size_t i = b; i < e; ++i) fn(i);
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(size_t(t - 1));
  for (int i = 1; i < t; ++i) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Field model (vector_field.hpp:14-24), generic over the scalar so that the
// FP32 restatement and the FP64 twin share one statement of the algorithm.
// ---------------------------------------------------------------------------
template <class T>
struct field {
  const T* data;
  int32_t dims[3];
  T offset[3], size[3], spacing[3];
};

field<float> view(const orc_field* f) {
  field<float> r;
  r.data = f->data;
  for (int i = 0; i < 3; ++i) {
    r.dims[i] = f->dims[i];
    r.offset[i] = f->offset[i];
    r.size[i] = f->size[i];
    r.spacing[i] = f->spacing[i];
  }
  return r;
}
field<double> view(const orc_field64* f) {
  field<double> r;
  r.data = f->data;
  for (int i = 0; i < 3; ++i) {
    r.dims[i] = f->dims[i];
    r.offset[i] = f->offset[i];
    r.size[i] = f->size[i];
    r.spacing[i] = f->spacing[i];
  }
  return r;
}

// vector_field::contains (vector_field.cpp:12-21): subscript = floor((p-o)/h)
// cast to size_t; inside iff subscript < shape-1.  A negative or NaN
// subscript becomes a huge size_t and fails; -0 becomes 0 and passes.
template <class T>
bool contains(const field<T>& f, const T p[4]) {
  for (int i = 0; i < 3; ++i) {
    const T s = std::floor((p[i] - f.offset[i]) / f.spacing[i]);
    if (!(s >= T(0) && s <= T(f.dims[i] - 2))) return false;
  }
  return true;
}

// linear_interpolate.hpp:6-10: (1 - w) * x + w * y, per component.
template <class T>
inline void lerp3(const T* a, const T* b, T w, T* out) {
  const T u = T(1) - w;
  for (int c = 0; c < 3; ++c) out[c] = u * a[c] + w * b[c];
}

// vector_field::interpolate (vector_field.cpp:22-48).  Cell index from
// floor((p-o)/h); weights from fmod(p-o, h)/h; lerp along z, then y, then x.
//
// The cell index comes from the rounded quotient and the weight from the
// exact remainder, so when (p-o)/h rounds up to an integer k while the true
// quotient is just below k, the reference reads cell k with weight ~1, i.e.
// one cell too far.  consistent = false keeps this (the reference);
// consistent = true evaluates the intended interpolant, w = q - floor(q).
template <class T>
void interpolate(const field<T>& f, const T p[4], T out[3], bool consistent) {
  int m[3];
  T w[3];
  for (int i = 0; i < 3; ++i) {
    const T q = (p[i] - f.offset[i]) / f.spacing[i];
    const T fl = std::floor(q);
    m[i] = int(fl);
    w[i] = consistent ? q - fl : std::fmod(p[i] - f.offset[i], f.spacing[i]) / f.spacing[i];
  }
  const int64_t d1 = f.dims[1], d2 = f.dims[2];
  auto at = [&](int x, int y, int z) -> const T* {
    return f.data + ((int64_t(x) * d1 + y) * d2 + z) * 3;
  };
  T c00[3], c01[3], c10[3], c11[3], c0[3], c1[3];
  lerp3(at(m[0], m[1], m[2]), at(m[0], m[1], m[2] + 1), w[2], c00);
  lerp3(at(m[0], m[1] + 1, m[2]), at(m[0], m[1] + 1, m[2] + 1), w[2], c01);
  lerp3(at(m[0] + 1, m[1], m[2]), at(m[0] + 1, m[1], m[2] + 1), w[2], c10);
  lerp3(at(m[0] + 1, m[1] + 1, m[2]), at(m[0] + 1, m[1] + 1, m[2] + 1), w[2], c11);
  lerp3(c00, c01, w[1], c0);
  lerp3(c10, c11, w[1], c1);
  lerp3(c0, c1, w[0], out);
}

// Stop test before the step.  Reference: vector.isZero() on a vector3f,
// i.e. |v_i| <= 1 * 1e-5f for all i (Eigen CwiseNullaryOp.h:481-489,
// NumTraits.h:157, MathFunctions.h:1297-1300).
template <class T>
bool stops(const T v[3], int rule, T eps) {
  if (rule == ORC_STOP_NORM) {
    const T n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    return n2 < eps * eps;
  }
  const T tol = T(1e-5f);
  return std::fabs(v[0]) <= tol && std::fabs(v[1]) <= tol && std::fabs(v[2]) <= tol;
}

// Frozen-derivative step (particle_tracer.cpp:315-324): the odeint system
// returns the same vector for every stage.  Boost.Odeint 1.69 is not
// vendored in the reference (a third-party dependency, Boost 1.69.0 per
// utility/shell/rwth/load_common_modules.zsh:5); its published algorithm:
//  - pa::runge_kutta_4_integrator is odeint::runge_kutta4<Vector4f, float,
//    Vector4f, float, vector_space_algebra> (integrators.hpp:12-17), i.e.
//    explicit_generic_rk<4, 4, ...> with the Butcher tableau of
//    boost/numeric/odeint/stepper/runge_kutta4.hpp: b = {1/6, 1/3, 1/3,
//    1/6}, each entry static_cast<Value>(1) / static_cast<Value>(6 or 3)
//    with Value = float;
//  - the last stage of detail/generic_rk_algorithm.hpp (calculate_stage,
//    stage_number == StageCount) calls the algebra with
//    generic_rk_scale_sum<4>(b, dt) (detail/generic_rk_operations.hpp),
//    which is default_operations::scale_sum5(1.0, b[0]*dt, b[1]*dt,
//    b[2]*dt, b[3]*dt): each coefficient is the float product b_k * dt;
//  - vector_space_algebra::for_each6 applies the operation to the whole
//    Eigen vectors: t1 = a1*t2 + a2*t3 + a3*t4 + a4*t5 + a5*t6, one Eigen
//    expression evaluated per component left to right, each product rounded
//    before its sum; pa is compiled with -msse -msse2 and no FMA
//    (pa/CMakeLists.txt:37-38), so nothing is contracted.
// With all four stage derivatives equal to d = (v0, v1, v2, 0) this is the
// five-term sum below.  euler is scale_sum2(1, dt).  runge_kutta4_classic
// (not used by pa) would round dt/6 and dt/3 instead: ORC_RK4_DT_OVER_6.
template <class T>
void step_frozen(int integrator, const T p[4], const T v[3], T dt, T out[4]) {
  const T d[4] = {v[0], v[1], v[2], T(0)};
  if (integrator == ORC_EULER) {
    for (int c = 0; c < 4; ++c) out[c] = T(1) * p[c] + dt * d[c];
    return;
  }
  T a0, a1, a2, a3;
  if (integrator == ORC_RK4_DT_OVER_6) {
    a0 = dt / T(6), a1 = dt / T(3), a2 = dt / T(3), a3 = dt / T(6);
  } else {
    const T b0 = T(1) / T(6), b1 = T(1) / T(3), b2 = T(1) / T(3), b3 = T(1) / T(6);
    a0 = b0 * dt, a1 = b1 * dt, a2 = b2 * dt, a3 = b3 * dt;
  }
  for (int c = 0; c < 4; ++c)
    out[c] = (((T(1) * p[c] + a0 * d[c]) + a1 * d[c]) + a2 * d[c]) + a3 * d[c];
}

// Textbook RK4 (build extension, no reference counterpart): the same odeint
// generic-RK stage sums, with the field evaluated at each stage point.
// Returns false when a stage point lies outside the block; the caller then
// takes the frozen-derivative step, so the particle still reaches the face
// and is handed off by the next contains() like in the reference.
template <class T>
bool step_classic(const field<T>& f, const T p[4], const T k1v[3], T dt, T out[4], bool cons) {
  const T half = T(1) / T(2);
  const T k1[4] = {k1v[0], k1v[1], k1v[2], T(0)};
  T x[4], k2[4], k3[4], k4[4], v[3];
  const T s21 = half * dt;
  for (int c = 0; c < 4; ++c) x[c] = T(1) * p[c] + s21 * k1[c];
  if (!contains(f, x)) return false;
  interpolate(f, x, v, cons);
  k2[0] = v[0]; k2[1] = v[1]; k2[2] = v[2]; k2[3] = T(0);
  const T s31 = T(0) * dt, s32 = half * dt;
  for (int c = 0; c < 4; ++c) x[c] = (T(1) * p[c] + s31 * k1[c]) + s32 * k2[c];
  if (!contains(f, x)) return false;
  interpolate(f, x, v, cons);
  k3[0] = v[0]; k3[1] = v[1]; k3[2] = v[2]; k3[3] = T(0);
  const T s41 = T(0) * dt, s42 = T(0) * dt, s43 = T(1) * dt;
  for (int c = 0; c < 4; ++c) x[c] = ((T(1) * p[c] + s41 * k1[c]) + s42 * k2[c]) + s43 * k3[c];
  if (!contains(f, x)) return false;
  interpolate(f, x, v, cons);
  k4[0] = v[0]; k4[1] = v[1]; k4[2] = v[2]; k4[3] = T(0);
  const T b0 = T(1) / T(6), b1 = T(1) / T(3), b2 = T(1) / T(3), b3 = T(1) / T(6);
  const T a0 = b0 * dt, a1 = b1 * dt, a2 = b2 * dt, a3 = b3 * dt;
  for (int c = 0; c < 4; ++c)
    out[c] = (((T(1) * p[c] + a0 * k1[c]) + a1 * k2[c]) + a2 * k3[c]) + a3 * k4[c];
  return true;
}

// Stage configuration.  Fields are held by value; has_nbr[k] marks the
// neighbours that exist (the std::optional of particle_tracer.hpp:66).
template <class T>
struct config {
  field<T> local;
  field<T> nbr[6];
  bool has_nbr[6];
  int32_t nbr_rank[6];
  int32_t nbr_mask[6];
  T dt;
  int integrator;
  int stop_rule;
  T stop_eps;
  bool consistent;
  const field<T>& block(int32_t vfi) const { return vfi < 0 ? local : nbr[vfi]; }
};

template <class C, class Cfg>
void fill_config(C& r, const Cfg* c) {
  r.local = view(c->local);
  for (int k = 0; k < 6; ++k) {
    r.has_nbr[k] = c->neighbors[k] != nullptr;
    if (r.has_nbr[k]) r.nbr[k] = view(c->neighbors[k]);
    r.nbr_rank[k] = c->neighbor_rank[k];
    r.nbr_mask[k] = c->neighbor_face_mask[k];
  }
  r.dt = c->step_size;
  r.integrator = c->integrator;
  r.stop_rule = c->stop_rule;
  r.stop_eps = c->stop_eps;
  r.consistent = sizeof(r.dt) == sizeof(double) || c->interp == ORC_INTERP_CONSISTENT;
}

const float kTermination[4] = {-2.0f, -2.0f, -2.0f, 0.0f};  // integral_curves.hpp:18
const float kInvalid[4] = {-1.0f, -1.0f, -1.0f, 0.0f};      // integral_curves.hpp:19

enum seg_end { SEG_END = 0, SEG_STOP = 1, SEG_EXIT = 2, SEG_NOLOOP = 3 };

// One particle's trace loop (particle_tracer.cpp:266-335).  emit(slot, v)
// receives every slot write in order: the termination vertex at slot it,
// then (if the step happens) the stepped vertex over it.  On SEG_EXIT,
// exit_it and last hold the iteration and the uncontained vertex.
template <class T, class Emit>
seg_end trace_loop(const field<T>& f, const config<T>& c, const T start[4], int32_t R, Emit&& emit,
                   int64_t& exit_it, T last[4], uint64_t& steps) {
  for (int i = 0; i < 4; ++i) last[i] = start[i];
  const T term[4] = {T(-2), T(-2), T(-2), T(0)};
  for (uint64_t it = 1; it < uint64_t(int64_t(R)); ++it) {
    emit(it, term);
    if (it == uint64_t(int64_t(R) - 1)) return SEG_END;
    if (!contains(f, last)) {
      exit_it = int64_t(it);
      return SEG_EXIT;
    }
    T v[3];
    interpolate(f, last, v, c.consistent);
    if (stops(v, c.stop_rule, c.stop_eps)) return SEG_STOP;
    T next[4];
    if (c.integrator == ORC_CLASSIC_RK4) {
      if (!step_classic(f, last, v, c.dt, next, c.consistent)) step_frozen(ORC_RK4, last, v, c.dt, next);
    } else {
      step_frozen(c.integrator, last, v, c.dt, next);
    }
    emit(it, next);
    for (int i = 0; i < 4; ++i) last[i] = next[i];
    ++steps;
  }
  return SEG_NOLOOP;
}

// Face priority of particle_tracer.cpp:288-293 (and :360-365,
// particle_advector.cpp:81-86): the first of x<min, x>max, y<min, y>max,
// z<min, z>max whose neighbour exists.  min = offset, max = offset + size.
template <class T>
int face_rule(const field<T>& f, const T p[4], int mask) {
  for (int axis = 0; axis < 3; ++axis) {
    const T lo = f.offset[axis];
    const T hi = f.offset[axis] + f.size[axis];
    if (p[axis] < lo && (mask >> (2 * axis) & 1)) return 2 * axis;
    if (p[axis] > hi && (mask >> (2 * axis + 1) & 1)) return 2 * axis + 1;
  }
  return -1;
}

// Careful: face_rule above interleaves the two tests per axis, which is the
// reference's order (x<min, x>max, y<min, ...).  A face whose neighbour is
// absent is skipped and the chain continues, as the `&& neighbors[k]` makes
// the else-if fall through.

template <class T>
int local_mask(const config<T>& c) {
  int m = 0;
  for (int k = 0; k < 6; ++k)
    if (c.nbr_rank[k] >= 0) m |= 1 << k;
  return m;
}

struct segment_rec {
  int64_t seed;
  int32_t block;
  std::vector<float> v;
};

// Union-of-7 continuation for one seed: the reference round semantics with
// each hand-off traced immediately in the receiving block.  Block -1 is the
// own block; block k the preloaded neighbour k.  A particle leaving block k
// goes back to the own block when block k's face rule picks the face
// opposite k (k ^ 1), and leaves the union otherwise.
template <class T, class OnSegment>
void union_walk(const config<T>& c, const T start_pos[4], int32_t R0, int32_t block0, uint64_t& steps,
                OnSegment&& on_segment, T endpoint[4]) {
  T start[4];
  for (int i = 0; i < 4; ++i) start[i] = start_pos[i];
  int32_t R = R0;
  int32_t b = block0;
  const int mask_local = local_mask(c);
  for (int i = 0; i < 4; ++i) endpoint[i] = start[i];
  for (;;) {
    const field<T>& f = c.block(b);
    std::vector<T> seg;
    seg.insert(seg.end(), start, start + 4);
    auto emit = [&](uint64_t slot, const T* v) {
      const size_t need = size_t(slot + 1) * 4;
      if (seg.size() < need) seg.resize(need);
      for (int i = 0; i < 4; ++i) seg[size_t(slot) * 4 + i] = v[i];
    };
    int64_t exit_it = 0;
    T last[4];
    const seg_end e = trace_loop(f, c, start, R, emit, exit_it, last, steps);
    for (int i = 0; i < 4; ++i) endpoint[i] = last[i];
    on_segment(b, seg);
    if (e != SEG_EXIT) return;
    const int32_t Rn = int32_t(int64_t(R) - exit_it);
    if (b < 0) {
      const int k = face_rule(f, last, mask_local);
      if (k < 0 || !c.has_nbr[k]) return;
      b = k;
    } else {
      const int d = face_rule(f, last, c.nbr_mask[b]);
      if (d != (b ^ 1)) return;
      b = -1;
    }
    R = Rn;
    for (int i = 0; i < 4; ++i) start[i] = last[i];
  }
}

}  // namespace

extern "C" {

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) { return thread_count(); }
int orc_hardware_concurrency(void) { return int(std::thread::hardware_concurrency()); }

// seed_generator.cpp:9-21 with unravel_index (index.hpp:9-19): count per
// axis = int(size / stride) (FP32 divide, truncation); position =
// offset + stride * float(index), w = 0.
int64_t orc_seed_count(const float size[3], const float stride[3], int32_t dims_out[3]) {
  int64_t n = 1;
  for (int i = 0; i < 3; ++i) {
    dims_out[i] = int32_t(size[i] / stride[i]);
    n *= dims_out[i];
  }
  return n;
}

void orc_seed_generate(const float offset[3], const float size[3], const float stride[3],
                       int32_t remaining_iterations, int32_t rank, orc_particle* out) {
  int32_t d[3];
  const int64_t n = orc_seed_count(size, stride, d);
  if (n <= 0) return;
  parallel_for(size_t(n), 4096, [&](size_t index) {
    size_t rest = index;
    int32_t ijk[3];
    for (int i = 2; i >= 0; --i) {
      ijk[i] = int32_t(rest % size_t(d[i]));
      rest = rest / size_t(d[i]);
    }
    orc_particle p;
    std::memset(&p, 0, sizeof p);
    for (int i = 0; i < 3; ++i) p.position[i] = offset[i] + stride[i] * float(ijk[i]);
    p.position[3] = 0.0f;
    p.remaining_iterations = remaining_iterations;
    p.vector_field_index = -1;
    p.original_rank = rank;
    for (int i = 0; i < 3; ++i) p.original_voxel[i] = ijk[i];
    out[index] = p;
  });
}

int orc_contains(const orc_field* f, const float position[4]) { return contains(view(f), position) ? 1 : 0; }
void orc_interpolate(const orc_field* f, const float position[4], float out[3]) {
  interpolate(view(f), position, out, false);
}
void orc_interpolate_consistent(const orc_field* f, const float position[4], float out[3]) {
  interpolate(view(f), position, out, true);
}
int orc_contains_f64(const orc_field64* f, const double position[4]) {
  return contains(view(f), position) ? 1 : 0;
}
void orc_interpolate_f64(const orc_field64* f, const double position[4], double out[3]) {
  interpolate(view(f), position, out, true);
}

void orc_step(int32_t integrator, const float position[4], const float vec[3], float dt, float out[4]) {
  step_frozen(integrator, position, vec, dt, out);
}

// particle_tracer.cpp:210-234.
int orc_round_info(const orc_particle* particles, size_t n, uint64_t curve_offset, uint64_t out[5]) {
  if (n == 0) return -1;
  int32_t maxr = particles[0].remaining_iterations;
  for (size_t i = 1; i < n; ++i) maxr = std::max(maxr, particles[i].remaining_iterations);
  if (maxr <= 0) return -1;
  const uint64_t max_vertices = uint64_t(std::numeric_limits<int32_t>::max()) / 16u;  // sizeof(vector4)
  const uint64_t ppic = max_vertices / uint64_t(maxr);
  out[0] = uint64_t(maxr);
  out[1] = ppic * uint64_t(maxr);
  out[2] = (n % ppic) * uint64_t(maxr);
  out[3] = curve_offset;
  out[4] = 1 + n / ppic;
  return 0;
}

void orc_fill_invalid(float* vertices, size_t n) {
  parallel_for(n, 1 << 16, [&](size_t i) { std::memcpy(vertices + 4 * i, kInvalid, sizeof kInvalid); });
}

// initialize (particle_tracer.cpp:243-252) + trace (:253-337).
int orc_trace(const orc_trace_config* cfg, const orc_particle* particles, size_t n, const uint64_t info[5],
              float* const* chunks, int32_t* vertex_counts, orc_handoff* handoffs, size_t* n_handoffs,
              uint64_t* total_steps) {
  config<float> c{};
  fill_config(c, cfg);
  const uint64_t maxr = info[0], vpic = info[1];
  for (size_t i = 0; i < n; ++i)
    if (particles[i].remaining_iterations < 0) return -1;
  for (size_t i = 0; i < n; ++i) {
    const int32_t vfi = particles[i].vector_field_index;
    if (vfi < -1 || vfi > 5 || (vfi >= 0 && !c.has_nbr[vfi])) return -1;
  }
  const int lmask = local_mask(c);
  std::atomic<uint64_t> steps_all{0};
  std::vector<orc_handoff> found(n);
  std::vector<char> has(n, 0);
  parallel_for(n, 1, [&](size_t pi) {
    const orc_particle& P = particles[pi];
    const uint64_t abs0 = uint64_t(pi) * maxr;
    float* base = chunks[abs0 / vpic] + 4 * (abs0 % vpic);
    std::memcpy(base, P.position, 16);  // initialize
    const field<float>& f = c.block(P.vector_field_index);
    int64_t written = 1;
    auto emit = [&](uint64_t slot, const float* v) {
      std::memcpy(base + 4 * slot, v, 16);
      written = int64_t(slot) + 1;
    };
    int64_t exit_it = 0;
    float last[4];
    uint64_t steps = 0;
    const seg_end e = trace_loop(f, c, P.position, P.remaining_iterations, emit, exit_it, last, steps);
    steps_all += steps;
    if (vertex_counts) vertex_counts[pi] = int32_t(written);
    if (e != SEG_EXIT) return;
    orc_handoff h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.particle.position, last, 16);
    h.particle.remaining_iterations = int32_t(int64_t(P.remaining_iterations) - exit_it);
    h.particle.vector_field_index = -1;
    h.particle.original_rank = 0;
    h.source = int64_t(pi);
    if (P.vector_field_index == -1) {
      const int k = face_rule(f, last, lmask);
      if (k < 0) return;  // no neighbour: dropped (find(-1) fails)
      h.kind = 0;
      h.slot = k;
    } else {
      h.kind = 1;
      h.slot = P.vector_field_index;
    }
    found[pi] = h;
    has[pi] = 1;
  });
  size_t m = 0;
  for (int kind = 0; kind < 2; ++kind)
    for (int slot = 0; slot < 6; ++slot)
      for (size_t i = 0; i < n; ++i)
        if (has[i] && found[i].kind == kind && found[i].slot == slot) handoffs[m++] = found[i];
  if (n_handoffs) *n_handoffs = m;
  if (total_steps) *total_steps = steps_all.load();
  return 0;
}

int orc_trace_union(const orc_trace_config* cfg, const orc_particle* particles, size_t n, float* out_vertices,
                    size_t vertex_capacity, size_t* n_vertices, int64_t* segment_seed, int32_t* segment_block,
                    int64_t* segment_begin, int32_t* segment_length, size_t segment_capacity, size_t* n_segments,
                    uint64_t* total_steps) {
  config<float> c{};
  fill_config(c, cfg);
  for (size_t i = 0; i < n; ++i)
    if (particles[i].remaining_iterations < 0) return -1;
  std::vector<std::vector<segment_rec>> per(n);
  std::atomic<uint64_t> steps_all{0};
  parallel_for(n, 1, [&](size_t pi) {
    const orc_particle& P = particles[pi];
    uint64_t steps = 0;
    float endpoint[4];
    union_walk(c, P.position, P.remaining_iterations, P.vector_field_index, steps,
               [&](int32_t b, const std::vector<float>& seg) { per[pi].push_back(segment_rec{int64_t(pi), b, seg}); },
               endpoint);
    steps_all += steps;
  });
  size_t nv = 0, ns = 0;
  for (size_t i = 0; i < n; ++i)
    for (auto& s : per[i]) {
      const size_t len = s.v.size() / 4;
      if (ns >= segment_capacity || nv + len > vertex_capacity) return -2;
      std::memcpy(out_vertices + 4 * nv, s.v.data(), len * 16);
      segment_seed[ns] = s.seed;
      segment_block[ns] = s.block;
      segment_begin[ns] = int64_t(nv);
      segment_length[ns] = int32_t(len);
      nv += len;
      ++ns;
    }
  *n_vertices = nv;
  *n_segments = ns;
  if (total_steps) *total_steps = steps_all.load();
  return 0;
}

int orc_endpoints(const orc_trace_config* cfg, const orc_particle* particles, size_t n, int union_mode,
                  float* endpoints, int32_t* steps_out) {
  config<float> c{};
  fill_config(c, cfg);
  parallel_for(n, 1, [&](size_t pi) {
    const orc_particle& P = particles[pi];
    uint64_t steps = 0;
    float endpoint[4];
    if (union_mode) {
      union_walk(c, P.position, P.remaining_iterations, P.vector_field_index, steps,
                 [](int32_t, const std::vector<float>&) {}, endpoint);
    } else {
      const field<float>& f = c.block(P.vector_field_index);
      int64_t exit_it = 0;
      trace_loop(f, c, P.position, P.remaining_iterations, [](uint64_t, const float*) {}, exit_it, endpoint, steps);
    }
    std::memcpy(endpoints + 4 * pi, endpoint, 16);
    steps_out[pi] = int32_t(steps);
  });
  return 0;
}

int orc_endpoints_f64(const orc_trace_config64* cfg, const double* positions, size_t n,
                      int32_t remaining_iterations, int union_mode, double* endpoints, int32_t* steps_out) {
  config<double> c{};
  fill_config(c, cfg);
  parallel_for(n, 1, [&](size_t pi) {
    uint64_t steps = 0;
    double endpoint[4];
    const double* p = positions + 4 * pi;
    if (union_mode) {
      union_walk(c, p, remaining_iterations, -1, steps, [](int32_t, const std::vector<double>&) {}, endpoint);
    } else {
      int64_t exit_it = 0;
      trace_loop(c.local, c, p, remaining_iterations, [](uint64_t, const double*) {}, exit_it, endpoint, steps);
    }
    std::memcpy(endpoints + 4 * pi, endpoint, 32);
    steps_out[pi] = int32_t(steps);
  });
  return 0;
}

// particle_tracer.cpp:402-409: std::remove of exact-equal invalid_vertex.
size_t orc_prune(const float* vertices, size_t n, float* out) {
  size_t m = 0;
  for (size_t i = 0; i < n; ++i) {
    const float* v = vertices + 4 * i;
    const bool inv = v[0] == kInvalid[0] && v[1] == kInvalid[1] && v[2] == kInvalid[2] && v[3] == kInvalid[3];
    if (!inv) {
      std::memcpy(out + 4 * m, v, 16);
      ++m;
    }
  }
  return m;
}

// index_generator.cpp:10-24.
int64_t orc_index_generate(const float* vertices, size_t n, int32_t* indices) {
  if (n == 0) return -1;
  auto is_term = [&](size_t i) {
    const float* v = vertices + 4 * i;
    return v[0] == kTermination[0] && v[1] == kTermination[1] && v[2] == kTermination[2] &&
           v[3] == kTermination[3];
  };
  int64_t m = 0;
  for (size_t i = 0; i + 1 < n; ++i)
    if (!is_term(i) && !is_term(i + 1)) indices[m++] = int32_t(i);
  return m;
}

}  // extern "C"

namespace {

// Eigen normalized() (Dot.h:122-133) on a Vector4f: squaredNorm is the SSE
// packet reduction (a0 + a2) + (a1 + a3) (arch/SSE/PacketMath.h:540-551),
// then each component is divided by sqrt(z) when z > 0.
void normalize4(const float d[4], float t[4]) {
  const float a0 = d[0] * d[0], a1 = d[1] * d[1], a2 = d[2] * d[2], a3 = d[3] * d[3];
  const float z = (a0 + a2) + (a1 + a3);
  if (z > 0.0f) {
    const float s = std::sqrt(z);
    for (int i = 0; i < 4; ++i) t[i] = d[i] / s;
  } else {
    for (int i = 0; i < 4; ++i) t[i] = d[i];
  }
}

// coordinates.hpp:8-16.  std::pow(float, int) promotes to double, so the
// radius is a double sum and sqrt, stored to float; atan2 and acos take the
// float overloads.
void to_spherical(const float t[4], float s[3]) {
  const double r = std::sqrt(std::pow(double(t[0]), 2) + std::pow(double(t[1]), 2) + std::pow(double(t[2]), 2));
  s[0] = float(r);
  s[1] = std::atan2(t[1], t[0]);
  s[2] = std::acos(t[2] / s[0]);
}

const double kPi = 3.14159265358979323846;

// color_generator.cpp:41-54.
void correct_range(float s[3]) {
  if (s[1] < 0.0f) s[1] += float(kPi);
  if (s[1] >= float(kPi)) s[1] -= float(kPi);
  s[1] = float(kPi) - s[1];
  if (s[2] < 0.0f) s[2] = std::fabs(s[2]);
  if (s[2] >= float(kPi / 2.0)) s[2] = float(kPi) - s[2];
  s[1] = float(double(s[1]) / kPi);
  s[2] = float(double(s[2]) / (kPi / 2.0f));
}

float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

// color_generator.cpp:55-64.
void hue_to_rgba(float h, float c[4]) {
  c[0] = clamp01(std::fabs(6.0f * h - 3.0f) - 1.0f);
  c[1] = clamp01(2.0f - std::fabs(6.0f * h - 2.0f));
  c[2] = clamp01(2.0f - std::fabs(6.0f * h - 4.0f));
  c[3] = 1.0f;
}

// color_generator.cpp:65-72.
void hsl_to_rgba(float h, float s, float l, float c[4]) {
  hue_to_rgba(h, c);
  const float pre = (1.0f - std::fabs(2.0f * l - 1.0f)) * s;
  for (int i = 0; i < 4; ++i) c[i] = (c[i] - 0.5f) * pre + l;
  c[3] = 1.0f;
}

// color_generator.cpp:73-79.
void hsv_to_rgba(float h, float s, float v, float c[4]) {
  hue_to_rgba(h, c);
  for (int i = 0; i < 4; ++i) c[i] = ((c[i] - 1.0f) * s + 1.0f) * v;
  c[3] = 1.0f;
}

void map_color(const float t[4], const float d[4], int32_t mode, float param, float c[4]) {
  if (mode == 4) {  // rgb (color_generator.cpp:95-98): y and z swapped
    c[0] = std::fabs(t[0]);
    c[1] = std::fabs(t[2]);
    c[2] = std::fabs(t[1]);
    c[3] = 1.0f;
    return;
  }
  if (mode == 100) {  // magnitude (extension): hue from segment length / parameter
    const float len = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    const float u = clamp01(len / param);
    hue_to_rgba(u * (2.0f / 3.0f), c);
    return;
  }
  float s[3];
  to_spherical(t, s);
  correct_range(s);
  switch (mode) {
    case 0: hsl_to_rgba(s[1], param, s[2], c); break;
    case 1: hsl_to_rgba(s[1], s[2], param, c); break;
    case 2: hsv_to_rgba(s[1], param, s[2], c); break;
    case 3: hsv_to_rgba(s[1], s[2], param, c); break;
    default: c[0] = c[1] = c[2] = c[3] = 0.0f; break;
  }
}

}  // namespace

extern "C" {

// color_generator.cpp:16-40.
void orc_color_generate(const float* vertices, size_t n_vertices, const int32_t* indices, size_t n_indices,
                        int32_t mode, float parameter, float* colors) {
  (void)n_vertices;
  parallel_for(n_indices, 1024, [&](size_t i) {
    const int64_t idx = indices[i];
    const float* a = vertices + 4 * idx;
    const float* b = vertices + 4 * (idx + 1);
    float d[4], t[4], c[4];
    for (int k = 0; k < 4; ++k) d[k] = b[k] - a[k];
    normalize4(d, t);
    map_color(t, d, mode, parameter, c);
    std::memcpy(colors + 4 * idx, c, 16);
    if (i != n_indices - 1 && int64_t(indices[i + 1]) != idx + 1) std::memcpy(colors + 4 * (idx + 1), c, 16);
  });
}

// particle_advector.cpp:61-141, one round over one block.
int orc_advect(const orc_trace_config* cfg, orc_particle* particles, size_t n, int32_t* state,
               int32_t* handoff_slot, uint64_t* total_steps) {
  config<float> c{};
  fill_config(c, cfg);
  const field<float>& f = c.local;
  const int lmask = local_mask(c);
  std::atomic<uint64_t> steps_all{0};
  parallel_for(n, 1, [&](size_t pi) {
    orc_particle& P = particles[pi];
    state[pi] = 3;
    handoff_slot[pi] = -1;
    uint64_t steps = 0;
    for (uint64_t it = 1; it < uint64_t(int64_t(P.remaining_iterations)); ++it) {
      if (!contains(f, P.position)) {
        P.remaining_iterations = int32_t(int64_t(P.remaining_iterations) - int64_t(it));
        const int k = face_rule(f, P.position, lmask);
        if (k < 0) {
          P.remaining_iterations = 0;
          state[pi] = 1;
        } else {
          state[pi] = 2;
          handoff_slot[pi] = k;
        }
        break;
      }
      float v[3];
      interpolate(f, P.position, v, c.consistent);
      if (stops(v, c.stop_rule, c.stop_eps)) {
        P.remaining_iterations = 0;
        state[pi] = 1;
        break;
      }
      float next[4];
      if (c.integrator == ORC_CLASSIC_RK4) {
        if (!step_classic(f, P.position, v, c.dt, next, c.consistent)) step_frozen(ORC_RK4, P.position, v, c.dt, next);
      } else {
        step_frozen(c.integrator, P.position, v, c.dt, next);
      }
      std::memcpy(P.position, next, 16);
      ++steps;
      if (it + 1 == uint64_t(int64_t(P.remaining_iterations))) {
        P.remaining_iterations = 0;
        state[pi] = 1;
        break;
      }
    }
    steps_all += steps;
  });
  if (total_steps) *total_steps = steps_all.load();
  return 0;
}

// flow_map_generator.cpp:12-52 (allocate, initialize, advect, assign).
int orc_flow_map(const orc_field* fld, const int32_t block_size[3], uint64_t iterations, float resolution_scale,
                 int32_t rank, int32_t integrator, float step_size, float* flow_map, int32_t dims_out[3],
                 uint64_t* total_steps) {
  field<float> f = view(fld);
  float spacing[3];
  for (int i = 0; i < 3; ++i) {
    dims_out[i] = int32_t(float(block_size[i]) * resolution_scale);
    spacing[i] = f.spacing[i] / resolution_scale;
  }
  const int64_t n_map = int64_t(dims_out[0]) * dims_out[1] * dims_out[2];
  // Voxels keep their node position unless a particle is assigned.
  parallel_for(size_t(n_map), 4096, [&](size_t v) {
    const int32_t z = int32_t(v % size_t(dims_out[2]));
    const int32_t y = int32_t((v / size_t(dims_out[2])) % size_t(dims_out[1]));
    const int32_t x = int32_t(v / (size_t(dims_out[2]) * size_t(dims_out[1])));
    const int32_t ijk[3] = {x, y, z};
    for (int i = 0; i < 3; ++i) flow_map[3 * v + i] = f.offset[i] + spacing[i] * float(ijk[i]);
  });
  int32_t sd[3];
  const int64_t ns = orc_seed_count(f.size, spacing, sd);
  std::vector<orc_particle> P(size_t(std::max<int64_t>(ns, 0)));
  orc_seed_generate(f.offset, f.size, spacing, int32_t(iterations), rank, P.data());
  orc_trace_config cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.local = fld;
  for (int k = 0; k < 6; ++k) cfg.neighbor_rank[k] = -1;
  cfg.step_size = step_size;
  cfg.integrator = integrator;
  std::vector<int32_t> state(P.size()), slot(P.size());
  orc_advect(&cfg, P.data(), P.size(), state.data(), slot.data(), total_steps);
  parallel_for(P.size(), 4096, [&](size_t i) {
    if (state[i] != 1) return;
    const int32_t* v = P[i].original_voxel;
    if (v[0] >= dims_out[0] || v[1] >= dims_out[1] || v[2] >= dims_out[2]) return;
    const size_t at = (size_t(v[0]) * size_t(dims_out[1]) + size_t(v[1])) * size_t(dims_out[2]) + size_t(v[2]);
    for (int k = 0; k < 3; ++k) flow_map[3 * at + k] = P[i].position[k];
  });
  return 0;
}

// Cyclic Jacobi eigenvalue iteration for a symmetric 3x3 matrix (Golub &
// Van Loan, Matrix Computations, 4th ed., Alg. 8.5.3), ascending order.
void orc_eigh3(const double a_in[6], double ev[3]) {
  double a[3][3] = {{a_in[0], a_in[1], a_in[2]}, {a_in[1], a_in[3], a_in[4]}, {a_in[2], a_in[4], a_in[5]}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off == 0.0 || off <= 1e-36 * diag) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0);
        const double sn = t * cs;
        for (int k = 0; k < 3; ++k) {  // A <- A J
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = cs * akp - sn * akq;
          a[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 3; ++k) {  // A <- J^T A
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = cs * apk - sn * aqk;
          a[q][k] = sn * apk + cs * aqk;
        }
      }
  }
  ev[0] = a[0][0];
  ev[1] = a[1][1];
  ev[2] = a[2][2];
  std::sort(ev, ev + 3);
}

// Textbook FTLE: J by central differences (one-sided at the lattice
// border), C = J^T J (ftle_map_generator.cpp:40), eigenvalues, then
// ln(sqrt(lambda_max)) / |T| (:50) or fractional anisotropy (:51-54).
void orc_ftle(const float* flow_map, const int32_t dims[3], const float spacing[3], double time, int32_t mode,
              float* ftle, double* lambda_max_out) {
  const int64_t d0 = dims[0], d1 = dims[1], d2 = dims[2];
  const int64_t n = d0 * d1 * d2;
  parallel_for(size_t(n), 256, [&](size_t v) {
    const int64_t z = int64_t(v) % d2, y = (int64_t(v) / d2) % d1, x = int64_t(v) / (d1 * d2);
    const int64_t ijk[3] = {x, y, z};
    const int64_t stride[3] = {d1 * d2, d2, 1};
    double J[3][3];
    for (int j = 0; j < 3; ++j) {
      const int64_t i = ijk[j];
      const double h = double(spacing[j]);
      int64_t lo = i, hi = i;
      double den = 0.0;
      if (dims[j] > 1) {
        if (i > 0 && i < dims[j] - 1) {
          lo = i - 1; hi = i + 1; den = 2.0 * h;
        } else if (i == 0) {
          hi = i + 1; den = h;
        } else {
          lo = i - 1; den = h;
        }
      }
      const float* fh = flow_map + 3 * (int64_t(v) + (hi - i) * stride[j]);
      const float* fl = flow_map + 3 * (int64_t(v) + (lo - i) * stride[j]);
      for (int c = 0; c < 3; ++c) J[c][j] = den > 0.0 ? (double(fh[c]) - double(fl[c])) / den : 0.0;
    }
    double C[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) C[a][b] = J[0][a] * J[0][b] + J[1][a] * J[1][b] + J[2][a] * J[2][b];
    const double sym[6] = {C[0][0], C[0][1], C[0][2], C[1][1], C[1][2], C[2][2]};
    double ev[3];
    orc_eigh3(sym, ev);
    double value;
    if (mode == 0) {
      value = std::log(std::sqrt(ev[2])) / std::fabs(time);
    } else {
      const double num = (ev[0] - ev[1]) * (ev[0] - ev[1]) + (ev[1] - ev[2]) * (ev[1] - ev[2]) +
                         (ev[2] - ev[0]) * (ev[2] - ev[0]);
      const double den = ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2];
      value = std::sqrt(0.5) * std::sqrt(num) / std::sqrt(den);
    }
    ftle[v] = float(value);
    if (lambda_max_out) lambda_max_out[v] = ev[2];
  });
}

void orc_abc_field_f64(const int32_t dims[3], const int32_t origin[3], double h, double A, double B, double C,
                       double* out) {
  const int64_t n = int64_t(dims[0]) * dims[1] * dims[2];
  parallel_for(size_t(n), 4096, [&](size_t v) {
    const int64_t z = int64_t(v) % dims[2], y = (int64_t(v) / dims[2]) % dims[1], x = int64_t(v) / (int64_t(dims[1]) * dims[2]);
    const double px = double(origin[0] + x) * h, py = double(origin[1] + y) * h, pz = double(origin[2] + z) * h;
    out[3 * v + 0] = A * std::sin(pz) + C * std::cos(py);
    out[3 * v + 1] = B * std::sin(px) + A * std::cos(pz);
    out[3 * v + 2] = C * std::sin(py) + B * std::cos(px);
  });
}

void orc_fourier_field_f64(const int32_t dims[3], const int32_t origin[3], double h, const double* modes,
                           int32_t n_modes, double* out) {
  const int64_t n = int64_t(dims[0]) * dims[1] * dims[2];
  parallel_for(size_t(n), 1024, [&](size_t v) {
    const int64_t z = int64_t(v) % dims[2], y = (int64_t(v) / dims[2]) % dims[1], x = int64_t(v) / (int64_t(dims[1]) * dims[2]);
    const double p[3] = {double(origin[0] + x) * h, double(origin[1] + y) * h, double(origin[2] + z) * h};
    double u[3] = {0.0, 0.0, 0.0};
    for (int32_t m = 0; m < n_modes; ++m) {
      const double* md = modes + 7 * m;
      const double arg = md[0] * p[0] + md[1] * p[1] + md[2] * p[2] + md[6];
      const double cs = std::cos(arg);
      u[0] += cs * md[3];
      u[1] += cs * md[4];
      u[2] += cs * md[5];
    }
    out[3 * v + 0] = u[0];
    out[3 * v + 1] = u[1];
    out[3 * v + 2] = u[2];
  });
}

}  // extern "C"
