// pa_amd/stages.hpp — C++ mirror of the reference's `pa` stage classes
// (acdemiralp/pars, pa/include/pa/stages/*.hpp) over the C ABI of
// libpa_gpu.so (include/pa_gpu.h).  Class and method names, argument order
// and meaning follow the reference so a pipeline written against `pa::`
// switches by namespace; host std::vectors are copied to and from HBM at
// the boundary exactly where the reference hands them between stages.
// Errors throw pa_amd::error (the reference has no status codes; its
// failures are exceptions from allocation or crashes).
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "pa_gpu.h"

namespace pa_amd {

using scalar = float;
using integer = std::int32_t;
using particle = pa_particle;  // pa::particle (particle.hpp:30-35)
struct vector3 { float x, y, z; };
struct vector4 { float x, y, z, w; };
static_assert(sizeof(vector4) == 16, "vector4 is 16 bytes");

// integral_curves.hpp:11-20
struct integral_curves {
  std::vector<vector4> vertices;
  std::vector<vector4> colors;
  std::vector<integer> indices;
};
inline constexpr vector4 termination_vertex{-2.0f, -2.0f, -2.0f, 0.0f};
inline constexpr vector4 invalid_vertex{-1.0f, -1.0f, -1.0f, 0.0f};
inline constexpr integer invalid_index = -1;

struct error : std::runtime_error {
  pa_status status;
  error(pa_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

class context {
 public:
  explicit context(int device = 0) {
    pa_status s = pa_gpu_ctx_create(device, &ctx_);
    if (s != PA_OK) throw error(s, "pa_gpu_ctx_create failed");
  }
  ~context() {
    if (ctx_) pa_gpu_ctx_destroy(ctx_);
  }
  context(const context&) = delete;
  context& operator=(const context&) = delete;
  pa_gpu_ctx* get() const { return ctx_; }
  void check(pa_status s, const char* what) const {
    if (s != PA_OK) throw error(s, std::string(what) + ": " + pa_gpu_last_error(ctx_));
  }

 private:
  pa_gpu_ctx* ctx_ = nullptr;
};

// Device buffer of T (RAII over pa_gpu_malloc / pa_gpu_free).
template <class T>
class device_vector {
 public:
  device_vector(context& c, size_t n) : c_(&c), n_(n) {
    void* p = nullptr;
    c.check(pa_gpu_malloc(c.get(), (n ? n : 1) * sizeof(T), &p), "pa_gpu_malloc");
    p_ = static_cast<T*>(p);
  }
  device_vector(context& c, const std::vector<T>& h) : device_vector(c, h.size()) {
    if (!h.empty()) c.check(pa_gpu_memcpy_h2d(c.get(), p_, h.data(), h.size() * sizeof(T)), "h2d");
  }
  ~device_vector() {
    if (p_) pa_gpu_free(c_->get(), p_);
  }
  device_vector(const device_vector&) = delete;
  device_vector& operator=(const device_vector&) = delete;
  T* data() const { return p_; }
  size_t size() const { return n_; }
  std::vector<T> to_host(size_t count) const {
    std::vector<T> h(count);
    if (count) c_->check(pa_gpu_memcpy_d2h(c_->get(), h.data(), p_, count * sizeof(T)), "d2h");
    return h;
  }

 private:
  context* c_;
  size_t n_;
  T* p_ = nullptr;
};

// pa::vector_field (vector_field.hpp:14-24), resident in HBM.
class vector_field {
 public:
  vector_field(context& c, const std::vector<vector3>& data, const std::array<int32_t, 3>& dims, vector3 offset,
               vector3 size, vector3 spacing, pa_storage storage = PA_STORAGE_FP32)
      : offset(offset), size(size), spacing(spacing), dims(dims), c_(&c) {
    const float o[3] = {offset.x, offset.y, offset.z}, s[3] = {size.x, size.y, size.z},
                h[3] = {spacing.x, spacing.y, spacing.z};
    c.check(pa_gpu_field_create(c.get(), &data[0].x, dims.data(), o, s, h, storage, &f_), "field_create");
  }
  ~vector_field() {
    if (f_) pa_gpu_field_destroy(c_->get(), f_);
  }
  vector_field(const vector_field&) = delete;
  vector_field& operator=(const vector_field&) = delete;
  const pa_gpu_field* get() const { return f_; }
  // Extension: also keep the 8-corner cell records (PA_LAYOUT_CELLS), one
  // cache line per trilinear gather.
  void set_layout(pa_layout layout) { c_->check(pa_gpu_field_set_layout(c_->get(), f_, layout), "set_layout"); }

  // data_io::load_vector_field (data_io.cpp:91-115): the ghosted hyperslab of
  // the file's "vectors" dataset at a rank's node offset, spacing normalised
  // by its largest component, offset and size as data_io.cpp:113-114 has
  // them; read into pinned memory and stored on the device.
  static std::unique_ptr<vector_field> load(context& c, const std::string& path, const std::array<int64_t, 3>& offset,
                                            const std::array<int32_t, 3>& ghosted,
                                            const std::array<int32_t, 3>& block,
                                            pa_storage storage = PA_STORAGE_FP32) {
    pa_gpu_field* f = nullptr;
    c.check(pa_gpu_field_load(c.get(), path.c_str(), offset.data(), ghosted.data(), block.data(), storage, &f),
            "field_load");
    return std::unique_ptr<vector_field>(new vector_field(c, f));
  }

  vector3 offset, size, spacing;
  std::array<int32_t, 3> dims;

 private:
  vector_field(context& c, pa_gpu_field* f) : c_(&c), f_(f) {
    float o[3], sz[3], h[3];
    c.check(pa_gpu_field_info(f, dims.data(), o, sz, h, nullptr), "field_info");
    offset = {o[0], o[1], o[2]};
    size = {sz[0], sz[1], sz[2]};
    spacing = {h[0], h[1], h[2]};
  }

  context* c_;
  pa_gpu_field* f_ = nullptr;
};

// pa::seed_generator (seed_generator.hpp:10-14).
class seed_generator {
 public:
  static std::vector<particle> generate(context& c, const vector3& offset, const vector3& size, const vector3& stride,
                                        integer remaining_iterations, integer rank) {
    const float o[3] = {offset.x, offset.y, offset.z}, s[3] = {size.x, size.y, size.z},
                t[3] = {stride.x, stride.y, stride.z};
    int32_t dims[3];
    size_t n = 0;
    c.check(pa_gpu_seed_count(s, t, dims, &n), "seed_count");
    device_vector<particle> d(c, n);
    size_t got = 0;
    c.check(pa_gpu_seed_generate(c.get(), o, s, t, remaining_iterations, rank, d.data(), n, &got), "seed_generate");
    return d.to_host(got);
  }
};

// pa::particle_tracer (particle_tracer.hpp:23-71) for one rank's round.
class particle_tracer {
 public:
  struct round_info : pa_round_info {
    std::array<std::vector<particle>, 6> out_of_bounds_particles;           // by neighbour slot
    std::array<std::vector<particle>, 6> neighbor_out_of_bounds_particles;  // by neighbour slot
  };

  explicit particle_tracer(context& c) : c_(&c) { c.check(pa_gpu_tracer_create(c.get(), &t_), "tracer_create"); }
  context& owner() const { return *c_; }
  ~particle_tracer() {
    if (t_) pa_gpu_tracer_destroy(t_);
  }
  particle_tracer(const particle_tracer&) = delete;
  particle_tracer& operator=(const particle_tracer&) = delete;

  void set_local_vector_field(const vector_field* f) {
    c_->check(pa_gpu_tracer_set_local_vector_field(t_, f ? f->get() : nullptr), "set_local_vector_field");
  }
  void set_neighbor_vector_fields(const std::array<const vector_field*, 6>& f, const std::array<int32_t, 6>& ranks) {
    const pa_gpu_field* h[6];
    for (int k = 0; k < 6; ++k) h[k] = f[k] ? f[k]->get() : nullptr;
    c_->check(pa_gpu_tracer_set_neighbor_vector_fields(t_, h, ranks.data()), "set_neighbor_vector_fields");
  }
  void set_integrator(pa_integrator i) { c_->check(pa_gpu_tracer_set_integrator(t_, i), "set_integrator"); }
  void set_step_size(scalar s) { c_->check(pa_gpu_tracer_set_step_size(t_, s), "set_step_size"); }
  // Extensions (defaults are the reference's behaviour).
  void set_stop_rule(pa_stop_rule r, float eps = 0.0f) {
    c_->check(pa_gpu_tracer_set_stop_rule(t_, r, eps), "set_stop_rule");
  }
  void set_interpolation(pa_interp i) { c_->check(pa_gpu_tracer_set_interpolation(t_, i), "set_interpolation"); }
  void set_engine(pa_engine e, int32_t waves = 0) { c_->check(pa_gpu_tracer_set_engine(t_, e, waves), "set_engine"); }
  // Extension: the hybrid engine's launch constants and the XCD workgroup
  // order (speed only; the results do not change).
  void set_launch(int32_t quad_switch, int32_t pair_live, int32_t xcd_order = 1) {
    c_->check(pa_gpu_tracer_set_launch(t_, quad_switch, pair_live), "set_launch");
    c_->check(pa_gpu_tracer_set_xcd_remap(t_, xcd_order), "set_xcd_remap");
  }
  void set_neighbor_face_masks(const std::array<int32_t, 6>& masks) {
    c_->check(pa_gpu_tracer_set_neighbor_face_masks(t_, masks.data()), "set_neighbor_face_masks");
  }
  pa_gpu_tracer* get() const { return t_; }

  // Extension: one rank's whole round over its 7 preloaded blocks, fused on
  // the device (union trace -> prune -> index -> colour,
  // pars/source/pipeline.cpp:186-260); the curves are those the stage chain
  // produces on the traced slots.  Particles whose segment would not fit in
  // region_slots come back in `continuation` for the next round.
  struct union_round {
    integral_curves curves;
    std::vector<particle> continuation;
    std::vector<int32_t> counts;    // vertices per seed (in seed order)
    std::vector<vector4> endpoints;  // last position per seed
    uint64_t total_steps = 0;
  };
  union_round trace_union_round(const std::vector<particle>& seeds, size_t region_slots,
                                pa_color_mode mode = PA_COLOR_HSV_CONSTANT_S, scalar parameter = 0.75f) {
    union_round out;
    const size_t n = seeds.size();
    if (n == 0) return out;
    device_vector<particle> d(*c_, seeds);
    const size_t groups = (n + 7) / 8;
    device_vector<vector4> slots(*c_, groups * 8 * region_slots);
    device_vector<int32_t> counts(*c_, n);
    device_vector<particle> cont(*c_, n);
    device_vector<vector4> ends(*c_, n);
    const size_t cap = n * region_slots;
    device_vector<vector4> vertices(*c_, cap);
    device_vector<int32_t> indices(*c_, cap);
    device_vector<vector4> colors(*c_, cap);
    size_t ncont = 0, nv = 0, ni = 0;
    c_->check(pa_gpu_tracer_trace_union_round(t_, d.data(), n, region_slots, reinterpret_cast<float*>(slots.data()),
                                              counts.data(), cont.data(), &ncont, mode, parameter,
                                              reinterpret_cast<float*>(vertices.data()), cap, &nv, indices.data(),
                                              cap, &ni, reinterpret_cast<float*>(colors.data()),
                                              reinterpret_cast<float*>(ends.data()), nullptr, &out.total_steps),
              "trace_union_round");
    out.curves.vertices = vertices.to_host(nv);
    out.curves.indices = indices.to_host(ni);
    out.curves.colors = colors.to_host(nv);
    out.continuation = cont.to_host(ncont);
    out.counts = counts.to_host(n);
    out.endpoints = ends.to_host(n);
    return out;
  }

  // compute_round_info + allocate + initialize + trace + prune of one round
  // (particle_tracer.cpp:210-337, 402-409): returns the pruned curves of the
  // round and fills the hand-off maps of `info`.
  std::vector<integral_curves> trace(const std::vector<particle>& particles, round_info* info = nullptr) {
    std::vector<integral_curves> curves;
    if (particles.empty()) return curves;
    device_vector<particle> d(*c_, particles);
    round_info ri;
    c_->check(pa_gpu_tracer_compute_round_info(t_, d.data(), particles.size(), 0, &ri), "compute_round_info");
    const size_t maxr = ri.maximum_remaining_iterations;
    const size_t slots = particles.size() * maxr;
    device_vector<vector4> chunk(*c_, slots);
    std::vector<float*> table;
    for (uint64_t k = 0; k < ri.integral_curve_count; ++k)
      table.push_back(reinterpret_cast<float*>(chunk.data() + k * ri.vertices_per_integral_curve));
    device_vector<int32_t> counts(*c_, particles.size());
    device_vector<particle> hand(*c_, particles.size());
    pa_handoff h{hand.data(), nullptr, particles.size(), {{0}}};
    uint64_t steps = 0;
    c_->check(pa_gpu_tracer_trace(t_, d.data(), particles.size(), &ri, table.data(), counts.data(), &h, &steps),
              "trace");
    // prune each chunk by the per-particle counts
    const size_t ppic = ri.vertices_per_integral_curve / maxr;
    for (uint64_t k = 0; k < ri.integral_curve_count; ++k) {
      const size_t first = k * ppic;
      const size_t np = first >= particles.size() ? 0 : std::min(ppic, particles.size() - first);
      integral_curves ic;
      if (np) {
        device_vector<vector4> out(*c_, np * maxr);
        size_t m = 0;
        c_->check(pa_gpu_prune_counted(c_->get(), reinterpret_cast<float*>(chunk.data() + first * maxr), maxr,
                                       counts.data() + first, np, reinterpret_cast<float*>(out.data()), np * maxr, &m,
                                       nullptr),
                  "prune");
        ic.vertices = out.to_host(m);
      }
      curves.push_back(std::move(ic));
    }
    if (info) {
      static_cast<pa_round_info&>(*info) = ri;
      std::vector<particle> all = hand.to_host(particles.size());
      size_t at = 0;
      for (int kind = 0; kind < 2; ++kind)
        for (int s = 0; s < 6; ++s) {
          auto& dst = kind == 0 ? info->out_of_bounds_particles[s] : info->neighbor_out_of_bounds_particles[s];
          dst.assign(all.begin() + at, all.begin() + at + h.counts[kind][s]);
          at += h.counts[kind][s];
        }
    }
    return curves;
  }

 private:
  context* c_;
  pa_gpu_tracer* t_ = nullptr;
};

// Extension: particle pool over up to 16 resident fields (batched seed sets,
// the step engine of continuous-batching serving; pars serves one request at
// a time, pars/source/service.cpp:20-69).  Each particle ends where a
// one-block trace of it on its field (vector_field_index) ends.
class particle_pool {
 public:
  particle_pool(context& c, size_t capacity, size_t requests = 0)
      : c_(&c), cap_(capacity), position_(c, capacity), iteration_(c, capacity), remaining_(c, capacity),
        steps_(c, capacity), status_(c, capacity), field_(c, capacity), request_(c, requests ? capacity : 0),
        request_done_(c, std::vector<int32_t>(requests, 0)) {
    // without request counters the pool reads and writes no request id
    pool_ = pa_pool{reinterpret_cast<float*>(position_.data()), iteration_.data(), remaining_.data(), steps_.data(),
                    status_.data(), field_.data(), requests ? request_.data() : nullptr,
                    requests ? request_done_.data() : nullptr, int32_t(requests)};
  }
  // Load particles (vector_field_index selects the field, original_rank the
  // request).
  void init(const std::vector<particle>& particles, int32_t n_fields) {
    if (particles.size() > cap_) throw error(PA_E_CAPACITY, "particle_pool: more particles than capacity");
    n_ = particles.size();
    if (n_ == 0) return;
    device_vector<particle> d(*c_, particles);
    c_->check(pa_gpu_pool_init(c_->get(), d.data(), n_, n_fields, &pool_), "pool_init");
  }
  // Advance by at most max_iterations (0: to the end); returns the steps taken.
  uint64_t run(particle_tracer& tracer, const std::vector<const vector_field*>& fields, int32_t max_iterations = 0,
               int32_t waves_per_cu = 0) {
    std::vector<const pa_gpu_field*> f;
    for (auto* v : fields) f.push_back(v->get());
    uint64_t steps = 0;
    c_->check(pa_gpu_pool_run(tracer.get(), f.data(), int32_t(f.size()), &pool_, n_, max_iterations, waves_per_cu,
                              &steps),
              "pool_run");
    return steps;
  }
  // Extension: run() that also keeps each entry's curve (pa_gpu_pool_run_vertices): entry i writes its
  // one-block round's slots from slots[base[i]] on and its vertex count to counts[i].
  uint64_t run_vertices(particle_tracer& tracer, const std::vector<const vector_field*>& fields,
                        const device_vector<int64_t>& base, device_vector<vector4>& slots,
                        device_vector<int32_t>& counts, int32_t max_iterations = 0, int32_t waves_per_cu = 0) {
    std::vector<const pa_gpu_field*> f;
    for (auto* v : fields) f.push_back(v->get());
    uint64_t steps = 0;
    c_->check(pa_gpu_pool_run_vertices(tracer.get(), f.data(), int32_t(f.size()), &pool_, n_, max_iterations,
                                       waves_per_cu, base.data(), reinterpret_cast<float*>(slots.data()),
                                       slots.size(), counts.data(), &steps),
              "pool_run_vertices");
    return steps;
  }
  std::vector<vector4> positions() const { return position_.to_host(n_); }
  std::vector<int32_t> steps() const { return steps_.to_host(n_); }
  std::vector<int32_t> status() const { return status_.to_host(n_); }
  std::vector<int32_t> request_done() const { return request_done_.to_host(request_done_.size()); }

 private:
  context* c_;
  size_t cap_, n_ = 0;
  device_vector<vector4> position_;
  device_vector<int32_t> iteration_, remaining_, steps_, status_, field_, request_, request_done_;
  pa_pool pool_{};
};

// Extension: rounds of a round loop (pars/source/pipeline.cpp:153-209) kept
// in flight: each submitted round traces on the context's stream and
// finishes (prune, index, colour) on its second stream while the next round
// traces (pa_gpu_tracer_trace_union_round_pipelined).  Two sets of device
// buffers rotate, so a round's result can be read until two more rounds are
// submitted.
class round_pipeline {
 public:
  round_pipeline(context& c, particle_tracer& tracer, size_t max_seeds, size_t region_slots,
                 pa_color_mode mode = PA_COLOR_HSV_CONSTANT_S, scalar parameter = 0.75f)
      : c_(&c), t_(&tracer), n_max_(max_seeds), region_(region_slots), mode_(mode), param_(parameter) {
    for (auto& b : sets_) b.reset(new buffers(c, max_seeds, region_slots));
  }
  // Queue one round over `seeds`; returns its ticket.
  size_t submit(const std::vector<particle>& seeds) {
    if (seeds.size() > n_max_) throw error(PA_E_CAPACITY, "round_pipeline: more seeds than max_seeds");
    buffers& b = *sets_[next_ % 2];
    b.n = seeds.size();
    if (b.n) c_->check(pa_gpu_memcpy_h2d(c_->get(), b.seeds.data(), seeds.data(), b.n * sizeof(particle)), "h2d");
    c_->check(pa_gpu_tracer_trace_union_round_pipelined(
                  t_->get(), b.seeds.data(), b.n, region_, reinterpret_cast<float*>(b.slots.data()),
                  b.counts.data(), b.cont.data(), mode_, param_, reinterpret_cast<float*>(b.vertices.data()), b.cap,
                  b.indices.data(), b.cap, reinterpret_cast<float*>(b.colors.data()),
                  reinterpret_cast<float*>(b.ends.data()), nullptr, b.result.data()),
              "trace_union_round_pipelined");
    return next_++;
  }
  // Wait for a round (one of the last two submitted) and copy its results.
  particle_tracer::union_round result(size_t ticket) {
    if (ticket >= next_ || ticket + 2 < next_) throw error(PA_E_STATE, "round_pipeline: round no longer held");
    buffers& b = *sets_[ticket % 2];
    c_->check(pa_gpu_synchronize(c_->get()), "synchronize");
    const pa_round_result r = b.result.to_host(1)[0];
    c_->check(pa_gpu_round_result_check(c_->get(), &r, b.cap, b.cap), "round");
    particle_tracer::union_round out;
    if (b.n == 0) return out;
    out.total_steps = r.total_steps;
    out.curves.vertices = b.vertices.to_host(size_t(r.n_vertices));
    out.curves.indices = b.indices.to_host(size_t(r.n_indices));
    out.curves.colors = b.colors.to_host(size_t(r.n_vertices));
    out.continuation = b.cont.to_host(r.n_continuation);
    out.counts = b.counts.to_host(b.n);
    out.endpoints = b.ends.to_host(b.n);
    return out;
  }

 private:
  struct buffers {
    buffers(context& c, size_t n, size_t region)
        : cap(n * region), seeds(c, n), slots(c, (n + 7) / 8 * 8 * region), counts(c, n), cont(c, n), ends(c, n),
          vertices(c, n * region), indices(c, n * region), colors(c, n * region), result(c, 1) {}
    size_t n = 0, cap;
    device_vector<particle> seeds;
    device_vector<vector4> slots;
    device_vector<int32_t> counts;
    device_vector<particle> cont;
    device_vector<vector4> ends;
    device_vector<vector4> vertices;
    device_vector<int32_t> indices;
    device_vector<vector4> colors;
    device_vector<pa_round_result> result;
  };
  context* c_;
  particle_tracer* t_;
  size_t n_max_, region_;
  pa_color_mode mode_;
  scalar param_;
  size_t next_ = 0;
  std::unique_ptr<buffers> sets_[2];
};

// Extension: the device-resident serving loop (pa_gpu_serve_*): requests are
// published into a pinned ring as they arrive and run on a persistent kernel;
// each particle ends where a one-block trace on its field ends (pars serves
// one request at a time, pars/source/service.cpp:20-69).
class resident_server {
 public:
  resident_server(particle_tracer& tracer, const std::vector<const vector_field*>& fields, size_t capacity,
                  int32_t request_slots = 1024, int32_t waves_per_cu = 0)
      : c_(&tracer.owner()) {
    std::vector<const pa_gpu_field*> f;
    for (auto* v : fields) f.push_back(v->get());
    c_->check(pa_gpu_serve_create(tracer.get(), f.data(), int32_t(f.size()), capacity, request_slots, waves_per_cu, 0,
                                  &s_),
              "serve_create");
  }
  ~resident_server() {
    if (s_) pa_gpu_serve_destroy(s_);
  }
  resident_server(const resident_server&) = delete;
  resident_server& operator=(const resident_server&) = delete;
  // Publish a request; false when the ring or the slots are full for now.
  bool submit(const std::vector<particle>& particles, int32_t* request) {
    const pa_status st = pa_gpu_serve_submit(s_, particles.data(), particles.size(), request);
    if (st == PA_E_CAPACITY && particles.size() > 0) return false;
    c_->check(st, "serve_submit");
    return true;
  }
  bool done(int32_t request) {
    int32_t d = 0;
    c_->check(pa_gpu_serve_poll(s_, request, &d), "serve_poll");
    return d != 0;
  }
  // A done request's endpoints and step counts; its slot and ring space are free again.
  void collect(int32_t request, size_t n, std::vector<vector4>* endpoints, std::vector<int32_t>* steps) {
    if (endpoints) endpoints->resize(n);
    if (steps) steps->resize(n);
    uint64_t sum = 0;
    c_->check(pa_gpu_serve_collect(s_, request, endpoints ? &(*endpoints)[0].x : nullptr,
                                   steps ? steps->data() : nullptr, &sum),
              "serve_collect");
  }
  // Drain and stop the kernel (before other GPU work of the process); total steps served.
  uint64_t stop() {
    uint64_t steps = 0, launches = 0;
    c_->check(pa_gpu_serve_stop(s_, &steps, &launches), "serve_stop");
    return steps;
  }

 private:
  context* c_;
  pa_gpu_server* s_ = nullptr;
};

// Extension: particles as the records boost::mpi sends (particle.hpp:16-28,
// 40 bytes each), converted on the device.
struct wire {
  static std::vector<uint8_t> to_wire(context& c, const std::vector<particle>& particles) {
    std::vector<uint8_t> out(particles.size() * 40);
    if (particles.empty()) return out;
    device_vector<particle> d(c, particles);
    device_vector<uint8_t> w(c, out.size());
    c.check(pa_gpu_particles_to_wire(c.get(), d.data(), particles.size(), w.data()), "to_wire");
    return w.to_host(out.size());
  }
  static std::vector<particle> from_wire(context& c, const std::vector<uint8_t>& records) {
    const size_t n = records.size() / 40;
    if (n == 0) return {};
    device_vector<uint8_t> w(c, records);
    device_vector<particle> d(c, n);
    c.check(pa_gpu_particles_from_wire(c.get(), w.data(), n, d.data()), "from_wire");
    return d.to_host(n);
  }
};

// pa::index_generator (index_generator.hpp:10-14).
class index_generator {
 public:
  static void generate(context& c, integral_curves* curves) {
    if (curves->vertices.empty()) throw error(PA_E_INVALID, "index generation of an empty vertex array");
    device_vector<vector4> v(c, curves->vertices);
    device_vector<int32_t> idx(c, curves->vertices.size());
    size_t m = 0;
    c.check(pa_gpu_index_generate(c.get(), reinterpret_cast<const float*>(v.data()), curves->vertices.size(),
                                  idx.data(), &m),
            "index_generate");
    curves->indices = idx.to_host(m);
  }
};

// pa::color_generator (color_generator.hpp:10-32).
class color_generator {
 public:
  static void generate(context& c, integral_curves* curves, pa_color_mode mode = PA_COLOR_RGB,
                       scalar constant_parameter = 0.5f) {
    const size_t n = curves->vertices.size();
    curves->colors.assign(n, vector4{0.0f, 0.0f, 0.0f, 0.0f});
    if (curves->indices.empty()) return;
    device_vector<vector4> v(c, curves->vertices);
    device_vector<int32_t> idx(c, curves->indices);
    device_vector<vector4> col(c, curves->colors);
    c.check(pa_gpu_color_generate(c.get(), reinterpret_cast<const float*>(v.data()), n, idx.data(),
                                  curves->indices.size(), mode, constant_parameter,
                                  reinterpret_cast<float*>(col.data())),
            "color_generate");
    curves->colors = col.to_host(n);
  }
};

// pa::flow_map_generator (flow_map_generator.hpp:12-28) + ftle_map_generator
// (ftle_map_generator.hpp:13-40) for one block without neighbours.
class flow_map_generator {
 public:
  explicit flow_map_generator(context& c) : c_(&c) { c.check(pa_gpu_tracer_create(c.get(), &t_), "tracer_create"); }
  ~flow_map_generator() {
    if (t_) pa_gpu_tracer_destroy(t_);
  }
  flow_map_generator(const flow_map_generator&) = delete;
  flow_map_generator& operator=(const flow_map_generator&) = delete;
  void set_vector_field(const vector_field* f) {
    field_ = f;
    c_->check(pa_gpu_tracer_set_local_vector_field(t_, f->get()), "set_vector_field");
  }
  void set_integrator(pa_integrator i) { c_->check(pa_gpu_tracer_set_integrator(t_, i), "set_integrator"); }
  void set_step_size(scalar s) { c_->check(pa_gpu_tracer_set_step_size(t_, s), "set_step_size"); }
  // Returns the flow map [d0][d1][d2] xyz on the host.
  std::vector<vector3> generate(const std::array<int32_t, 3>& block_size, size_t iterations,
                                scalar resolution_scale, std::array<int32_t, 3>* dims) {
    std::array<int32_t, 3> d;
    for (int i = 0; i < 3; ++i) d[i] = int32_t(float(block_size[i]) * resolution_scale);
    const size_t n = size_t(d[0]) * size_t(d[1]) * size_t(d[2]);
    device_vector<vector3> fm(*c_, n);
    uint64_t steps = 0;
    c_->check(pa_gpu_flow_map_generate(t_, block_size.data(), iterations, resolution_scale, 0,
                                       reinterpret_cast<float*>(fm.data()), d.data(), &steps),
              "flow_map_generate");
    if (dims) *dims = d;
    return fm.to_host(n);
  }

 private:
  context* c_;
  pa_gpu_tracer* t_ = nullptr;
  const vector_field* field_ = nullptr;
};

class ftle_map_generator {
 public:
  enum class mode { regular = PA_FTLE_REGULAR, fractional_anisotropy = PA_FTLE_FRACTIONAL_ANISOTROPY };
  explicit ftle_map_generator(context& c, mode m = mode::regular) : c_(&c), mode_(m) {}
  void set_flow_map(const std::vector<vector3>* flow_map, const std::array<int32_t, 3>& dims, vector3 spacing) {
    flow_map_ = flow_map;
    dims_ = dims;
    spacing_ = spacing;
  }
  void set_time(scalar t) { time_ = t; }
  std::vector<float> generate() {
    const size_t n = flow_map_->size();
    device_vector<vector3> fm(*c_, *flow_map_);
    device_vector<float> out(*c_, n);
    const float h[3] = {spacing_.x, spacing_.y, spacing_.z};
    c_->check(pa_gpu_ftle_generate(c_->get(), reinterpret_cast<const float*>(fm.data()), dims_.data(), h, time_,
                                   pa_ftle_mode(mode_), out.data(), nullptr),
              "ftle_generate");
    return out.to_host(n);
  }

 private:
  context* c_;
  mode mode_;
  const std::vector<vector3>* flow_map_ = nullptr;
  std::array<int32_t, 3> dims_{};
  vector3 spacing_{};
  scalar time_ = 1.0f;
};

}  // namespace pa_amd
