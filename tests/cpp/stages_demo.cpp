// Drives one streamline pass through the C++ stage mirror (pa_amd/stages.hpp)
// the way pars/source/pipeline.cpp:107-260 drives pa::: seeds over the own
// block, one trace round, prune, index, colour; then the same round fused on
// the device, the same seeds through the particle pool in bounded passes,
// two rounds kept in flight (round_pipeline), the seeds as requests of the
// resident server, their wire records, and (given an HDF5 file) a block
// loaded the way data_io loads it.  Used by tests/test_gpu_cpp_mirror.py,
// which compares the files it writes against the oracle and each other.
//
// usage: stages_demo <field.f32> <n> <R> <dt> <out_prefix> [vectors.h5]
//   field.f32: n^3 x 3 float32 (C order, z fastest), spacing 2pi/(n-1)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "pa_amd/stages.hpp"

template <class T>
static void dump(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

int main(int argc, char** argv) {
  if (argc != 6 && argc != 7) {
    std::fprintf(stderr, "usage: %s field.f32 n R dt out_prefix [vectors.h5]\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[2]);
  const int R = std::atoi(argv[3]);
  const float dt = float(std::atof(argv[4]));
  const std::string out = argv[5];
  std::vector<pa_amd::vector3> data(size_t(n) * n * n);
  {
    std::ifstream f(argv[1], std::ios::binary);
    f.read(reinterpret_cast<char*>(data.data()), std::streamsize(data.size() * sizeof(pa_amd::vector3)));
    if (!f) {
      std::fprintf(stderr, "cannot read %s\n", argv[1]);
      return 2;
    }
  }
  try {
    pa_amd::context ctx(0);
    const float h = float(2.0 * 3.14159265358979323846 / (n - 1));
    const float s = float(n) * h;
    pa_amd::vector_field field(ctx, data, {n, n, n}, {0, 0, 0}, {s, s, s}, {h, h, h});
    const float stride = 3.0f * h;
    auto seeds = pa_amd::seed_generator::generate(ctx, {0, 0, 0}, {s, s, s}, {stride, stride, stride}, R, 0);
    pa_amd::particle_tracer tracer(ctx);
    tracer.set_local_vector_field(&field);
    tracer.set_integrator(PA_INTEGRATOR_RUNGE_KUTTA_4);
    tracer.set_step_size(dt);
    pa_amd::particle_tracer::round_info info;
    auto curves = tracer.trace(seeds, &info);
    auto& c = curves.at(0);
    pa_amd::index_generator::generate(ctx, &c);
    pa_amd::color_generator::generate(ctx, &c, PA_COLOR_HSV_CONSTANT_S, 0.75f);
    dump(out + ".seeds", seeds);
    dump(out + ".vertices", c.vertices);
    dump(out + ".indices", c.indices);
    dump(out + ".colors", c.colors);
    // the same round fused on the device (extension), and the particle pool
    pa_amd::particle_tracer fused(ctx);
    fused.set_local_vector_field(&field);
    fused.set_integrator(PA_INTEGRATOR_RUNGE_KUTTA_4);
    fused.set_step_size(dt);
    auto round = fused.trace_union_round(seeds, size_t(R), PA_COLOR_HSV_CONSTANT_S, 0.75f);
    dump(out + ".round_vertices", round.curves.vertices);
    dump(out + ".round_indices", round.curves.indices);
    dump(out + ".round_colors", round.curves.colors);
    dump(out + ".round_endpoints", round.endpoints);
    std::vector<pa_amd::particle> pool_seeds = seeds;
    for (auto& p : pool_seeds) p.vector_field_index = 0;  // field 0 of the pool
    pa_amd::particle_pool pool(ctx, pool_seeds.size());
    pool.init(pool_seeds, 1);
    uint64_t pool_steps = 0;
    while (true) {  // bounded passes of 40 iterations, as a serving loop runs them
      const uint64_t s = pool.run(fused, {&field}, 40);
      pool_steps += s;
      if (s == 0) break;
    }
    dump(out + ".pool_positions", pool.positions());
    // two rounds in flight: all seeds, then the first half (a round's curves are per seed, so the
    // second round's are a prefix of the first's)
    pa_amd::round_pipeline rounds(ctx, fused, seeds.size(), size_t(R), PA_COLOR_HSV_CONSTANT_S, 0.75f);
    const size_t t1 = rounds.submit(seeds);
    const size_t t2 = rounds.submit(std::vector<pa_amd::particle>(seeds.begin(), seeds.begin() + seeds.size() / 2));
    auto r1 = rounds.result(t1);
    auto r2 = rounds.result(t2);
    dump(out + ".pipe1_vertices", r1.curves.vertices);
    dump(out + ".pipe2_vertices", r2.curves.vertices);
    dump(out + ".pipe2_indices", r2.curves.indices);
    // the seeds as two requests of the resident server
    pa_amd::resident_server server(fused, {&field}, 4 * pool_seeds.size(), 8);
    const size_t half = pool_seeds.size() / 2;
    std::vector<pa_amd::particle> qa(pool_seeds.begin(), pool_seeds.begin() + half);
    std::vector<pa_amd::particle> qb(pool_seeds.begin() + half, pool_seeds.end());
    int32_t ra = -1, rb = -1;
    if (!server.submit(qa, &ra) || !server.submit(qb, &rb)) throw pa_amd::error(PA_E_CAPACITY, "server full");
    while (!server.done(ra) || !server.done(rb)) {
    }
    std::vector<pa_amd::vector4> ea, eb;
    server.collect(ra, qa.size(), &ea, nullptr);
    server.collect(rb, qb.size(), &eb, nullptr);
    const uint64_t served = server.stop();
    ea.insert(ea.end(), eb.begin(), eb.end());
    dump(out + ".served_endpoints", ea);
    // wire records and back
    const auto w = pa_amd::wire::to_wire(ctx, seeds);
    dump(out + ".wire", w);
    const auto back = pa_amd::wire::from_wire(ctx, w);
    dump(out + ".from_wire", back);
    if (argc == 7) {  // a block of the HDF5 fixture: rank offset (2, 1, 3), 9 x 8 x 7 ghosted nodes
      auto loaded = pa_amd::vector_field::load(ctx, argv[6], {2, 1, 3}, {9, 8, 7}, {8, 7, 6});
      const std::vector<float> info = {float(loaded->dims[0]), float(loaded->dims[1]), float(loaded->dims[2]),
                                       loaded->offset.x, loaded->offset.y, loaded->offset.z,
                                       loaded->size.x, loaded->size.y, loaded->size.z,
                                       loaded->spacing.x, loaded->spacing.y, loaded->spacing.z};
      dump(out + ".h5_info", info);
    }
    std::printf("round_pipeline: %zu + %zu vertices; server: %llu steps\n", r1.curves.vertices.size(),
                r2.curves.vertices.size(), static_cast<unsigned long long>(served));
    std::printf("stages_demo: %zu seeds, %zu vertices, %zu indices; round %zu vertices, %llu steps; pool %llu steps\n",
                seeds.size(), c.vertices.size(), c.indices.size(), round.curves.vertices.size(),
                static_cast<unsigned long long>(round.total_steps), static_cast<unsigned long long>(pool_steps));
  } catch (const pa_amd::error& e) {
    std::fprintf(stderr, "pa_amd error %d: %s\n", int(e.status), e.what());
    return 1;
  }
  return 0;
}
