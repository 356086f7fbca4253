// dam_break.cpp -- the reference's simple example (examples/simple.cpp:12-62) on
// the drop-in C++ API without rendering: a 1 x 1 world of n x n cells, one fluid
// source filling x in [2/n, 0.35], y in [2/n, 1 - 2/n] once at rest, density 0.01,
// and `steps` solver steps of dt = 0.01, each preceded by FluidDomain::update.
//
// usage: fsg_dam_break n steps scheme out.bin [pic_ratio] [solver] [cg_tol] [cg_max_iter]
//                      [frame.ppm]
//   scheme: 0 semi-Lagrangian, 1 PIC, 2 FLIP, 3 PIC/FLIP; solver, cg_tol and
//   cg_max_iter set the fsg_params fields of those names (defaults: the reference's);
//   frame.ppm: the final frame as the example renders it (a 512 x 512 canvas, cells
//   then particles, examples/simple.cpp:75-80)
//
// out.bin (little endian): int32 n, steps, scheme, n_initial, n_final; float32
// initial particles (n_initial x 4: x, y, vx, vy); float32 final particles; float32
// front u and v (n x n each, row-major i + j*n); uint8 labels (n x n).
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "FluidDomain.h"
#include "FluidSolver.h"
#include "Renderer.h"

namespace {
void dumpParticles(const MarkerParticleSet& set, std::vector<float>& out) {
  out.clear();
  for (const MarkerParticle& p : set) {
    out.push_back(p.posX());
    out.push_back(p.posY());
    out.push_back(p.velX());
    out.push_back(p.velY());
  }
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr,
                 "usage: %s n steps scheme out.bin [pic_ratio] [solver] [cg_tol] [cg_max_iter] "
                 "[frame.ppm]\n",
                 argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[1]);
  const int steps = std::atoi(argv[2]);
  const int scheme = std::atoi(argv[3]);
  const char* out_path = argv[4];
  const MyFloat pic_ratio = argc > 5 ? (MyFloat)std::atof(argv[5]) : (MyFloat)0.05;
  const int solver = argc > 6 ? std::atoi(argv[6]) : FSG_SOLVER_PLAIN;
  if (n < 8 || steps < 0 || scheme < 0 || scheme > 3) {
    std::fprintf(stderr, "bad arguments\n");
    return 2;
  }
  const double delta = 1.0 / double(n);
  try {
    FluidDomain fluid_domain(n, n, 1, 1, 0.01, pic_ratio);
    FluidSolverMemoryPool mem_pool(fluid_domain);
    FluidSolver fluid_solver(mem_pool);
    fsg_params prm = fluid_solver.params();
    prm.solver = solver;
    if (argc > 7) prm.cg_tol = std::atof(argv[7]);
    if (argc > 8) prm.cg_max_iter = std::atoi(argv[8]);
    fluid_solver.setParams(prm);

    const BBox<MyFloat> area = {MyFloat(2.0 / n), MyFloat(0.35), MyFloat(2.0 / n),
                                MyFloat(1 - 2.0 / n)};
    fluid_domain.addFluidSource(FluidSource(area, delta, delta, 0.0, 0.0, 0.0, 1));
    std::vector<float> initial, final_state;
    const MyFloat dt = 0.01;
    for (int s = 0; s < steps; ++s) {
      fluid_domain.update(dt);
      if (s == 0) dumpParticles(fluid_domain.markerParticleSet(), initial);
      switch (scheme) {
        case 0: fluid_solver.stepSemiLagrangian(fluid_domain, dt); break;
        case 1: fluid_solver.stepPIC(fluid_domain, dt); break;
        case 2: fluid_solver.stepFLIP(fluid_domain, dt); break;
        default: fluid_solver.stepPICFLIP(fluid_domain, dt); break;
      }
    }
    dumpParticles(fluid_domain.markerParticleSet(), final_state);
    const MacGrid& g = fluid_domain.macGrid();
    std::vector<float> u((size_t)n * n), v((size_t)n * n);
    std::vector<unsigned char> lab((size_t)n * n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        u[(size_t)i + (size_t)j * n] = g.velXHalfIndexed(i, j);
        v[(size_t)i + (size_t)j * n] = g.velYHalfIndexed(i, j);
        lab[(size_t)i + (size_t)j * n] = (unsigned char)g.cellType(i, j);
      }
    FILE* f = std::fopen(out_path, "wb");
    if (!f) throw std::runtime_error("cannot open output file");
    const int hdr[5] = {n, steps, scheme, (int)initial.size() / 4, (int)final_state.size() / 4};
    std::fwrite(hdr, sizeof(int), 5, f);
    std::fwrite(initial.data(), sizeof(float), initial.size(), f);
    std::fwrite(final_state.data(), sizeof(float), final_state.size(), f);
    std::fwrite(u.data(), sizeof(float), u.size(), f);
    std::fwrite(v.data(), sizeof(float), v.size(), f);
    std::fwrite(lab.data(), 1, lab.size(), f);
    std::fclose(f);
    if (argc > 9) {
      Renderer renderer({0, 1, 0, 1});
      Canvas canvas(512, 512);
      renderer.clearCanvas(canvas);
      renderer.renderGridCellsToCanvas(fluid_domain.macGrid(), canvas);
      renderer.renderParticlesToCanvas(fluid_domain.markerParticleSet(), canvas);
      renderer.writeCanvasToPpm(argv[9], canvas);
    }
    std::printf("dam_break n=%d steps=%d scheme=%d particles=%d\n", n, steps, scheme, hdr[4]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
