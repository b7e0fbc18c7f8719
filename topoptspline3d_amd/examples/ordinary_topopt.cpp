// ordinary_topopt.cpp -- the reference's ordinary SIMP loop (TestSuit::testOrdinaryTopopt,
// utils/test_utils.cpp:604-706) written against the reference-named facade.
//
//   ordinary_topopt <library.so> <symbol prefix> <nx> <ny> <nz> <iterations> [recycle depth]
//
// Cantilever: x = 0 fixed, unit total -z load on the bottom edge at x = nx.  Prints one
// line per iteration: "itn c v pcg_iterations rel_res".  A recycle depth m > 0 starts
// every solve at the Galerkin projection onto the last m solutions (an addition of this
// library; the reference starts from the previous displacement).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tos_facade.hpp"

int main(int argc, char** argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s lib prefix nx ny nz iters\n", argv[0]);
    return 2;
  }
  try {
    tos::Api::load(argv[1], argv[2]);
    const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), nz = std::atoi(argv[5]), iters = std::atoi(argv[6]);
    const int recycle = argc > 7 ? std::atoi(argv[7]) : 0;
    const size_t ne = (size_t)nx * ny * nz, nn = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    std::vector<uint8_t> solid(ne, 1), support(nn, 0);
    for (int z = 0; z <= nz; ++z)
      for (int y = 0; y <= ny; ++y) support[(size_t)(nx + 1) * (y + (size_t)(ny + 1) * z)] = 1;
    std::vector<double> f(3 * nn, 0.0);
    for (int y = 0; y <= ny; ++y) f[2 * nn + nx + (size_t)(nx + 1) * y] = -1.0 / (ny + 1);

    Parameter p;
    p.solve_rtol = 1e-8;
    p.solve_precision = 0;
    setParameters(p);
    buildGrids(nx, ny, nz, solid.data(), support.data(), nullptr);
    uploadTemplateMatrix();
    getGrids().setLoads(1, f.data());
    if (recycle > 0) getGrids().set_start_recycling(recycle);
    initDensities(params().volume_ratio);
    const float Vgoal = params().volume_ratio;
    for (int itn = 1; itn <= iters; ++itn) {
      update_stencil();
      const double rr = getGrids().solve(params().solve_rtol, params().solve_precision);
      const double c = getGrids()[0].compliance();
      const double v = getGrids()[0].volumeRatio();
      std::printf("%d %.12e %.12f %d %.3e\n", itn, c, v, getGrids().last_iterations(), rr);
      computeSensitivity();
      updateDensities(Vgoal);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
