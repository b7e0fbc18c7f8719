// self_supporting_topopt.cpp -- the reference's self-supporting spline loop
// (TestSuit::testMMAOptimizationSelfSupport with ENABLE_SELFSUPPORT,
// utils/test_utils.cpp:1004-1200) written against the reference-named facade: Heaviside-
// projected spline densities, MMA on the coefficients with three constraints (volume,
// self-support, drip) evaluated on the spline's surface points each iteration.
//
//   self_supporting_topopt <library.so> <symbol prefix> <nx> <ny> <nz> <iterations> <coeffs.f32>
//
// Cantilever: x = 0 fixed, unit total -z load on the bottom edge at x = nx; cubic uniform
// spline with a coefficient every 4 voxels, initial coefficients read from <coeffs.f32>
// (raw float32, one per coefficient).  Prints one line per iteration:
// "itn c v ss drip beta pcg_iterations".
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tos_facade.hpp"

int main(int argc, char** argv) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: %s lib prefix nx ny nz iters coeffs.f32\n", argv[0]);
    return 2;
  }
  try {
    tos::Api::load(argv[1], argv[2]);
    const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), nz = std::atoi(argv[5]), iters = std::atoi(argv[6]);
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
    auto& G = getGrids();
    G.setLoads(1, f.data());
    G.set_spline_lattice(4, 4);
    const int n = G[0].n_cijk();
    std::vector<float> c0(n);
    FILE* fp = std::fopen(argv[7], "rb");
    if (!fp || std::fread(c0.data(), sizeof(float), n, fp) != (size_t)n) {
      std::fprintf(stderr, "cannot read %d coefficients from %s\n", n, argv[7]);
      return 2;
    }
    std::fclose(fp);
    G.setCoeffs(c0.data());

    MMA::mma_t mma(n, 3);
    mma.init(p.min_rho, 1.0, 0.06);
    mma.set_x(std::vector<double>(c0.begin(), c0.end()));
    const double volScale = 1e3;
    double sensScale = 0;
    float beta = 4.f;
    for (int itn = 1; itn <= iters; ++itn) {
      G[0].coeff2density();
      projectDensities(beta);
      const double v = G[0].volumeRatio();
      update_stencil();
      G.solve(params().solve_rtol, params().solve_precision);
      const double c = G[0].compliance();
      const double ssScale = itn > 80 ? 1e1 : itn > 30 ? 1e0 : 1e-1;
      const double dripScale = itn > 80 ? 1e4 : itn > 30 ? 1e3 : 1e2;
      deal_surface_points(beta);
      std::printf("%d %.12e %.12f %.12e %.12e %g %d\n", itn, c, v, G[0].global_selfsupp_constraint(),
                  G[0].global_drip_constraint(), beta, G.last_iterations());
      computeSensitivity2(beta);
      G[0].ddensity2dcoeff_update();
      G[0].dvol2dcoeff();
      if (sensScale == 0) sensScale = 1.0 / c;
      mma.update_on_surface(G, sensScale, volScale, params().volume_ratio, ssScale, dripScale);
      if (itn % 10 == 0 && itn > 2 && beta < 32) beta *= 2;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
