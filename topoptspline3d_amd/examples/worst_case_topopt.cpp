// worst_case_topopt.cpp -- the reference's default optimisation loop (optimization(),
// optimization/optimization.cpp:673-767): per iteration update_stencil, the worst-case
// load by the modified power method (modifiedPM), its sensitivity, OC update; written
// against the reference-named facade.
//
//   worst_case_topopt <library.so> <symbol prefix> <nx> <ny> <nz> <iterations> [out_dir]
//
// Box with x = 0 fixed; load region: the nodes of the x = nx face, force along the face
// normal (+x).  Iteration i draws its start force with seed i.  Prints one line per
// iteration: "itn c_worst v power_steps".  With out_dir, writes the reference's records
// there at the end (optimization.cpp:755-765): cworst, vrec (float64 per iteration),
// flast (worst force on the load nodes), ulast (its displacement).
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
    const size_t ne = (size_t)nx * ny * nz, nn = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    std::vector<uint8_t> solid(ne, 1), support(nn, 0);
    std::vector<int64_t> load;
    std::vector<double> normals;
    for (int z = 0; z <= nz; ++z)
      for (int y = 0; y <= ny; ++y) {
        const size_t row = (size_t)(nx + 1) * (y + (size_t)(ny + 1) * z);
        support[row] = 1;
        load.push_back((int64_t)(row + nx));
        normals.insert(normals.end(), {1.0, 0.0, 0.0});
      }

    Parameter p;
    setParameters(p);
    buildGrids(nx, ny, nz, solid.data(), support.data(), nullptr);
    uploadTemplateMatrix();
    std::vector<double> f0(3 * nn, 0.0);
    getGrids().setLoads(1, f0.data());
    getGrids().setLoadNodes(load, normals);
    initDensities(params().volume_ratio);
    const float Vgoal = params().volume_ratio;
    std::vector<double> cRecord, volRecord;
    for (int itn = 1; itn <= iters; ++itn) {
      update_stencil();
      const double c = modifiedPM((uint64_t)itn);
      const double v = getGrids()[0].volumeRatio();
      std::printf("%d %.12e %.12f %d\n", itn, c, v, getGrids().last_iterations());
      cRecord.push_back(c);
      volRecord.push_back(v);
      computeSensitivity();
      updateDensities(Vgoal);
    }
    if (argc > 7) {
      const std::string dir = argv[7];
      bio::write_vector(dir + "/cworst", cRecord);
      bio::write_vector(dir + "/vrec", volRecord);
      getGrids().writeSupportForce(dir + "/flast");
      getGrids().writeDisplacement(dir + "/ulast");
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
