// sanitize_main.cpp -- TEST INFRASTRUCTURE ONLY: drives every part of the FP64 oracle once
// on small domains, built together with oracle.cpp under AddressSanitizer and
// UndefinedBehaviorSanitizer (`make -C oracle sanitize`), so out-of-bounds accesses, leaks
// and undefined arithmetic in the checker show up as a failing exit status.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../include/tos_api.h"

extern "C" {
int orc_ctx_create(const tos_grid_desc*, tos_ctx**);
int orc_ctx_destroy(tos_ctx*);
int orc_set_material(tos_ctx*, double, double, double, double);
int orc_set_loads(tos_ctx*, int, const double*);
int orc_set_density(tos_ctx*, const float*);
int orc_update_operator(tos_ctx*);
int orc_solve(tos_ctx*, const tos_solve_opts*, tos_solve_stats*);
int orc_vcycle(tos_ctx*, const double*, double*, int, int, int);
int orc_apply_operator(tos_ctx*, const double*, double*);
int orc_compliance(tos_ctx*, double*);
int orc_sensitivity(tos_ctx*, int);
int orc_init_vol_sensitivity(tos_ctx*);
int orc_filter(tos_ctx*, int, double);
int orc_oc_update(tos_ctx*, double, double, double, double*, int*);
int orc_project_density(tos_ctx*, double);
int orc_projection_sensitivity(tos_ctx*, int, double);
int orc_oc_update_projected(tos_ctx*, double, double, double, double, double*, int*, double*);
int orc_spline_init(tos_ctx*, int, int, int);
int orc_spline_set_coeffs(tos_ctx*, const float*);
int orc_spline_eval(tos_ctx*);
int orc_spline_adjoint(tos_ctx*, int);
int orc_oc_update_coeffs(tos_ctx*, double, double, double, double, double, double*, int*);
int orc_get_info(tos_ctx*, tos_info*);
int orc_mma_create(int, int, const double*, const double*, double, tos_mma**);
int orc_mma_destroy(tos_mma*);
int orc_mma_set_x(tos_mma*, const double*);
int orc_ctx_mma_step(tos_ctx*, tos_mma*, double, double, double, int*);
int orc_ctx_mma_step_surface(tos_ctx*, tos_mma*, double, double, double, double, double, int*);
int orc_surface_constraints(tos_ctx*, const tos_surface_opts*, tos_surface_stats*);
int orc_worst_compliance(tos_ctx*, const tos_worst_opts*, tos_worst_stats*);
}

#define CHECK(x)                                              \
  do {                                                        \
    const int rc_ = (x);                                      \
    if (rc_ != 0) {                                           \
      std::fprintf(stderr, "%s failed: %d\n", #x, rc_);       \
      return 1;                                               \
    }                                                         \
  } while (0)

int main() {
  const int nx = 12, ny = 6, nz = 6;
  std::vector<uint8_t> solid(nx * ny * nz, 1), support((nx + 1) * (ny + 1) * (nz + 1), 0), shell(solid.size(), 0);
  for (int z = 0; z <= nz; ++z)
    for (int y = 0; y <= ny; ++y) support[(size_t)(nx + 1) * (y + (size_t)(ny + 1) * z)] = 1;
  solid[5] = 0;
  shell[7] = 1;
  tos_grid_desc d{nx, ny, nz, 1.0, solid.data(), support.data(), shell.data(), 0, 0};
  tos_ctx* c = nullptr;
  CHECK(orc_ctx_create(&d, &c));
  CHECK(orc_set_material(c, 1e6, 0.4, 3.0, 1e-3));
  const size_t nn = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
  std::vector<double> f(3 * nn, 0.0), u(3 * nn), z(3 * nn);
  for (int y = 0; y <= ny; ++y) f[2 * nn + nx + (size_t)(nx + 1) * y] = -1.0 / (ny + 1);
  CHECK(orc_set_loads(c, 1, f.data()));
  std::vector<float> rho(solid.size(), 0.4f);
  CHECK(orc_set_density(c, rho.data()));
  CHECK(orc_update_operator(c));
  tos_solve_opts o{1e-8, 200, 0, 1, 1, 0, 1};
  tos_solve_stats st{};
  CHECK(orc_solve(c, &o, &st));
  CHECK(orc_apply_operator(c, f.data(), u.data()));
  CHECK(orc_vcycle(c, f.data(), z.data(), 0, 1, 1));
  double comp = 0, g = 0, ch[2];
  int steps = 0;
  CHECK(orc_compliance(c, &comp));
  CHECK(orc_sensitivity(c, 0));
  CHECK(orc_init_vol_sensitivity(c));
  CHECK(orc_filter(c, 0, 2.0));
  CHECK(orc_oc_update(c, 0.4, 0.03, 0.5, &g, &steps));
  CHECK(orc_project_density(c, 8.0));
  CHECK(orc_projection_sensitivity(c, 0, 8.0));
  CHECK(orc_oc_update_projected(c, 0.4, 0.03, 0.5, 8.0, &g, &steps, ch));
  std::vector<int64_t> nodes;
  std::vector<double> normals;
  for (int y = 1; y < ny; ++y) {
    nodes.push_back(nx + (int64_t)(nx + 1) * (y + (int64_t)(ny + 1) * nz));
    normals.insert(normals.end(), {0.0, 0.0, 1.0});
  }
  tos_worst_opts wo{(int)nodes.size(), nodes.data(), normals.data(), 1, 0, 1, 1, 500, 1e-4, 1e-2};
  tos_worst_stats ws{};
  CHECK(orc_update_operator(c));
  CHECK(orc_worst_compliance(c, &wo, &ws));
  // spline + MMA with and without the surface constraints
  CHECK(orc_spline_init(c, 1, 4, 4));
  tos_info info{};
  CHECK(orc_get_info(c, &info));
  const int n = (int)info.n_coeffs;
  std::vector<float> cf(n);
  for (int j = 0; j < n; ++j) cf[j] = 0.45f + 0.3f * std::sin(0.7f * j);
  CHECK(orc_spline_set_coeffs(c, cf.data()));
  CHECK(orc_spline_eval(c));
  CHECK(orc_update_operator(c));
  CHECK(orc_solve(c, &o, &st));
  CHECK(orc_sensitivity(c, 0));
  CHECK(orc_init_vol_sensitivity(c));
  CHECK(orc_spline_adjoint(c, 0));
  CHECK(orc_spline_adjoint(c, 1));
  std::vector<double> lo(n, 1e-3), hi(n, 1.0), x(cf.begin(), cf.end());
  tos_mma* m1 = nullptr;
  tos_mma* m3 = nullptr;
  CHECK(orc_mma_create(n, 1, lo.data(), hi.data(), 0.06, &m1));
  CHECK(orc_mma_set_x(m1, x.data()));
  int conv = 0;
  CHECK(orc_ctx_mma_step(c, m1, 1.0, 1e3, 0.4, &conv));
  tos_surface_opts so{8, 8.0, 0.75 * M_PI, 4, 7, 8.0, 10.0, 0.09, 4.0, 2.0, 160.0 / 180.0 * M_PI, 4.0, 1.0, 0};
  tos_surface_stats ss{};
  CHECK(orc_surface_constraints(c, &so, &ss));
  CHECK(orc_mma_create(n, 3, lo.data(), hi.data(), 0.06, &m3));
  CHECK(orc_mma_set_x(m3, x.data()));
  CHECK(orc_ctx_mma_step_surface(c, m3, 1.0, 1e3, 0.4, 0.1, 100.0, &conv));
  CHECK(orc_oc_update_coeffs(c, 0.4, 0.05, 0.5, 1e-3, 1.0, &g, &steps));
  CHECK(orc_mma_destroy(m1));
  CHECK(orc_mma_destroy(m3));
  CHECK(orc_ctx_destroy(c));
  std::printf("sanitized oracle run ok: compliance %.6e, %lld surface points\n", comp, (long long)ss.n_points);
  return 0;
}
