// tos_facade.hpp -- header-only C++ surface with the reference's names over the C ABI of
// tos_api.h, so host code written against the reference's optimisation loop compiles
// against this library:
//   grid::HierarchyGrid / grid::Grid   (grid/Grid.h:214-1037)
//   setParameters, buildGrids, uploadTemplateMatrix, initDensities, initCoeffs,
//   update_stencil, computeSensitivity[2], updateDensities[2], updateCoeff,
//   projectDensities, deal_surface_points, getGrids
//                                      (optimization/optimization.h:41-126,
//                                       optimization/optimization.cpp:1103-1128)
//   MMA::mma_t                         (optimization/mma_t.h:159-266)
// The entry points are resolved at run time from a shared library (`tos::Api::load`,
// default prefix "tos_"), so one driver binary can select the library it drives.
// Differences from the reference: the domain comes in as voxel bits plus support/shell
// masks (tos_grid_desc) instead of a mesh file; there is no global device state -- the
// globals below hold one context; errors throw tos::Error instead of exiting.
#pragma once
#include <dlfcn.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "tos_api.h"
#include "tos_bio.hpp"

namespace tos {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// function table of the C ABI
struct Api {
#define TOS_FN(ret, name, ...) ret (*name)(__VA_ARGS__) = nullptr;
  TOS_FN(int, ctx_create, const tos_grid_desc*, tos_ctx**)
  TOS_FN(int, ctx_destroy, tos_ctx*)
  TOS_FN(const char*, last_error, const tos_ctx*)
  TOS_FN(int, get_info, tos_ctx*, tos_info*)
  TOS_FN(int, set_material, tos_ctx*, double, double, double, double)
  TOS_FN(int, set_loads, tos_ctx*, int, const double*)
  TOS_FN(int, set_density, tos_ctx*, const float*)
  TOS_FN(int, get_density, tos_ctx*, float*)
  TOS_FN(int, update_operator, tos_ctx*)
  TOS_FN(int, solve, tos_ctx*, const tos_solve_opts*, tos_solve_stats*)
  TOS_FN(int, get_displacement, tos_ctx*, int, double*)
  TOS_FN(int, compliance, tos_ctx*, double*)
  TOS_FN(int, volume, tos_ctx*, double*)
  TOS_FN(int, sensitivity, tos_ctx*, int)
  TOS_FN(int, get_sensitivity, tos_ctx*, int, float*)
  TOS_FN(int, init_vol_sensitivity, tos_ctx*)
  TOS_FN(int, filter, tos_ctx*, int, double)
  TOS_FN(int, oc_update, tos_ctx*, double, double, double, double*, int*)
  TOS_FN(int, project_density, tos_ctx*, double)
  TOS_FN(int, projection_sensitivity, tos_ctx*, int, double)
  TOS_FN(int, spline_init, tos_ctx*, int, int, int)
  TOS_FN(int, spline_set_coeffs, tos_ctx*, const float*)
  TOS_FN(int, spline_get_coeffs, tos_ctx*, float*)
  TOS_FN(int, spline_eval, tos_ctx*)
  TOS_FN(int, spline_adjoint, tos_ctx*, int)
  TOS_FN(int, oc_update_coeffs, tos_ctx*, double, double, double, double, double, double*, int*)
  TOS_FN(int, mma_create, int, int, const double*, const double*, double, tos_mma**)
  TOS_FN(int, mma_destroy, tos_mma*)
  TOS_FN(int, mma_set_x, tos_mma*, const double*)
  TOS_FN(int, mma_get_x, tos_mma*, double*)
  TOS_FN(int, mma_update, tos_mma*, const double*, const double*, const double*, int*)
  TOS_FN(int, ctx_mma_step, tos_ctx*, tos_mma*, double, double, double, int*)
  TOS_FN(int, worst_compliance, tos_ctx*, const tos_worst_opts*, tos_worst_stats*)
  TOS_FN(int, balance_loads, tos_ctx*, int, int, const int64_t*, const double*)
  TOS_FN(int, remove_rigid, tos_ctx*, int)
  TOS_FN(int, get_loads, tos_ctx*, int, double*)
  TOS_FN(int, oc_update_projected, tos_ctx*, double, double, double, double, double*, int*, double*)
  TOS_FN(int, surface_constraints, tos_ctx*, const tos_surface_opts*, tos_surface_stats*)
  TOS_FN(int, spline_get_coeff_sens, tos_ctx*, int, float*)
  TOS_FN(int, ctx_mma_step_surface, tos_ctx*, tos_mma*, double, double, double, double, double, int*)
  TOS_FN(int, set_start_recycling, tos_ctx*, int)
#undef TOS_FN
  void* handle = nullptr;

  static Api& get() {
    static Api a;
    return a;
  }
  // Resolve every entry point from `path` with symbol prefix `prefix`.
  static Api& load(const std::string& path, const std::string& prefix = "tos_") {
    Api& a = get();
    a.handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!a.handle) throw Error(TOS_E_ARG, std::string("cannot load ") + path + ": " + dlerror());
    auto sym = [&](const char* n) {
      void* f = dlsym(a.handle, (prefix + n).c_str());
      if (!f) throw Error(TOS_E_ARG, prefix + n + " missing in " + path);
      return f;
    };
#define TOS_SYM(name) a.name = reinterpret_cast<decltype(a.name)>(sym(#name));
    TOS_SYM(ctx_create) TOS_SYM(ctx_destroy) TOS_SYM(last_error) TOS_SYM(get_info) TOS_SYM(set_material)
    TOS_SYM(set_loads) TOS_SYM(set_density) TOS_SYM(get_density) TOS_SYM(update_operator) TOS_SYM(solve)
    TOS_SYM(get_displacement) TOS_SYM(compliance) TOS_SYM(volume) TOS_SYM(sensitivity) TOS_SYM(get_sensitivity)
    TOS_SYM(init_vol_sensitivity) TOS_SYM(filter) TOS_SYM(oc_update) TOS_SYM(project_density)
    TOS_SYM(projection_sensitivity) TOS_SYM(spline_init) TOS_SYM(spline_set_coeffs) TOS_SYM(spline_get_coeffs)
    TOS_SYM(spline_eval) TOS_SYM(spline_adjoint) TOS_SYM(oc_update_coeffs) TOS_SYM(mma_create)
    TOS_SYM(mma_destroy) TOS_SYM(mma_set_x) TOS_SYM(mma_get_x) TOS_SYM(mma_update) TOS_SYM(ctx_mma_step)
    TOS_SYM(worst_compliance) TOS_SYM(balance_loads) TOS_SYM(remove_rigid) TOS_SYM(get_loads) TOS_SYM(oc_update_projected) TOS_SYM(surface_constraints)
    TOS_SYM(spline_get_coeff_sens) TOS_SYM(ctx_mma_step_surface) TOS_SYM(set_start_recycling)
#undef TOS_SYM
    return a;
  }
};

inline void check(int rc, tos_ctx* c, const char* where) {
  if (rc != TOS_OK) {
    std::string msg = std::string(where) + " failed (" + std::to_string(rc) + ")";
    if (c) msg += ": " + std::string(Api::get().last_error(c));
    throw Error(rc, msg);
  }
}

}  // namespace tos

// ------------------------------------------------------------------------------------------
// Parameters (optimization/optimization.h:15-35)
// ------------------------------------------------------------------------------------------
struct Parameter {
  float volume_ratio = 0.4f, filter_radius = 2.f, damp_ratio = 0.5f, design_step = 0.03f;
  float youngs_modulu = 1e6f, poisson_ratio = 0.4f, power_penalty = 3.f, min_rho = 1e-3f;
  float min_cijk = 1e-3f, max_cijk = 1.f;
  int spline_order = 3, spline_partition = 10;
  double solve_rtol = 1e-6;   // flexible PCG tolerance (the reference iterates V-cycles to 1e-3)
  int solve_precision = 1;    // 1: FP32 V-cycle inside FP64 PCG; 0: FP64 throughout
};

namespace grid {

// One multigrid level as seen by the reference's drivers: only the fine level (layer 0)
// carries densities and sensitivities, as in the reference.
class Grid {
 public:
  explicit Grid(tos_ctx* c, int layer) : c_(c), layer_(layer) {}
  double compliance() const {
    std::vector<double> cs(n_cases());
    tos::check(tos::Api::get().compliance(c_, cs.data()), c_, "compliance");
    return cs[0];
  }
  double volumeRatio() const {
    double v;
    tos::check(tos::Api::get().volume(c_, &v), c_, "volume");
    return v;
  }
  void filterSensitivity(double r) { tos::check(tos::Api::get().filter(c_, 0, r), c_, "filter"); }
  void filterVolSensitivity(double r) { tos::check(tos::Api::get().filter(c_, 1, r), c_, "filter"); }
  void coeff2density() { tos::check(tos::Api::get().spline_eval(c_), c_, "spline_eval"); }
  void ddensity2dcoeff_update() { tos::check(tos::Api::get().spline_adjoint(c_, 0), c_, "spline_adjoint"); }
  void dvol2dcoeff() { tos::check(tos::Api::get().spline_adjoint(c_, 1), c_, "spline_adjoint"); }
  void init_rho(float r) {
    std::vector<float> v(n_rho(), r);
    tos::check(tos::Api::get().set_density(c_, v.data()), c_, "set_density");
  }
  std::vector<float> getRho() const {
    std::vector<float> v(n_rho());
    tos::check(tos::Api::get().get_density(c_, v.data()), c_, "get_density");
    return v;
  }
  std::vector<float> getSens() const {
    std::vector<float> v(n_rho());
    tos::check(tos::Api::get().get_sensitivity(c_, 0, v.data()), c_, "get_sensitivity");
    return v;
  }
  // surface-constraint values of the last deal_surface_points (grid/Grid.cu:6101-6145,
  // 6485-6525; the reference returns them as float, here the double the library computed)
  double global_selfsupp_constraint() const { return surf_.ss_value; }
  double global_drip_constraint() const { return surf_.drip_value; }
  std::vector<float> getSSCSens() const { return coeff_sens(2); }
  std::vector<float> getDripCSens() const { return coeff_sens(3); }
  std::vector<float> getCSens() const { return coeff_sens(0); }
  std::vector<float> getVolCSens() const { return coeff_sens(1); }
  tos_surface_stats surf_{0, 0, 1.0, 1.0, 0.0};
  int n_rho() const { return (int)info().n_elements; }
  int n_nodes() const { return (int)info().n_nodes; }
  int n_cijk() const { return (int)info().n_coeffs; }
  int n_cases() const { return info().n_cases; }
  int layer() const { return layer_; }
  tos_ctx* ctx() const { return c_; }

 private:
  std::vector<float> coeff_sens(int which) const {
    std::vector<float> v(n_cijk());
    tos::check(tos::Api::get().spline_get_coeff_sens(c_, which, v.data()), c_, "spline_get_coeff_sens");
    return v;
  }
  tos_info info() const {
    tos_info i;
    tos::check(tos::Api::get().get_info(c_, &i), c_, "get_info");
    return i;
  }
  tos_ctx* c_;
  int layer_;
};

class HierarchyGrid {
 public:
  ~HierarchyGrid() {
    if (c_) tos::Api::get().ctx_destroy(c_);
  }
  // genFromMesh(bits, reso) equivalent (grid/Grid.cpp:977-1236) on a box of voxels
  void genFromBits(int nx, int ny, int nz, const uint8_t* solid, const uint8_t* support, const uint8_t* shell,
                   double h = 1.0) {
    tos_grid_desc d{nx, ny, nz, h, solid, support, shell, 0, 0, skip_ ? 1 : 0};
    tos::check(tos::Api::get().ctx_create(&d, &c_), nullptr, "ctx_create");
    tos_info i;
    tos::check(tos::Api::get().get_info(c_, &i), c_, "get_info");
    for (int l = 0; l < i.n_levels; ++l) grids_.emplace_back(c_, l);
  }
  // the reference's set_skip_layer (optimization/optimization.cpp:20), before genFromBits
  void set_skip_layer(bool skip) { skip_ = skip; }
  void setLoads(int n_cases, const double* f) { tos::check(tos::Api::get().set_loads(c_, n_cases, f), c_, "set_loads"); }
  // spline knots over the box (set_spline_partition, grid/Grid.cu:3134-3203)
  void set_spline_partition(int order, int partition) {
    tos::check(tos::Api::get().spline_init(c_, 0, order, partition), c_, "spline_init");
  }
  // uniform lattice: a coefficient every `spacing` voxels (no counterpart in the reference)
  void set_spline_lattice(int order, int spacing) {
    tos::check(tos::Api::get().spline_init(c_, 1, order, spacing), c_, "spline_init");
  }
  void setCoeffs(const float* c) { tos::check(tos::Api::get().spline_set_coeffs(c_, c), c_, "spline_set_coeffs"); }
  void update_stencil() { tos::check(tos::Api::get().update_operator(c_), c_, "update_operator"); }
  // solve to `rtol`; returns the final relative residual (v_cycle returns the residual of
  // one cycle in the reference, grid/Grid.cpp:2399-2444)
  double solve(double rtol, int precision = 1, int maxit = 400) {
    tos_solve_opts o{rtol, maxit, precision, 1, 1, 1, 1};
    tos_solve_stats st;
    tos::check(tos::Api::get().solve(c_, &o, &st), c_, "solve");
    last_iterations_ = st.iterations[0];
    return st.rel_res[0];
  }
  // no counterpart in the reference (its solves start from the previous displacement):
  // every later solve starts at the Galerkin projection of the solution onto the current
  // displacement and the depth - 1 previous solutions (tos_set_start_recycling); call
  // after the grids are built, 0 switches it off
  void set_start_recycling(int depth) {
    tos::check(tos::Api::get().set_start_recycling(c_, depth), c_, "set_start_recycling");
  }
  double v_cycle(int pre = 1, int post = 1) {
    tos_solve_opts o{1e-300, 1, 0, pre, post, 1, 0};
    tos_solve_stats st;
    const int rc = tos::Api::get().solve(c_, &o, &st);
    if (rc != TOS_OK && rc != TOS_E_NOCONV) tos::check(rc, c_, "v_cycle");
    return st.rel_res[0];
  }
  // load region of the worst-case search (setLoadNodes, optimization/projection.cpp:129-268):
  // lexicographic node ids and, unless the force direction is free, one normal per node
  void setLoadNodes(const std::vector<int64_t>& nodes, const std::vector<double>& normals) {
    load_nodes_ = nodes;
    load_normals_ = normals;
  }
  // worst-case compliance by the modified power method (modifiedPM,
  // optimization/optimization.cpp:213-351); load case 0 then holds the worst force
  double worstCompliance(uint64_t seed = 0, int precision = 0) {
    tos_worst_opts o{(int)load_nodes_.size(), load_nodes_.data(),
                     load_normals_.empty() ? nullptr : load_normals_.data(), seed, precision, 1, 1, 500, 1e-4, 1e-2};
    tos_worst_stats st;
    tos::check(tos::Api::get().worst_compliance(c_, &o, &st), c_, "worst_compliance");
    last_iterations_ = st.iterations;
    return st.compliance;
  }
  // records in the reference's bio format (grid/Grid.cpp:1238-1262, 2136-2170): x, y, z
  // interleaved per load node (in the order given to setLoadNodes) or per node
  // (lexicographic), coefficients x fastest
  void writeSupportForce(const std::string& file) {
    const std::vector<double> f = nodeVector(tos::Api::get().get_loads, "get_loads");
    const size_t nn = f.size() / 3;
    std::vector<double> fs[3];
    for (int k = 0; k < 3; ++k)
      for (int64_t n : load_nodes_) fs[k].push_back(f[k * nn + (size_t)n]);
    if (!bio::write_vectors<double, 3>(file, fs)) throw tos::Error(TOS_E_ARG, "cannot write " + file);
  }
  void writeDisplacement(const std::string& file) {
    const std::vector<double> u = nodeVector(tos::Api::get().get_displacement, "get_displacement");
    const size_t nn = u.size() / 3;
    std::vector<double> us[3];
    for (int k = 0; k < 3; ++k) us[k].assign(u.begin() + k * nn, u.begin() + (k + 1) * nn);
    if (!bio::write_vectors<double, 3>(file, us)) throw tos::Error(TOS_E_ARG, "cannot write " + file);
  }
  void writeCoeff(const std::string& file) {
    tos_info i;
    tos::check(tos::Api::get().get_info(c_, &i), c_, "get_info");
    std::vector<float> cf((size_t)i.n_coeffs);
    tos::check(tos::Api::get().spline_get_coeffs(c_, cf.data()), c_, "spline_get_coeffs");
    if (!bio::write_vector(file, cf)) throw tos::Error(TOS_E_ARG, "cannot write " + file);
  }
  int n_grid() const { return (int)grids_.size(); }
  Grid& operator[](int i) { return grids_.at(i); }
  int last_iterations() const { return last_iterations_; }
  // Grid::hasSupport (grid/Grid.h:730): false for a floating body
  bool hasSupport() const {
    tos_info i;
    tos::check(tos::Api::get().get_info(c_, &i), c_, "get_info");
    return i.has_support != 0;
  }
  // forceProject (optimization/projection.cpp:289-350) of load case 0 over the load region
  // given to setLoadNodes
  void forceProject() {
    tos::check(tos::Api::get().balance_loads(c_, 0, (int)load_nodes_.size(), load_nodes_.data(),
                                             load_normals_.empty() ? nullptr : load_normals_.data()),
               c_, "balance_loads");
  }
  // displacementProject (optimization/projection.cpp:374-395), floating bodies only
  void displacementProject() { tos::check(tos::Api::get().remove_rigid(c_, 0), c_, "remove_rigid"); }
  tos_ctx* ctx() const { return c_; }

 private:
  tos_ctx* c_ = nullptr;
  std::vector<Grid> grids_;
  int last_iterations_ = 0;
  bool skip_ = false;
  std::vector<int64_t> load_nodes_;
  std::vector<double> load_normals_;
  template <class Get>
  std::vector<double> nodeVector(Get get, const char* what) {
    tos_info i;
    tos::check(tos::Api::get().get_info(c_, &i), c_, "get_info");
    std::vector<double> v(3 * (size_t)i.n_nodes);
    tos::check(get(c_, 0, v.data()), c_, what);
    return v;
  }
};

}  // namespace grid

// ------------------------------------------------------------------------------------------
// Free functions (optimization/optimization.h:41-126) over one global hierarchy
// ------------------------------------------------------------------------------------------
inline Parameter& params() {
  static Parameter p;
  return p;
}
inline grid::HierarchyGrid& getGrids() {
  static grid::HierarchyGrid g;
  return g;
}
inline void setParameters(const Parameter& p) { params() = p; }
inline void buildGrids(int nx, int ny, int nz, const uint8_t* solid, const uint8_t* support, const uint8_t* shell) {
  getGrids().genFromBits(nx, ny, nz, solid, support, shell);
}
inline void uploadTemplateMatrix() {
  const Parameter& p = params();
  tos::check(tos::Api::get().set_material(getGrids().ctx(), p.youngs_modulu, p.poisson_ratio, p.power_penalty,
                                          p.min_rho),
             getGrids().ctx(), "set_material");
}
inline void initDensities(float v) { getGrids()[0].init_rho(v); }
inline void initCoeffs(float v) {
  auto c = getGrids().ctx();
  const Parameter& p = params();
  tos::check(tos::Api::get().spline_init(c, 0, p.spline_order, p.spline_partition), c, "spline_init");
  std::vector<float> cf(getGrids()[0].n_cijk(), v);
  tos::check(tos::Api::get().spline_set_coeffs(c, cf.data()), c, "spline_set_coeffs");
}
inline void update_stencil() { getGrids().update_stencil(); }
inline double modifiedPM(uint64_t seed = 0) { return getGrids().worstCompliance(seed); }
inline void computeSensitivity() {
  auto c = getGrids().ctx();
  tos::check(tos::Api::get().sensitivity(c, 0), c, "sensitivity");
  tos::check(tos::Api::get().filter(c, 0, params().filter_radius), c, "filter");
}
inline void computeSensitivity2(float beta) {
  auto c = getGrids().ctx();
  auto& A = tos::Api::get();
  tos::check(A.sensitivity(c, 0), c, "sensitivity");
  tos::check(A.init_vol_sensitivity(c), c, "init_vol_sensitivity");
  tos::check(A.projection_sensitivity(c, 0, beta), c, "projection_sensitivity");
  tos::check(A.projection_sensitivity(c, 1, beta), c, "projection_sensitivity");
  tos::check(A.filter(c, 0, params().filter_radius), c, "filter");
  tos::check(A.filter(c, 1, params().filter_radius), c, "filter");
}
inline float updateDensities(float Vgoal) {
  auto c = getGrids().ctx();
  double g;
  int n;
  tos::check(tos::Api::get().oc_update(c, Vgoal, params().design_step, params().damp_ratio, &g, &n), c, "oc_update");
  return (float)g;
}
inline float updateCoeff(float Vgoal) {
  auto c = getGrids().ctx();
  double g;
  int n;
  tos::check(tos::Api::get().oc_update_coeffs(c, Vgoal, params().design_step, params().damp_ratio, params().min_cijk,
                                               params().max_cijk, &g, &n),
             c, "oc_update_coeffs");
  return (float)g;
}
inline void projectDensities(float beta) {
  auto c = getGrids().ctx();
  tos::check(tos::Api::get().project_density(c, beta), c, "project_density");
}
// OC on the design variable behind the projection (updateDensities2, optimization.cu:462-557);
// change[0], change[1]: max and mean design change
inline float updateDensities2(float Vgoal, float beta, float change[2]) {
  auto c = getGrids().ctx();
  double g, ch[2];
  int n;
  tos::check(tos::Api::get().oc_update_projected(c, Vgoal, params().design_step, params().damp_ratio, beta, &g, &n, ch),
             c, "oc_update_projected");
  change[0] = (float)ch[0];
  change[1] = (float)ch[1];
  return (float)g;
}
// Surface points of the spline, their self-supporting and drip constraints and coefficient
// gradients (deal_surface_points, optimization/optimization.cpp:1103-1128), with the
// reference's defaults (64 background cells, modes "oh", print angle 3 pi/4); read the
// results with getGrids()[0].global_selfsupp_constraint() / global_drip_constraint() /
// getSSCSens() / getDripCSens().
inline void deal_surface_points(float beta, int ss_mode = 4, int drip_mode = 4) {
  auto c = getGrids().ctx();
  tos_surface_opts o{64, beta, 0.75 * M_PI, ss_mode, drip_mode, 8.0, 10.0, 0.09, 128.0, 32.0,
                     160.0 / 180.0 * M_PI, 4.0, 10.0, 0};
  tos::check(tos::Api::get().surface_constraints(c, &o, &getGrids()[0].surf_), c, "surface_constraints");
}

// ------------------------------------------------------------------------------------------
// MMA::mma_t (optimization/mma_t.h:159-266)
// ------------------------------------------------------------------------------------------
namespace MMA {
class mma_t {
 public:
  mma_t(int ndim, int ncon) : n_(ndim), m_(ncon) {}
  ~mma_t() {
    if (h_) tos::Api::get().mma_destroy(h_);
  }
  void init(double lower, double upper, double move = 0.5) {
    std::vector<double> lo(n_, lower), hi(n_, upper);
    tos::check(tos::Api::get().mma_create(n_, m_, lo.data(), hi.data(), move, &h_), nullptr, "mma_create");
  }
  void set_x(const std::vector<double>& x) { tos::check(tos::Api::get().mma_set_x(h_, x.data()), nullptr, "mma_set_x"); }
  std::vector<double> get_x() const {
    std::vector<double> x(n_);
    tos::check(tos::Api::get().mma_get_x(h_, x.data()), nullptr, "mma_get_x");
    return x;
  }
  // df: n, dg: m*n constraint-major, g: m; returns convergence (mma_t.cpp:18-51)
  bool update(const double* df, const double* dg, const double* g) {
    int cv = 0;
    tos::check(tos::Api::get().mma_update(h_, df, dg, g, &cv), nullptr, "mma_update");
    return cv != 0;
  }
  // one step of the spline driver on the hierarchy's coefficients
  bool update_on(grid::HierarchyGrid& G, double sens_scale, double vol_scale, double vfrac) {
    int cv = 0;
    tos::check(tos::Api::get().ctx_mma_step(G.ctx(), h_, sens_scale, vol_scale, vfrac, &cv), G.ctx(), "ctx_mma_step");
    return cv != 0;
  }
  // the self-supporting spline driver's step (m = 3: volume, self-support, drip; utils/test_utils.cpp:1135-1188)
  bool update_on_surface(grid::HierarchyGrid& G, double sens_scale, double vol_scale, double vfrac, double ss_scale,
                         double drip_scale) {
    int cv = 0;
    tos::check(tos::Api::get().ctx_mma_step_surface(G.ctx(), h_, sens_scale, vol_scale, vfrac, ss_scale, drip_scale,
                                                    &cv),
               G.ctx(), "ctx_mma_step_surface");
    return cv != 0;
  }
  int n_dim() const { return n_; }
  int n_constrain() const { return m_; }

 private:
  int n_, m_;
  tos_mma* h_ = nullptr;
};
}  // namespace MMA
