// oracle_core.hpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of the reference's topology-optimisation hot path.  It is the
// checker for the HIP library and the CPU baseline of bench.py; nothing in the product
// (topoptspline3d_amd/, include/) links, loads or calls it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Every function cites the reference file:line whose behaviour it restates (paths are
// relative to the reference repository).  Where this build deliberately deviates from
// the reference (symmetric V-cycle inside flexible PCG, elimination of fixed DOFs on
// every level, dense Cholesky coarsest solve, deterministic adjoint, volume normalised
// by active elements) the same choice is made in the HIP library, so the oracle checks
// the algorithm the product runs.
#pragma once
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

namespace orc {

struct Level {
  int ex = 0, ey = 0, ez = 0;  // element dims
  int nx = 0, ny = 0, nz = 0;  // node dims = element dims + 1
  int64_t ne = 0, nn = 0;
  std::vector<uint8_t> eact;   // element active
  std::vector<uint8_t> fre;    // node free (carries an unknown)
  std::vector<double> st;      // Galerkin stencil, levels >= 1: 243*nn
  std::vector<double> u, f, r; // V-cycle work vectors, 3*nn
  int64_t nid(int x, int y, int z) const { return x + (int64_t)nx * (y + (int64_t)ny * z); }
  int64_t eid(int x, int y, int z) const { return x + (int64_t)ex * (y + (int64_t)ey * z); }
};

struct SplineAxis {
  int order = 0;
  std::vector<int> idx;     // n_elem * order coefficient index per term
  std::vector<double> wt;   // n_elem * order basis value per term
  std::vector<uint8_t> ok;  // n_elem: element centre inside the knot range
};

struct Mma {
  int n = 0, m = 0;
  double move = 0.5;
  std::vector<double> x, xmin, xmax, L, U, alpha, beta, lastdx, xi, eta;
  std::vector<double> y, lam, mu, s, a, c, d;
  double z = 1, zeta = 1, a0 = 1, eps = 1;
  int stop_counter = 0;
  // subproblem data
  std::vector<std::vector<double>> p, q;  // (m+1) x n
  std::vector<double> plam, qlam, b;
};

struct Ctx {
  int nx, ny, nz;
  double h;
  std::vector<uint8_t> solid, support, shell;
  std::vector<Level> L;
  double E = 1e6, nu = 0.4, penal = 3, rho_min = 1e-3;
  double Ke[24][24];
  double Mc[8][24][24];     // Galerkin element products of the 8 child positions
  std::vector<float> rho, rho_pre, sens, vsens;
  std::vector<double> w;    // rho^p on active elements, 0 elsewhere
  int K = 0;
  std::vector<double> F, U; // K * 3 * nn0
  // recycled start: per case the last rc_depth solutions, newest at rc_head
  int rc_depth = 0;
  std::vector<std::vector<double>> rc_ring;  // [case * rc_depth + slot]
  std::vector<int> rc_head, rc_cnt;
  int64_t n_active = 0;
  bool op_ready = false;
  bool floating = false;    // no supported fine node: the rigid motions are K's null space
  bool skip = false;        // skip-layer hierarchy: level 1 is 4x coarser than level 0
  // coarsest dense system
  std::vector<int64_t> cnode;   // free coarsest nodes
  std::vector<int64_t> cmap;    // node -> compact index or -1
  std::vector<double> chol;     // n x n lower Cholesky factor
  int cn = 0;
  // spline
  int sp_mode = -1, sp_order = 0, sp_param = 0;
  int ncx = 0, ncy = 0, ncz = 0;
  SplineAxis ax[3];
  std::vector<float> coef, csens, cvsens;
  std::vector<float> sscsens, dripcsens;        // self-support / drip coefficient gradients
  std::vector<double> surf_p, surf_n;          // surface points and normals (3 per point)
  double ss_value = 1.0, drip_value = 0.0;     // last surface-constraint values
  double surf_dir = 1.0;                        // orientation of the surface normals
  std::string err;
};

}  // namespace orc
