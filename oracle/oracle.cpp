// oracle.cpp -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// FP64 CPU restatement of the reference's per-iteration hot path, exported with the
// same C ABI as the HIP library (include/tos_api.h) under the prefix `orc_`, so that
// tests drive both backends with identical calls.  Plain loops; OpenMP over
// independent outputs only; every reduction sums fixed-size chunks in a fixed order so
// results do not depend on the thread count.
#include "oracle_core.hpp"
#include "../include/tos_api.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace orc {

// ------------------------------------------------------------------------------------
// Deterministic reductions
// ------------------------------------------------------------------------------------
static constexpr int64_t kChunk = 8192;

template <class F>
static double chunked_sum(int64_t n, F term) {
  int64_t nch = (n + kChunk - 1) / kChunk;
  std::vector<double> part(std::max<int64_t>(nch, 1), 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nch; ++c) {
    double s = 0;
    int64_t e = std::min(n, (c + 1) * kChunk);
    for (int64_t i = c * kChunk; i < e; ++i) s += term(i);
    part[c] = s;
  }
  double s = 0;
  for (int64_t c = 0; c < nch; ++c) s += part[c];
  return s;
}

static double dot3(const Level& L, const double* a, const double* b) {
  return chunked_sum(3 * L.nn, [&](int64_t i) { return a[i] * b[i]; });
}

// ------------------------------------------------------------------------------------
// Element template: element/templateMatrix.cpp:15-99.
// Ke = sum over the 2x2x2 Gauss points (1 +- 1/sqrt3)/2 of B^T D B (h/2)^3, local node
// i <-> (i&1, (i>>1)&1, i>>2), DOF 3i+axis, engineering shear rows yz, xz, xy.
// ------------------------------------------------------------------------------------
static void element_matrix(double h, double E, double nu, double Ke[24][24]) {
  double D[6][6] = {};
  double f = E / ((1 + nu) * (1 - 2 * nu));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D[i][j] = f * (i == j ? 1 - nu : nu);
  for (int i = 3; i < 6; ++i) D[i][i] = f * (1 - 2 * nu) / 2;
  double g = 1.0 / std::sqrt(3.0);
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) Ke[i][j] = 0;
  for (int gp = 0; gp < 8; ++gp) {
    double xi[3] = {(1 + (2 * (gp & 1) - 1) * g) / 2, (1 + (2 * ((gp >> 1) & 1) - 1) * g) / 2,
                    (1 + (2 * (gp >> 2) - 1) * g) / 2};
    double dN[8][3];
    for (int a = 0; a < 8; ++a) {
      int id[3] = {a & 1, (a >> 1) & 1, a >> 2};
      double r[3];
      for (int k = 0; k < 3; ++k) r[k] = id[k] ? xi[k] : 1 - xi[k];
      for (int k = 0; k < 3; ++k) {
        double s = id[k] ? 1.0 / h : -1.0 / h;
        dN[a][k] = s * r[(k + 1) % 3] * r[(k + 2) % 3];
      }
    }
    double B[6][24] = {};
    for (int a = 0; a < 8; ++a) {
      for (int k = 0; k < 3; ++k) B[k][3 * a + k] = dN[a][k];
      B[3][3 * a + 1] = dN[a][2]; B[3][3 * a + 2] = dN[a][1];  // yz
      B[4][3 * a + 0] = dN[a][2]; B[4][3 * a + 2] = dN[a][0];  // xz
      B[5][3 * a + 0] = dN[a][1]; B[5][3 * a + 1] = dN[a][0];  // xy
    }
    double DB[6][24];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 24; ++j) {
        double s = 0;
        for (int k = 0; k < 6; ++k) s += D[i][k] * B[k][j];
        DB[i][j] = s;
      }
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        double s = 0;
        for (int k = 0; k < 6; ++k) s += B[k][i] * DB[k][j];
        Ke[i][j] += s;
      }
  }
  double sc = (h / 2) * (h / 2) * (h / 2);
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) Ke[i][j] *= sc;
}

// Trilinear prolongation weight of fine local node i (of the fine element at child
// position c inside a coarse element) from coarse local node m: per axis t = c + i in
// {0,1,2}, weight (2-t)/2 for the low coarse node and t/2 for the high one
// (prolongate_correction_kernel, grid/Grid.cu:2594-2632).
static double pweight(int c, int i, int m) {
  double w = 1;
  for (int a = 0; a < 3; ++a) {
    int t = ((c >> a) & 1) + ((i >> a) & 1);
    int ma = (m >> a) & 1;
    w *= ma ? t * 0.5 : (2 - t) * 0.5;
  }
  return w;
}

static void galerkin_child_matrices(const double Ke[24][24], double Mc[8][24][24]) {
  for (int c = 0; c < 8; ++c) {
    double P[8][8];
    for (int i = 0; i < 8; ++i)
      for (int m = 0; m < 8; ++m) P[i][m] = pweight(c, i, m);
    for (int m = 0; m < 8; ++m)
      for (int n = 0; n < 8; ++n)
        for (int r = 0; r < 3; ++r)
          for (int s = 0; s < 3; ++s) {
            double acc = 0;
            for (int i = 0; i < 8; ++i)
              for (int j = 0; j < 8; ++j) acc += P[i][m] * Ke[3 * i + r][3 * j + s] * P[j][n];
            Mc[c][3 * m + r][3 * n + s] = acc;
          }
  }
}

// ------------------------------------------------------------------------------------
// Hierarchy: coarse element active iff any child active (setSolidElementFromFineGrid,
// grid/Grid.cu:4453-4473); node active iff an incident element is active
// (cubeGridSetSolidVertices, grid/Grid.cu:4394-4422); coarsen while the level holds
// more than `threshold` active elements (grid/Grid.cpp:747).
// Free nodes: fine = active and not supported; coarse = some free fine node lies in
// the support of its trilinear basis function (the eliminated Galerkin row is then
// nonzero).
// ------------------------------------------------------------------------------------
static void level_dims(Level& L, int ex, int ey, int ez) {
  L.ex = ex; L.ey = ey; L.ez = ez;
  L.nx = ex + 1; L.ny = ey + 1; L.nz = ez + 1;
  L.ne = (int64_t)ex * ey * ez;
  L.nn = (int64_t)L.nx * L.ny * L.nz;
}

static void build_hierarchy(Ctx& c, int n_levels, int threshold, bool skip) {
  c.L.clear();
  Level L0;
  level_dims(L0, c.nx, c.ny, c.nz);
  L0.eact.assign(c.solid.begin(), c.solid.end());
  for (auto& v : L0.eact) v = v ? 1 : 0;
  c.L.push_back(std::move(L0));
  auto count_active = [](const Level& L) {
    int64_t n = 0;
    for (auto v : L.eact) n += v;
    return n;
  };
  while (true) {
    const Level& F = c.L.back();
    bool more = n_levels > 0 ? (int)c.L.size() < n_levels : count_active(F) > threshold;
    if (!more || (F.ex == 1 && F.ey == 1 && F.ez == 1)) break;
    Level C;
    level_dims(C, (F.ex + 1) / 2, (F.ey + 1) / 2, (F.ez + 1) / 2);
    C.eact.assign(C.ne, 0);
    for (int z = 0; z < F.ez; ++z)
      for (int y = 0; y < F.ey; ++y)
        for (int x = 0; x < F.ex; ++x)
          if (F.eact[F.eid(x, y, z)]) C.eact[C.eid(x / 2, y / 2, z / 2)] = 1;
    c.L.push_back(std::move(C));
  }
  // skip-layer (set_skip_layer, optimization/optimization.cpp:20; grid/Grid.cpp:849-897):
  // the dyadic level 1 is dropped, level 2 follows level 0 with a factor-4 transfer
  c.skip = skip && c.L.size() >= 3;
  if (c.skip) c.L.erase(c.L.begin() + 1);
  // free masks
  for (size_t l = 0; l < c.L.size(); ++l) {
    Level& L = c.L[l];
    L.fre.assign(L.nn, 0);
    if (l == 0) {
      for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
          for (int x = 0; x < L.nx; ++x) {
            bool act = false;
            for (int a = 0; a < 8 && !act; ++a) {
              int ex = x - 1 + (a & 1), ey = y - 1 + ((a >> 1) & 1), ez = z - 1 + (a >> 2);
              if (ex < 0 || ey < 0 || ez < 0 || ex >= L.ex || ey >= L.ey || ez >= L.ez) continue;
              act = L.eact[L.eid(ex, ey, ez)];
            }
            int64_t n = L.nid(x, y, z);
            L.fre[n] = act && !(c.support.size() && c.support[n]);
          }
    } else if (l == 1 && c.skip) {
      // every node of an active element is an unknown: the supports enter the operator as
      // the Dirichlet penalty of the skip-layer Galerkin product
      for (int z = 0; z < L.ez; ++z)
        for (int y = 0; y < L.ey; ++y)
          for (int x = 0; x < L.ex; ++x)
            if (L.eact[L.eid(x, y, z)])
              for (int a = 0; a < 8; ++a) L.fre[L.nid(x + (a & 1), y + ((a >> 1) & 1), z + (a >> 2))] = 1;
    } else {
      const Level& F = c.L[l - 1];
      for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
          for (int x = 0; x < L.nx; ++x) {
            bool fr = false;
            for (int dz = -1; dz <= 1 && !fr; ++dz)
              for (int dy = -1; dy <= 1 && !fr; ++dy)
                for (int dx = -1; dx <= 1 && !fr; ++dx) {
                  int fx = 2 * x + dx, fy = 2 * y + dy, fz = 2 * z + dz;
                  if (fx < 0 || fy < 0 || fz < 0 || fx >= F.nx || fy >= F.ny || fz >= F.nz) continue;
                  fr = F.fre[F.nid(fx, fy, fz)];
                }
            L.fre[L.nid(x, y, z)] = fr;
          }
    }
    L.u.assign(3 * L.nn, 0);
    L.f.assign(3 * L.nn, 0);
    L.r.assign(3 * L.nn, 0);
  }
  // a body with no supported node floats (Grid::hasSupport, grid/Grid.h:730)
  const Level& F0 = c.L[0];
  bool any_free = false, any_support = false;
  for (int64_t n = 0; n < F0.nn; ++n) {
    any_free |= F0.fre[n] != 0;
    if (c.support.size() && c.support[n]) {
      // a supported node counts when it touches an active element
      const int x = (int)(n % F0.nx), y = (int)((n / F0.nx) % F0.ny), z = (int)(n / ((int64_t)F0.nx * F0.ny));
      for (int a = 0; a < 8; ++a) {
        int ex = x - 1 + (a & 1), ey = y - 1 + ((a >> 1) & 1), ez = z - 1 + (a >> 2);
        if (ex < 0 || ey < 0 || ez < 0 || ex >= F0.ex || ey >= F0.ey || ez >= F0.ez) continue;
        any_support |= F0.eact[F0.eid(ex, ey, ez)] != 0;
      }
    }
  }
  c.floating = any_free && !any_support;
}

// ------------------------------------------------------------------------------------
// Rigid motions (optimization/projection.cpp).  Column k of the rigid basis at a node at
// p: k < 3 the unit translation e_k, k >= 3 column k-3 of p^ = [[0,-p2,p1],[p2,0,-p0],
// [-p1,p0,0]] (projection.cpp:70-77, 212-219).
// ------------------------------------------------------------------------------------
static void rigid_column(int k, const double p[3], double out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (k < 3) { out[k] = 1; return; }
  const double P[3][3] = {{0, -p[2], p[1]}, {p[2], 0, -p[0]}, {-p[1], p[0], 0}};
  for (int r = 0; r < 3; ++r) out[r] = P[r][k - 3];
}

// Gram-Schmidt over the columns of R (n rows, 6 columns, column-major) in order; with
// `tol` a column whose residual norm is below 1e-6 of its norm, or whose norm is below
// 1e-6, is dropped (projection.cpp:234-262), else every column is kept (:90-96).
// Returns the kept columns, orthonormal, compacted to the front.
static int gram_schmidt(std::vector<double>& R, int64_t n, bool tol) {
  bool invalid[6] = {false, false, false, false, false, false};
  for (int i = 0; i < 6; ++i) {
    double* ci = &R[(size_t)i * n];
    double old2 = 0;
    for (int64_t t = 0; t < n; ++t) old2 += ci[t] * ci[t];
    for (int j = 0; j < i; ++j) {
      if (invalid[j]) continue;
      const double* cj = &R[(size_t)j * n];
      double w = 0;
      for (int64_t t = 0; t < n; ++t) w += ci[t] * cj[t];
      for (int64_t t = 0; t < n; ++t) ci[t] -= w * cj[t];
    }
    double r2 = 0;
    for (int64_t t = 0; t < n; ++t) r2 += ci[t] * ci[t];
    const double res = std::sqrt(r2), oldres = std::sqrt(old2);
    if (tol && (res / oldres < 1e-6 || oldres < 1e-6)) { invalid[i] = true; continue; }
    for (int64_t t = 0; t < n; ++t) ci[t] /= res;
  }
  int m = 0;
  for (int i = 0; i < 6; ++i) {
    if (invalid[i]) continue;
    if (m != i) std::copy(R.begin() + (size_t)i * n, R.begin() + (size_t)(i + 1) * n, R.begin() + (size_t)m * n);
    ++m;
  }
  return m;
}

// displacementProject (projection.cpp:374-395) with the basis of setNodes (:45-103): the
// rigid motions at the lattice coordinates of every active fine node, orthonormalised,
// removed from u one column after another
static void displacement_project(const Ctx& c, double* u) {
  const Level& L = c.L[0];
  std::vector<int64_t> ids;
  for (int64_t i = 0; i < L.nn; ++i)
    if (L.fre[i]) ids.push_back(i);
  const int64_t n = 3 * (int64_t)ids.size();
  std::vector<double> R(6 * (size_t)n);
  for (size_t q = 0; q < ids.size(); ++q) {
    const int64_t i = ids[q];
    const double p[3] = {(double)(i % L.nx), (double)((i / L.nx) % L.ny), (double)(i / ((int64_t)L.nx * L.ny))};
    for (int k = 0; k < 6; ++k) {
      double col[3];
      rigid_column(k, p, col);
      for (int r = 0; r < 3; ++r) R[(size_t)k * n + 3 * q + r] = col[r];
    }
  }
  const int m = gram_schmidt(R, n, false);
  for (int k = 0; k < m; ++k) {
    const double* ck = &R[(size_t)k * n];
    double s = 0;
    for (size_t q = 0; q < ids.size(); ++q)
      for (int r = 0; r < 3; ++r) s += u[r * L.nn + ids[q]] * ck[3 * q + r];
    for (size_t q = 0; q < ids.size(); ++q)
      for (int r = 0; r < 3; ++r) u[r * L.nn + ids[q]] -= s * ck[3 * q + r];
  }
}

// forceProject on the load-node forces fs (node-major, 3 per node) (projection.cpp:129-268
// for the bases, :300-345 for the projection): Nd holds two unit tangents per node when
// the direction is constrained; Rd the rigid motions at the load nodes, minus their
// tangential part when constrained, orthonormalised with the drop rule; then
//   free force, no support:         f -= Rd Rd^T f
//   free force, with support:       f unchanged
//   constrained, no support:        f -= Nd^T Nd f + Rd Rd^T f
//   constrained, with support:      f -= Nd^T Nd f
static void force_project(const Ctx& c, int n, const int64_t* nodes, const double* nrm, double* fs) {
  const Level& L = c.L[0];
  const bool constrained = nrm != nullptr, support = !c.floating;
  std::vector<double> T(6 * (size_t)n, 0.0);  // two tangents per node
  if (constrained)
    for (int i = 0; i < n; ++i) {
      const double* nv = nrm + 3 * (size_t)i;
      // first tangent: the axis least aligned with the normal, orthogonalised; second: n x t0
      int a = 0;
      for (int k = 1; k < 3; ++k)
        if (std::fabs(nv[k]) < std::fabs(nv[a])) a = k;
      double t0[3] = {0, 0, 0};
      t0[a] = 1;
      const double d = t0[0] * nv[0] + t0[1] * nv[1] + t0[2] * nv[2];
      for (int k = 0; k < 3; ++k) t0[k] -= d * nv[k];
      const double l0 = std::sqrt(t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2]);
      for (int k = 0; k < 3; ++k) t0[k] /= l0;
      const double t1[3] = {nv[1] * t0[2] - nv[2] * t0[1], nv[2] * t0[0] - nv[0] * t0[2], nv[0] * t0[1] - nv[1] * t0[0]};
      for (int k = 0; k < 3; ++k) {
        T[6 * (size_t)i + k] = t0[k];
        T[6 * (size_t)i + 3 + k] = t1[k];
      }
    }
  auto tangential = [&](int i, const double* v, double out[3]) {  // Nd^T Nd v at node i
    const double* t0 = &T[6 * (size_t)i];
    const double* t1 = t0 + 3;
    const double w0 = t0[0] * v[0] + t0[1] * v[1] + t0[2] * v[2];
    const double w1 = t1[0] * v[0] + t1[1] * v[1] + t1[2] * v[2];
    for (int k = 0; k < 3; ++k) out[k] = w0 * t0[k] + w1 * t1[k];
  };
  const int64_t m3 = 3 * (int64_t)n;
  std::vector<double> Rd;
  int m = 0;
  if (!support) {
    Rd.assign(6 * (size_t)m3, 0.0);
    for (int i = 0; i < n; ++i) {
      const int64_t id = nodes[i];
      const double p[3] = {(double)(id % L.nx), (double)((id / L.nx) % L.ny), (double)(id / ((int64_t)L.nx * L.ny))};
      for (int k = 0; k < 6; ++k) {
        double col[3];
        rigid_column(k, p, col);
        if (constrained) {
          double tg[3];
          tangential(i, col, tg);
          for (int r = 0; r < 3; ++r) col[r] -= tg[r];
        }
        for (int r = 0; r < 3; ++r) Rd[(size_t)k * m3 + 3 * i + r] = col[r];
      }
    }
    m = gram_schmidt(Rd, m3, true);
  }
  std::vector<double> w(m, 0.0);
  for (int k = 0; k < m; ++k)
    for (int64_t t = 0; t < m3; ++t) w[k] += Rd[(size_t)k * m3 + t] * fs[t];
  std::vector<double> out(fs, fs + m3);
  if (constrained)
    for (int i = 0; i < n; ++i) {
      double tg[3];
      tangential(i, fs + 3 * (size_t)i, tg);
      for (int r = 0; r < 3; ++r) out[3 * (size_t)i + r] -= tg[r];
    }
  for (int k = 0; k < m; ++k)
    for (int64_t t = 0; t < m3; ++t) out[t] -= Rd[(size_t)k * m3 + t] * w[k];
  std::copy(out.begin(), out.end(), fs);
}

// ------------------------------------------------------------------------------------
// Fine-level matrix-free operator (update_residual_OTFA_WS_kernel grid/Grid.cu:2277-2333,
// applyK_OTFA_kernel :4278-4356): y_i = sum_{e incident} rho_e^p sum_{j in e}
// Ke[loc_i][loc_j] u_j; supported / inactive nodes carry no unknown (row and column
// eliminated), so their u is treated as 0 and their output is 0.
// ------------------------------------------------------------------------------------
static void apply_fine(const Ctx& c, const double* u, double* y) {
  const Level& L = c.L[0];
  const int64_t nn = L.nn;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L.nz; ++z)
    for (int yy = 0; yy < L.ny; ++yy)
      for (int x = 0; x < L.nx; ++x) {
        int64_t i = L.nid(x, yy, z);
        double acc[3] = {0, 0, 0};
        if (L.fre[i]) {
          for (int a = 0; a < 8; ++a) {
            int ex = x - 1 + (a & 1), ey = yy - 1 + ((a >> 1) & 1), ez = z - 1 + (a >> 2);
            if (ex < 0 || ey < 0 || ez < 0 || ex >= L.ex || ey >= L.ey || ez >= L.ez) continue;
            double w = c.w[L.eid(ex, ey, ez)];
            if (w == 0) continue;
            int li = 7 - a;
            double t[3] = {0, 0, 0};
            for (int j = 0; j < 8; ++j) {
              int64_t nj = L.nid(ex + (j & 1), ey + ((j >> 1) & 1), ez + (j >> 2));
              if (!L.fre[nj]) continue;
              for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s) t[r] += c.Ke[3 * li + r][3 * j + s] * u[s * nn + nj];
            }
            for (int r = 0; r < 3; ++r) acc[r] += w * t[r];
          }
        }
        for (int r = 0; r < 3; ++r) y[r * nn + i] = acc[r];
      }
}

static void apply_stencil(const Level& L, const double* u, double* y) {
  const int64_t nn = L.nn;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L.nz; ++z)
    for (int yy = 0; yy < L.ny; ++yy)
      for (int x = 0; x < L.nx; ++x) {
        int64_t i = L.nid(x, yy, z);
        double acc[3] = {0, 0, 0};
        if (L.fre[i]) {
          for (int k = 0; k < 27; ++k) {
            int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
            int jx = x + dx, jy = yy + dy, jz = z + dz;
            if (jx < 0 || jy < 0 || jz < 0 || jx >= L.nx || jy >= L.ny || jz >= L.nz) continue;
            int64_t j = L.nid(jx, jy, jz);
            if (!L.fre[j]) continue;
            for (int r = 0; r < 3; ++r)
              for (int s = 0; s < 3; ++s)
                acc[r] += L.st[(k * 9 + r * 3 + s) * nn + i] * u[s * nn + j];
          }
        }
        for (int r = 0; r < 3; ++r) y[r * nn + i] = acc[r];
      }
}

static void apply_level(const Ctx& c, int l, const double* u, double* y) {
  if (l == 0) apply_fine(c, u, y);
  else apply_stencil(c.L[l], u, y);
}

// ------------------------------------------------------------------------------------
// 8-colour block Gauss-Seidel (gs_relax_OTFA_WS_kernel grid/Grid.cu:1700-1842,
// gs_relax_kernel :1455-1566): colour x%2 + 2(y%2) + 4(z%2) (compute_gscolor
// :1267-1342); per node the 3x3 diagonal block is relaxed component by component
// (x, y, z) with the updated components.  The backward sweep (colours 7..0, components
// z, y, x) is the adjoint of the forward one, which makes the V-cycle symmetric.
// ------------------------------------------------------------------------------------
static void node_block(const Ctx& c, int l, int x, int y, int z, const double* u,
                       double s[3], double D[9]) {
  const Level& L = c.L[l];
  const int64_t nn = L.nn;
  for (int r = 0; r < 3; ++r) s[r] = 0;
  for (int r = 0; r < 9; ++r) D[r] = 0;
  if (l == 0) {
    for (int a = 0; a < 8; ++a) {
      int ex = x - 1 + (a & 1), ey = y - 1 + ((a >> 1) & 1), ez = z - 1 + (a >> 2);
      if (ex < 0 || ey < 0 || ez < 0 || ex >= L.ex || ey >= L.ey || ez >= L.ez) continue;
      double w = c.w[L.eid(ex, ey, ez)];
      if (w == 0) continue;
      int li = 7 - a;
      for (int j = 0; j < 8; ++j) {
        if (j == li) {
          for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) D[r * 3 + q] += w * c.Ke[3 * li + r][3 * li + q];
          continue;
        }
        int64_t nj = L.nid(ex + (j & 1), ey + ((j >> 1) & 1), ez + (j >> 2));
        if (!L.fre[nj]) continue;
        for (int r = 0; r < 3; ++r)
          for (int q = 0; q < 3; ++q) s[r] += w * c.Ke[3 * li + r][3 * j + q] * u[q * nn + nj];
      }
    }
  } else {
    int64_t i = L.nid(x, y, z);
    for (int k = 0; k < 27; ++k) {
      if (k == 13) {
        for (int e = 0; e < 9; ++e) D[e] = L.st[(13 * 9 + e) * nn + i];
        continue;
      }
      int jx = x + k % 3 - 1, jy = y + (k / 3) % 3 - 1, jz = z + k / 9 - 1;
      if (jx < 0 || jy < 0 || jz < 0 || jx >= L.nx || jy >= L.ny || jz >= L.nz) continue;
      int64_t j = L.nid(jx, jy, jz);
      if (!L.fre[j]) continue;
      for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) s[r] += L.st[(k * 9 + r * 3 + q) * nn + i] * u[q * nn + j];
    }
  }
}

static void gs_sweep(const Ctx& c, int l, double* u, const double* f, bool backward) {
  const Level& L = c.L[l];
  const int64_t nn = L.nn;
  for (int ci = 0; ci < 8; ++ci) {
    int col = backward ? 7 - ci : ci;
    int px = col & 1, py = (col >> 1) & 1, pz = col >> 2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = pz; z < L.nz; z += 2)
      for (int y = py; y < L.ny; y += 2)
        for (int x = px; x < L.nx; x += 2) {
          int64_t i = L.nid(x, y, z);
          if (!L.fre[i]) continue;
          double s[3], D[9];
          node_block(c, l, x, y, z, u, s, D);
          double v[3] = {u[i], u[nn + i], u[2 * nn + i]};
          double b[3] = {f[i] - s[0], f[nn + i] - s[1], f[2 * nn + i] - s[2]};
          if (!backward) {
            v[0] = (b[0] - D[1] * v[1] - D[2] * v[2]) / D[0];
            v[1] = (b[1] - D[3] * v[0] - D[5] * v[2]) / D[4];
            v[2] = (b[2] - D[6] * v[0] - D[7] * v[1]) / D[8];
          } else {
            v[2] = (b[2] - D[6] * v[0] - D[7] * v[1]) / D[8];
            v[1] = (b[1] - D[3] * v[0] - D[5] * v[2]) / D[4];
            v[0] = (b[0] - D[1] * v[1] - D[2] * v[2]) / D[0];
          }
          u[i] = v[0]; u[nn + i] = v[1]; u[2 * nn + i] = v[2];
        }
  }
}

// ------------------------------------------------------------------------------------
// Transfers (restrict_residual_kernel grid/Grid.cu:2456-2507, prolongate_correction_kernel
// :2594-2632): trilinear P with weights 1, 1/2, 1/4, 1/8; restriction is P^T.  Both are
// masked to free nodes (P with the rows of eliminated fine nodes removed).
// ------------------------------------------------------------------------------------
// factor q = 2 (dyadic) or 4 (skip-layer, restrict_residual_nondyadic_kernel
// grid/Grid.cu:2509-2570): f_c[I] = sum over fine nodes at q I + d, |d| < q per axis, of
// prod (q - |d|) / q times r_f
static void restrict_to(const Level& F, Level& C, int q = 2) {
  const int64_t nf = F.nn, nc = C.nn;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < C.nz; ++z)
    for (int y = 0; y < C.ny; ++y)
      for (int x = 0; x < C.nx; ++x) {
        int64_t I = C.nid(x, y, z);
        double acc[3] = {0, 0, 0};
        if (C.fre[I]) {
          for (int dz = 1 - q; dz < q; ++dz)
            for (int dy = 1 - q; dy < q; ++dy)
              for (int dx = 1 - q; dx < q; ++dx) {
                int fx = q * x + dx, fy = q * y + dy, fz = q * z + dz;
                if (fx < 0 || fy < 0 || fz < 0 || fx >= F.nx || fy >= F.ny || fz >= F.nz) continue;
                int64_t a = F.nid(fx, fy, fz);
                if (!F.fre[a]) continue;
                double w = (double)(q - std::abs(dx)) * (q - std::abs(dy)) * (q - std::abs(dz)) / ((double)q * q * q);
                for (int r = 0; r < 3; ++r) acc[r] += w * F.r[r * nf + a];
              }
        }
        for (int r = 0; r < 3; ++r) C.f[r * nc + I] = acc[r];
      }
}

// factor-4 prolongation (prolongate_correction_nondyadic_kernel grid/Grid.cu:2635-2671): a
// fine node at 4 X + t takes (4 - t) / 4 of coarse node X and t / 4 of X + 1 per axis
static void prolong_add4(const Level& C, Level& F) {
  const int64_t nf = F.nn, nc = C.nn;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < F.nz; ++z)
    for (int y = 0; y < F.ny; ++y)
      for (int x = 0; x < F.nx; ++x) {
        int64_t a = F.nid(x, y, z);
        if (!F.fre[a]) continue;
        const int p[3] = {x, y, z};
        double acc[3] = {0, 0, 0};
        for (int k = 0; k < 8; ++k) {
          int cc[3];
          double w = 1;
          for (int ax = 0; ax < 3; ++ax) {
            const int t = p[ax] % 4, up = (k >> ax) & 1;
            cc[ax] = p[ax] / 4 + up;
            w *= up ? t / 4.0 : (4 - t) / 4.0;
          }
          if (w == 0) continue;
          const int64_t I = C.nid(cc[0], cc[1], cc[2]);
          for (int r = 0; r < 3; ++r) acc[r] += w * C.u[r * nc + I];
        }
        for (int r = 0; r < 3; ++r) F.u[r * nf + a] += acc[r];
      }
}

static void prolong_add(const Level& C, Level& F) {
  const int64_t nf = F.nn, nc = C.nn;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < F.nz; ++z)
    for (int y = 0; y < F.ny; ++y)
      for (int x = 0; x < F.nx; ++x) {
        int64_t a = F.nid(x, y, z);
        if (!F.fre[a]) continue;
        int cx[2], cy[2], cz[2];
        double wx[2], wy[2], wz[2];
        int nxp = (x & 1) ? 2 : 1, nyp = (y & 1) ? 2 : 1, nzp = (z & 1) ? 2 : 1;
        if (x & 1) { cx[0] = x / 2; cx[1] = x / 2 + 1; wx[0] = wx[1] = 0.5; } else { cx[0] = x / 2; wx[0] = 1; }
        if (y & 1) { cy[0] = y / 2; cy[1] = y / 2 + 1; wy[0] = wy[1] = 0.5; } else { cy[0] = y / 2; wy[0] = 1; }
        if (z & 1) { cz[0] = z / 2; cz[1] = z / 2 + 1; wz[0] = wz[1] = 0.5; } else { cz[0] = z / 2; wz[0] = 1; }
        double acc[3] = {0, 0, 0};
        for (int k = 0; k < nzp; ++k)
          for (int j = 0; j < nyp; ++j)
            for (int i = 0; i < nxp; ++i) {
              int64_t I = C.nid(cx[i], cy[j], cz[k]);
              double w = wx[i] * wy[j] * wz[k];
              for (int r = 0; r < 3; ++r) acc[r] += w * C.u[r * nc + I];
            }
        for (int r = 0; r < 3; ++r) F.u[r * nf + a] += acc[r];
      }
}

// ------------------------------------------------------------------------------------
// Galerkin coarse operators K_c = P^T K P with eliminated DOFs.  The reference assembles
// them per coarse node directly from fine rho (restrict_stencil_nondyadic_OTFA_WS_kernel
// grid/Grid.cu:1112-1225, restrict_stencil_dyadic_kernel :840-896).
// Level 1: per coarse element sum over its active children of rho^p P_c^T Ke P_c, with
// the rows/columns of eliminated fine nodes removed.  Deeper levels: triple product of
// the stored 27-point stencil with trilinear P.
// ------------------------------------------------------------------------------------
static void galerkin_first(Ctx& c) {
  const Level& F = c.L[0];
  Level& C = c.L[1];
  const int64_t nc = C.nn;
  C.st.assign(243 * nc, 0.0);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int Z = 0; Z < C.nz; ++Z)
    for (int Y = 0; Y < C.ny; ++Y)
      for (int X = 0; X < C.nx; ++X) {
        int64_t I = C.nid(X, Y, Z);
        if (!C.fre[I]) continue;
        double acc[27][9] = {};
        for (int a = 0; a < 8; ++a) {  // coarse element E = I - 1 + bits(a)
          int Ex = X - 1 + (a & 1), Ey = Y - 1 + ((a >> 1) & 1), Ez = Z - 1 + (a >> 2);
          if (Ex < 0 || Ey < 0 || Ez < 0 || Ex >= C.ex || Ey >= C.ey || Ez >= C.ez) continue;
          int mI = 7 - a;
          for (int ch = 0; ch < 8; ++ch) {
            int fx = 2 * Ex + (ch & 1), fy = 2 * Ey + ((ch >> 1) & 1), fz = 2 * Ez + (ch >> 2);
            if (fx >= F.ex || fy >= F.ey || fz >= F.ez) continue;
            double w = c.w[F.eid(fx, fy, fz)];
            if (w == 0) continue;
            bool fr[8];
            bool all = true;
            for (int i = 0; i < 8; ++i) {
              fr[i] = F.fre[F.nid(fx + (i & 1), fy + ((i >> 1) & 1), fz + (i >> 2))];
              all = all && fr[i];
            }
            for (int n = 0; n < 8; ++n) {
              int dx = (a & 1) + (n & 1) - 1, dy = ((a >> 1) & 1) + ((n >> 1) & 1) - 1,
                  dz = (a >> 2) + (n >> 2) - 1;
              int k = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
              for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s) {
                  double v;
                  if (all) {
                    v = c.Mc[ch][3 * mI + r][3 * n + s];
                  } else {
                    v = 0;
                    for (int i = 0; i < 8; ++i) {
                      if (!fr[i]) continue;
                      double pi = pweight(ch, i, mI);
                      if (pi == 0) continue;
                      for (int j = 0; j < 8; ++j) {
                        if (!fr[j]) continue;
                        v += pi * c.Ke[3 * i + r][3 * j + s] * pweight(ch, j, n);
                      }
                    }
                  }
                  acc[k][r * 3 + s] += w * v;
                }
            }
          }
        }
        for (int k = 0; k < 27; ++k) {
          int jx = X + k % 3 - 1, jy = Y + (k / 3) % 3 - 1, jz = Z + k / 9 - 1;
          bool ok = jx >= 0 && jy >= 0 && jz >= 0 && jx < C.nx && jy < C.ny && jz < C.nz &&
                    C.fre[C.nid(jx, jy, jz)];
          for (int e = 0; e < 9; ++e) C.st[(k * 9 + e) * nc + I] = ok ? acc[k][e] : 0.0;
        }
      }
}

// Skip-layer level 1 (restrict_stencil_nondyadic_OTFA_WS_kernel grid/Grid.cu:1112-1225):
// K_c = P4^T K~ P4 per coarse node, K~ the fine element matrices rho^p Ke with the rows and
// columns of supported nodes dropped and 1e6 (not scaled by rho^p) on their diagonal, one
// per element (grid/Grid.cu:14, 1199-1207); summed over the 8^3 fine elements around the
// node, weight prod (4 - |d|) / 4 of a fine node at offset d from a coarse node.
static void galerkin_skip(Ctx& c) {
  const Level& F = c.L[0];
  Level& C = c.L[1];
  const int64_t nc = C.nn;
  C.st.assign(243 * nc, 0.0);
  auto w4 = [](int d) { d = std::abs(d); return d < 4 ? (4 - d) / 4.0 : 0.0; };
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int Z = 0; Z < C.nz; ++Z)
    for (int Y = 0; Y < C.ny; ++Y)
      for (int X = 0; X < C.nx; ++X) {
        const int64_t I = C.nid(X, Y, Z);
        if (!C.fre[I]) continue;
        double acc[27][9] = {};
        for (int fz = 4 * Z - 4; fz < 4 * Z + 4; ++fz)
          for (int fy = 4 * Y - 4; fy < 4 * Y + 4; ++fy)
            for (int fx = 4 * X - 4; fx < 4 * X + 4; ++fx) {
              if (fx < 0 || fy < 0 || fz < 0 || fx >= F.ex || fy >= F.ey || fz >= F.ez) continue;
              if (!F.eact[F.eid(fx, fy, fz)]) continue;
              const double w = c.w[F.eid(fx, fy, fz)];
              bool fix[8];
              double wi[8];
              for (int i = 0; i < 8; ++i) {
                const int x = fx + (i & 1), y = fy + ((i >> 1) & 1), z = fz + (i >> 2);
                fix[i] = c.support.size() && c.support[F.nid(x, y, z)];
                wi[i] = w4(x - 4 * X) * w4(y - 4 * Y) * w4(z - 4 * Z);
              }
              // t[j][r][s] = sum_i w_I(i) K~[3i + r][3j + s]
              double t[8][3][3] = {};
              for (int i = 0; i < 8; ++i) {
                if (wi[i] == 0) continue;
                for (int j = 0; j < 8; ++j)
                  for (int r = 0; r < 3; ++r)
                    for (int s = 0; s < 3; ++s) {
                      double v;
                      if (fix[i] || fix[j])
                        v = (i == j && r == s) ? 1e6 : 0.0;
                      else
                        v = w * c.Ke[3 * i + r][3 * j + s];
                      t[j][r][s] += wi[i] * v;
                    }
              }
              // column node J = I + (dx, dy, dz): weight of fine node j from J
              for (int k = 0; k < 27; ++k) {
                const int JX = X + k % 3 - 1, JY = Y + (k / 3) % 3 - 1, JZ = Z + k / 9 - 1;
                for (int j = 0; j < 8; ++j) {
                  const int x = fx + (j & 1), y = fy + ((j >> 1) & 1), z = fz + (j >> 2);
                  const double wj = w4(x - 4 * JX) * w4(y - 4 * JY) * w4(z - 4 * JZ);
                  if (wj == 0) continue;
                  for (int r = 0; r < 3; ++r)
                    for (int s = 0; s < 3; ++s) acc[k][r * 3 + s] += t[j][r][s] * wj;
                }
              }
            }
        for (int k = 0; k < 27; ++k) {
          int jx = X + k % 3 - 1, jy = Y + (k / 3) % 3 - 1, jz = Z + k / 9 - 1;
          bool ok = jx >= 0 && jy >= 0 && jz >= 0 && jx < C.nx && jy < C.ny && jz < C.nz &&
                    C.fre[C.nid(jx, jy, jz)];
          for (int e = 0; e < 9; ++e) C.st[(k * 9 + e) * nc + I] = ok ? acc[k][e] : 0.0;
        }
      }
}

static void galerkin_next(const Level& F, Level& C) {
  const int64_t nf = F.nn, nc = C.nn;
  C.st.assign(243 * nc, 0.0);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int Z = 0; Z < C.nz; ++Z)
    for (int Y = 0; Y < C.ny; ++Y)
      for (int X = 0; X < C.nx; ++X) {
        int64_t I = C.nid(X, Y, Z);
        if (!C.fre[I]) continue;
        double acc[27][9] = {};
        for (int az = -1; az <= 1; ++az)
          for (int ay = -1; ay <= 1; ++ay)
            for (int axx = -1; axx <= 1; ++axx) {
              int fx = 2 * X + axx, fy = 2 * Y + ay, fz = 2 * Z + az;
              if (fx < 0 || fy < 0 || fz < 0 || fx >= F.nx || fy >= F.ny || fz >= F.nz) continue;
              int64_t A = F.nid(fx, fy, fz);
              if (!F.fre[A]) continue;
              double wa = (axx ? 0.5 : 1.0) * (ay ? 0.5 : 1.0) * (az ? 0.5 : 1.0);
              for (int k = 0; k < 27; ++k) {
                int bx = fx + k % 3 - 1, by = fy + (k / 3) % 3 - 1, bz = fz + k / 9 - 1;
                if (bx < 0 || by < 0 || bz < 0 || bx >= F.nx || by >= F.ny || bz >= F.nz) continue;
                if (!F.fre[F.nid(bx, by, bz)]) continue;
                const double* blk[9];
                double bv[9];
                for (int e = 0; e < 9; ++e) { bv[e] = F.st[(k * 9 + e) * nf + A]; (void)blk; }
                // parents of fine node b
                int px[2], py[2], pz[2];
                double wx[2], wy[2], wz[2];
                int npx = (bx & 1) ? 2 : 1, npy = (by & 1) ? 2 : 1, npz = (bz & 1) ? 2 : 1;
                if (bx & 1) { px[0] = bx / 2; px[1] = bx / 2 + 1; wx[0] = wx[1] = 0.5; } else { px[0] = bx / 2; wx[0] = 1; }
                if (by & 1) { py[0] = by / 2; py[1] = by / 2 + 1; wy[0] = wy[1] = 0.5; } else { py[0] = by / 2; wy[0] = 1; }
                if (bz & 1) { pz[0] = bz / 2; pz[1] = bz / 2 + 1; wz[0] = wz[1] = 0.5; } else { pz[0] = bz / 2; wz[0] = 1; }
                for (int q = 0; q < npz; ++q)
                  for (int p = 0; p < npy; ++p)
                    for (int o = 0; o < npx; ++o) {
                      int dx = px[o] - X, dy = py[p] - Y, dz = pz[q] - Z;
                      int ko = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                      double wgt = wa * wx[o] * wy[p] * wz[q];
                      for (int e = 0; e < 9; ++e) acc[ko][e] += wgt * bv[e];
                    }
              }
            }
        for (int k = 0; k < 27; ++k) {
          int jx = X + k % 3 - 1, jy = Y + (k / 3) % 3 - 1, jz = Z + k / 9 - 1;
          bool ok = jx >= 0 && jy >= 0 && jz >= 0 && jx < C.nx && jy < C.ny && jz < C.nz &&
                    C.fre[C.nid(jx, jy, jz)];
          for (int e = 0; e < 9; ++e) C.st[(k * 9 + e) * nc + I] = ok ? acc[k][e] : 0.0;
        }
      }
}

// ------------------------------------------------------------------------------------
// Coarsest level: dense SPD system over the free nodes, Cholesky-factored once per
// operator update.  The reference uses Eigen BDCSVD on the host (buildCoarsestSystem
// grid/Grid.cpp:3614-3669, solve_fem_host :4040-4102); for an SPD matrix both give the
// same solution.
// ------------------------------------------------------------------------------------
static int coarse_factor(Ctx& c) {
  const Level& L = c.L.back();
  if (c.L.size() == 1) return TOS_OK;  // single level: never used
  c.cmap.assign(L.nn, -1);
  c.cnode.clear();
  for (int64_t i = 0; i < L.nn; ++i)
    if (L.fre[i]) { c.cmap[i] = (int64_t)c.cnode.size(); c.cnode.push_back(i); }
  int n = 3 * (int)c.cnode.size();
  c.cn = n;
  c.chol.assign((size_t)n * n, 0.0);
  for (size_t p = 0; p < c.cnode.size(); ++p) {
    int64_t i = c.cnode[p];
    int x = i % L.nx, y = (i / L.nx) % L.ny, z = i / ((int64_t)L.nx * L.ny);
    for (int k = 0; k < 27; ++k) {
      int jx = x + k % 3 - 1, jy = y + (k / 3) % 3 - 1, jz = z + k / 9 - 1;
      if (jx < 0 || jy < 0 || jz < 0 || jx >= L.nx || jy >= L.ny || jz >= L.nz) continue;
      int64_t q = c.cmap[L.nid(jx, jy, jz)];
      if (q < 0) continue;
      for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s)
          c.chol[(size_t)(3 * p + r) * n + 3 * q + s] = L.st[(k * 9 + r * 3 + s) * L.nn + i];
    }
  }
  // symmetrise (the stencil is symmetric up to roundoff)
  double* A = c.chol.data();
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      double v = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
      A[(size_t)i * n + j] = v; A[(size_t)j * n + i] = v;
    }
  std::vector<double> diag0(n);
  for (int j = 0; j < n; ++j) diag0[j] = A[(size_t)j * n + j];
  if (c.floating) {
    // null space of a floating body: factor A + alpha R R^T, R the orthonormalised rigid
    // motions of the coarsest nodes, alpha the mean diagonal; for a right-hand side
    // orthogonal to R this gives A's minimum-norm solution, which the reference's BDCSVD
    // solve returns (grid/Grid.cpp:3654-3661, 4072)
    std::vector<double> R(6 * (size_t)n, 0.0);
    for (size_t q = 0; q < c.cnode.size(); ++q) {
      const int64_t i = c.cnode[q];
      const double p[3] = {(double)(i % L.nx), (double)((i / L.nx) % L.ny), (double)(i / ((int64_t)L.nx * L.ny))};
      for (int k = 0; k < 6; ++k) {
        double col[3];
        rigid_column(k, p, col);
        for (int r = 0; r < 3; ++r) R[(size_t)k * n + 3 * q + r] = col[r];
      }
    }
    const int m = gram_schmidt(R, n, true);
    double alpha = 0;
    for (int j = 0; j < n; ++j) alpha += diag0[j] / n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double t = 0;
        for (int k = 0; k < m; ++k) t += R[(size_t)k * n + i] * R[(size_t)k * n + j];
        A[(size_t)i * n + j] += alpha * t;
      }
  }
  for (int j = 0; j < n; ++j) {
    double d = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    // a pivot below 1e-12 of the unfactored diagonal: floating part / no supports
    if (!(d > 1e-12 * diag0[j])) { c.err = "coarsest Cholesky: non-positive pivot"; return TOS_E_SINGULAR; }
    d = std::sqrt(d);
    A[(size_t)j * n + j] = d;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < n; ++i) {
      double s = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = s / d;
    }
  }
  return TOS_OK;
}

static void coarse_solve(Ctx& c) {
  Level& L = c.L.back();
  int n = c.cn;
  std::vector<double> b(n);
  for (size_t p = 0; p < c.cnode.size(); ++p)
    for (int r = 0; r < 3; ++r) b[3 * p + r] = L.f[r * L.nn + c.cnode[p]];
  const double* A = c.chol.data();
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= A[(size_t)i * n + k] * b[k];
    b[i] = s / A[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < n; ++k) s -= A[(size_t)k * n + i] * b[k];
    b[i] = s / A[(size_t)i * n + i];
  }
  std::fill(L.u.begin(), L.u.end(), 0.0);
  for (size_t p = 0; p < c.cnode.size(); ++p)
    for (int r = 0; r < 3; ++r) L.u[r * L.nn + c.cnode[p]] = b[3 * p + r];
}

// ------------------------------------------------------------------------------------
// V-cycle (HierarchyGrid::v_cycle grid/Grid.cpp:2399-2444): on each level relax from
// zero, restrict the residual, recurse, prolongate, relax with the adjoint sweep.
// Reads level.f, writes level.u.
// ------------------------------------------------------------------------------------
static void vcycle_level(Ctx& c, size_t l, int pre, int post) {
  Level& L = c.L[l];
  if (l + 1 == c.L.size()) {
    if (c.L.size() == 1) {  // single level: smooth only
      std::fill(L.u.begin(), L.u.end(), 0.0);
      for (int s = 0; s < std::max(1, pre); ++s) gs_sweep(c, l, L.u.data(), L.f.data(), false);
      for (int s = 0; s < std::max(1, post); ++s) gs_sweep(c, l, L.u.data(), L.f.data(), true);
      return;
    }
    coarse_solve(c);
    return;
  }
  std::fill(L.u.begin(), L.u.end(), 0.0);
  for (int s = 0; s < pre; ++s) gs_sweep(c, l, L.u.data(), L.f.data(), false);
  apply_level(c, (int)l, L.u.data(), L.r.data());
  for (int64_t i = 0; i < 3 * L.nn; ++i) L.r[i] = L.f[i] - L.r[i];
  const bool four = l == 0 && c.skip;
  restrict_to(L, c.L[l + 1], four ? 4 : 2);
  vcycle_level(c, l + 1, pre, post);
  if (four)
    prolong_add4(c.L[l + 1], L);
  else
    prolong_add(c.L[l + 1], L);
  for (int s = 0; s < post; ++s) gs_sweep(c, l, L.u.data(), L.f.data(), true);
}

static void precondition(Ctx& c, const double* r, double* z, int pre, int post) {
  Level& L = c.L[0];
  std::copy(r, r + 3 * L.nn, L.f.begin());
  vcycle_level(c, 0, pre, post);
  std::copy(L.u.begin(), L.u.end(), z);
}

static void mask_free(const Level& L, double* v) {
  for (int r = 0; r < 3; ++r)
    for (int64_t i = 0; i < L.nn; ++i)
      if (!L.fre[i]) v[r * L.nn + i] = 0;
}

// ------------------------------------------------------------------------------------
// Solve K u = f per load case.  use_pcg = 0 repeats u += B(f - K u) like the drivers'
// V-cycle loop (utils/test_utils.cpp:659-663); use_pcg = 1 runs flexible CG (Polak-
// Ribiere beta) preconditioned by one V-cycle.  Warm start from the stored u.
// ------------------------------------------------------------------------------------
static int solve_case(Ctx& c, int k, const tos_solve_opts& o, int* its, double* relres) {
  Level& L = c.L[0];
  const int64_t n3 = 3 * L.nn;
  double* u = c.U.data() + k * n3;
  const double* f = c.F.data() + k * n3;
  if (!o.warm_start) std::fill(u, u + n3, 0.0);
  mask_free(L, u);
  double nf = std::sqrt(dot3(L, f, f));
  *its = 0;
  if (nf == 0) { std::fill(u, u + n3, 0.0); *relres = 0; return TOS_OK; }
  std::vector<double> r(n3), z(n3), p(n3), q(n3);
  // ORC_PROGRESS set: one line per FCG iteration on stderr (long solves of large domains)
  const bool progress = std::getenv("ORC_PROGRESS") != nullptr;
  apply_fine(c, u, q.data());
  for (int64_t i = 0; i < n3; ++i) r[i] = f[i] - q[i];
  double rn = std::sqrt(dot3(L, r.data(), r.data()));
  if (!std::isfinite(rn)) { c.err = "NaN in residual"; return TOS_E_NAN; }
  int it = 0;
  if (!o.use_pcg) {
    while (rn / nf > o.rtol && it < o.maxit) {
      precondition(c, r.data(), z.data(), o.pre, o.post);
      for (int64_t i = 0; i < n3; ++i) u[i] += z[i];
      apply_fine(c, u, q.data());
      for (int64_t i = 0; i < n3; ++i) r[i] = f[i] - q[i];
      rn = std::sqrt(dot3(L, r.data(), r.data()));
      ++it;
    }
  } else if (rn / nf > o.rtol) {
    precondition(c, r.data(), z.data(), o.pre, o.post);
    p = z;
    double rz = dot3(L, r.data(), z.data());
    while (it < o.maxit) {
      apply_fine(c, p.data(), q.data());
      double pq = dot3(L, p.data(), q.data());
      double alpha = rz / pq;
      for (int64_t i = 0; i < n3; ++i) { u[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
      ++it;
      rn = std::sqrt(dot3(L, r.data(), r.data()));
      if (!std::isfinite(rn)) { c.err = "NaN in PCG"; return TOS_E_NAN; }
      if (progress) std::fprintf(stderr, "[oracle] FCG iteration %d rel_res %.3e\n", it, rn / nf);
      if (rn / nf <= o.rtol) break;
      precondition(c, r.data(), z.data(), o.pre, o.post);
      double zr = dot3(L, z.data(), r.data());
      double zq = dot3(L, z.data(), q.data());
      double beta = -alpha * zq / rz;
      rz = zr;
      for (int64_t i = 0; i < n3; ++i) p[i] = z[i] + beta * p[i];
    }
  }
  apply_fine(c, u, q.data());
  for (int64_t i = 0; i < n3; ++i) r[i] = f[i] - q[i];
  *relres = std::sqrt(dot3(L, r.data(), r.data())) / nf;
  *its = it;
  return *relres <= o.rtol ? TOS_OK : TOS_E_NOCONV;
}

// ------------------------------------------------------------------------------------
// Sensitivity (computeSensitivity_kernel optimization/optimization.cu:96-159):
// s_e = -p rho^(p-1) u_e^T Ke u_e, one deterministic value per element.
// ------------------------------------------------------------------------------------
static double element_energy(const Ctx& c, const double* u, int ex, int ey, int ez) {
  const Level& L = c.L[0];
  double ue[24];
  for (int j = 0; j < 8; ++j) {
    int64_t n = L.nid(ex + (j & 1), ey + ((j >> 1) & 1), ez + (j >> 2));
    for (int s = 0; s < 3; ++s) ue[3 * j + s] = L.fre[n] ? u[s * L.nn + n] : 0.0;
  }
  double e = 0;
  for (int i = 0; i < 24; ++i) {
    double t = 0;
    for (int j = 0; j < 24; ++j) t += c.Ke[i][j] * ue[j];
    e += ue[i] * t;
  }
  return e;
}

// Filter weight 1 - 6r^2 + 8r^3 - 3r^4 with r = d/R over the window floor(R+1/2)
// (filterSensitivity_kernel grid/Grid.cu:4131-4216, weight lambda :4226-4229).
static void filter_field(const Ctx& c, std::vector<float>& v, double R) {
  const Level& L = c.L[0];
  std::vector<float> src(v);
  int Ri = (int)(R + 0.5);
  double R2 = R * R;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L.ez; ++z)
    for (int y = 0; y < L.ey; ++y)
      for (int x = 0; x < L.ex; ++x) {
        int64_t e = L.eid(x, y, z);
        if (!L.eact[e]) { v[e] = 0; continue; }
        double gs = 0, ws = 0;
        for (int dz = -Ri; dz <= Ri; ++dz)
          for (int dy = -Ri; dy <= Ri; ++dy)
            for (int dx = -Ri; dx <= Ri; ++dx) {
              int qx = x + dx, qy = y + dy, qz = z + dz;
              if (qx < 0 || qy < 0 || qz < 0 || qx >= L.ex || qy >= L.ey || qz >= L.ez) continue;
              double r2 = dx * dx + dy * dy + dz * dz;
              if (r2 > R2) continue;
              int64_t q = L.eid(qx, qy, qz);
              if (!L.eact[q]) continue;
              double r = std::sqrt(r2 / R2);
              double w = 1 - 6 * r * r + 8 * r * r * r - 3 * r * r * r * r;
              gs += w * src[q];
              ws += w;
            }
        v[e] = (float)(gs / ws);
      }
}

// ------------------------------------------------------------------------------------
// OC bisection (updateDensities / trySensMultiplier_kernel optimization/optimization.cu
// :355-460): g in [0, max|s|], rho' = clamp(clamp(rho (max(0,-s)/g)^eta, rho-m, rho+m),
// rho_min, 1), shell -> 1; stop at |V - Vgoal| <= 1e-4 or after 31 trials.
// ------------------------------------------------------------------------------------
static double oc_trial(const Ctx& c, double g, double move, double damp, std::vector<float>* out) {
  const Level& L = c.L[0];
  double vs = chunked_sum(L.ne, [&](int64_t e) -> double {
    if (!L.eact[e]) return 0.0;
    double s = c.sens[e];
    double gg = s > 0 ? 0.0 : -s;
    double r0 = c.rho[e];
    double rn = r0 * std::pow(gg / g, damp);
    rn = std::min(std::max(rn, r0 - move), r0 + move);
    rn = std::min(std::max(rn, c.rho_min), 1.0);
    if (c.shell.size() && c.shell[e]) rn = 1;
    float rf = (float)rn;
    if (out) (*out)[e] = rf;
    return (double)rf;
  });
  return vs / (double)c.n_active;
}

// ------------------------------------------------------------------------------------
// Spline tables.
// ------------------------------------------------------------------------------------
// Cox-de Boor (The NURBS Book, algorithm A2.2) as in SplineBasisX grid/Grid.cu:384-410.
static void basis_funs(int span, double x, int p, const std::vector<double>& U, double* N) {
  double left[16], right[16];
  N[0] = 1.0;
  for (int j = 1; j <= p; ++j) {
    left[j] = x - U[span + 1 - j];
    right[j] = U[span + j] - x;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      double temp = N[r] / (right[r + 1] + left[j - r]);
      N[r] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    N[j] = saved;
  }
}

// Cardinal B-spline of order o on [0, o] (uniform knots).
static double cardinal(int o, double t) {
  if (t < 0 || t >= o) return 0.0;
  if (o == 1) return 1.0;
  return (t * cardinal(o - 1, t) + (o - t) * cardinal(o - 1, t - 1)) / (o - 1);
}

static void spline_axis(SplineAxis& A, int mode, int order, int param, int ne, double h, int* ncoef) {
  A.order = order;
  A.idx.assign((size_t)ne * order, 0);
  A.wt.assign((size_t)ne * order, 0.0);
  A.ok.assign(ne, 1);
  if (mode == 0) {
    // clamped knots on the box enlarged by 2*box/(P+1) per side
    // (Grid::set_spline_knot_series grid/Grid.cpp:3253-3377)
    int P = param;
    double bmin = 0, bmax = ne * h, delta = bmax - bmin;
    double lo = bmin - 2.0 * delta / (P + 1), hi = bmax + 2.0 * delta / (P + 1);
    double step = (hi - lo) / (P + 1);
    std::vector<double> U;
    for (int i = 0; i < order; ++i) U.push_back(lo);
    double t = lo;
    for (int i = 0; i < P; ++i) { t += step; U.push_back(t); }
    for (int i = 0; i < order; ++i) U.push_back(hi);
    int nb = order + P;
    *ncoef = nb;
    for (int e = 0; e < ne; ++e) {
      double x = (e + 0.5) * h;
      int i = (int)((x - lo) / step) + order;  // coeff2density calc_node, grid/Grid.cu:3393-3402
      if (i < order || i > nb) { A.ok[e] = 0; continue; }
      double N[16];
      basis_funs(i - 1, x, order - 1, U, N);
      for (int q = 0; q < order; ++q) {
        A.idx[(size_t)e * order + q] = i - order + q;
        A.wt[(size_t)e * order + q] = N[q];
      }
    }
  } else {
    // uniform lattice, coefficient i at node param*i, boundary coefficients replicated
    int s = param;
    int nc = (ne + s - 1) / s + 1;
    *ncoef = nc;
    for (int e = 0; e < ne; ++e) {
      double t = (e + 0.5) / s;                     // lattice coordinate of the centre
      int i0 = (int)std::floor(t - order / 2.0) + 1;
      for (int q = 0; q < order; ++q) {
        int i = i0 + q;
        double w = cardinal(order, t - i + order / 2.0);
        A.idx[(size_t)e * order + q] = std::min(std::max(i, 0), nc - 1);
        A.wt[(size_t)e * order + q] = w;
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// MMA (optimization/mma_t.cu, mma_t.cpp), FP64.
// ------------------------------------------------------------------------------------
static void mma_clamp_asymptotes(Mma& M) {  // mma_t.cu:747-761
  for (int j = 0; j < M.n; ++j) {
    double st = M.xmax[j] - M.xmin[j];
    M.L[j] = std::min(std::max(M.L[j], M.x[j] - 10 * st), M.x[j] - 0.01 * st);
    M.U[j] = std::min(std::max(M.U[j], M.x[j] + 0.01 * st), M.x[j] + 10 * st);
    M.alpha[j] = std::max(M.xmin[j], std::max(M.L[j] + 0.1 * (M.x[j] - M.L[j]), M.x[j] - M.move * st));
    M.beta[j] = std::min(M.xmax[j], std::min(M.U[j] - 0.1 * (M.U[j] - M.x[j]), M.x[j] + M.move * st));
  }
}

static void mma_init_vars(Mma& M) {  // mma_t.cu:797-807 and the constructor mma_t.h:202-214
  M.y.assign(M.m, 1); M.lam.assign(M.m, 1); M.s.assign(M.m, 1); M.mu.assign(M.m, 1);
  M.a.assign(M.m, 0); M.c.assign(M.m, 10); M.d.assign(M.m, 1);
  M.z = 1; M.zeta = 1; M.a0 = 1; M.eps = 1;
  M.xi.assign(M.n, 1); M.eta.assign(M.n, 1);
  for (int j = 0; j < M.n; ++j) {
    M.x[j] = 0.5 * (M.alpha[j] + M.beta[j]);
    M.xi[j] = std::max(1.0 / (M.x[j] - M.alpha[j]), 1.0);
    M.eta[j] = std::max(1.0 / (M.beta[j] - M.x[j]), 1.0);
  }
  for (int i = 0; i < M.m; ++i) M.mu[i] = std::max(M.c[i] / 2, 1.0);
}

struct MmaW {  // w = (x, y, z, lambda, xi, eta, mu, zeta, s)
  std::vector<double> x, y, lam, xi, eta, mu, s;
  double z, zeta;
};

static void mma_pqlam(Mma& M, const std::vector<double>& lam) {  // update_pqlambda mma_t.cu:681-694
  for (int j = 0; j < M.n; ++j) {
    double pl = M.p[0][j], ql = M.q[0][j];
    for (int i = 0; i < M.m; ++i) { pl += lam[i] * M.p[i + 1][j]; ql += lam[i] * M.q[i + 1][j]; }
    M.plam[j] = pl; M.qlam[j] = ql;
  }
}

static std::pair<double, double> mma_kkt(Mma& M, const MmaW& w, double damp) {  // mma_t.cu:574-678
  double sum = 0, mx = -10;
  mma_pqlam(M, w.lam);
  auto acc = [&](double v) { sum += v * v; mx = std::max(mx, std::fabs(v)); };
  double sx = 0, mxx = 0;
  for (int j = 0; j < M.n; ++j) {
    double ux = M.U[j] - w.x[j], xl = w.x[j] - M.L[j];
    double v = M.plam[j] / (ux * ux) - M.qlam[j] / (xl * xl) - w.xi[j] + w.eta[j];
    sx += v * v; mxx = std::max(mxx, std::fabs(v));
  }
  sum += sx; mx = std::max(mx, mxx);
  for (int i = 0; i < M.m; ++i) acc(M.c[i] + M.d[i] * w.y[i] - w.lam[i] - w.mu[i]);
  double la = 0;
  for (int i = 0; i < M.m; ++i) la += w.lam[i] * M.a[i];
  acc(M.a0 - w.zeta - la);
  for (int i = 0; i < M.m; ++i) {
    double g = 0;
    for (int j = 0; j < M.n; ++j)
      g += M.p[i + 1][j] / (M.U[j] - w.x[j]) + M.q[i + 1][j] / (w.x[j] - M.L[j]);
    double v = (g - M.a[i] * w.z - w.y[i] + w.s[i] - M.b[i]) / damp;
    acc(v);
  }
  double s1 = 0, m1 = 0, s2 = 0, m2 = 0;
  for (int j = 0; j < M.n; ++j) {
    double v1 = w.xi[j] * (w.x[j] - M.alpha[j]) - M.eps;
    double v2 = w.eta[j] * (M.beta[j] - w.x[j]) - M.eps;
    s1 += v1 * v1; m1 = std::max(m1, std::fabs(v1));
    s2 += v2 * v2; m2 = std::max(m2, std::fabs(v2));
  }
  sum += s1; mx = std::max(mx, m1);
  sum += s2; mx = std::max(mx, m2);
  for (int i = 0; i < M.m; ++i) acc(w.mu[i] * w.y[i] - M.eps);
  acc(w.zeta * w.z - M.eps);
  for (int i = 0; i < M.m; ++i) acc(w.lam[i] * w.s[i] - M.eps);
  return {std::sqrt(sum), mx};
}

static void solve_small(int n, std::vector<double>& A, std::vector<double>& b) {
  // Gaussian elimination with partial pivoting (the reference uses Eigen
  // colPivHouseholderQr, optimization/mma_t.cpp:83-118; same solution when regular).
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > std::fabs(A[piv * n + k])) piv = i;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < n; ++i) {
      double f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
    b[i] = s / A[i * n + i];
  }
}

static bool mma_update(Mma& M, const double* df, const double* dg, const double* g) {
  const int n = M.n, m = M.m;
  std::vector<double> curx = M.x;
  // mma_subproblem_t::init, mma_t.cu:15-118
  M.p.assign(m + 1, std::vector<double>(n));
  M.q.assign(m + 1, std::vector<double>(n));
  for (int i = 0; i <= m; ++i) {
    const double* d = i == 0 ? df : dg + (size_t)(i - 1) * n;
    for (int j = 0; j < n; ++j) {
      double ux = M.U[j] - M.x[j], xl = M.x[j] - M.L[j];
      double dp = std::max(d[j], 0.0), dm = std::max(-d[j], 0.0);
      double r = 1e-5 / (M.xmax[j] - M.xmin[j]);
      M.p[i][j] = ux * ux * (1.001 * dp + 0.001 * dm + r);
      M.q[i][j] = xl * xl * (0.001 * dp + 1.001 * dm + r);
    }
  }
  M.b.assign(m, 0);
  for (int i = 0; i < m; ++i) {
    double gp = 0;
    for (int j = 0; j < n; ++j) gp += M.p[i + 1][j] / (M.U[j] - M.x[j]) + M.q[i + 1][j] / (M.x[j] - M.L[j]);
    M.b[i] = gp - g[i];
  }
  // reset, mma_t.cu:702-713
  for (int j = 0; j < n; ++j) {
    M.x[j] = 0.5 * (M.alpha[j] + M.beta[j]);
    M.xi[j] = std::max(1.0 / (M.x[j] - M.alpha[j]), 1.0);
    M.eta[j] = std::max(1.0 / (M.beta[j] - M.x[j]), 1.0);
  }
  M.eps = 1; M.zeta = 1; M.z = 1;
  for (int i = 0; i < m; ++i) { M.y[i] = 1; M.lam[i] = 1; M.s[i] = 1; M.mu[i] = std::max(M.c[i] / 2, 1.0); }
  M.plam.assign(n, 0); M.qlam.assign(n, 0);
  mma_pqlam(M, M.lam);

  // mma_subproblem_t::solve, mma_t.cu:330-499
  double lasterr = 0;
  int nsub = 0;
  bool overflow = false;
  std::vector<double> Dx(n), dx(n), dxi(n), deta(n), G((size_t)m * n), gp(m), dlx(n);
  for (int lev = 0; lev < 6 && !overflow; ++lev) {  // eps = 1 .. 1e-5
    double err = 1e30, damp = 1;
    int stays = 0;
    while (err > 0.9 * M.eps) {
      for (int j = 0; j < n; ++j) {
        double ux = M.U[j] - M.x[j], xl = M.x[j] - M.L[j];
        double psi = 2 * M.plam[j] / (ux * ux * ux) + 2 * M.qlam[j] / (xl * xl * xl);
        Dx[j] = psi + M.xi[j] / (M.x[j] - M.alpha[j]) + M.eta[j] / (M.beta[j] - M.x[j]);
        for (int i = 0; i < m; ++i)
          G[(size_t)i * n + j] = M.p[i + 1][j] / (ux * ux) - M.q[i + 1][j] / (xl * xl);
        double phi = M.plam[j] / (ux * ux) - M.qlam[j] / (xl * xl);
        dlx[j] = -(M.eps / (M.x[j] - M.alpha[j]) - M.eps / (M.beta[j] - M.x[j]) - phi);
      }
      for (int i = 0; i < m; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += M.p[i + 1][j] / (M.U[j] - M.x[j]) + M.q[i + 1][j] / (M.x[j] - M.L[j]);
        gp[i] = s;
      }
      std::vector<double> Dy(m), dly(m), dll(m);
      for (int i = 0; i < m; ++i) {
        Dy[i] = M.d[i] + M.mu[i] / M.y[i];
        dly[i] = -(M.lam[i] + M.eps / M.y[i] - M.c[i] - M.d[i] * M.y[i]);
        dll[i] = -(M.z * M.a[i] + M.y[i] + M.b[i] - M.eps / M.lam[i] - gp[i]);
      }
      double la = 0;
      for (int i = 0; i < m; ++i) la += M.lam[i] * M.a[i];
      double dlz = -(la + M.eps / M.z - M.a0);
      // reduced (m+1) system, mma_t.cu:228-328
      std::vector<double> A((m + 1) * (m + 1), 0.0), rhs(m + 1);
      for (int i = 0; i < m; ++i) {
        A[i * (m + 1) + i] += M.s[i] / M.lam[i] + 1.0 / Dy[i];
        for (int k = 0; k < m; ++k) {
          double s = 0;
          for (int j = 0; j < n; ++j) s += G[(size_t)i * n + j] * (G[(size_t)k * n + j] / Dx[j]);
          A[i * (m + 1) + k] += s;
        }
        double s = 0;
        for (int j = 0; j < n; ++j) s += G[(size_t)i * n + j] * (dlx[j] / Dx[j]);
        rhs[i] = dll[i] + dly[i] / Dy[i] - s;
        A[i * (m + 1) + m] = M.a[i];
        A[m * (m + 1) + i] = M.a[i];
      }
      A[m * (m + 1) + m] = -M.zeta / M.z;
      rhs[m] = dlz;
      solve_small(m + 1, A, rhs);
      std::vector<double> dlam(rhs.begin(), rhs.begin() + m);
      double dz = rhs[m];
      for (int j = 0; j < n; ++j) {
        double v = -dlx[j] / Dx[j];
        for (int i = 0; i < m; ++i) v -= dlam[i] * G[(size_t)i * n + j] / Dx[j];
        dx[j] = v;
      }
      std::vector<double> dy(m), dmu(m), ds(m);
      for (int i = 0; i < m; ++i) dy[i] = dlam[i] / Dy[i] - dly[i] / Dy[i];
      for (int j = 0; j < n; ++j) {
        double xa = M.x[j] - M.alpha[j], bx = M.beta[j] - M.x[j];
        dxi[j] = M.eps / xa - dx[j] / xa * M.xi[j] - M.xi[j];
        deta[j] = M.eps / bx + M.eta[j] / bx * dx[j] - M.eta[j];
      }
      for (int i = 0; i < m; ++i) {
        dmu[i] = M.eps / M.y[i] - dy[i] * M.mu[i] / M.y[i] - M.mu[i];
        ds[i] = M.eps / M.lam[i] - M.s[i] / M.lam[i] * dlam[i] - M.s[i];
      }
      double dzeta = M.eps / M.z - M.zeta - (M.zeta / M.z) * dz;
      // step bound (fraction 0.99 to the boundary)
      const double cf = 1.0 / 0.99;
      double t1 = -1e300, t2 = -1e300, t3 = -1e300;
      for (int j = 0; j < n; ++j) {
        t1 = std::max(t1, -cf * (dx[j] / (M.x[j] - M.alpha[j])));
        t2 = std::max(t2, cf * (dx[j] / (M.beta[j] - M.x[j])));
      }
      auto art = [&](double dv, double v) { t3 = std::max(t3, -cf * dv / v); };
      for (int i = 0; i < m; ++i) art(dy[i], M.y[i]);
      art(dz, M.z);
      for (int i = 0; i < m; ++i) art(dlam[i], M.lam[i]);
      for (int j = 0; j < n; ++j) art(dxi[j], M.xi[j]);
      for (int j = 0; j < n; ++j) art(deta[j], M.eta[j]);
      for (int i = 0; i < m; ++i) art(dmu[i], M.mu[i]);
      art(dzeta, M.zeta);
      for (int i = 0; i < m; ++i) art(ds[i], M.s[i]);
      double step = 1.0 / std::max(std::max(t1, t2), std::max(t3, 1.0));
      // backtracking line search on the KKT 2-norm, mma_t.cu:501-571
      MmaW cur{M.x, M.y, M.lam, M.xi, M.eta, M.mu, M.s, M.z, M.zeta};
      double old_err = mma_kkt(M, cur, 1.0).first;
      double new_err = 1e30, new_inf = 1e30;
      while (step > 1e-12 && new_err > old_err) {
        MmaW t;
        t.x.resize(n); t.xi.resize(n); t.eta.resize(n);
        t.y.resize(m); t.lam.resize(m); t.mu.resize(m); t.s.resize(m);
        for (int j = 0; j < n; ++j) {
          t.x[j] = M.x[j] + dx[j] * step;
          t.xi[j] = M.xi[j] + dxi[j] * step;
          t.eta[j] = M.eta[j] + deta[j] * step;
        }
        for (int i = 0; i < m; ++i) {
          t.y[i] = M.y[i] + dy[i] * step;
          t.lam[i] = M.lam[i] + dlam[i] * step;
          t.mu[i] = M.mu[i] + dmu[i] * step;
          t.s[i] = M.s[i] + ds[i] * step;
        }
        t.z = M.z + dz * step;
        t.zeta = M.zeta + dzeta * step;
        auto e2 = mma_kkt(M, t, damp);
        new_err = e2.first; new_inf = e2.second;
        if (new_err < old_err) {
          M.x = t.x; M.y = t.y; M.z = t.z; M.lam = t.lam; M.xi = t.xi; M.eta = t.eta;
          M.mu = t.mu; M.zeta = t.zeta; M.s = t.s;
          break;
        }
        step /= 2;
      }
      err = new_inf;
      if ((err - lasterr) / err < 1e-5) ++stays; else stays = 0;
      lasterr = err;
      if (stays > 6) { damp *= 10; stays = 0; }
      if (nsub++ > 80) { overflow = true; break; }
    }
    M.eps *= 0.1;
  }
  // adjust_asym, mma_t.cu:164-226
  std::vector<double> cdx(n);
  for (int j = 0; j < n; ++j) cdx[j] = M.x[j] - curx[j];
  for (int j = 0; j < n; ++j) {
    double gam = 1;
    double d1 = cdx[j], d2 = M.lastdx[j];
    if (d1 * d2 > 0) gam = 1.2;
    if (d1 * d2 < 0) gam = 0.7;
    double st = M.xmax[j] - M.xmin[j];
    double xv = M.x[j];
    double al = xv - gam * (xv - d1 - M.L[j]);
    double au = xv + gam * (M.U[j] - xv + d1);
    al = std::min(std::max(al, xv - 10 * st), xv - 0.01 * st);
    au = std::min(std::max(au, xv + 0.01 * st), xv + 10 * st);
    M.L[j] = al; M.U[j] = au;
    double na = M.xmin[j];
    na = std::max(na, al + 0.1 * (xv - al));
    na = std::max(na, xv - M.move * st);
    double nb = M.xmax[j];
    nb = std::min(nb, au - 0.1 * (au - xv));
    nb = std::min(nb, xv + M.move * st);
    M.alpha[j] = na; M.beta[j] = nb;
  }
  M.lastdx = cdx;
  double nrm = 0;
  for (double v : cdx) nrm += v * v;
  if (std::sqrt(nrm) < 1e-3) M.stop_counter++;
  return M.stop_counter > 10;
}

}  // namespace orc

using namespace orc;

// ======================================================================================
// C ABI (same signatures as include/tos_api.h, prefix orc_)
// ======================================================================================
struct tos_ctx : orc::Ctx {};
struct tos_mma : orc::Mma {};

#define GUARD(ctx)                        \
  if (!(ctx)) return TOS_E_ARG;           \
  try {
#define END_GUARD(ctx)                    \
  } catch (const std::exception& ex) {    \
    (ctx)->err = ex.what();               \
    return TOS_E_ARG;                     \
  }

ORC_API int orc_ctx_create(const tos_grid_desc* d, tos_ctx** out) {
  if (!d || !out || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || !d->solid || !(d->h > 0)) return TOS_E_ARG;
  tos_ctx* c = new tos_ctx();
  c->nx = d->nx; c->ny = d->ny; c->nz = d->nz; c->h = d->h;
  int64_t ne = (int64_t)d->nx * d->ny * d->nz;
  int64_t nn = (int64_t)(d->nx + 1) * (d->ny + 1) * (d->nz + 1);
  c->solid.assign(d->solid, d->solid + ne);
  if (d->support) c->support.assign(d->support, d->support + nn);
  if (d->shell) c->shell.assign(d->shell, d->shell + ne);
  build_hierarchy(*c, d->n_levels, d->coarsen_threshold > 0 ? d->coarsen_threshold : 400, d->skip_layer != 0);
  c->n_active = 0;
  for (auto v : c->L[0].eact) c->n_active += v;
  c->rho.assign(ne, 0.0f);
  c->rho_pre.assign(ne, 0.0f);
  c->sens.assign(ne, 0.0f);
  c->vsens.assign(ne, 0.0f);
  c->w.assign(ne, 0.0);
  element_matrix(c->h, c->E, c->nu, c->Ke);
  galerkin_child_matrices(c->Ke, c->Mc);
  c->K = 1;
  c->F.assign(3 * nn, 0.0);
  c->U.assign(3 * nn, 0.0);
  *out = c;
  return TOS_OK;
}

ORC_API int orc_ctx_destroy(tos_ctx* c) { delete c; return TOS_OK; }
ORC_API const char* orc_last_error(const tos_ctx* c) { return c ? c->err.c_str() : "null context"; }

ORC_API int orc_get_info(tos_ctx* c, tos_info* o) {
  if (!c || !o) return TOS_E_ARG;
  std::memset(o, 0, sizeof(*o));
  o->n_levels = (int)c->L.size();
  for (size_t l = 0; l < c->L.size() && l < 16; ++l) {
    o->level_dims[l][0] = c->L[l].ex; o->level_dims[l][1] = c->L[l].ey; o->level_dims[l][2] = c->L[l].ez;
    int64_t na = 0, nf = 0;
    for (auto v : c->L[l].eact) na += v;
    for (auto v : c->L[l].fre) nf += v;
    o->n_active_elements[l] = na; o->n_free_nodes[l] = nf;
  }
  o->n_elements = c->L[0].ne; o->n_nodes = c->L[0].nn;
  int64_t nfc = 0;
  for (auto v : c->L.back().fre) nfc += v;
  o->coarse_dofs = c->L.size() > 1 ? (int)(3 * nfc) : 0;
  o->n_cases = c->K;
  o->n_coeffs = (int64_t)c->ncx * c->ncy * c->ncz;
  o->has_support = c->floating ? 0 : 1;
  return TOS_OK;
}

ORC_API int orc_get_node_flags(tos_ctx* c, int level, uint8_t* out) {
  if (!c || !out || level < 0 || level >= (int)c->L.size()) return TOS_E_ARG;
  const Level& L = c->L[level];
  for (int z = 0; z < L.nz; ++z)
    for (int y = 0; y < L.ny; ++y)
      for (int x = 0; x < L.nx; ++x) {
        int64_t n = L.nid(x, y, z);
        bool act = false;
        for (int a = 0; a < 8 && !act; ++a) {
          int ex = x - 1 + (a & 1), ey = y - 1 + ((a >> 1) & 1), ez = z - 1 + (a >> 2);
          if (ex < 0 || ey < 0 || ez < 0 || ex >= L.ex || ey >= L.ey || ez >= L.ez) continue;
          act = L.eact[L.eid(ex, ey, ez)];
        }
        bool sup = level == 0 && c->support.size() && c->support[n];
        int col = (x & 1) + 2 * (y & 1) + 4 * (z & 1);
        out[n] = (uint8_t)((act ? 1 : 0) | (sup ? 2 : 0) | (L.fre[n] ? 4 : 0) | (col << 3));
      }
  return TOS_OK;
}

ORC_API int orc_set_material(tos_ctx* c, double E, double nu, double p, double rmin) {
  if (!c || !(E > 0) || !(nu > -1 && nu < 0.5) || !(p >= 1) || !(rmin > 0)) return TOS_E_ARG;
  c->E = E; c->nu = nu; c->penal = p; c->rho_min = rmin;
  element_matrix(c->h, E, nu, c->Ke);
  galerkin_child_matrices(c->Ke, c->Mc);
  c->op_ready = false;
  return TOS_OK;
}

ORC_API int orc_get_template_matrix(tos_ctx* c, double* ke) {
  if (!c || !ke) return TOS_E_ARG;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) ke[i * 24 + j] = c->Ke[i][j];
  return TOS_OK;
}

ORC_API int orc_set_loads(tos_ctx* c, int K, const double* f) {
  if (!c || K <= 0 || K > 64 || !f) return TOS_E_ARG;
  const Level& L = c->L[0];
  int64_t n3 = 3 * L.nn;
  c->K = K;
  c->F.assign(f, f + K * n3);
  for (int k = 0; k < K; ++k) mask_free(L, c->F.data() + k * n3);
  c->U.assign(K * n3, 0.0);
  return TOS_OK;
}

ORC_API int orc_get_loads(tos_ctx* c, int k, double* f) {
  if (!c || !f || k < 0 || k >= c->K) return TOS_E_ARG;
  int64_t n3 = 3 * c->L[0].nn;
  std::copy(c->F.begin() + k * n3, c->F.begin() + (k + 1) * n3, f);
  return TOS_OK;
}

ORC_API int orc_set_density(tos_ctx* c, const float* r) {
  if (!c || !r) return TOS_E_ARG;
  for (int64_t e = 0; e < c->L[0].ne; ++e) {
    if (!std::isfinite(r[e])) { c->err = "non-finite density"; return TOS_E_NAN; }
    c->rho[e] = c->L[0].eact[e] ? r[e] : 0.0f;
  }
  c->op_ready = false;
  return TOS_OK;
}

ORC_API int orc_get_density(tos_ctx* c, float* r) {
  if (!c || !r) return TOS_E_ARG;
  std::copy(c->rho.begin(), c->rho.end(), r);
  return TOS_OK;
}

ORC_API int orc_update_operator(tos_ctx* c) {
  if (!c) return TOS_E_ARG;
  const Level& L0 = c->L[0];
  for (int64_t e = 0; e < L0.ne; ++e)
    c->w[e] = L0.eact[e] ? std::pow((double)c->rho[e], c->penal) : 0.0;
  if (c->L.size() > 1) {
    if (c->skip)
      galerkin_skip(*c);
    else
      galerkin_first(*c);
  }
  for (size_t l = 2; l < c->L.size(); ++l) galerkin_next(c->L[l - 1], c->L[l]);
  int rc = coarse_factor(*c);
  if (rc) return rc;
  c->op_ready = true;
  return TOS_OK;
}

ORC_API int orc_get_stencil(tos_ctx* c, int level, double* out) {
  if (!c || !out || level < 1 || level >= (int)c->L.size() || !c->op_ready) return TOS_E_ARG;
  std::copy(c->L[level].st.begin(), c->L[level].st.end(), out);
  return TOS_OK;
}

ORC_API int orc_apply_operator(tos_ctx* c, const double* u, double* y) {
  if (!c || !u || !y) return TOS_E_ARG;
  if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
  std::vector<double> uu(u, u + 3 * c->L[0].nn);
  mask_free(c->L[0], uu.data());
  apply_fine(*c, uu.data(), y);
  return TOS_OK;
}

ORC_API int orc_vcycle(tos_ctx* c, const double* r, double* z, int precision, int pre, int post) {
  (void)precision;
  if (!c || !r || !z) return TOS_E_ARG;
  if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
  std::vector<double> rr(r, r + 3 * c->L[0].nn);
  mask_free(c->L[0], rr.data());
  precondition(*c, rr.data(), z, pre, post);
  return TOS_OK;
}

// ------------------------------------------------------------------------------------
// Recycled start (no reference counterpart): with depth m the warm start u of case k is
// replaced by the Galerkin projection of the solution onto span{u, s_2, ..., s_m}, s_i
// the i-th newest stored solution of the case: u0 = V c, (V^T K V) c = V^T f.  The Gram
// system is scaled to unit diagonal and solved through its eigenvalues (cyclic Jacobi,
// eigenvalues below 1e-10 of the largest dropped).
// ------------------------------------------------------------------------------------
static bool rc_coefficients(int m, const double* A, const double* b, double* cf) {
  double d[8], S[8][8], Q[8][8], bs[8];
  for (int i = 0; i < m; ++i) d[i] = A[i * m + i] > 0 ? 1.0 / std::sqrt(A[i * m + i]) : 0.0;
  for (int i = 0; i < m; ++i) {
    bs[i] = b[i] * d[i];
    for (int j = 0; j < m; ++j) {
      S[i][j] = A[i * m + j] * d[i] * d[j];
      Q[i][j] = i == j;
    }
  }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0, dia = 0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) (i == j ? dia : off) += S[i][j] * S[i][j];
    if (off <= 1e-30 * dia) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        if (S[p][q] == 0) continue;
        // rotation in the (p, q) plane that zeroes S[p][q]: t = tan of its angle, the
        // smaller root of t^2 + 2 th t - 1 = 0
        const double th = (S[q][q] - S[p][p]) / (2 * S[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
        for (int k = 0; k < m; ++k) {
          const double a = S[k][p], e = S[k][q];
          S[k][p] = cs * a - sn * e;
          S[k][q] = sn * a + cs * e;
        }
        for (int k = 0; k < m; ++k) {
          const double a = S[p][k], e = S[q][k];
          S[p][k] = cs * a - sn * e;
          S[q][k] = sn * a + cs * e;
        }
        for (int k = 0; k < m; ++k) {
          const double a = Q[k][p], e = Q[k][q];
          Q[k][p] = cs * a - sn * e;
          Q[k][q] = sn * a + cs * e;
        }
      }
  }
  double lmax = 0;
  for (int k = 0; k < m; ++k) lmax = std::max(lmax, S[k][k]);
  if (!(lmax > 0)) return false;
  double ch[8] = {0};
  for (int k = 0; k < m; ++k) {
    if (!(S[k][k] > 1e-10 * lmax)) continue;
    double w = 0;
    for (int i = 0; i < m; ++i) w += Q[i][k] * bs[i];
    w /= S[k][k];
    for (int i = 0; i < m; ++i) ch[i] += w * Q[i][k];
  }
  for (int i = 0; i < m; ++i) {
    cf[i] = ch[i] * d[i];
    if (!std::isfinite(cf[i])) return false;
  }
  return true;
}

static void rc_reserve(Ctx& c) {
  if ((int)c.rc_ring.size() == c.K * c.rc_depth) return;
  c.rc_ring.assign((size_t)c.K * c.rc_depth, std::vector<double>(3 * c.L[0].nn, 0.0));
  c.rc_head.assign(c.K, 0);
  c.rc_cnt.assign(c.K, 0);
}

static const std::vector<double>& rc_slot(const Ctx& c, int k, int i) {  // i = 1: newest
  return c.rc_ring[(size_t)k * c.rc_depth + (c.rc_head[k] - (i - 1) + c.rc_depth) % c.rc_depth];
}

static void rc_start(Ctx& c, int k) {
  rc_reserve(c);
  const Level& L = c.L[0];
  const int64_t n3 = 3 * L.nn;
  double* u = c.U.data() + k * n3;
  const double* f = c.F.data() + k * n3;
  mask_free(L, u);
  std::vector<const double*> V{u};
  for (int i = 2; i <= c.rc_cnt[k] && (int)V.size() < c.rc_depth; ++i) V.push_back(rc_slot(c, k, i).data());
  const int m = (int)V.size();
  std::vector<std::vector<double>> KV(m, std::vector<double>(n3));
  for (int a = 0; a < m; ++a) apply_fine(c, V[a], KV[a].data());
  double A[64], b[8], cf[8];
  for (int a = 0; a < m; ++a) {
    for (int e = a; e < m; ++e) A[a * m + e] = A[e * m + a] = dot3(L, V[a], KV[e].data());
    b[a] = dot3(L, V[a], f);
  }
  if (!rc_coefficients(m, A, b, cf)) return;
  std::vector<double> u0(n3, 0.0);
  for (int64_t i = 0; i < n3; ++i) {
    double s = 0;
    for (int a = 0; a < m; ++a) s = std::fma(cf[a], V[a][i], s);
    u0[i] = s;
  }
  std::copy(u0.begin(), u0.end(), u);
}

static void rc_push(Ctx& c, int k) {
  rc_reserve(c);
  const int64_t n3 = 3 * c.L[0].nn;
  c.rc_head[k] = (c.rc_head[k] + 1) % c.rc_depth;
  c.rc_cnt[k] = std::min(c.rc_cnt[k] + 1, c.rc_depth);
  std::vector<double>& s = c.rc_ring[(size_t)k * c.rc_depth + c.rc_head[k]];
  std::copy(c.U.begin() + k * n3, c.U.begin() + (k + 1) * n3, s.begin());
}

ORC_API int orc_set_start_recycling(tos_ctx* c, int depth) {
  if (!c || depth < 0 || depth > 8) return TOS_E_ARG;
  if (depth != c->rc_depth) {
    c->rc_depth = depth;
    c->rc_ring.clear();
  }
  return TOS_OK;
}

ORC_API int orc_solve(tos_ctx* c, const tos_solve_opts* o, tos_solve_stats* st) {
  if (!c || !o) return TOS_E_ARG;
  if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
  auto t0 = std::chrono::steady_clock::now();
  int rc_all = TOS_OK;
  for (int k = 0; k < c->K; ++k) {
    int its; double rr;
    if (c->rc_depth > 0 && o->warm_start) rc_start(*c, k);
    int rc = solve_case(*c, k, *o, &its, &rr);
    if (st && k < 64) { st->iterations[k] = its; st->rel_res[k] = rr; }
    if (rc == TOS_E_NAN) return rc;
    if (rc) rc_all = rc;
    if (c->rc_depth > 0 && rc == TOS_OK) rc_push(*c, k);
  }
  if (st) st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc_all;
}

ORC_API int orc_get_displacement(tos_ctx* c, int k, double* u) {
  if (!c || !u || k < 0 || k >= c->K) return TOS_E_ARG;
  int64_t n3 = 3 * c->L[0].nn;
  std::copy(c->U.begin() + k * n3, c->U.begin() + (k + 1) * n3, u);
  return TOS_OK;
}

ORC_API int orc_set_displacement(tos_ctx* c, int k, const double* u) {
  if (!c || !u || k < 0 || k >= c->K) return TOS_E_ARG;
  int64_t n3 = 3 * c->L[0].nn;
  std::copy(u, u + n3, c->U.begin() + k * n3);
  mask_free(c->L[0], c->U.data() + k * n3);
  return TOS_OK;
}

ORC_API int orc_compliance(tos_ctx* c, double* out) {
  if (!c || !out) return TOS_E_ARG;
  const Level& L = c->L[0];
  int64_t n3 = 3 * L.nn;
  for (int k = 0; k < c->K; ++k) out[k] = dot3(L, c->F.data() + k * n3, c->U.data() + k * n3);
  return TOS_OK;
}

ORC_API int orc_volume(tos_ctx* c, double* v) {
  if (!c || !v) return TOS_E_ARG;
  const Level& L = c->L[0];
  double s = chunked_sum(L.ne, [&](int64_t e) { return L.eact[e] ? (double)c->rho[e] : 0.0; });
  *v = s / (double)c->n_active;
  return TOS_OK;
}

// ------------------------------------------------------------------------------------
// Worst-case load by the modified power method: modifiedPM (optimization/
// optimization.cpp:213-351) with forceProject (optimization/projection.cpp:289-350) on a
// grid with supports.  Start force: random on the load nodes (Grid::randForce,
// grid/Grid.cu:3018-3021, here splitmix64 instead of a time-seeded cuRAND), projected,
// unit 2-norm (Grid::unitizeForce :3023-3029), u = 0.  Step: one V-cycle on u
// (HierarchyGrid::v_cycle relaxes u against f, i.e. u += B(f - K u)), relative residual,
// new force = u on the load nodes, projected, unit norm; force change
// |f_new - f_old| / |f_old| (Grid::supportForceCh / supportForceNorm :3031-3061).
// Projection with a normal n (unit): f <- (n . f) n, which is what removing the two
// tangential components with an orthonormal tangent pair (f -= N^T N f) leaves.
// ------------------------------------------------------------------------------------
static double splitmix_pm1(uint64_t& s) {  // splitmix64 (Steele, Lea, Flood 2014) -> [-1, 1)
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

ORC_API int orc_worst_compliance(tos_ctx* c, const tos_worst_opts* o, tos_worst_stats* st) {
  if (!c || !o || o->n_load <= 0 || !o->nodes || o->max_iter < 0 || o->pre < 0 || o->post < 0 ||
      o->precision < 0 || o->precision > 2 || o->fch_tol < 0 || o->res_tol < 0)
    return TOS_E_ARG;
  if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
  Level& L = c->L[0];
  const int64_t nn = L.nn, n3 = 3 * nn;
  const int n = o->n_load;
  std::vector<uint8_t> seen(nn, 0);
  std::vector<double> nrm(3 * (size_t)n, 0.0), fs(3 * (size_t)n), g(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const int64_t id = o->nodes[i];
    if (id < 0 || id >= nn || !L.fre[id] || seen[id]) { c->err = "bad load node"; return TOS_E_ARG; }
    seen[id] = 1;
    if (o->normals) {
      const double* a = o->normals + 3 * (size_t)i;
      const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      if (!(len > 0) || !std::isfinite(len)) { c->err = "bad normal"; return TOS_E_ARG; }
      for (int k = 0; k < 3; ++k) nrm[3 * (size_t)i + k] = a[k] / len;
    }
  }
  auto t0 = std::chrono::steady_clock::now();
  double* u = c->U.data();  // load case 0
  double* f = c->F.data();
  // g <- forceProject(g), then f = g / |g| on the load nodes (0 elsewhere); returns |g|^2
  auto set_force = [&]() {
    force_project(*c, n, o->nodes, o->normals ? nrm.data() : nullptr, g.data());
    double g2 = 0;
    for (int i = 0; i < n; ++i) {
      const double* gi = &g[3 * (size_t)i];
      g2 += gi[0] * gi[0] + gi[1] * gi[1] + gi[2] * gi[2];
    }
    if (!(g2 > 0) || !std::isfinite(g2)) return g2;
    const double inv = 1.0 / std::sqrt(g2);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) f[k * nn + o->nodes[i]] = g[3 * (size_t)i + k] * inv;
    return g2;
  };
  uint64_t seed = o->seed;
  for (double& x : g) x = splitmix_pm1(seed);
  std::fill(f, f + n3, 0.0);
  std::fill(u, u + n3, 0.0);
  const double g0 = set_force();
  if (!(g0 > 0) || !std::isfinite(g0)) { c->err = "the start force is zero after the projection"; return TOS_E_ARG; }
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) fs[3 * (size_t)i + k] = f[k * nn + o->nodes[i]];
  std::vector<double> r(n3), z(n3), q(n3);
  int it = 0;
  double fch = 1, rel = 1;
  while (it < o->max_iter && (fch > o->fch_tol || rel > o->res_tol)) {
    apply_fine(*c, u, q.data());
    for (int64_t i = 0; i < n3; ++i) r[i] = f[i] - q[i];
    precondition(*c, r.data(), z.data(), o->pre, o->post);
    for (int64_t i = 0; i < n3; ++i) u[i] += z[i];
    if (c->floating) displacement_project(*c, u);  // optimization/optimization.cpp:282-285
    apply_fine(*c, u, q.data());
    for (int64_t i = 0; i < n3; ++i) r[i] = f[i] - q[i];
    rel = std::sqrt(dot3(L, r.data(), r.data()) / dot3(L, f, f));
    ++it;
    if (!std::isfinite(rel)) { c->err = "NaN in the power iteration"; return TOS_E_NAN; }
    if (rel > 1e4) { c->err = "power iteration diverged"; return TOS_E_NOCONV; }
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) g[3 * (size_t)i + k] = u[k * nn + o->nodes[i]];
    const double g2 = set_force();
    if (!(g2 > 0) || !std::isfinite(g2)) { c->err = "displacement vanished on the load nodes"; return TOS_E_NAN; }
    double dd = 0, oo = 0;
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        const double fn = f[k * nn + o->nodes[i]], fo = fs[3 * (size_t)i + k];
        dd += (fn - fo) * (fn - fo);
        oo += fo * fo;
        fs[3 * (size_t)i + k] = fn;
      }
    fch = std::sqrt(dd) / std::sqrt(oo);
  }
  const double comp = dot3(L, f, u);
  if (st) {
    st->iterations = it;
    st->compliance = comp;
    st->fch = fch;
    st->rel_res = rel;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return TOS_OK;
}

ORC_API int orc_balance_loads(tos_ctx* c, int k, int n, const int64_t* nodes, const double* normals) {
  if (!c || k < 0 || k >= c->K || n <= 0 || !nodes) return TOS_E_ARG;
  const Level& L = c->L[0];
  std::vector<uint8_t> seen(L.nn, 0);
  std::vector<double> nrm(3 * (size_t)n), fs(3 * (size_t)n);
  double* f = c->F.data() + (size_t)k * 3 * L.nn;
  for (int i = 0; i < n; ++i) {
    const int64_t id = nodes[i];
    if (id < 0 || id >= L.nn || !L.fre[id] || seen[id]) { c->err = "bad load node"; return TOS_E_ARG; }
    seen[id] = 1;
    if (normals) {
      const double* a = normals + 3 * (size_t)i;
      const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      if (!(len > 0) || !std::isfinite(len)) { c->err = "bad normal"; return TOS_E_ARG; }
      for (int r = 0; r < 3; ++r) nrm[3 * (size_t)i + r] = a[r] / len;
    }
    for (int r = 0; r < 3; ++r) fs[3 * (size_t)i + r] = f[r * L.nn + id];
  }
  force_project(*c, n, nodes, normals ? nrm.data() : nullptr, fs.data());
  std::fill(f, f + 3 * L.nn, 0.0);
  for (int i = 0; i < n; ++i)
    for (int r = 0; r < 3; ++r) f[r * L.nn + nodes[i]] = fs[3 * (size_t)i + r];
  return TOS_OK;
}

ORC_API int orc_remove_rigid(tos_ctx* c, int k) {
  if (!c || k < 0 || k >= c->K) return TOS_E_ARG;
  if (!c->floating) { c->err = "the context has supports: no rigid motions to remove"; return TOS_E_STATE; }
  displacement_project(*c, c->U.data() + (size_t)k * 3 * c->L[0].nn);
  return TOS_OK;
}

ORC_API int orc_sensitivity(tos_ctx* c, int k) {
  if (!c || k < 0 || k >= c->K) return TOS_E_ARG;
  const Level& L = c->L[0];
  const double* u = c->U.data() + k * 3 * L.nn;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L.ez; ++z)
    for (int y = 0; y < L.ey; ++y)
      for (int x = 0; x < L.ex; ++x) {
        int64_t e = L.eid(x, y, z);
        if (!L.eact[e]) { c->sens[e] = 0; continue; }
        double q = element_energy(*c, u, x, y, z);
        c->sens[e] = (float)(-c->penal * std::pow((double)c->rho[e], c->penal - 1) * q);
      }
  return TOS_OK;
}

ORC_API int orc_element_compliance(tos_ctx* c, int k, double* out) {
  if (!c || !out || k < 0 || k >= c->K) return TOS_E_ARG;
  const Level& L = c->L[0];
  const double* u = c->U.data() + k * 3 * L.nn;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L.ez; ++z)
    for (int y = 0; y < L.ey; ++y)
      for (int x = 0; x < L.ex; ++x) {
        int64_t e = L.eid(x, y, z);
        out[e] = L.eact[e] ? c->w[e] * element_energy(*c, u, x, y, z) : 0.0;
      }
  return TOS_OK;
}

static std::vector<float>* sens_buf(tos_ctx* c, int which) {
  return which == 0 ? &c->sens : which == 1 ? &c->vsens : nullptr;
}

ORC_API int orc_get_sensitivity(tos_ctx* c, int which, float* out) {
  auto* b = c ? sens_buf(c, which) : nullptr;
  if (!b || !out) return TOS_E_ARG;
  std::copy(b->begin(), b->end(), out);
  return TOS_OK;
}

ORC_API int orc_set_sensitivity(tos_ctx* c, int which, const float* in) {
  auto* b = c ? sens_buf(c, which) : nullptr;
  if (!b || !in) return TOS_E_ARG;
  for (int64_t e = 0; e < c->L[0].ne; ++e) (*b)[e] = c->L[0].eact[e] ? in[e] : 0.0f;
  return TOS_OK;
}

ORC_API int orc_init_vol_sensitivity(tos_ctx* c) {
  if (!c) return TOS_E_ARG;
  float v = (float)(1.0 / (double)c->n_active);
  for (int64_t e = 0; e < c->L[0].ne; ++e) c->vsens[e] = c->L[0].eact[e] ? v : 0.0f;
  return TOS_OK;
}

ORC_API int orc_filter(tos_ctx* c, int which, double R) {
  auto* b = c ? sens_buf(c, which) : nullptr;
  if (!b || !(R > 0)) return TOS_E_ARG;
  filter_field(*c, *b, R);
  return TOS_OK;
}

ORC_API int orc_oc_update(tos_ctx* c, double vgoal, double move, double damp, double* g_out, int* steps) {
  if (!c) return TOS_E_ARG;
  const Level& L = c->L[0];
  double gmax = 0;
  for (int64_t e = 0; e < L.ne; ++e)
    if (L.eact[e]) gmax = std::max(gmax, (double)std::fabs(c->sens[e]));
  double lo = 0, hi = gmax, g = 0, V = 2;
  int itn = 0, trials = 0;
  do {
    g = (lo + hi) / 2;
    V = oc_trial(*c, g, move, damp, nullptr);
    ++trials;
    if (V > vgoal) lo = g;
    else if (V < vgoal) hi = g;
  } while (std::fabs(V - vgoal) > 1e-4 && itn++ < 30);
  oc_trial(*c, g, move, damp, &c->rho);
  for (int64_t e = 0; e < L.ne; ++e)
    if (!L.eact[e]) c->rho[e] = 0;
  if (g_out) *g_out = g;
  if (steps) *steps = trials;
  c->op_ready = false;
  return TOS_OK;
}

// Heaviside projection of one design value (computeProjection_kernel
// optimization/optimization.cu:313-337): eta = 0.5, clamp [rho_min, 1], shell -> 1.
static double heaviside(double x, double beta, double rmin, bool shell) {
  const double eta = 0.5;
  double r = (std::tanh(beta * eta) + std::tanh(beta * (x - eta))) / (std::tanh(beta * eta) + std::tanh(beta * (1 - eta)));
  r = std::min(std::max(r, rmin), 1.0);
  return shell ? 1.0 : r;
}

// OC update of the projected design (updateDensities2 optimization/optimization.cu:462-557):
// the trials update the design variable rho_pre (InitRho) and are projected before the
// volume sum; the bisection stops at |V - vgoal| <= 1e-4, at a relative bracket
// (hi - lo)/(lo + hi) <= 1e-3, or after 51 trials; the final trial becomes the new design
// variable and rho its projection (projectDensities).  change = max and mean |x_new - x_old|
// over the active elements (the reference's mean runs over its padded element count).
ORC_API int orc_oc_update_projected(tos_ctx* c, double vgoal, double move, double damp, double beta, double* g_out,
                                    int* steps, double* change) {
  if (!c || !(move > 0) || !(damp > 0) || !(beta > 0)) return TOS_E_ARG;
  const Level& L = c->L[0];
  const bool has_shell = !c->shell.empty();
  auto trial = [&](int64_t e, double g) {
    double s = c->sens[e];
    double gg = s > 0 ? 0.0 : -s;
    double r0 = c->rho_pre[e];
    double rn = r0 * std::pow(gg / g, damp);
    rn = std::min(std::max(rn, r0 - move), r0 + move);
    rn = std::min(std::max(rn, c->rho_min), 1.0);
    if (has_shell && c->shell[e]) rn = 1;
    return (float)rn;
  };
  double gmax = 0;
  for (int64_t e = 0; e < L.ne; ++e)
    if (L.eact[e]) gmax = std::max(gmax, (double)std::fabs(c->sens[e]));
  double lo = 0, hi = gmax, g = 0, V = 2;
  int itn = 0, trials = 0;
  do {
    g = (lo + hi) / 2;
    V = chunked_sum(L.ne, [&](int64_t e) -> double {
          if (!L.eact[e]) return 0.0;
          return (double)(float)heaviside(trial(e, g), beta, c->rho_min, has_shell && c->shell[e]);
        }) / (double)c->n_active;
    ++trials;
    if (V > vgoal) lo = g;
    else if (V < vgoal) hi = g;
  } while (std::fabs(V - vgoal) > 1e-4 && (hi - lo) / (lo + hi) > 1e-3 && itn++ < 50);
  double cmax = 0, csum = 0;
  for (int64_t e = 0; e < L.ne; ++e) {
    if (!L.eact[e]) { c->rho[e] = 0; continue; }
    const float xn = trial(e, g);
    const double d = std::fabs((double)xn - (double)c->rho_pre[e]);
    cmax = std::max(cmax, d);
    csum += d;
    c->rho_pre[e] = xn;
    c->rho[e] = (float)heaviside(xn, beta, c->rho_min, has_shell && c->shell[e]);
  }
  if (g_out) *g_out = g;
  if (steps) *steps = trials;
  if (change) { change[0] = cmax; change[1] = csum / (double)c->n_active; }
  c->op_ready = false;
  return TOS_OK;
}

// Heaviside projection (computeProjection_kernel optimization/optimization.cu:313-337).
ORC_API int orc_project_density(tos_ctx* c, double beta) {
  if (!c) return TOS_E_ARG;
  const double eta = 0.5;
  double den = std::tanh(beta * eta) + std::tanh(beta * (1 - eta));
  c->rho_pre = c->rho;
  for (int64_t e = 0; e < c->L[0].ne; ++e) {
    if (!c->L[0].eact[e]) continue;
    double r = (std::tanh(beta * eta) + std::tanh(beta * (c->rho_pre[e] - eta))) / den;
    r = std::min(std::max(r, c->rho_min), 1.0);
    if (c->shell.size() && c->shell[e]) r = 1;
    c->rho[e] = (float)r;
  }
  c->op_ready = false;
  return TOS_OK;
}

// computeProjectionSensitivity_kernel optimization/optimization.cu:161-183.
ORC_API int orc_projection_sensitivity(tos_ctx* c, int which, double beta) {
  auto* b = c ? sens_buf(c, which) : nullptr;
  if (!b) return TOS_E_ARG;
  const double eta = 0.5;
  double den = std::tanh(beta * eta) + std::tanh(beta * (1 - eta));
  for (int64_t e = 0; e < c->L[0].ne; ++e) {
    if (!c->L[0].eact[e]) continue;
    double t = std::tanh(beta * (c->rho_pre[e] - eta));
    (*b)[e] = (float)((*b)[e] * (beta * (1 - t * t) / den));
  }
  return TOS_OK;
}

ORC_API int orc_spline_init(tos_ctx* c, int mode, int order, int param) {
  if (!c || (mode != 0 && mode != 1) || order < 2 || order > 8 || param < 1) return TOS_E_ARG;
  c->sp_mode = mode; c->sp_order = order; c->sp_param = param;
  int dims[3] = {c->nx, c->ny, c->nz};
  int nc[3];
  for (int a = 0; a < 3; ++a) spline_axis(c->ax[a], mode, order, param, dims[a], c->h, &nc[a]);
  c->ncx = nc[0]; c->ncy = nc[1]; c->ncz = nc[2];
  size_t n = (size_t)nc[0] * nc[1] * nc[2];
  c->coef.assign(n, 0.0f); c->csens.assign(n, 0.0f); c->cvsens.assign(n, 0.0f);
  return TOS_OK;
}

ORC_API int orc_spline_set_coeffs(tos_ctx* c, const float* v) {
  if (!c || !v || c->sp_mode < 0) return TOS_E_ARG;
  std::copy(v, v + c->coef.size(), c->coef.begin());
  return TOS_OK;
}

ORC_API int orc_spline_get_coeffs(tos_ctx* c, float* v) {
  if (!c || !v || c->sp_mode < 0) return TOS_E_ARG;
  std::copy(c->coef.begin(), c->coef.end(), v);
  return TOS_OK;
}

// coeff2density (grid/Grid.cu:3325-3430): tensor-product sum at the element centre,
// -0.2 outside the knot range, clamp to [rho_min, 1], shell -> 1.
ORC_API int orc_spline_eval(tos_ctx* c) {
  if (!c || c->sp_mode < 0) return TOS_E_ARG;
  const Level& L = c->L[0];
  const int o = c->sp_order;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < L.ez; ++z)
    for (int y = 0; y < L.ey; ++y)
      for (int x = 0; x < L.ex; ++x) {
        int64_t e = L.eid(x, y, z);
        if (!L.eact[e]) { c->rho[e] = 0; continue; }
        double v;
        if (!c->ax[0].ok[x] || !c->ax[1].ok[y] || !c->ax[2].ok[z]) {
          v = -0.2;
        } else {
          v = 0;
          for (int k = 0; k < o; ++k)
            for (int j = 0; j < o; ++j)
              for (int i = 0; i < o; ++i) {
                int64_t ci = c->ax[0].idx[x * o + i] +
                             (int64_t)c->ncx * (c->ax[1].idx[y * o + j] + (int64_t)c->ncy * c->ax[2].idx[z * o + k]);
                v += (double)c->coef[ci] * c->ax[0].wt[x * o + i] * c->ax[1].wt[y * o + j] * c->ax[2].wt[z * o + k];
              }
        }
        v = std::min(std::max(v, c->rho_min), 1.0);
        if (c->shell.size() && c->shell[e]) v = 1;
        c->rho[e] = (float)v;
      }
  c->op_ready = false;
  return TOS_OK;
}

// ddensity2dcoeff_update / dvol2dcoeff (grid/Grid.cu:3643-3859): dJ/dc = sum over active,
// non-shell elements inside the knot range of g_e N_i N_j N_k (sequential scatter here).
ORC_API int orc_spline_adjoint(tos_ctx* c, int which) {
  auto* b = c ? sens_buf(c, which) : nullptr;
  if (!b || c->sp_mode < 0) return TOS_E_ARG;
  const Level& L = c->L[0];
  const int o = c->sp_order;
  std::vector<double> acc(c->coef.size(), 0.0);
  for (int z = 0; z < L.ez; ++z)
    for (int y = 0; y < L.ey; ++y)
      for (int x = 0; x < L.ex; ++x) {
        int64_t e = L.eid(x, y, z);
        if (!L.eact[e] || (c->shell.size() && c->shell[e])) continue;
        if (!c->ax[0].ok[x] || !c->ax[1].ok[y] || !c->ax[2].ok[z]) continue;
        double g = (*b)[e];
        for (int k = 0; k < o; ++k)
          for (int j = 0; j < o; ++j)
            for (int i = 0; i < o; ++i) {
              int64_t ci = c->ax[0].idx[x * o + i] +
                           (int64_t)c->ncx * (c->ax[1].idx[y * o + j] + (int64_t)c->ncy * c->ax[2].idx[z * o + k]);
              acc[ci] += g * c->ax[0].wt[x * o + i] * c->ax[1].wt[y * o + j] * c->ax[2].wt[z * o + k];
            }
      }
  auto& out = which == 0 ? c->csens : c->cvsens;
  for (size_t i = 0; i < acc.size(); ++i) out[i] = (float)acc[i];
  return TOS_OK;
}

ORC_API int orc_spline_get_coeff_sens(tos_ctx* c, int which, float* out) {
  if (!c || !out || c->sp_mode < 0 || which < 0 || which > 3) return TOS_E_ARG;
  auto& v = which == 0 ? c->csens : which == 1 ? c->cvsens : which == 2 ? c->sscsens : c->dripcsens;
  if (v.size() != c->coef.size()) {
    std::fill(out, out + c->coef.size(), 0.0f);
    return TOS_OK;
  }
  std::copy(v.begin(), v.end(), out);
  return TOS_OK;
}

// ------------------------------------------------------------------------------------
// Self-supporting and drip constraints on spline surface points (deal_surface_points,
// optimization/optimization.cpp:1103-1128).
// ------------------------------------------------------------------------------------
// Basis values and first / second derivatives (NURBS Book algorithm A2.3, as
// SplineBasisDeriX grid/Grid.cu:469-530) of the p+1 functions nonzero on knot span `span`.
static void ders_basis_funs(int span, double x, int p, const std::vector<double>& U, double ders[3][16]) {
  double ndu[16][16], left[16], right[16], a[2][16];
  ndu[0][0] = 1.0;
  for (int j = 1; j <= p; ++j) {
    left[j] = x - U[span + 1 - j];
    right[j] = U[span + j] - x;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      ndu[j][r] = right[r + 1] + left[j - r];
      double temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    ndu[j][j] = saved;
  }
  for (int j = 0; j <= p; ++j) ders[0][j] = ndu[j][p];
  for (int k = 1; k <= 2; ++k)
    for (int j = 0; j <= p; ++j) ders[k][j] = 0.0;
  for (int r = 0; r <= p; ++r) {
    int s1 = 0, s2 = 1;
    a[0][0] = 1.0;
    for (int k = 1; k <= 2 && k <= p; ++k) {
      double d = 0.0;
      int rk = r - k, pk = p - k;
      if (r >= k) {
        a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
        d = a[s2][0] * ndu[rk][pk];
      }
      int j1 = rk >= -1 ? 1 : -rk, j2 = (r - 1 <= pk) ? k - 1 : p - r;
      for (int j = j1; j <= j2; ++j) {
        a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
        d += a[s2][j] * ndu[rk + j][pk];
      }
      if (r <= pk) {
        a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
        d += a[s2][k] * ndu[r][pk];
      }
      ders[k][r] = d;
      std::swap(s1, s2);
    }
  }
  int r = p;
  for (int k = 1; k <= 2; ++k) {
    for (int j = 0; j <= p; ++j) ders[k][j] *= r;
    r *= (p - k);
  }
}

// spline basis along one axis at coordinate x: first coefficient index per term, values
// and d/dx, d2/dx2; false outside the spline's range (mode 0: the knot range)
struct AxisDers {
  int idx[16];
  double N[3][16];
};
static bool axis_ders(const Ctx& c, int axis, double x, AxisDers& D) {
  const int k = c.sp_order;
  const int ne = axis == 0 ? c.nx : axis == 1 ? c.ny : c.nz;
  const int nc = axis == 0 ? c.ncx : axis == 1 ? c.ncy : c.ncz;
  if (c.sp_mode == 1) {  // uniform lattice every param voxels, boundary coefficients replicated
    const double sh = c.sp_param * c.h, t = x / sh;
    const int i0 = (int)std::floor(t - k / 2.0) + 1;
    for (int q = 0; q < k; ++q) {
      const int i = i0 + q;
      const double u = t - i + k / 2.0;
      D.idx[q] = std::min(std::max(i, 0), nc - 1);
      D.N[0][q] = cardinal(k, u);
      D.N[1][q] = k >= 2 ? (cardinal(k - 1, u) - cardinal(k - 1, u - 1)) / sh : 0.0;
      D.N[2][q] = k >= 3 ? (cardinal(k - 2, u) - 2 * cardinal(k - 2, u - 1) + cardinal(k - 2, u - 2)) / (sh * sh) : 0.0;
    }
    return true;
  }
  // clamped knots on the box enlarged by 2*box/(P+1) per side (Grid.cpp:3253-3377)
  const int P = c.sp_param;
  const double delta = ne * c.h, lo = -2.0 * delta / (P + 1), hi = delta + 2.0 * delta / (P + 1);
  const double step = (hi - lo) / (P + 1);
  std::vector<double> U;
  for (int i = 0; i < k; ++i) U.push_back(lo);
  double tk = lo;
  for (int i = 0; i < P; ++i) { tk += step; U.push_back(tk); }
  for (int i = 0; i < k; ++i) U.push_back(hi);
  const int span1 = (int)((x - lo) / step) + k;  // calc_node grid/Grid.cu:5550-5552
  if (span1 < k || span1 > nc) return false;
  double ders[3][16];
  ders_basis_funs(span1 - 1, x, k - 1, U, ders);
  for (int q = 0; q < k; ++q) {
    D.idx[q] = span1 - k + q;
    for (int d = 0; d < 3; ++d) D.N[d][q] = ders[d][q];
  }
  return true;
}

// spline value, gradient and Hessian (row-major 3x3) at p
static bool spline_point(const Ctx& c, const double p[3], double& phi, double g[3], double H[9]) {
  AxisDers A[3];
  for (int a = 0; a < 3; ++a)
    if (!axis_ders(c, a, p[a], A[a])) return false;
  const int k = c.sp_order;
  phi = 0;
  for (int i = 0; i < 3; ++i) g[i] = 0;
  for (int i = 0; i < 9; ++i) H[i] = 0;
  for (int qz = 0; qz < k; ++qz)
    for (int qy = 0; qy < k; ++qy)
      for (int qx = 0; qx < k; ++qx) {
        const double cv = c.coef[A[0].idx[qx] + (size_t)c.ncx * (A[1].idx[qy] + (size_t)c.ncy * A[2].idx[qz])];
        const double X[3] = {A[0].N[0][qx], A[0].N[1][qx], A[0].N[2][qx]};
        const double Y[3] = {A[1].N[0][qy], A[1].N[1][qy], A[1].N[2][qy]};
        const double Z[3] = {A[2].N[0][qz], A[2].N[1][qz], A[2].N[2][qz]};
        phi += cv * X[0] * Y[0] * Z[0];
        g[0] += cv * X[1] * Y[0] * Z[0];
        g[1] += cv * X[0] * Y[1] * Z[0];
        g[2] += cv * X[0] * Y[0] * Z[1];
        H[0] += cv * X[2] * Y[0] * Z[0];
        H[4] += cv * X[0] * Y[2] * Z[0];
        H[8] += cv * X[0] * Y[0] * Z[2];
        H[1] += cv * X[1] * Y[1] * Z[0];
        H[2] += cv * X[1] * Y[0] * Z[1];
        H[5] += cv * X[0] * Y[1] * Z[1];
      }
  H[3] = H[1]; H[6] = H[2]; H[7] = H[5];
  return true;
}

static double sigm(double x) { return 1.0 / (1.0 + std::exp(-x)); }

// per-point self-support value and d/dinner (grid/Grid.cu:117-160, 6015-6100)
static void ss_point(const tos_surface_opts& o, double inner, bool down, double& v, double& dv) {
  v = dv = 0;
  const int m = o.ss_mode;
  if (m % 2 == 0 && !down) return;
  if (m <= 1) { v = std::pow(inner, o.p_norm); dv = o.p_norm * std::pow(inner, o.p_norm - 1); }
  else if (m <= 3) { v = inner >= 1 ? o.hfunction_c * (inner - 1) * (inner - 1) : 0; dv = inner >= 1 ? 2 * o.hfunction_c * (inner - 1) : 0; }
  else { double s = sigm((inner - 1.5) / o.sigmoid_c); v = s + 1; dv = s * (1 - s) / o.sigmoid_c; }
}

// Surface constraints and their coefficient gradients (points fixed); see tos_api.h.
// Constraint values at the stored surface points with the stored orientation, and with
// `grad` their coefficient gradients (the points held fixed).
static void surface_eval(Ctx& c, const tos_surface_opts& o, bool grad, tos_surface_stats* st) {
  const int64_t n = (int64_t)c.surf_p.size() / 3;
  const double dir = c.surf_dir;
  std::vector<double> HH(9 * n);
  c.surf_n.assign(3 * n, 0.0);
  for (int64_t pi = 0; pi < n; ++pi) {
    double phi, g[3], H[9];
    if (!spline_point(c, &c.surf_p[3 * pi], phi, g, H)) { for (double& v : g) v = 0; for (double& v : H) v = 0; }
    for (int a = 0; a < 3; ++a) c.surf_n[3 * pi + a] = dir * g[a];
    for (int a = 0; a < 9; ++a) HH[9 * pi + a] = H[a];
  }
  // per-point values and gradient weights: d(value)/dc_j = w . grad N_j + wxx N_j,xx +
  // wyy N_j,yy + wxy N_j,xy
  const double cs = std::cos(o.print_angle), cd = std::cos(o.drip_angle);
  std::vector<double> wss(3 * n, 0.0), wdr(6 * n, 0.0);
  double ss_sum = 0, dr_sum = 0;
  int64_t n_down = 0;
  const bool drip_on = o.drip_mode >= 6;
  for (int64_t pi = 0; pi < n; ++pi) {
    const double* nv = &c.surf_n[3 * pi];
    const double* H = &HH[9 * pi];
    const double nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    const bool down = nv[2] < 0;
    n_down += down;
    if (nn == 0) continue;
    // d inner / d n = (e_z |n| - n_z n / |n|) / (|n|^2 cos); d n / d c_j = dir grad N_j
    double dinner[3];
    for (int a = 0; a < 3; ++a) dinner[a] = dir * ((a == 2 ? nn : 0.0) - nv[2] * nv[a] / nn) / (nn * nn);
    double v, dv;
    ss_point(o, nv[2] / nn / cs, down, v, dv);
    ss_sum += v;
    for (int a = 0; a < 3; ++a) wss[3 * pi + a] = dv * dinner[a] / cs;
    if (drip_on && (o.drip_mode == 7 || down)) {
      const double hin = nv[2] / nn / cd;
      double u, du[3];
      if (o.drip_version == 1) {
        const double px = nv[0] / nv[2], py = nv[1] / nv[2];
        u = std::exp(-o.drip_alpha * (px * px + py * py));
        // d px / d n = (e_x n_z - n_x e_z) / n_z^2, likewise py
        for (int a = 0; a < 3; ++a) {
          const double dpx = ((a == 0 ? nv[2] : 0.0) - (a == 2 ? nv[0] : 0.0)) / (nv[2] * nv[2]);
          const double dpy = ((a == 1 ? nv[2] : 0.0) - (a == 2 ? nv[1] : 0.0)) / (nv[2] * nv[2]);
          du[a] = -2 * o.drip_alpha * (px * dpx + py * dpy) * u * dir;
        }
      } else {
        double s, ds;
        if (o.drip_angle > 0.5 * M_PI) { s = sigm((hin - 1.2) * o.drip_mu); ds = s * (1 - s) * o.drip_mu; }
        else { s = sigm(-(hin + 1.2) * o.drip_mu); ds = -s * (1 - s) * o.drip_mu; }
        u = s;
        for (int a = 0; a < 3; ++a) du[a] = ds * dinner[a] / cd;
      }
      const double x1 = H[0], x2 = H[0] * H[4] - H[1] * H[3];
      const double e1 = std::exp(o.drip_beta * x1), e2 = std::exp(-o.drip_beta * x2);
      const double q1 = o.drip_beta_scale * x1 * x1 / (1 + e1);
      const double q2 = o.drip_beta_scale * x2 * x2 / (1 + e2);
      const double dq1 = o.drip_beta_scale * (2 * x1 / (1 + e1) - x1 * x1 * o.drip_beta * e1 / ((1 + e1) * (1 + e1)));
      const double dq2 = o.drip_beta_scale * (2 * x2 / (1 + e2) + x2 * x2 * o.drip_beta * e2 / ((1 + e2) * (1 + e2)));
      dr_sum += u * q1 * q2;
      for (int a = 0; a < 3; ++a) wdr[6 * pi + a] = du[a] * q1 * q2;
      // d x1 = N_xx; d x2 = N_xx H_yy + H_xx N_yy - 2 H_xy N_xy (H symmetric)
      wdr[6 * pi + 3] = u * (dq1 * q2 + q1 * dq2 * H[4]);
      wdr[6 * pi + 4] = u * q1 * dq2 * H[0];
      wdr[6 * pi + 5] = u * q1 * dq2 * (-2 * H[1]);
    }
  }
  const double cnt_ss = o.ss_mode % 2 == 0 ? (double)n_down : (double)n;
  const double cnt_dr = o.drip_mode == 6 ? (double)n_down : o.drip_mode % 2 == 0 ? (double)n_down : (double)n;
  double ss_val = 0, ss_fac = 0, dr_val = 0, dr_fac = 0;
  if (cnt_ss > 0) {
    const double mean = ss_sum / cnt_ss;
    if (o.ss_mode <= 1) {
      ss_val = std::pow(mean, 1.0 / o.p_norm) - 1;
      ss_fac = mean > 0 ? std::pow(mean, 1.0 / o.p_norm - 1) / o.p_norm / cnt_ss : 0.0;
    } else {
      ss_val = mean;
      ss_fac = 1.0 / cnt_ss;
    }
  }
  if (cnt_dr > 0) {
    dr_val = o.drip_mode <= 1 ? -1.0 : dr_sum / cnt_dr;
    dr_fac = drip_on ? 1.0 / cnt_dr : 0.0;
  }
  c.ss_value = ss_val;
  c.drip_value = dr_val;
  if (!grad) {
    if (st) { st->n_points = n; st->n_down = n_down; st->direction = dir; st->ss_value = ss_val; st->drip_value = dr_val; }
    return;
  }
  // coefficient gradients: every point scatters into its order^3 coefficients in point order
  const size_t nc = c.coef.size();
  std::vector<double> gss(nc, 0.0), gdr(nc, 0.0);
  const int k = c.sp_order;
  for (int64_t pi = 0; pi < n; ++pi) {
    AxisDers A[3];
    bool ok = true;
    for (int a = 0; a < 3; ++a) ok = ok && axis_ders(c, a, c.surf_p[3 * pi + a], A[a]);
    if (!ok) continue;
    const double* ws = &wss[3 * pi];
    const double* wd = &wdr[6 * pi];
    for (int qz = 0; qz < k; ++qz)
      for (int qy = 0; qy < k; ++qy)
        for (int qx = 0; qx < k; ++qx) {
          const size_t j = A[0].idx[qx] + (size_t)c.ncx * (A[1].idx[qy] + (size_t)c.ncy * A[2].idx[qz]);
          const double X0 = A[0].N[0][qx], X1 = A[0].N[1][qx], X2 = A[0].N[2][qx];
          const double Y0 = A[1].N[0][qy], Y1 = A[1].N[1][qy], Y2 = A[1].N[2][qy];
          const double Z0 = A[2].N[0][qz], Z1 = A[2].N[1][qz];
          const double gx = X1 * Y0 * Z0, gy = X0 * Y1 * Z0, gz = X0 * Y0 * Z1;
          gss[j] += ws[0] * gx + ws[1] * gy + ws[2] * gz;
          gdr[j] += wd[0] * gx + wd[1] * gy + wd[2] * gz + wd[3] * X2 * Y0 * Z0 + wd[4] * X0 * Y2 * Z0 +
                    wd[5] * X1 * Y1 * Z0;
        }
  }
  c.sscsens.assign(nc, 0.0f);
  c.dripcsens.assign(nc, 0.0f);
  for (size_t j = 0; j < nc; ++j) {
    c.sscsens[j] = (float)(ss_fac * gss[j]);
    c.dripcsens[j] = (float)(dr_fac * gdr[j]);
  }
  c.ss_value = ss_val;
  c.drip_value = dr_val;
  if (st) {
    st->n_points = n;
    st->n_down = n_down;
    st->direction = dir;
    st->ss_value = ss_val;
    st->drip_value = dr_val;
  }
}


ORC_API int orc_surface_constraints(tos_ctx* c, const tos_surface_opts* o, tos_surface_stats* st) {
  if (!c || !o || c->sp_mode < 0 || o->cells < 1 || o->ss_mode < 0 || o->ss_mode > 5 || o->drip_mode < 0 ||
      o->drip_mode > 7 || !(o->beta > 0))
    return TOS_E_ARG;
  // all coefficients equal: no surface; the driver's constants (utils/test_utils.cpp:1106-1113)
  bool equal = true;
  for (float v : c->coef) equal = equal && v == c->coef[0];
  if (equal) {
    c->surf_p.clear();
    c->surf_n.clear();
    c->sscsens.assign(c->coef.size(), 0.0f);
    c->dripcsens.assign(c->coef.size(), 0.0f);
    c->ss_value = 1.0;
    c->drip_value = 0.0;
    if (st) { st->n_points = 0; st->n_down = 0; st->direction = 1.0; st->ss_value = 1.0; st->drip_value = 0.0; }
    return TOS_OK;
  }
  const int dims[3] = {c->nx, c->ny, c->nz};
  const int B = o->cells, NB = B + 2;
  double eh[3];
  for (int a = 0; a < 3; ++a) eh[a] = dims[a] * c->h / B;
  auto pos = [&](int a, int i) { return -eh[a] + (i + 0.5) * eh[a]; };
  // lattice values: spline - 1/2 inside the box, -1/2 outside
  std::vector<double> val((size_t)NB * NB * NB);
  for (int k = 0; k < NB; ++k)
    for (int j = 0; j < NB; ++j)
      for (int i = 0; i < NB; ++i) {
        const double p[3] = {pos(0, i), pos(1, j), pos(2, k)};
        bool in = true;
        for (int a = 0; a < 3; ++a) in = in && p[a] >= 0 && p[a] <= dims[a] * c->h;
        double phi = 0, g[3], H[9];
        if (in && !spline_point(*c, p, phi, g, H)) phi = -0.2;
        val[i + (size_t)NB * (j + (size_t)NB * k)] = (in ? phi : 0.0) - 0.5;
      }
  // marching-cubes vertices on sign-changing edges, kept inside active elements
  c->surf_p.clear();
  const Level& L = c->L[0];
  for (int k = 0; k < NB; ++k)
    for (int j = 0; j < NB; ++j)
      for (int i = 0; i < NB; ++i) {
        const int ijk[3] = {i, j, k};
        const double va = val[i + (size_t)NB * (j + (size_t)NB * k)];
        for (int d = 0; d < 3; ++d) {
          int q[3] = {i, j, k};
          if (++q[d] >= NB) continue;
          const double vb = val[q[0] + (size_t)NB * (q[1] + (size_t)NB * q[2])];
          if ((va < 0) == (vb < 0)) continue;
          const double t = (0.0 - va) / (vb - va);
          double p[3];
          for (int a = 0; a < 3; ++a) p[a] = pos(a, ijk[a]) + (a == d ? t * eh[a] : 0.0);
          int e[3];
          bool ok = true;
          for (int a = 0; a < 3; ++a) {
            e[a] = (int)std::floor(p[a] / c->h);
            ok = ok && e[a] >= 0 && e[a] < dims[a];
          }
          if (!ok || !L.eact[L.eid(e[0], e[1], e[2])]) continue;
          for (int a = 0; a < 3; ++a) c->surf_p.push_back(p[a]);
        }
      }
  const int64_t n = (int64_t)c->surf_p.size() / 3;
  // gradient, Hessian and the orientation test (correct_spline_surface_point_normal_direction)
  std::vector<double> G(3 * n), HH(9 * n);
  double count = 0;
  const double eta = 0.5, hden = std::tanh(o->beta * eta) + std::tanh(o->beta * (1 - eta));
  for (int64_t pi = 0; pi < n; ++pi) {
    double phi, g[3], H[9];
    if (!spline_point(*c, &c->surf_p[3 * pi], phi, g, H)) { for (double& v : g) v = 0; for (double& v : H) v = 0; }
    for (int a = 0; a < 3; ++a) G[3 * pi + a] = g[a];
    for (int a = 0; a < 9; ++a) HH[9 * pi + a] = H[a];
    double pd[3], v2 = -0.1, g2[3], H2[9];
    for (int a = 0; a < 3; ++a) pd[a] = c->surf_p[3 * pi + a] + 1e-3 * g[a];
    if (!spline_point(*c, pd, v2, g2, H2)) v2 = -0.1;
    const double hs = (std::tanh(o->beta * eta) + std::tanh(o->beta * (v2 - eta))) / hden;
    count += hs < 0.5 ? 1.0 : -1.0;
  }
  c->surf_dir = count > 0 ? 1.0 : -1.0;
  surface_eval(*c, *o, true, st);
  return TOS_OK;
}

// Test access (no counterpart in the product ABI): the constraint values at the points and
// orientation of the last orc_surface_constraints call, for the current coefficients.
ORC_API int orc_surface_eval_fixed(tos_ctx* c, const tos_surface_opts* o, double* ss, double* drip) {
  if (!c || !o || c->sp_mode < 0) return TOS_E_ARG;
  tos_surface_stats st;
  surface_eval(*c, *o, false, &st);
  if (ss) *ss = st.ss_value;
  if (drip) *drip = st.drip_value;
  return TOS_OK;
}

ORC_API int orc_surface_points(tos_ctx* c, int64_t cap, double* xyz, double* normals) {
  if (!c || cap < 0 || (cap > 0 && (!xyz || !normals))) return TOS_E_ARG;
  const int64_t n = std::min<int64_t>(cap, (int64_t)c->surf_p.size() / 3);
  std::copy(c->surf_p.begin(), c->surf_p.begin() + 3 * n, xyz);
  std::copy(c->surf_n.begin(), c->surf_n.begin() + 3 * n, normals);
  return TOS_OK;
}

// updateCoeff / tryCSensMultiplier_kernel (optimization/optimization.cu:559-695): OC
// bisection on the coefficients, each trial evaluated to densities; the coefficients end
// at the last trial.
ORC_API int orc_oc_update_coeffs(tos_ctx* c, double vgoal, double move, double damp, double cmin, double cmax,
                                  double* g_out, int* steps) {
  if (!c || c->sp_mode < 0) return TOS_E_ARG;
  const size_t n = c->coef.size();
  double gmax = 0;
  for (size_t j = 0; j < n; ++j) gmax = std::max(gmax, (double)std::fabs(c->csens[j]));
  std::vector<float> old(c->coef);
  double lo = 0, hi = gmax, g = 0, V = 2;
  int itn = 0, trials = 0;
  do {
    g = (lo + hi) / 2;
    for (size_t j = 0; j < n; ++j) {
      double s = c->csens[j];
      double gg = s > 0 ? 0.0 : -s;
      double c0 = old[j];
      double cn = c0 * std::pow(gg / g, damp);
      cn = std::min(std::max(cn, c0 - move), c0 + move);
      cn = std::min(std::max(cn, cmin), cmax);
      c->coef[j] = (float)cn;
    }
    orc_spline_eval(c);
    orc_volume(c, &V);
    ++trials;
    if (V > vgoal) lo = g;
    else if (V < vgoal) hi = g;
  } while (std::fabs(V - vgoal) > 1e-4 && itn++ < 30);
  if (g_out) *g_out = g;
  if (steps) *steps = trials;
  return TOS_OK;
}

ORC_API int orc_mma_create(int n, int m, const double* xmin, const double* xmax, double move, tos_mma** out) {
  if (n <= 0 || m <= 0 || m > 8 || !xmin || !xmax || !out || !(move > 0)) return TOS_E_ARG;
  tos_mma* M = new tos_mma();
  M->n = n; M->m = m; M->move = move;
  M->xmin.assign(xmin, xmin + n); M->xmax.assign(xmax, xmax + n);
  // set_bound + init_subproblem_variable (mma_t.cu:797-855)
  M->alpha = M->xmin; M->beta = M->xmax;
  M->L = M->alpha; M->U = M->beta;
  M->x.resize(n);
  for (int j = 0; j < n; ++j) M->x[j] = 0.5 * (M->alpha[j] + M->beta[j]);
  mma_clamp_asymptotes(*M);
  mma_init_vars(*M);
  M->lastdx.assign(n, 0.0);
  *out = M;
  return TOS_OK;
}

ORC_API int orc_mma_destroy(tos_mma* M) { delete M; return TOS_OK; }

ORC_API int orc_mma_set_x(tos_mma* M, const double* x) {
  if (!M || !x) return TOS_E_ARG;
  for (int j = 0; j < M->n; ++j) {
    double st = M->xmax[j] - M->xmin[j];
    M->x[j] = x[j];
    M->L[j] = x[j] - 0.5 * st;
    M->U[j] = x[j] + 0.5 * st;
  }
  mma_clamp_asymptotes(*M);
  return TOS_OK;
}

ORC_API int orc_mma_get_x(tos_mma* M, double* x) {
  if (!M || !x) return TOS_E_ARG;
  std::copy(M->x.begin(), M->x.end(), x);
  return TOS_OK;
}

// Test access to the MMA state (no counterpart in the product ABI): asymptotes L, U and
// move bounds alpha, beta after the last update -- the subproblem data of the next one.
ORC_API int orc_mma_state(tos_mma* M, double* L, double* U, double* alpha, double* beta) {
  if (!M || !L || !U || !alpha || !beta) return TOS_E_ARG;
  std::copy(M->L.begin(), M->L.end(), L);
  std::copy(M->U.begin(), M->U.end(), U);
  std::copy(M->alpha.begin(), M->alpha.end(), alpha);
  std::copy(M->beta.begin(), M->beta.end(), beta);
  return TOS_OK;
}

ORC_API int orc_mma_update(tos_mma* M, const double* df, const double* dg, const double* g, int* conv) {
  if (!M || !df || !dg || !g) return TOS_E_ARG;
  bool cv = mma_update(*M, df, dg, g);
  if (conv) *conv = cv ? 1 : 0;
  return TOS_OK;
}

ORC_API int orc_ctx_mma_step(tos_ctx* c, tos_mma* M, double sscale, double vscale, double vfrac, int* conv) {
  if (!c || !M || c->sp_mode < 0 || M->n != (int)c->coef.size() || M->m != 1) return TOS_E_ARG;
  double vol;
  orc_volume(c, &vol);
  int n = M->n;
  std::vector<double> df(n), dg(n);
  for (int j = 0; j < n; ++j) { df[j] = sscale * (double)c->csens[j]; dg[j] = vscale * (double)c->cvsens[j]; }
  double g = vscale * (vol - vfrac);
  bool cv = mma_update(*M, df.data(), dg.data(), &g);
  for (int j = 0; j < n; ++j) c->coef[j] = (float)M->x[j];
  if (conv) *conv = cv ? 1 : 0;
  return TOS_OK;
}

// MMA step with the volume, self-supporting and drip constraints (utils/test_utils.cpp:1135-1188).
ORC_API int orc_ctx_mma_step_surface(tos_ctx* c, tos_mma* M, double sscale, double vscale, double vfrac,
                                     double ss_scale, double drip_scale, int* conv) {
  if (!c || !M || c->sp_mode < 0 || M->n != (int)c->coef.size() || M->m != 3 ||
      c->sscsens.size() != c->coef.size())
    return TOS_E_ARG;
  double vol;
  orc_volume(c, &vol);
  int n = M->n;
  std::vector<double> df(n), dg(3 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    df[j] = sscale * (double)c->csens[j];
    dg[j] = vscale * (double)c->cvsens[j];
    dg[n + j] = ss_scale * (double)c->sscsens[j];
    dg[2 * (size_t)n + j] = drip_scale * (double)c->dripcsens[j];
  }
  double g[3] = {vscale * (vol - vfrac), ss_scale * (c->ss_value - 1.0), drip_scale * c->drip_value};
  bool cv = mma_update(*M, df.data(), dg.data(), g);
  for (int j = 0; j < n; ++j) c->coef[j] = (float)M->x[j];
  if (conv) *conv = cv ? 1 : 0;
  return TOS_OK;
}

// Kernel timing has no meaning for the CPU oracle: report no classes.
ORC_API int orc_profile(tos_ctx* c, int enable, tos_kernel_stats* out) {
  (void)enable;
  if (!c) return TOS_E_ARG;
  if (out) std::memset(out, 0, sizeof(*out));
  return TOS_OK;
}

// The batched solve is the same mathematics per case: the oracle solves the cases in turn.
ORC_API int orc_solve_batched(tos_ctx* c, const tos_solve_opts* o, tos_solve_stats* st) {
  if (!c) return TOS_E_ARG;
  const int depth = c->rc_depth;  // the batched solve starts without the recycled projection
  c->rc_depth = 0;
  const int rc = orc_solve(c, o, st);
  c->rc_depth = depth;
  return rc;
}

ORC_API int orc_batch_tuning(tos_ctx* c, char* buf, int len) {
  if (!c || !buf || len <= 0) return TOS_E_ARG;
  buf[0] = 0;
  return TOS_OK;
}

// The oracle always relaxes colour by colour (the GPU's wavefront schedule performs the
// same updates).
ORC_API int orc_set_sweep_mode(tos_ctx* c, int mode) {
  if (!c || mode < 0 || mode > 3) return TOS_E_ARG;
  return TOS_OK;
}

// The oracle builds and keeps FP64 stencils only; it accepts and ignores the storage
// precision of the library's FP32 cycles.
ORC_API int orc_set_galerkin_precision(tos_ctx* c, int prec) {
  if (!c || prec < 0 || prec > 1) return TOS_E_ARG;
  return TOS_OK;
}

// Graph replay is a GPU launch detail; the oracle accepts and ignores the switch.
ORC_API int orc_set_graph_replay(tos_ctx* c, int enable) {
  (void)enable;
  return c ? TOS_OK : TOS_E_ARG;
}

// a measurement of matrix-core kernels: nothing to measure on the host
ORC_API int orc_probe_element_mfma(tos_ctx* c, int reps, double* out) {
  (void)reps;
  (void)out;
  if (!c) return TOS_E_ARG;
  c->err = "the matrix-core measurement runs on the GPU library only";
  return TOS_E_STATE;
}
