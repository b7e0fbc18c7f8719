// tos_api.hip -- host implementation of include/tos_api.h for gfx950.
//
// One context = one HIP stream; every call enqueues kernels on that stream and only
// synchronises where a result must reach the host (convergence checks, returned
// scalars, copies out).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "tos_internal.hpp"

#include "kernels_mg.hip"
#include "kernels_batch.hip"
#include "kernels_nn.hip"
#include "kernels_opt.hip"
#include "kernels_surface.hip"
#include "kernels_probe.hip"
#include "mma.hip"

#define TOS_EXPORT extern "C" __attribute__((visibility("default")))

using namespace tos;

struct tos_ctx : tos::Ctx {};

struct tos_mma {
  int dev = 0;
  hipStream_t stream = nullptr;
  MmaVecs V{};
  int nb = 1;                 // blocks of the n-length passes
  double* partial = nullptr;  // MMA_CNT x nb
  double* red = nullptr;      // folded columns (device)
  double* h_red = nullptr;    // pinned host mirror
  double* stage = nullptr;    // df / dg staging
  std::vector<void*> allocs;
  // scalar state of the subproblem (host)
  double y[MMAX], lam[MMAX], mu[MMAX], s[MMAX], a[MMAX], c[MMAX], d[MMAX], b[MMAX];
  double z = 1, zeta = 1, eps = 1, move = 0.5;
  int stop_counter = 0;
  int last_newton = 0;
};

namespace {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH_CHECK() HIPCHK(hipGetLastError())

void* raw_alloc(std::vector<void*>& list, hipStream_t s, size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 8;
  HIPCHK(hipMalloc(&p, bytes));
  HIPCHK(hipMemsetAsync(p, 0, bytes, s));
  list.push_back(p);
  return p;
}

template <class T>
T* dalloc(Ctx& c, size_t n) {
  ++c.alloc_epoch;
  return (T*)raw_alloc(c.allocs, c.stream, n * sizeof(T));
}

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// launch shape of the stored-stencil paired GS (one wave per segment, no block-wide
// reduction): four waves per workgroup, or on small levels (fewer than 1024 segments) one,
// so that the few waves of a latency-bound pass spread over as many CUs
constexpr int kSpreadBelow = 1024;
inline dim3 spread_grid(int64_t waves) {
  return dim3(waves < kSpreadBelow ? (unsigned)waves : (unsigned)((waves + 3) / 4));
}
inline dim3 spread_block(int64_t waves) { return dim3(waves < kSpreadBelow ? 64 : 256); }

}  // namespace

#include "unet_host.hpp"

namespace {

// ---------------------------------------------------------------------------------------
// Element stiffness of the 8-node trilinear hex (element/templateMatrix.cpp:31-99):
// 2x2x2 Gauss points at (1 +- 1/sqrt3)/2, isotropic elasticity, weight (h/2)^3.
// Written in Lame form: K_ab(r,s) = lam g_a,r g_b,s + mu g_a,s g_b,r + mu d_rs g_a.g_b.
// ---------------------------------------------------------------------------------------
void hex_stiffness(double h, double E, double nu, double* K) {
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
  std::fill(K, K + 576, 0.0);
  const double gq[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
  for (int g = 0; g < 8; ++g) {
    const double pt[3] = {gq[g & 1], gq[(g >> 1) & 1], gq[g >> 2]};
    double grad[8][3];
    for (int a = 0; a < 8; ++a) {
      const int b[3] = {a & 1, (a >> 1) & 1, a >> 2};
      double N1[3], dN1[3];
      for (int k = 0; k < 3; ++k) {
        N1[k] = b[k] ? pt[k] : 1 - pt[k];
        dN1[k] = (b[k] ? 1.0 : -1.0) / h;
      }
      grad[a][0] = dN1[0] * N1[1] * N1[2];
      grad[a][1] = N1[0] * dN1[1] * N1[2];
      grad[a][2] = N1[0] * N1[1] * dN1[2];
    }
    const double wq = (h * h * h) / 8.0;
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) {
        const double dd = grad[a][0] * grad[b][0] + grad[a][1] * grad[b][1] + grad[a][2] * grad[b][2];
        for (int r = 0; r < 3; ++r)
          for (int s = 0; s < 3; ++s)
            K[(3 * a + r) * 24 + 3 * b + s] +=
                wq * (lam * grad[a][r] * grad[b][s] + mu * grad[a][s] * grad[b][r] + (r == s ? mu * dd : 0.0));
      }
  }
}

double pweight_h(int c, int i, int m) {
  double w = 1;
  for (int a = 0; a < 3; ++a) {
    const int t = ((c >> a) & 1) + ((i >> a) & 1);
    w *= ((m >> a) & 1) ? t * 0.5 : (2 - t) * 0.5;
  }
  return w;
}

// Mc[c] = P_c^T Ke P_c for the 8 child positions c of a fine element in its coarse element
void child_products(const double* Ke, double* Mc) {
  for (int c = 0; c < 8; ++c) {
    double P[8][8];
    for (int i = 0; i < 8; ++i)
      for (int m = 0; m < 8; ++m) P[i][m] = pweight_h(c, i, m);
    static thread_local double T[24][24];
    for (int i = 0; i < 24; ++i)
      for (int n = 0; n < 8; ++n)
        for (int s = 0; s < 3; ++s) {
          double acc = 0;
          for (int j = 0; j < 8; ++j) acc += Ke[i * 24 + 3 * j + s] * P[j][n];
          T[i][3 * n + s] = acc;
        }
    for (int m = 0; m < 8; ++m)
      for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 24; ++col) {
          double acc = 0;
          for (int i = 0; i < 8; ++i) acc += P[i][m] * T[3 * i + r][col];
          Mc[c * 576 + (3 * m + r) * 24 + col] = acc;
        }
  }
}

// The same for a factor-4 coarsening (skip-layer): the 64 fine elements of a coarse
// element at positions c = cx + 4 cy + 16 cz; fine node i of child c sits at t = c + i
// (0..4 per axis) inside the coarse element, with weight prod (m ? t : 4 - t) / 4 from
// coarse corner m (the factor-4 hat of grid/Grid.cu:1130-1136, 2635-2671):
// Mc4[c] = P_c^T Ke P_c.
void child_products4(const double* Ke, double* Mc4) {
  for (int c = 0; c < 64; ++c) {
    const int cp[3] = {c & 3, (c >> 2) & 3, c >> 4};
    double P[8][8];
    for (int i = 0; i < 8; ++i)
      for (int m = 0; m < 8; ++m) {
        double w = 1;
        for (int a = 0; a < 3; ++a) {
          const int t = cp[a] + ((i >> a) & 1);
          w *= ((m >> a) & 1) ? t / 4.0 : (4 - t) / 4.0;
        }
        P[i][m] = w;
      }
    for (int m = 0; m < 8; ++m)
      for (int r = 0; r < 3; ++r)
        for (int n = 0; n < 8; ++n)
          for (int s2 = 0; s2 < 3; ++s2) {
            double acc = 0;
            for (int i = 0; i < 8; ++i) {
              if (P[i][m] == 0) continue;
              double row = 0;
              for (int j = 0; j < 8; ++j) row += Ke[(3 * i + r) * 24 + 3 * j + s2] * P[j][n];
              acc += P[i][m] * row;
            }
            Mc4[c * 576 + (3 * m + r) * 24 + 3 * n + s2] = acc;
          }
  }
}

LevelView make_view(int ex, int ey, int ez) {
  LevelView v{};
  v.ex = ex; v.ey = ey; v.ez = ez;
  v.nx = ex + 1; v.ny = ey + 1; v.nz = ez + 1;
  v.NX = v.nx + 3; v.NY = v.ny + 3;
  v.NP = ((int64_t)v.NX * v.NY * (v.nz + 3) + 63) / 64 * 64;
  v.EX = ex + 2; v.EY = ey + 2;
  v.NE = ((int64_t)v.EX * v.EY * (ez + 2) + 63) / 64 * 64;
  return v;
}

// Hierarchy on the host: coarse element active iff a child is active
// (setSolidElementFromFineGrid_kernel grid/Grid.cu:4453-4473), coarsen while more than
// `threshold` active elements (grid/Grid.cpp:747).  Free nodes: fine = touches an active
// element and is not supported; coarse = some free fine node lies inside the support of
// its trilinear basis function.
// ---------------------------------------------------------------------------------------
// Rigid motions (floating bodies).  R^T v at one node, host side (the device's rigid_dot).
// ---------------------------------------------------------------------------------------
void rigid_dot_h(double px, double py, double pz, const double v[3], double b[6]) {
  b[0] += v[0];
  b[1] += v[1];
  b[2] += v[2];
  b[3] += pz * v[1] - py * v[2];
  b[4] += px * v[2] - pz * v[0];
  b[5] += py * v[0] - px * v[1];
}

// Gram-Schmidt over the six columns given their Gram matrix G, in column order, dropping
// a column whose residual norm is below 1e-6 of its norm or whose norm is below 1e-6
// (the reference's orthonormalisation, optimization/projection.cpp:234-255): the kept
// columns q_j = R t_j are orthonormal and M = sum_j t_j t_j^T projects, R M R^T v.
RigidProj rigid_projector(const double G[36], const double o[3]) {
  RigidProj P{};
  double T[6][6] = {};
  bool keep[6] = {};
  for (int i = 0; i < 6; ++i) {
    double t[6] = {};
    t[i] = 1.0;
    for (int j = 0; j < i; ++j) {
      if (!keep[j]) continue;
      double d = 0;  // (R e_i) . q_j
      for (int k = 0; k < 6; ++k) d += T[j][k] * G[i * 6 + k];
      for (int k = 0; k < 6; ++k) t[k] -= d * T[j][k];
    }
    double r2 = 0;
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) r2 += t[a] * G[a * 6 + b] * t[b];
    const double old2 = G[i * 6 + i];
    keep[i] = old2 >= 1e-12 && r2 > 0 && r2 >= 1e-12 * old2;
    if (!keep[i]) continue;
    const double inv = 1.0 / std::sqrt(r2);
    for (int k = 0; k < 6; ++k) T[i][k] = t[k] * inv;
    ++P.m;
  }
  for (int j = 0; j < 6; ++j)
    if (keep[j])
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) P.M[a * 6 + b] += T[j][a] * T[j][b];
  for (int k = 0; k < 3; ++k) P.o[k] = o[k];
  return P;
}

// projector over the marked nodes of an nx*ny*nz lattice (no normals): with coordinates
// relative to their centroid the Gram matrix is [[N I, 0], [0, sum(|p|^2 I - p p^T)]]
RigidProj rigid_projector_nodes(const std::vector<uint8_t>& mark, int nx, int ny, int nz) {
  double s0 = 0, s[3] = {0, 0, 0};
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        if (mark[x + (size_t)nx * (y + (size_t)ny * z)]) {
          s0 += 1;
          s[0] += x; s[1] += y; s[2] += z;
        }
  const double o[3] = {s[0] / s0, s[1] / s0, s[2] / s0};
  double S[3][3] = {};
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        if (mark[x + (size_t)nx * (y + (size_t)ny * z)]) {
          const double p[3] = {x - o[0], y - o[1], z - o[2]};
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) S[a][b] += p[a] * p[b];
        }
  double G[36] = {};
  const double tr = S[0][0] + S[1][1] + S[2][2];
  for (int a = 0; a < 3; ++a) {
    G[a * 6 + a] = s0;
    for (int b = 0; b < 3; ++b) G[(3 + a) * 6 + 3 + b] = (a == b ? tr : 0.0) - S[a][b];
  }
  return rigid_projector(G, o);
}

// projector over a load region: node coordinates `pos` (3 per node, any origin), the
// 3-vectors taken along the unit normals when given; writes the coordinates relative to
// the projector's origin into `rel`
RigidProj rigid_projector_region(int n, const double* pos, const double* nrm, std::vector<double>& rel) {
  double o[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) o[k] += pos[3 * (size_t)i + k] / n;
  rel.resize(3 * (size_t)n);
  double G[36] = {};
  for (int i = 0; i < n; ++i) {
    const double p[3] = {pos[3 * (size_t)i] - o[0], pos[3 * (size_t)i + 1] - o[1], pos[3 * (size_t)i + 2] - o[2]};
    for (int k = 0; k < 3; ++k) rel[3 * (size_t)i + k] = p[k];
    // G = sum_i R_i^T N_i R_i: with N_i = n n^T one vector a = R_i^T n, else the three R_i^T e_c
    for (int c = 0; c < (nrm ? 1 : 3); ++c) {
      double v[3] = {0, 0, 0};
      if (nrm) {
        for (int k = 0; k < 3; ++k) v[k] = nrm[3 * (size_t)i + k];
      } else {
        v[c] = 1.0;
      }
      double a[6] = {0, 0, 0, 0, 0, 0};
      rigid_dot_h(p[0], p[1], p[2], v, a);
      for (int r = 0; r < 6; ++r)
        for (int q = 0; q < 6; ++q) G[r * 6 + q] += a[r] * a[q];
    }
  }
  return rigid_projector(G, o);
}

void build_levels(Ctx& c, int n_levels, int threshold, bool skip) {
  struct HL {
    int ex, ey, ez;
    std::vector<uint8_t> act;
  };
  std::vector<HL> hl;
  hl.push_back({c.nx, c.ny, c.nz, c.h_solid});
  for (auto& a : hl[0].act) a = a ? 1 : 0;
  auto nact = [](const HL& l) {
    int64_t s = 0;
    for (auto a : l.act) s += a;
    return s;
  };
  for (;;) {
    const HL& f = hl.back();
    const bool more = n_levels > 0 ? (int)hl.size() < n_levels : nact(f) > threshold;
    if (!more || (f.ex == 1 && f.ey == 1 && f.ez == 1) || hl.size() >= 16) break;
    HL C{(f.ex + 1) / 2, (f.ey + 1) / 2, (f.ez + 1) / 2, {}};
    C.act.assign((size_t)C.ex * C.ey * C.ez, 0);
    for (int z = 0; z < f.ez; ++z)
      for (int y = 0; y < f.ey; ++y)
        for (int x = 0; x < f.ex; ++x)
          if (f.act[x + (size_t)f.ex * (y + (size_t)f.ey * z)])
            C.act[(x >> 1) + (size_t)C.ex * ((y >> 1) + (size_t)C.ey * (z >> 1))] = 1;
    hl.push_back(std::move(C));
  }
  // skip-layer: drop dyadic level 1 (level 2's elements are the 4x4x4 blocks of level 0)
  c.skip = skip && hl.size() >= 3;
  if (c.skip) hl.erase(hl.begin() + 1);
  c.L.resize(hl.size());
  for (size_t l = 0; l < hl.size(); ++l) {
    Level& L = c.L[l];
    L.v = make_view(hl[l].ex, hl[l].ey, hl[l].ez);
    L.h_eact = std::move(hl[l].act);
    int64_t na = 0;
    for (auto a : L.h_eact) na += a;
    L.n_active = na;
    const LevelView& v = L.v;
    L.h_free.assign((size_t)v.nx * v.ny * v.nz, 0);
    if (l == 0) {
      for (int z = 0; z < v.ez; ++z)
        for (int y = 0; y < v.ey; ++y)
          for (int x = 0; x < v.ex; ++x) {
            if (!L.h_eact[x + (size_t)v.ex * (y + (size_t)v.ey * z)]) continue;
            for (int j = 0; j < 8; ++j)
              L.h_free[(x + (j & 1)) + (size_t)v.nx * ((y + ((j >> 1) & 1)) + (size_t)v.ny * (z + (j >> 2)))] = 1;
          }
      int64_t n_act_nodes = 0;
      for (auto a : L.h_free) n_act_nodes += a;
      if (!c.h_support.empty())
        for (size_t i = 0; i < L.h_free.size(); ++i)
          if (c.h_support[i]) L.h_free[i] = 0;
      int64_t n_free0 = 0;
      for (auto a : L.h_free) n_free0 += a;
      c.floating = n_free0 > 0 && n_free0 == n_act_nodes;
      if (c.floating) c.rig_all = rigid_projector_nodes(L.h_free, v.nx, v.ny, v.nz);
    } else if (l == 1 && c.skip) {
      // every node of an active element carries an unknown (supports enter the Galerkin
      // operator as a penalty, not as eliminated nodes)
      for (int z = 0; z < v.ez; ++z)
        for (int y = 0; y < v.ey; ++y)
          for (int x = 0; x < v.ex; ++x) {
            if (!L.h_eact[x + (size_t)v.ex * (y + (size_t)v.ey * z)]) continue;
            for (int j = 0; j < 8; ++j)
              L.h_free[(x + (j & 1)) + (size_t)v.nx * ((y + ((j >> 1) & 1)) + (size_t)v.ny * (z + (j >> 2)))] = 1;
          }
    } else {
      const Level& F = c.L[l - 1];
      const LevelView& fv = F.v;
      for (int z = 0; z < fv.nz; ++z)
        for (int y = 0; y < fv.ny; ++y)
          for (int x = 0; x < fv.nx; ++x) {
            if (!F.h_free[x + (size_t)fv.nx * (y + (size_t)fv.ny * z)]) continue;
            for (int Z = z >> 1; Z <= (z + 1) >> 1; ++Z)
              for (int Y = y >> 1; Y <= (y + 1) >> 1; ++Y)
                for (int X = x >> 1; X <= (x + 1) >> 1; ++X)
                  if (X < v.nx && Y < v.ny && Z < v.nz) L.h_free[X + (size_t)v.nx * (Y + (size_t)v.ny * Z)] = 1;
          }
    }
    int64_t nf = 0;
    for (auto a : L.h_free) nf += a;
    L.n_free = nf;
  }
}

std::vector<uint8_t> pad_nodes(const LevelView& v, const std::vector<uint8_t>& src) {
  std::vector<uint8_t> out(v.NP, 0);
  for (int z = 0; z < v.nz; ++z)
    for (int y = 0; y < v.ny; ++y)
      for (int x = 0; x < v.nx; ++x) out[v.pnode(x, y, z)] = src[x + (size_t)v.nx * (y + (size_t)v.ny * z)];
  return out;
}

std::vector<uint8_t> pad_elems(const LevelView& v, const uint8_t* src) {
  std::vector<uint8_t> out(v.NE, 0);
  if (!src) return out;
  for (int z = 0; z < v.ez; ++z)
    for (int y = 0; y < v.ey; ++y)
      for (int x = 0; x < v.ex; ++x) out[v.pelem(x, y, z)] = src[x + (size_t)v.ex * (y + (size_t)v.ey * z)] ? 1 : 0;
  return out;
}

void upload(hipStream_t s, void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));  // src is pageable host memory owned by the caller
}
void download(hipStream_t s, void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
}

void* stage(Ctx& c, size_t bytes) {
  if (c.stage_bytes < bytes) {
    if (c.stage) HIPCHK(hipFree(c.stage));
    c.stage = nullptr;
    ++c.alloc_epoch;
    HIPCHK(hipMalloc(&c.stage, bytes));
    c.stage_bytes = bytes;
  }
  return c.stage;
}

// ---------------------------------------------------------------------------------------
// Chunk lists (see kernels_mg.hip): runs of free nodes along x, at most 64 per chunk
// ---------------------------------------------------------------------------------------
void build_chunks(Ctx& c, Level& L, bool with_slots) {
  const LevelView& v = L.v;
  auto freeat = [&](int x, int y, int z) { return L.h_free[x + (size_t)v.nx * (y + (size_t)v.ny * z)] != 0; };
  auto emit = [&](std::vector<int4>& out, int x0, int y, int z, int cnt) {
    out.push_back(make_int4((int)v.pnode(x0, y, z), x0 | (cnt << 16), y, z));
  };
  std::vector<int4> lin, col;
  for (int z = 0; z < v.nz; ++z)
    for (int y = 0; y < v.ny; ++y) {
      int x = 0;
      while (x < v.nx) {
        if (!freeat(x, y, z)) { ++x; continue; }
        int x1 = x;
        while (x1 < v.nx && freeat(x1, y, z) && x1 - x < 64) ++x1;
        emit(lin, x, y, z, x1 - x);
        x = x1;
      }
    }
  for (int cidx = 0; cidx < 8; ++cidx) {
    L.col_off[cidx] = (int)col.size();
    const int px = cidx & 1, py = (cidx >> 1) & 1, pz = cidx >> 2;
    for (int z = pz; z < v.nz; z += 2)
      for (int y = py; y < v.ny; y += 2) {
        int x = px;
        while (x < v.nx) {
          if (!freeat(x, y, z)) { x += 2; continue; }
          int x1 = x, cnt = 0;
          while (x1 < v.nx && freeat(x1, y, z) && cnt < 64) { x1 += 2; ++cnt; }
          emit(col, x, y, z, cnt);
          x = x1;
        }
      }
  }
  L.col_off[8] = (int)col.size();
  // row lists of the two-colour GS launches (fine level)
  if (!with_slots) {
    for (int cls = 0; cls < 4; ++cls) {
      const int A = 2 * cls, B = A + 1;  // colours (even x, odd x) of class (y%2, z%2) = (cls&1, cls>>1)
      int ia = L.col_off[A], ea = L.col_off[A + 1], ib = L.col_off[B], eb = L.col_off[B + 1];
      std::vector<int4> fwd, bwd;
      auto key = [&](int i) { return (int64_t)col[i].w * (1 << 20) + col[i].z; };
      while (ia < ea || ib < eb) {
        const int64_t k = std::min(ia < ea ? key(ia) : INT64_MAX, ib < eb ? key(ib) : INT64_MAX);
        const int a0 = ia, b0 = ib;
        while (ia < ea && key(ia) == k) ++ia;
        while (ib < eb && key(ib) == k) ++ib;
        fwd.push_back(make_int4(a0, ia - a0, b0, ib - b0));
        bwd.push_back(make_int4(b0, ib - b0, a0, ia - a0));
      }
      L.n_pair[cls] = (int)fwd.size();
      L.pair_fwd[cls] = dalloc<int4>(c, std::max<size_t>(fwd.size(), 1));
      L.pair_bwd[cls] = dalloc<int4>(c, std::max<size_t>(bwd.size(), 1));
      if (!fwd.empty()) {
        upload(c.stream, L.pair_fwd[cls], fwd.data(), fwd.size() * sizeof(int4));
        upload(c.stream, L.pair_bwd[cls], bwd.data(), bwd.size() * sizeof(int4));
      }
    }
  }
  // row segments of the in-wave two-colour GS: per row of class (y%2, z%2), the free
  // nodes' x range [xmin, xe], first-colour parity P; segments start at xs (xmin or
  // xmin - 1, of parity P) and every 126 nodes after it until the last c2 node is covered.
  // Stencil levels use them only when every row is a single segment.
  if (v.nx + 1 < 0x7fff) {
    bool single = true;
    for (int cls = 0; cls < 4; ++cls)
      for (int P = 0; P < 2; ++P) {
        const int py = cls & 1, pz = cls >> 1;
        std::vector<int4> sg;
        for (int z = pz; z < v.nz; z += 2)
          for (int y = py; y < v.ny; y += 2) {
            int xmin = -1, xmax = -1;
            for (int x = 0; x < v.nx; ++x)
              if (freeat(x, y, z)) {
                if (xmin < 0) xmin = x;
                xmax = x;
              }
            if (xmin < 0) continue;
            const int xs = ((xmin & 1) == P) ? xmin : xmin - 1;
            const int nseg = (xmax - xs + 1 + 125) / 126;
            if (nseg > 1) single = false;
            for (int q = 0; q < nseg; ++q) {
              const int x0 = xs + 126 * q;
              sg.push_back(make_int4((int)v.pnode(x0, y, z), (x0 + 1) | ((xmax + 1) << 15) | ((q > 0 ? 1 : 0) << 30),
                                     y, z));
            }
          }
        L.n_segw[cls][P] = (int)sg.size();
        L.segw[cls][P] = dalloc<int4>(c, std::max<size_t>(sg.size(), 1));
        if (!sg.empty()) upload(c.stream, L.segw[cls][P], sg.data(), sg.size() * sizeof(int4));
      }
    L.pair_ok = with_slots;
    L.pair_single = single;
    std::vector<int4> sp;  // all rows, even x first
    for (int z = 0; z < v.nz; ++z)
      for (int y = 0; y < v.ny; ++y) {
        int xmin = -1, xmax = -1;
        for (int x = 0; x < v.nx; ++x)
          if (freeat(x, y, z)) {
            if (xmin < 0) xmin = x;
            xmax = x;
          }
        if (xmin < 0) continue;
        const int xs = xmin & ~1;
        const int nseg = (xmax - xs + 1 + 125) / 126;
        for (int q = 0; q < nseg; ++q) {
          const int x0 = xs + 126 * q;
          sp.push_back(make_int4((int)v.pnode(x0, y, z), (x0 + 1) | ((xmax + 1) << 15) | ((q > 0 ? 1 : 0) << 30), y, z));
        }
      }
    L.n_segp = (int)sp.size();
    L.segp = dalloc<int4>(c, std::max<size_t>(sp.size(), 1));
    if (!sp.empty()) upload(c.stream, L.segp, sp.data(), sp.size() * sizeof(int4));
  } else if (!with_slots) {
    throw ArgError("grid too wide for the row segments (nx < 32766)");
  }
  // colour-slab wavefront of the fine GS: ~32 z-slabs so that the slabs in flight during a
  // few consecutive steps fit in the memory-side cache
  if (!with_slots) {
    L.wf_T = std::max(1, (v.nz + 31) / 32);
    L.wf_S = (v.nz + L.wf_T - 1) / L.wf_T;
    L.slab_off.assign(8 * (L.wf_S + 1), 0);
    for (int cidx = 0; cidx < 8; ++cidx) {
      int g = L.col_off[cidx];
      for (int sl = 0; sl <= L.wf_S; ++sl) {
        while (g < L.col_off[cidx + 1] && col[g].w < sl * L.wf_T) ++g;
        L.slab_off[cidx * (L.wf_S + 1) + sl] = g;
      }
    }
  }
  L.n_lin = (int)lin.size();
  L.n_col = (int)col.size();
  L.cl_lin = dalloc<int4>(c, std::max<size_t>(lin.size(), 1));
  L.cl_col = dalloc<int4>(c, std::max<size_t>(col.size(), 1));
  if (!lin.empty()) upload(c.stream, L.cl_lin, lin.data(), lin.size() * sizeof(int4));
  if (!col.empty()) upload(c.stream, L.cl_col, col.data(), col.size() * sizeof(int4));
  L.NS = 64 * (int64_t)L.n_col;
  if (with_slots) {
    L.h_slot.assign(v.NP, -1);
    for (size_t g = 0; g < col.size(); ++g) {
      const int cnt = col[g].y >> 16;
      for (int l = 0; l < cnt; ++l) L.h_slot[col[g].x + 2 * l] = (int)(g * 64 + l);
    }
    L.slot = dalloc<int>(c, v.NP);
    upload(c.stream, L.slot, L.h_slot.data(), v.NP * sizeof(int));
  }
}

inline unsigned cblk(int nch) { return (unsigned)std::max(1, (nch + 3) / 4); }

// ---------------------------------------------------------------------------------------
// Reductions (results stay on the device in c.scal)
// ---------------------------------------------------------------------------------------
const double* read_scal(Ctx& c) {
  download(c.stream, c.h_scal, c.scal, sizeof(double) * S_COUNT);
  return c.h_scal;
}

// ---------------------------------------------------------------------------------------
// Kernel timing: events recorded around launches while c.prof_on
// ---------------------------------------------------------------------------------------
enum {
  P_FINE_APPLY64 = 0, P_FINE_RESID64, P_FINE_RESID32, P_FINE_GS64, P_FINE_GS32, P_ST_GS64, P_ST_GS32,
  P_ST_RESID, P_TRANSFER, P_COARSE, P_GALERKIN, P_COARSE_FACTOR, P_VECTOR, P_SENS_FILTER, P_SPLINE, P_OPT,
  P_ST_GS32_L1, P_LEARNED, P_COUNT
};
static_assert(P_COUNT <= TOS_PROF_CLASSES, "timing classes");
const char* kProfNames[P_COUNT] = {"fine_apply_f64", "fine_resid_f64", "fine_resid_f32", "fine_gs_f64",
                                   "fine_gs_f32",    "st_gs_f64",      "st_gs_f32",      "st_resid",
                                   "transfer",       "coarse_solve",   "galerkin",       "coarse_factor",
                                   "vector",         "sens_filter",    "spline",         "oc_mma",
                                   "st_gs_f32_l1",   "learned"};

hipEvent_t next_event(Ctx& c) {
  if (c.ev_used == c.ev_pool.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c.ev_pool.push_back(e);
  }
  return c.ev_pool[c.ev_used++];
}

struct Prof {
  Ctx& c;
  int cls;
  hipEvent_t a{};
  Prof(Ctx& c_, int cls_) : c(c_), cls(cls_) {
    if (c.prof_on) {
      a = next_event(c);
      HIPCHK(hipEventRecord(a, c.stream));
    }
  }
  ~Prof() {
    if (c.prof_on) {
      hipEvent_t b = next_event(c);
      (void)hipEventRecord(b, c.stream);
      c.prof.push_back({cls, a, b});
    }
  }
};

// finisher over the partials of a launch over the fine chunk list
// (two stages when there are many blocks: one long single-block loop is latency bound)
void finish0(Ctx& c, int two, int s0, int s1, int op) {
  if (c.nb0 > 4 * kStage) {
    double* st = c.partial + 2 * (size_t)std::max(kRedBlocks, c.nb0);
    hipLaunchKernelGGL(k_reduce_partials, dim3(kStage), dim3(256), 0, c.stream, (const double*)c.partial, c.nb0,
                       two, st);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, c.stream, (const double*)st, kStage, two, s0, s1, op,
                       c.scal);
    return;
  }
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, c.stream, (const double*)c.partial, c.nb0, two, s0, s1, op,
                     c.scal);
}

// scal[s0], scal[s1] = sums of the two sets of np partials part[0 .. np), part[np .. 2 np)
void finish_parts(Ctx& c, const double* part, int64_t np, int s0, int s1, int op) {
  double* st = c.partial + 2 * (size_t)std::max(kRedBlocks, c.nb0);
  hipLaunchKernelGGL(k_reduce_partials, dim3(kStage), dim3(256), 0, c.stream, part, (int)np, 1, st);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, c.stream, (const double*)st, kStage, 1, s0, s1, op, c.scal);
}

// scal[s0] = a.b (and scal[s1] = a2.b2) over the free fine nodes
void dot_to(Ctx& c, const double* a, const double* b, const double* a2, const double* b2, int s0, int s1, int op) {
  Prof pr(c, P_VECTOR);
  Level& L0 = c.L[0];
  hipLaunchKernelGGL(k_dot_nodes, dim3(c.nb0), dim3(256), 0, c.stream, (const int4*)L0.cl_lin, L0.n_lin, L0.v.NP, a,
                     b, a2, b2, c.partial);
  finish0(c, a2 ? 1 : 0, s0, s1, op);
  LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Multigrid (V-cycle, HierarchyGrid::v_cycle grid/Grid.cpp:2399-2444)
// ---------------------------------------------------------------------------------------
template <class T>
struct Prec;
template <>
struct Prec<double> {
  static double* u(Level& L) { return L.u64; }
  static double* f(Level& L) { return L.f64; }
  static double* r(Level& L) { return L.r64; }
  static const double* st(Level& L) { return L.st64; }
  static const double* w(Ctx& c) { return c.w64; }
  static const double* Ke(Ctx& c) { return c.Kc64; }
};
template <>
struct Prec<float> {
  static float* u(Level& L) { return L.u32; }
  static float* f(Level& L) { return L.f32; }
  static float* r(Level& L) { return L.r32; }
  static const float* st(Level& L) { return L.st32; }
  static const float* w(Ctx& c) { return c.w32; }
  static const float* Ke(Ctx& c) { return c.Kc32; }
};

// calls fn with the reader of level L's stencil: the FP16 copy in an FP32 cycle of a
// solve with precision 2, else the stencil at the cycle's precision
template <class T, class Fn>
void st_dispatch(Ctx& c, Level& L, Fn&& fn) {
  if constexpr (std::is_same<T, float>::value) {
    if (c.half_st) {
      fn(StHalf{L.st16, L.st_m, L.NS});
      return;
    }
  }
  fn(StFull<T>{Prec<T>::st(L), L.NS});
}

// FP16 copies of the FP32 stencils of every stored-stencil level (allocated on first use),
// on stream s (c.stream, or c.stream2 beside the coarsest factor once allocated)
void pack_half_stencils(Ctx& c, hipStream_t s) {
  for (size_t l = 1; l < c.L.size(); ++l) {
    Level& L = c.L[l];
    if (!L.n_col) continue;
    if (!L.st16) {
      L.st16 = dalloc<uint32_t>(c, (size_t)kHalfWords * L.NS);
      L.st_m = dalloc<float>(c, L.NS);
    }
    Prof pr(c, P_GALERKIN);
    hipLaunchKernelGGL(k_st_pack16, dim3(cblk(L.n_col)), dim3(256), 0, s, (const int4*)L.cl_col, L.n_col,
                       (const float*)L.st32, L.NS, L.st16, L.st_m);
  }
  LAUNCH_CHECK();
  c.st16_valid = true;
}

// the FP16 copies can be packed on the second stream: allocated, and no event timing
bool pack_beside_ok(const Ctx& c) {
  if (c.prof_on || !c.want16 || c.gal_prec == 0) return false;
  for (size_t l = 1; l < c.L.size(); ++l)
    if (c.L[l].n_col && !c.L[l].st16) return false;
  return true;
}

int ensure_fp64_operator(Ctx& c);

// precision 2 = FP32 cycle on FP16 stored stencils: make sure the copies are current;
// precision 0 = FP64 cycle: make sure the FP64 stencils are (Galerkin precision 1)
int select_stencils(Ctx& c, int precision) {
  c.half_st = precision == 2;
  c.want16 = c.half_st;
  if (c.half_st && !c.st16_valid) pack_half_stencils(c, c.stream);
  return precision == 0 ? ensure_fp64_operator(c) : TOS_OK;
}

// coarsest-level factor for a cycle of type T (FP32 cycles use the one of the FP32-built
// stencils when the operator was built in FP32)
template <class T>
const double* coarse_inverse(const Ctx& c) {
  return (std::is_same<T, double>::value || c.gal_prec == 0) ? c.cinv : c.cinv32;
}

// Fine-level sweep as a colour-slab wavefront: step t relaxes colour c (in sweep order
// c' = 0..7) on z-slab s = t - 2 c'.  Neighbouring nodes differ by at most one slab, so a
// node of colour c' is relaxed at least one step after every neighbour of an earlier
// colour and before every neighbour of a later one: the same updates as eight full colour
// passes (bit-identical).  Meant to reuse each slab's data across its 8 colours; on the
// 511x320x320 workload the L2 misses per sweep stayed the same as with colour passes and
// the 5x more launches made it slower, so it is opt-in (tos_set_sweep_mode).
template <class T>
void gs_sweep_wavefront(Ctx& c, T* u, const T* f, bool backward) {
  Level& L = c.L[0];
  const int S = L.wf_S;
  for (int t = 0; t < S + 14; ++t) {
    SegTab tab{};
    int nb = 0;
    for (int cc = 0; cc < 8; ++cc) {
      const int sl = t - 2 * cc;
      if (sl < 0 || sl >= S) continue;
      const int col = backward ? 7 - cc : cc;
      const int* so = L.slab_off.data() + col * (S + 1);
      const int b = so[sl], e = so[sl + 1];
      if (e <= b) continue;
      tab.beg[tab.n] = b;
      tab.cnt[tab.n] = e - b;
      tab.blk[tab.n] = nb;
      nb += (e - b + 3) / 4;
      ++tab.n;
    }
    if (!tab.n) continue;
    tab.blk[tab.n] = nb;
    Prof pr(c, sizeof(T) == 8 ? P_FINE_GS64 : P_FINE_GS32);
    if (backward)
      hipLaunchKernelGGL((k_fine_gs_wf<T, true>), dim3(nb), dim3(256), 0, c.stream, L.v, (const int4*)L.cl_col, tab,
                         Prec<T>::w(c), Prec<T>::Ke(c), u, f);
    else
      hipLaunchKernelGGL((k_fine_gs_wf<T, false>), dim3(nb), dim3(256), 0, c.stream, L.v, (const int4*)L.cl_col, tab,
                         Prec<T>::w(c), Prec<T>::Ke(c), u, f);
  }
  LAUNCH_CHECK();
}

#define TOS_COLOUR_SWITCH(col, GO) \
  switch (col) {                   \
    case 0: GO(0); break;          \
    case 1: GO(1); break;          \
    case 2: GO(2); break;          \
    case 3: GO(3); break;          \
    case 4: GO(4); break;          \
    case 5: GO(5); break;          \
    case 6: GO(6); break;          \
    default: GO(7); break;         \
  }

// Fine-level sweep in four two-colour launches (k_fine_gs_pair): classes (y%2, z%2) in
// colour order, forward (0,0) (1,0) (0,1) (1,1), backward the reverse.
template <class T>
void gs_sweep_pairs(Ctx& c, T* u, const T* f, bool backward, bool from_zero) {
  Level& L = c.L[0];
  const LevelView& v = L.v;
  for (int i = 0; i < 4; ++i) {
    const int cls = backward ? 3 - i : i;
    const int n = L.n_pair[cls];
    if (!n) continue;
    Prof pr(c, sizeof(T) == 8 ? P_FINE_GS64 : P_FINE_GS32);
    const dim3 g(n), b(256);
    const int4* cl = L.cl_col;
    if (backward) {
      hipLaunchKernelGGL((k_fine_gs_pair<T, true, -1, -1>), g, b, 0, c.stream, v, cl, (const int4*)L.pair_bwd[cls],
                         n, Prec<T>::w(c), Prec<T>::Ke(c), u, f);
    } else if (!from_zero) {
      hipLaunchKernelGGL((k_fine_gs_pair<T, false, -1, -1>), g, b, 0, c.stream, v, cl, (const int4*)L.pair_fwd[cls],
                         n, Prec<T>::w(c), Prec<T>::Ke(c), u, f);
    } else {
#define GO(A) hipLaunchKernelGGL((k_fine_gs_pair<T, false, A, A + 1>), g, b, 0, c.stream, v, cl, \
                                 (const int4*)L.pair_fwd[cls], n, Prec<T>::w(c), Prec<T>::Ke(c), u, f)
      switch (cls) {
        case 0: GO(0); break;
        case 1: GO(2); break;
        case 2: GO(4); break;
        default: GO(6); break;
      }
#undef GO
    }
  }
  LAUNCH_CHECK();
}

// One colour-ordered sweep.  from_zero: u is 0 on entry (the first pre-smoothing sweep of a
// V-cycle level), so colour c reads only the neighbours of colours < c; same updates as
// a sweep over a zeroed u, without zeroing it first or reading the zeros.
// Fine-level FP32 sweep in four in-wave two-colour passes (k_fine_gs_pairw), each followed
// by the copy of its deferred nodes; classes in colour order, forward (0,0) (1,0) (0,1)
// (1,1) with the even-x colour first, backward the reverse with the odd-x colour first.
void gs_sweep_pairw(Ctx& c, float* u, const float* f, bool backward, bool from_zero) {
  Level& L = c.L[0];
  const LevelView& v = L.v;
  if (!L.ub32) L.ub32 = dalloc<float>(c, 3 * (size_t)v.NP);
  const int P = backward ? 1 : 0;
  // the FCG's z.r and z.q from this sweep's final values (one sum per wave and pass)
  // (c.gsdots is sized by the solve before any capture: one zr and one zq per segment)
  int64_t nd = 0;
  for (int i = 0; i < 4; ++i) nd += L.n_segw[i][1];
  const bool dots = backward && c.dots_last && c.dots_q && c.gsdots && c.gsdots_n == nd;
  int64_t doff = 0;
  for (int i = 0; i < 4; ++i) {
    const int cls = backward ? 3 - i : i;
    const int n = L.n_segw[cls][P];
    if (!n) continue;
    Prof pr(c, P_FINE_GS32);
    const dim3 g((n + 3) / 4), b(256);
    const int4* sg = L.segw[cls][P];
    const uint8_t* fr = L.fre;
    if (dots) {
      hipLaunchKernelGGL((k_fine_gs_pairw<float, true, -1, -1, true>), g, b, 0, c.stream, v, sg, n, fr, c.w32,
                         c.Kc32, u, f, L.ub32, c.dots_q, c.gsdots + doff, nd);
      doff += n;
    } else if (backward) {
      hipLaunchKernelGGL((k_fine_gs_pairw<float, true, -1, -1>), g, b, 0, c.stream, v, sg, n, fr, c.w32, c.Kc32, u, f,
                         L.ub32, (const float*)nullptr, (double*)nullptr, (int64_t)0);
    } else if (!from_zero) {
      hipLaunchKernelGGL((k_fine_gs_pairw<float, false, -1, -1>), g, b, 0, c.stream, v, sg, n, fr, c.w32, c.Kc32, u,
                         f, L.ub32, (const float*)nullptr, (double*)nullptr, (int64_t)0);
    } else {
#define GO(A) hipLaunchKernelGGL((k_fine_gs_pairw<float, false, A, A + 1>), g, b, 0, c.stream, v, sg, n, fr, c.w32, \
                                 c.Kc32, u, f, L.ub32, (const float*)nullptr, (double*)nullptr, (int64_t)0)
      switch (cls) {
        case 0: GO(0); break;
        case 1: GO(2); break;
        case 2: GO(4); break;
        default: GO(6); break;
      }
#undef GO
    }
    hipLaunchKernelGGL((k_fine_pairw_copy<float>), dim3(nblk(3 * (int64_t)n)), dim3(256), 0, c.stream, v, sg, n, fr,
                       (const float*)L.ub32, u);
  }
  if (dots) c.dots_done = true;
  LAUNCH_CHECK();
}

// Stored-stencil level sweep in four in-wave two-colour passes (k_st_gs_pairw), each
// followed by the copy of its deferred nodes when rows span several segments; same class
// order as gs_sweep_pairw.
template <class T>
void gs_sweep_st_pairw(Ctx& c, int l, T* u, const T* f, bool backward, bool from_zero) {
  Level& L = c.L[l];
  const LevelView& v = L.v;
  const int P = backward ? 1 : 0;
  if (!L.pair_single && !L.ubuf) L.ubuf = dalloc<double>(c, 3 * (size_t)v.NP);
  T* ub = L.pair_single ? nullptr : (T*)L.ubuf;
  for (int i = 0; i < 4; ++i) {
    const int cls = backward ? 3 - i : i;
    const int n = L.n_segw[cls][P];
    if (!n) continue;
    Prof pr(c, sizeof(T) == 8 ? P_ST_GS64 : (l == 1 ? P_ST_GS32_L1 : P_ST_GS32));
    // small levels: one wave per workgroup, so the few waves of a pass spread over as many
    // CUs (a latency-bound pass then shares no CU's load pipeline among its waves)
    const dim3 g = spread_grid(n), b = spread_block(n);
    const int4* sg = L.segw[cls][P];
    st_dispatch<T>(c, L, [&](auto st) {
      using ST = decltype(st);
      if (backward) {
        hipLaunchKernelGGL((k_st_gs_pairw<T, true, -1, -1, ST>), g, b, 0, c.stream, v, sg, n, (const uint8_t*)L.fre,
                           (const int*)L.slot, st, u, f, ub);
      } else if (!from_zero) {
        hipLaunchKernelGGL((k_st_gs_pairw<T, false, -1, -1, ST>), g, b, 0, c.stream, v, sg, n, (const uint8_t*)L.fre,
                           (const int*)L.slot, st, u, f, ub);
      } else {
#define GO(A) hipLaunchKernelGGL((k_st_gs_pairw<T, false, A, A + 1, ST>), g, b, 0, c.stream, v, sg, n, \
                                 (const uint8_t*)L.fre, (const int*)L.slot, st, u, f, ub)
        switch (cls) {
          case 0: GO(0); break;
          case 1: GO(2); break;
          case 2: GO(4); break;
          default: GO(6); break;
        }
#undef GO
      }
    });
    if (ub)
      hipLaunchKernelGGL((k_fine_pairw_copy<T>), dim3(nblk(3 * (int64_t)n)), dim3(256), 0, c.stream, v, sg, n,
                         (const uint8_t*)L.fre, (const T*)ub, u);
  }
  LAUNCH_CHECK();
}

template <class T>
void gs_sweep(Ctx& c, int l, T* u, const T* f, bool backward, bool from_zero = false) {
  Level& L = c.L[l];
  const LevelView& v = L.v;
  if constexpr (std::is_same<T, float>::value) {
    if (l == 0 && c.sweep_mode == 3) {
      gs_sweep_pairw(c, u, f, backward, from_zero);
      return;
    }
  }
  // (FP32 cycles only: with FP64 values the paired kernel's registers halve its occupancy)
  if (std::is_same<T, float>::value && l > 0 && c.sweep_mode == 3 && L.pair_ok) {
    gs_sweep_st_pairw<T>(c, l, u, f, backward, from_zero);
    return;
  }
  if (l == 0 && c.sweep_mode == 2) {
    gs_sweep_pairs<T>(c, u, f, backward, from_zero);
    return;
  }
  if (l == 0 && L.wf_S > 0 && c.sweep_mode == 1) {
    if (from_zero)
      hipLaunchKernelGGL((k_zero_nodes<T>), dim3(cblk(L.n_lin)), dim3(256), 0, c.stream, (const int4*)L.cl_lin,
                         L.n_lin, L.v.NP, u);
    gs_sweep_wavefront<T>(c, u, f, backward);
    return;
  }
  for (int ci = 0; ci < 8; ++ci) {
    const int col = backward ? 7 - ci : ci;
    const int off = L.col_off[col], nch = L.col_off[col + 1] - off;
    if (nch == 0) continue;
    const int4* cl = L.cl_col + off;
    Prof pr(c, l == 0 ? (sizeof(T) == 8 ? P_FINE_GS64 : P_FINE_GS32)
                      : (sizeof(T) == 8 ? P_ST_GS64 : (l == 1 ? P_ST_GS32_L1 : P_ST_GS32)));
    const dim3 g(cblk(nch)), b(256);
    if (l == 0) {
      if (backward) {
        hipLaunchKernelGGL((k_fine_gs<T, true>), g, b, 0, c.stream, v, cl, nch, Prec<T>::w(c), Prec<T>::Ke(c), u, f);
      } else if (!from_zero) {
        hipLaunchKernelGGL((k_fine_gs<T, false>), g, b, 0, c.stream, v, cl, nch, Prec<T>::w(c), Prec<T>::Ke(c), u,
                           f);
      } else {
#define GO(ZC) hipLaunchKernelGGL((k_fine_gs<T, false, ZC>), g, b, 0, c.stream, v, cl, nch, Prec<T>::w(c), \
                                  Prec<T>::Ke(c), u, f)
        TOS_COLOUR_SWITCH(col, GO);
#undef GO
      }
    } else {
      const int64_t s0 = 64 * (int64_t)off;
      st_dispatch<T>(c, L, [&](auto st) {
        using ST = decltype(st);
        if (backward) {
          hipLaunchKernelGGL((k_st_gs<T, true, -1, ST>), g, b, 0, c.stream, v, cl, nch, s0, st, u, f);
        } else if (!from_zero) {
          hipLaunchKernelGGL((k_st_gs<T, false, -1, ST>), g, b, 0, c.stream, v, cl, nch, s0, st, u, f);
        } else {
#define GO(ZC) hipLaunchKernelGGL((k_st_gs<T, false, ZC, ST>), g, b, 0, c.stream, v, cl, nch, s0, st, u, f)
          TOS_COLOUR_SWITCH(col, GO);
#undef GO
        }
      });
    }
  }
  LAUNCH_CHECK();
}

// level l: MODE 0 y = K u (fine level: partial sums of u.y when `partial`), MODE 1
// y = f - K u (fine level: float copy y32 when given)
template <class T, int MODE>
void apply_level(Ctx& c, int l, const T* u, const T* f, T* y, double* partial = nullptr, float* y32 = nullptr) {
  Level& L = c.L[l];
  Prof pr(c, l > 0 ? P_ST_RESID
                   : (MODE == 0 ? P_FINE_APPLY64 : (sizeof(T) == 8 ? P_FINE_RESID64 : P_FINE_RESID32)));
  if (l == 0) {
    if (L.n_lin) {
      // y = K u in FP64: two halves of the element loop (102 VGPRs, 4 waves per SIMD; 852
      // against 954 us at the 511 workload); the residual form measured slower that way
      // (1249 against 1140 us at 256 VGPRs) and keeps k_fine_apply
      if constexpr (sizeof(T) == 8 && MODE == 0)
        hipLaunchKernelGGL(k_fine_apply_u64, dim3(cblk(L.n_lin)), dim3(256), 0, c.stream, L.v,
                           (const int4*)L.cl_lin, L.n_lin, Prec<T>::w(c), Prec<T>::Ke(c), u, y, partial);
      else
        hipLaunchKernelGGL((k_fine_apply<T, MODE>), dim3(cblk(L.n_lin)), dim3(256), 0, c.stream, L.v,
                           (const int4*)L.cl_lin, L.n_lin, Prec<T>::w(c), Prec<T>::Ke(c), u, f, y, partial, y32);
    }
  } else if (L.n_col) {
    st_dispatch<T>(c, L, [&](auto st) {
      hipLaunchKernelGGL((k_st_apply<T, MODE, decltype(st)>), dim3(cblk(L.n_col)), dim3(256), 0, c.stream, L.v,
                         (const int4*)L.cl_col, L.n_col, st, u, f, y);
    });
  }
  LAUNCH_CHECK();
}

// q = K p on the fine level for an FP32 search direction p (FP64 arithmetic), with the
// partial sums of p.q
void apply_fine_dir(Ctx& c, const float* p, double* q, double* partial) {
  Level& L = c.L[0];
  Prof pr(c, P_FINE_APPLY64);
  if (L.n_lin)
    hipLaunchKernelGGL((k_fine_apply_dir<double>), dim3(cblk(L.n_lin)), dim3(256), 0, c.stream, L.v,
                       (const int4*)L.cl_lin, L.n_lin, Prec<double>::w(c), Prec<double>::Ke(c), p, partial, q);
  LAUNCH_CHECK();
}

// r = f - A u right after one forward sweep from u = 0 (colour by colour, later-colour
// neighbours only; see k_fine_resid_z)
template <class T>
void resid_after_zero_sweep(Ctx& c, int l, const T* u, T* r) {
  Level& L = c.L[l];
  const LevelView& v = L.v;
  Prof pr(c, l > 0 ? P_ST_RESID : (sizeof(T) == 8 ? P_FINE_RESID64 : P_FINE_RESID32));
  if constexpr (std::is_same<T, float>::value) {
    if (l == 0 && c.sweep_mode == 3) {  // two colours per pass on the forward segments
      for (int cls = 0; cls < 4; ++cls) {
        const int n = L.n_segw[cls][0];
        if (!n) continue;
        const dim3 g((n + 3) / 4), b(256);
#define GO(A) hipLaunchKernelGGL((k_fine_resid_pairw<float, A, A + 1>), g, b, 0, c.stream, v, (const int4*)L.segw[cls][0], \
                                 n, (const uint8_t*)L.fre, (const float*)c.w32, (const float*)c.Kc32, u, r)
        switch (cls) {
          case 0: GO(0); break;
          case 1: GO(2); break;
          case 2: GO(4); break;
          default: GO(6); break;
        }
#undef GO
      }
      LAUNCH_CHECK();
      return;
    }
    if (l > 0 && c.sweep_mode == 3 && L.pair_ok) {  // stored stencils: two colours per pass likewise
      st_dispatch<T>(c, L, [&](auto st) {
        using ST = decltype(st);
        for (int cls = 0; cls < 4; ++cls) {
          const int n = L.n_segw[cls][0];
          if (!n) continue;
          const dim3 g((n + 3) / 4), b(256);
#define GO(A) hipLaunchKernelGGL((k_st_resid_pairw<float, A, A + 1, ST>), g, b, 0, c.stream, v, \
                                 (const int4*)L.segw[cls][0], n, (const uint8_t*)L.fre, (const int*)L.slot, st, u, r)
          switch (cls) {
            case 0: GO(0); break;
            case 1: GO(2); break;
            case 2: GO(4); break;
            default: GO(6); break;
          }
#undef GO
        }
      });
      LAUNCH_CHECK();
      return;
    }
  }
  for (int col = 0; col < 8; ++col) {
    const int off = L.col_off[col], nch = L.col_off[col + 1] - off;
    if (nch == 0) continue;
    const int4* cl = L.cl_col + off;
    const dim3 g(cblk(nch)), b(256);
    if (l == 0) {
#define GO(ZC) hipLaunchKernelGGL((k_fine_resid_z<T, ZC>), g, b, 0, c.stream, v, cl, nch, Prec<T>::w(c), \
                                  Prec<T>::Ke(c), u, r)
      TOS_COLOUR_SWITCH(col, GO);
#undef GO
    } else {
      const int64_t s0 = 64 * (int64_t)off;
      st_dispatch<T>(c, L, [&](auto st) {
#define GO(ZC) hipLaunchKernelGGL((k_st_resid_z<T, ZC, decltype(st)>), g, b, 0, c.stream, v, cl, nch, s0, st, u, r)
        TOS_COLOUR_SWITCH(col, GO);
#undef GO
      });
    }
  }
  LAUNCH_CHECK();
}

constexpr int64_t kResidZMinNodes = 1 << 20;

void learned_step(Ctx& c, int l, const float* f, float* u, int pre, int post);
bool learned_here(const Ctx& c, int l);

template <class T>
void vcycle_level(Ctx& c, int l, const T* f, T* u, int pre, int post) {
  Level& L = c.L[l];
  const int last = (int)c.L.size() - 1;
  if (l == last && last > 0) {
    if (c.cn == 0) return;  // no unknowns on the coarsest level (u stays 0)
    Prof pr(c, P_COARSE);
    hipLaunchKernelGGL((k_coarse_solve<T>), dim3(nblk((int64_t)c.cn * 64)), dim3(256), 0, c.stream, L.v,
                       (const int64_t*)c.cidx, c.cn, coarse_inverse<T>(c), f, u);
    LAUNCH_CHECK();
    return;
  }
  if (L.n_lin == 0) return;
  // u = 0 on entry: the first forward sweep starts from zero without a zeroing pass
  if (l == last) {  // a single level: smoothing only
    for (int s = 0; s < std::max(1, pre); ++s) gs_sweep<T>(c, l, u, f, false, s == 0);
    for (int s = 0; s < std::max(1, post); ++s) gs_sweep<T>(c, l, u, f, true);
    return;
  }
  if (pre == 0)
    hipLaunchKernelGGL((k_zero_nodes<T>), dim3(cblk(L.n_lin)), dim3(256), 0, c.stream, (const int4*)L.cl_lin,
                       L.n_lin, L.v.NP, u);
  for (int s = 0; s < pre; ++s) gs_sweep<T>(c, l, u, f, false, s == 0);
  T* r = Prec<T>::r(L);
  // after a single sweep from zero the residual needs only the later-colour half of each
  // neighbourhood, but costs 8 launches: worth it on the large levels only
  if (pre == 1 && L.n_free >= kResidZMinNodes)
    resid_after_zero_sweep<T>(c, l, u, r);
  else
    apply_level<T, 1>(c, l, u, f, r);
  Level& C = c.L[l + 1];
  const bool four = l == 0 && c.skip;  // skip-layer: factor-4 transfers
  if (C.n_lin) {
    Prof pr(c, P_TRANSFER);
    if (four)
      hipLaunchKernelGGL((k_restrict4<T>), dim3(cblk(C.n_lin)), dim3(256), 0, c.stream, L.v, C.v,
                         (const int4*)C.cl_lin, C.n_lin, (const T*)r, Prec<T>::f(C));
    else
      hipLaunchKernelGGL((k_restrict<T>), dim3(cblk(C.n_lin)), dim3(256), 0, c.stream, L.v, C.v,
                         (const int4*)C.cl_lin, C.n_lin, (const T*)r, Prec<T>::f(C));
  }
  vcycle_level<T>(c, l + 1, Prec<T>::f(C), Prec<T>::u(C), pre, post);
  {
    Prof pr(c, P_TRANSFER);
    if (four)
      hipLaunchKernelGGL((k_prolong4<T>), dim3(cblk(L.n_lin)), dim3(256), 0, c.stream, L.v, C.v,
                         (const int4*)L.cl_lin, L.n_lin, (const T*)Prec<T>::u(C), u);
    else if (L.n_segp)
      hipLaunchKernelGGL((k_prolong_add_pairw<T>), dim3((L.n_segp + 3) / 4), dim3(256), 0, c.stream, L.v, C.v,
                         (const int4*)L.segp, L.n_segp, (const uint8_t*)L.fre, (const T*)Prec<T>::u(C), u);
    else
      hipLaunchKernelGGL((k_prolong_add<T>), dim3(cblk(L.n_lin)), dim3(256), 0, c.stream, L.v, C.v,
                         (const int4*)L.cl_lin, L.n_lin, (const T*)Prec<T>::u(C), u);
  }
  LAUNCH_CHECK();
  if constexpr (std::is_same<T, float>::value)
    if (learned_here(c, l)) learned_step(c, l, f, u, pre, post);
  for (int s = 0; s < post; ++s) {
    c.dots_last = l == 0 && s == post - 1;
    gs_sweep<T>(c, l, u, f, true);
  }
  c.dots_last = false;
}

// z = B r.  precision 0: FP64 cycle from c.r into c.z.  precision 1: FP32 cycle from
// L0.f32 (a float copy of r, written here unless `have_r32`) into L0.u32; with
// `want_z64` the result is also cast into c.z.
void precondition(Ctx& c, int precision, int pre, int post, bool have_r32, bool want_z64) {
  Level& L0 = c.L[0];
  if (precision == 0) {
    vcycle_level<double>(c, 0, c.r, c.z, pre, post);
    return;
  }
  if (!have_r32)
    hipLaunchKernelGGL((k_cast_nodes<double, float>), dim3(cblk(L0.n_lin)), dim3(256), 0, c.stream,
                       (const int4*)L0.cl_lin, L0.n_lin, L0.v.NP, (const double*)c.r, L0.f32);
  vcycle_level<float>(c, 0, L0.f32, L0.u32, pre, post);
  if (want_z64)
    hipLaunchKernelGGL((k_cast_nodes<float, double>), dim3(cblk(L0.n_lin)), dim3(256), 0, c.stream,
                       (const int4*)L0.cl_lin, L0.n_lin, L0.v.NP, (const float*)L0.u32, c.z);
  LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Learned coarse-level correction (no counterpart in the reference).  At level `level` of
// an FP32 V-cycle, after the coarse-grid correction has been prolongated and before the
// post-smoothing, the residual r = f - A u is formed and either
//   logged: the pair (features of r, target e = A^{-1} r from `n_inner` nested V-cycles
//           on this level) is appended to a device dataset, or
//   used:   the network proposes e; it is kept (u += e) only when the scaled residual
//           |D^{-1/2}(r - A e)| drops below theta |D^{-1/2} r|, so plain multigrid takes
//           over whenever the network does not help.  The decision stays on the device.
// ---------------------------------------------------------------------------------------
}  // namespace

namespace tos {
struct LearnState {
  tos_unet* net = nullptr;
  int level = -1;
  int mode = 0;
  tos_dataset* ds = nullptr;
  int n_inner = 8, stride = 1;
  int64_t calls = 0;
  bool busy = false;
  float* lrho = nullptr;
  float *e = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
  bool paused = false;  // set by the solver after repeated rejections, until the next solve
  int dry = 0;           // consecutive solves whose corrections were all rejected
  int skip = 0;          // solves still to run without the correction (back-off)
  NetActs* acts = nullptr;
  std::vector<void*> mem;
};
}  // namespace tos

namespace {

constexpr double kNNStepMin = 0.2, kNNStepMax = 5.0;  // accepted energy-optimal steps

bool learned_here(const Ctx& c, int l) {
  const LearnState* S = c.learn;
  return S && l == S->level && !S->busy && ((S->mode == 1 && S->net && !S->paused) || S->ds);
}

void learn_logrho(Ctx& c) {
  LearnState& S = *c.learn;
  const Level& L = c.L[S.level];
  const int64_t V = (int64_t)L.v.nx * L.v.ny * L.v.nz;
  hipLaunchKernelGGL(k_node_logrho, dim3(nblk(V)), dim3(256), 0, c.stream, c.L[0].v, (const uint8_t*)c.eact,
                     (const float*)c.rho, c.nx, c.ny, c.nz, 1 << S.level, (float)c.rho_min, L.v.nx, L.v.ny, L.v.nz,
                     S.lrho);
  LAUNCH_CHECK();
}

// set up the per-level buffers once the level is known
void learn_prepare(Ctx& c, int level) {
  if (!c.learn) c.learn = new LearnState();
  LearnState& S = *c.learn;
  if (S.level == level) return;
  for (void* p : S.mem) (void)hipFree(p);
  S.mem.clear();
  ++c.alloc_epoch;
  S.acts = nullptr;
  S.level = level;
  const Level& L = c.L[level];
  const int64_t V = (int64_t)L.v.nx * L.v.ny * L.v.nz;
  S.lrho = (float*)raw_alloc(S.mem, c.stream, V * sizeof(float));
  S.e = (float*)raw_alloc(S.mem, c.stream, 3 * L.v.NP * sizeof(float));
  S.t1 = (float*)raw_alloc(S.mem, c.stream, 3 * L.v.NP * sizeof(float));
  S.t2 = (float*)raw_alloc(S.mem, c.stream, 3 * L.v.NP * sizeof(float));
  S.t3 = (float*)raw_alloc(S.mem, c.stream, 3 * L.v.NP * sizeof(float));
  learn_logrho(c);
  HIPCHK(hipStreamSynchronize(c.stream));
}

void learn_release(Ctx& c) {
  if (!c.learn) return;
  for (void* p : c.learn->mem) (void)hipFree(p);
  ++c.alloc_epoch;
  delete c.learn;
  c.learn = nullptr;
}

// scaled squared norm sum_c r_c^2 / D_c over the free nodes of level l -> scal[slot]
void scaled_norm(Ctx& c, int l, const float* r, int slot) {
  Level& L = c.L[l];
  const unsigned nb = cblk(L.n_lin);
  hipLaunchKernelGGL(k_nn_norm, dim3(nb), dim3(256), 0, c.stream, L.v, (const int4*)L.cl_lin, L.n_lin, r,
                     (const float*)L.st32, L.NS, (const int*)L.slot, c.partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c.stream, (const double*)c.partial, (int)nb,
                     c.scal + slot);
  LAUNCH_CHECK();
}

void learned_step(Ctx& c, int l, const float* f, float* u, int pre, int post) {
  LearnState& S = *c.learn;
  Level& L = c.L[l];
  const int64_t V = (int64_t)L.v.nx * L.v.ny * L.v.nz;
  const double nfree = (double)std::max<int64_t>(L.n_free, 1);
  float* r = L.r32;
  apply_level<float, 1>(c, l, u, f, r);
  scaled_norm(c, l, r, S_NN0);
  tos_dataset* ds = S.ds;
  if (ds && ds->n < ds->cap && (S.calls++ % S.stride) == 0) {
    // target e = A^{-1} r by nested V-cycles on this level (u, f of the running cycle untouched)
    const int64_t n3 = 3 * L.v.NP;
    HIPCHK(hipMemcpyAsync(S.t1, r, n3 * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_zero<float>, dim3(nblk(n3)), dim3(256), 0, c.stream, S.e, n3);
    S.busy = true;
    for (int it = 0; it < S.n_inner; ++it) {
      apply_level<float, 1>(c, l, S.e, S.t1, S.t2);
      vcycle_level<float>(c, l, S.t2, S.t3, pre, post);
      hipLaunchKernelGGL(k_add_nodes_f, dim3(cblk(L.n_lin)), dim3(256), 0, c.stream, (const int4*)L.cl_lin, L.n_lin,
                         L.v.NP, S.e, (const float*)S.t3);
    }
    S.busy = false;
    const int i = ds->n;
    hipLaunchKernelGGL(k_nn_features, dim3(nblk(V)), dim3(256), 0, c.stream, L.v, (const float*)S.t1,
                       (const float*)L.st32, L.NS, (const int*)L.slot, (const float*)S.lrho,
                       (const double*)(c.scal + S_NN0), nfree, ds->feat + (int64_t)i * V * 8);
    hipLaunchKernelGGL(k_nn_target, dim3(nblk(V)), dim3(256), 0, c.stream, L.v, (const float*)S.e,
                       (const float*)L.st32, L.NS, (const int*)L.slot, (const double*)(c.scal + S_NN0), nfree,
                       ds->tgt + (int64_t)i * V * 3, ds->mask + (int64_t)i * V);
    hipLaunchKernelGGL(k_target_norm, dim3(256), dim3(256), 0, c.stream, (const float*)(ds->tgt + (int64_t)i * V * 3),
                       (const uint8_t*)(ds->mask + (int64_t)i * V), V, c.partial);
    hipLaunchKernelGGL(k_inv_den, dim3(1), dim3(256), 0, c.stream, (const double*)c.partial, 256, ds->inv_den + i);
    LAUNCH_CHECK();
    ds->n += 1;
    return;
  }
  if (S.mode != 1 || !S.net) return;
  {
    Prof pr(c, P_LEARNED);
    NetActs* A = S.acts;
    hipLaunchKernelGGL(k_nn_features, dim3(nblk(V)), dim3(256), 0, c.stream, L.v, (const float*)r,
                       (const float*)L.st32, L.NS, (const int*)L.slot, (const float*)S.lrho,
                       (const double*)(c.scal + S_NN0), nfree, A->feat);
    unet_forward(S.net, A, c.stream);
    hipLaunchKernelGGL(k_nn_correction, dim3(nblk(V)), dim3(256), 0, c.stream, L.v, (const float*)A->y,
                       (const float*)L.st32, L.NS, (const int*)L.slot, (const double*)(c.scal + S_NN0), nfree, S.e);
    LAUNCH_CHECK();
  }
  apply_level<float, 1>(c, l, S.e, r, S.t2);  // t2 = r - A e
  const unsigned nb = cblk(L.n_lin);
  hipLaunchKernelGGL(k_nn_dots, dim3(nb), dim3(256), 0, c.stream, L.v, (const int4*)L.cl_lin, L.n_lin,
                     (const float*)S.e, (const float*)r, (const float*)S.t2, c.partial);
  hipLaunchKernelGGL(k_sum2_partials, dim3(1), dim3(256), 0, c.stream, (const double*)c.partial, (int)nb,
                     c.scal + S_NN1);
  hipLaunchKernelGGL(k_nn_accept, dim3(nb), dim3(256), 0, c.stream, L.v, (const int4*)L.cl_lin, L.n_lin,
                     (const double*)(c.scal + S_NN1), kNNStepMin, kNNStepMax, (const float*)S.e, u,
                     c.scal + S_NN_TRIED);
  LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// K load cases in one solve (no counterpart in the reference, which solves its load
// cases one after another).  Case-interleaved vectors (one stencil / density read serves
// all K cases), one FCG per case with per-case scalars and convergence flags on the
// device, FP32 V-cycle with FP64 outer vectors.  The body of an FCG iteration (operator,
// updates, V-cycle, dots) is captured once into a hipGraph and replayed; the host reads
// K residual norms per iteration.  Block sizes (nodes x cases per block) of every kernel
// class and level are picked by timing the candidates on the context's own data.
// ---------------------------------------------------------------------------------------
}  // namespace

namespace tos {
struct BatchState {
  int K = 0, KP = 0;
  std::vector<float*> u32, f32, r32;  // per level, 3 * NP * K
  double *U = nullptr, *R = nullptr, *Q = nullptr, *F = nullptr;
  float* P = nullptr;       // FCG search direction (FP32 storage, FP64 arithmetic)
  double* part = nullptr;   // per-case block partials [2][nb][K]
  double* fold = nullptr;   // [2][kFoldG][K]
  double* sc = nullptr;     // [B_COUNT][K]
  double* h_sc = nullptr;
  // tuned block sizes
  int bs_fine_gs = 256, bs_fine_res = 256, bs_transfer = 256;
  std::vector<int> bs_st_gs, bs_st_res;
  bool tuned = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_gal = -1;  // Galerkin precision and smoothing counts the graph was captured with
  int graph_pre = -1, graph_post = -1;
  std::vector<void*> mem;
};
}  // namespace tos

namespace {

constexpr int kFoldG = 512;  // first-stage groups of the per-case folds
constexpr int kVecBS = 256;  // block size of the kernels that write per-case partials

inline unsigned bgrid(int nch, int KP, int bs) { return (unsigned)(((int64_t)nch * 64 * KP + bs - 1) / bs); }

void batch_release(Ctx& c) {
  if (!c.batch) return;
  BatchState& B = *c.batch;
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  for (void* p : B.mem) (void)hipFree(p);
  if (B.h_sc) (void)hipHostFree(B.h_sc);
  delete c.batch;
  c.batch = nullptr;
}

BatchState& batch_prepare(Ctx& c) {
  if (c.batch && c.batch->K == c.K) return *c.batch;
  batch_release(c);
  c.batch = new BatchState();
  BatchState& B = *c.batch;
  B.K = c.K;
  B.KP = 1;
  while (B.KP < B.K) B.KP <<= 1;
  const int nl = (int)c.L.size();
  B.u32.assign(nl, nullptr);
  B.f32.assign(nl, nullptr);
  B.r32.assign(nl, nullptr);
  for (int l = 0; l < nl; ++l) {
    const size_t n = 3 * (size_t)c.L[l].v.NP * B.K;
    B.u32[l] = (float*)raw_alloc(B.mem, c.stream, n * sizeof(float));
    B.f32[l] = (float*)raw_alloc(B.mem, c.stream, n * sizeof(float));
    if (l + 1 < nl) B.r32[l] = (float*)raw_alloc(B.mem, c.stream, n * sizeof(float));
  }
  const size_t n0 = 3 * (size_t)c.L[0].v.NP * B.K;
  B.U = (double*)raw_alloc(B.mem, c.stream, n0 * sizeof(double));
  B.R = (double*)raw_alloc(B.mem, c.stream, n0 * sizeof(double));
  B.P = (float*)raw_alloc(B.mem, c.stream, n0 * sizeof(float));
  B.Q = (double*)raw_alloc(B.mem, c.stream, n0 * sizeof(double));
  B.F = (double*)raw_alloc(B.mem, c.stream, n0 * sizeof(double));
  const size_t nb = bgrid(c.L[0].n_lin, B.KP, kVecBS);
  B.part = (double*)raw_alloc(B.mem, c.stream, 2 * std::max<size_t>(nb, 1) * B.K * sizeof(double));
  B.fold = (double*)raw_alloc(B.mem, c.stream, 2 * (size_t)kFoldG * B.K * sizeof(double));
  B.sc = (double*)raw_alloc(B.mem, c.stream, (size_t)B_COUNT * B.K * sizeof(double));
  HIPCHK(hipHostMalloc((void**)&B.h_sc, (size_t)B_COUNT * B.K * sizeof(double)));
  B.bs_st_gs.assign(nl, 256);
  B.bs_st_res.assign(nl, 256);
  HIPCHK(hipStreamSynchronize(c.stream));
  return B;
}

#define TOS_BS_SWITCH(bs, GO) \
  switch (bs) {               \
    case 64: GO(64); break;   \
    case 128: GO(128); break; \
    default: GO(256); break;  \
  }

// fine level, colour `col` of a first forward sweep from zero / of the residual after it
template <int BS>
void fine_zero_colour_b(Ctx& c, BatchState& B, const LevelView& v, const int4* cl, int nch, int col, float* u,
                        const float* f, float* r) {
  const unsigned nb = bgrid(nch, B.KP, BS);
  if (!r) {
#define GO(ZC) hipLaunchKernelGGL((k_fine_gs_b<float, false, BS, ZC>), dim3(nb), dim3(BS), 0, c.stream, v, cl, nch, \
                                  B.K, B.KP, (const float*)c.w32, (const float*)c.Kc32, u, f)
    TOS_COLOUR_SWITCH(col, GO);
#undef GO
  } else {
#define GO(ZC) hipLaunchKernelGGL((k_fine_resid_z_b<float, ZC, BS>), dim3(nb), dim3(BS), 0, c.stream, v, cl, nch, \
                                  B.K, B.KP, (const float*)c.w32, (const float*)c.Kc32, (const float*)u, r)
    TOS_COLOUR_SWITCH(col, GO);
#undef GO
  }
}

// fine level: the first forward sweep from zero (r = null), or the residual after it into
// r (later-colour neighbours only), colour by colour
void fine_zero_sweep_b(Ctx& c, BatchState& B, float* u, const float* f, float* r) {
  Level& L = c.L[0];
  const int bs = r ? B.bs_fine_res : B.bs_fine_gs;
  for (int col = 0; col < 8; ++col) {
    const int off = L.col_off[col], nch = L.col_off[col + 1] - off;
    if (nch == 0) continue;
#define GO(BS) fine_zero_colour_b<BS>(c, B, L.v, L.cl_col + off, nch, col, u, f, r)
    TOS_BS_SWITCH(bs, GO)
#undef GO
  }
  LAUNCH_CHECK();
}

// one colour-ordered GS sweep of level l on the interleaved vectors
void gs_b(Ctx& c, BatchState& B, int l, float* u, const float* f, bool backward) {
  Level& L = c.L[l];
  const LevelView& v = L.v;
  const int bs = l == 0 ? B.bs_fine_gs : B.bs_st_gs[l];
  for (int ci = 0; ci < 8; ++ci) {
    const int col = backward ? 7 - ci : ci;
    const int off = L.col_off[col], nch = L.col_off[col + 1] - off;
    if (nch == 0) continue;
    const int4* cl = L.cl_col + off;
    const unsigned nb = bgrid(nch, B.KP, bs);
    if (l == 0) {
#define GO(BS)                                                                                                     \
  if (backward)                                                                                                    \
    hipLaunchKernelGGL((k_fine_gs_b<float, true, BS>), dim3(nb), dim3(BS), 0, c.stream, v, cl, nch, B.K, B.KP,     \
                       (const float*)c.w32, (const float*)c.Kc32, u, f);                                           \
  else                                                                                                             \
    hipLaunchKernelGGL((k_fine_gs_b<float, false, BS>), dim3(nb), dim3(BS), 0, c.stream, v, cl, nch, B.K, B.KP,    \
                       (const float*)c.w32, (const float*)c.Kc32, u, f);
      TOS_BS_SWITCH(bs, GO)
#undef GO
    } else {
#define GO(BS)                                                                                                     \
  if (backward)                                                                                                    \
    hipLaunchKernelGGL((k_st_gs_b<float, true, BS>), dim3(nb), dim3(BS), 0, c.stream, v, cl, nch,                  \
                       64 * (int64_t)off, B.K, B.KP, (const float*)L.st32, L.NS, u, f);                            \
  else                                                                                                             \
    hipLaunchKernelGGL((k_st_gs_b<float, false, BS>), dim3(nb), dim3(BS), 0, c.stream, v, cl, nch,                 \
                       64 * (int64_t)off, B.K, B.KP, (const float*)L.st32, L.NS, u, f);
      TOS_BS_SWITCH(bs, GO)
#undef GO
    }
  }
  LAUNCH_CHECK();
}

// r = f - A u on level l (FP32, interleaved)
void resid_b(Ctx& c, BatchState& B, int l, const float* u, const float* f, float* r) {
  Level& L = c.L[l];
  if (l == 0) {
    if (!L.n_lin) return;
    const int bs = B.bs_fine_res;
    const unsigned nb = bgrid(L.n_lin, B.KP, bs);
#define GO(BS)                                                                                                      \
  hipLaunchKernelGGL((k_fine_apply_b<float, 1, BS>), dim3(nb), dim3(BS), 0, c.stream, L.v, (const int4*)L.cl_lin, \
                     L.n_lin, B.K, B.KP, (const float*)c.w32, (const float*)c.Kc32, u, f, r, (double*)nullptr,      \
                     (float*)nullptr);
    TOS_BS_SWITCH(bs, GO)
#undef GO
  } else {
    if (!L.n_col) return;
    const int bs = B.bs_st_res[l];
    const unsigned nb = bgrid(L.n_col, B.KP, bs);
#define GO(BS)                                                                                                    \
  hipLaunchKernelGGL((k_st_apply_b<float, 1, BS>), dim3(nb), dim3(BS), 0, c.stream, L.v, (const int4*)L.cl_col, \
                     L.n_col, B.K, B.KP, (const float*)L.st32, L.NS, u, f, r);
    TOS_BS_SWITCH(bs, GO)
#undef GO
  }
  LAUNCH_CHECK();
}

void restrict_b(Ctx& c, BatchState& B, int l, const float* rf, float* fc) {
  Level& F = c.L[l];
  Level& C = c.L[l + 1];
  if (!C.n_lin) return;
  const int bs = B.bs_transfer;
  const unsigned nb = bgrid(C.n_lin, B.KP, bs);
#define GO(BS)                                                                                                     \
  hipLaunchKernelGGL((k_restrict_b<float, BS>), dim3(nb), dim3(BS), 0, c.stream, F.v, C.v, (const int4*)C.cl_lin, \
                     C.n_lin, B.K, B.KP, rf, fc);
  TOS_BS_SWITCH(bs, GO)
#undef GO
}

void prolong_b(Ctx& c, BatchState& B, int l, const float* uc, float* uf) {
  Level& F = c.L[l];
  Level& C = c.L[l + 1];
  if (!F.n_lin) return;
  const int bs = B.bs_transfer;
  const unsigned nb = bgrid(F.n_lin, B.KP, bs);
#define GO(BS)                                                                                                         \
  hipLaunchKernelGGL((k_prolong_add_b<float, BS>), dim3(nb), dim3(BS), 0, c.stream, F.v, C.v, (const int4*)F.cl_lin, \
                     F.n_lin, B.K, B.KP, uc, uf);
  TOS_BS_SWITCH(bs, GO)
#undef GO
}

void vcycle_b(Ctx& c, BatchState& B, int l, const float* f, float* u, int pre, int post) {
  Level& L = c.L[l];
  const int last = (int)c.L.size() - 1;
  if (l == last && last > 0) {
    if (c.cn == 0) return;
    hipLaunchKernelGGL((k_coarse_solve_b<float>), dim3(c.cn), dim3(256), 0, c.stream, L.v,
                       (const int64_t*)c.cidx, c.cn, B.K, coarse_inverse<float>(c), f, u);
    LAUNCH_CHECK();
    return;
  }
  if (L.n_lin == 0) return;
  if (l == 0 && l < last && pre >= 1) {  // fine level: first sweep from zero without clearing u
    fine_zero_sweep_b(c, B, u, f, nullptr);
    for (int s = 1; s < pre; ++s) gs_b(c, B, l, u, f, false);
    if (pre == 1)
      fine_zero_sweep_b(c, B, u, f, B.r32[0]);
    else
      resid_b(c, B, l, u, f, B.r32[l]);
    restrict_b(c, B, l, B.r32[l], B.f32[l + 1]);
    vcycle_b(c, B, l + 1, B.f32[l + 1], B.u32[l + 1], pre, post);
    prolong_b(c, B, l, B.u32[l + 1], u);
    LAUNCH_CHECK();
    for (int s = 0; s < post; ++s) gs_b(c, B, l, u, f, true);
    return;
  }
  hipLaunchKernelGGL((k_zero_nodes_b<float, kVecBS>), dim3(bgrid(L.n_lin, B.KP, kVecBS)), dim3(kVecBS), 0, c.stream,
                     (const int4*)L.cl_lin, L.n_lin, L.v.NP, B.K, B.KP, u);
  if (l == last) {
    for (int s = 0; s < std::max(1, pre); ++s) gs_b(c, B, l, u, f, false);
    for (int s = 0; s < std::max(1, post); ++s) gs_b(c, B, l, u, f, true);
    return;
  }
  for (int s = 0; s < pre; ++s) gs_b(c, B, l, u, f, false);
  resid_b(c, B, l, u, f, B.r32[l]);
  restrict_b(c, B, l, B.r32[l], B.f32[l + 1]);
  vcycle_b(c, B, l + 1, B.f32[l + 1], B.u32[l + 1], pre, post);
  prolong_b(c, B, l, B.u32[l + 1], u);
  LAUNCH_CHECK();
  for (int s = 0; s < post; ++s) gs_b(c, B, l, u, f, true);
}

// per-case fold of `ncol` partial columns of a launch with nb blocks into sc
void foldb(Ctx& c, BatchState& B, unsigned nb, int ncol, int s0, int s1, int op) {
  hipLaunchKernelGGL(k_foldb1, dim3(kFoldG), dim3(256), 0, c.stream, (const double*)B.part, (int)nb, B.K, ncol,
                     B.fold);
  hipLaunchKernelGGL(k_foldb2, dim3(1), dim3(256), 0, c.stream, (const double*)B.fold, kFoldG, B.K, ncol, s0, s1, op,
                     B.sc);
}

// one FCG iteration (captured): q = A p, alpha, u/r update, |r|^2, V-cycle, beta, p
void batch_iteration(Ctx& c, BatchState& B, int pre, int post) {
  Level& L0 = c.L[0];
  const int4* cl = L0.cl_lin;
  const int nch = L0.n_lin;
  const int64_t NP = L0.v.NP;
  const unsigned nb = bgrid(nch, B.KP, kVecBS);
  hipLaunchKernelGGL((k_fine_apply_b<double, 0, kVecBS, float>), dim3(nb), dim3(kVecBS), 0, c.stream, L0.v, cl, nch,
                     B.K, B.KP, (const double*)c.w64, (const double*)c.Kc64, (const float*)B.P, (const double*)nullptr,
                     B.Q, B.part, (float*)nullptr);
  foldb(c, B, nb, 1, B_PQ, 0, OPB_ALPHA);
  hipLaunchKernelGGL((k_pcgb_ur<kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, cl, nch, NP, B.K, B.KP, B.U, B.R,
                     (const float*)B.P, (const double*)B.Q, (const double*)B.sc, B.f32[0], B.part);
  foldb(c, B, nb, 1, B_RN2, 0, OPB_RN2);
  vcycle_b(c, B, 0, B.f32[0], B.u32[0], pre, post);
  hipLaunchKernelGGL((k_pcgb_zdots<kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, cl, nch, NP, B.K, B.KP,
                     (const float*)B.u32[0], (const double*)B.R, (const double*)B.Q, B.part);
  foldb(c, B, nb, 2, B_ZR, B_ZQ, OPB_BETA);
  hipLaunchKernelGGL((k_pcgb_p<kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, cl, nch, NP, B.K, B.KP, B.P,
                     (const float*)B.u32[0], (const double*)B.sc, 0);
  LAUNCH_CHECK();
}

// time `f` (reps launches after one warm-up) on the context stream, ms per launch
template <class F>
float time_ms(Ctx& c, F&& f, int reps = 3) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  f();
  HIPCHK(hipEventRecord(a, c.stream));
  for (int i = 0; i < reps; ++i) f();
  HIPCHK(hipEventRecord(b, c.stream));
  HIPCHK(hipEventSynchronize(b));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

// pick the block size of every tuned kernel class by timing 64 / 128 / 256 on this
// context's levels (the V-cycle vectors are scratch at this point)
void batch_tune(Ctx& c, BatchState& B) {
  const int cand[3] = {64, 128, 256};
  auto pick = [&](auto&& run) {
    int best = 256;
    float tb = 1e30f;
    for (int bs : cand) {
      const float t = time_ms(c, [&]() { run(bs); });
      if (t < tb) { tb = t; best = bs; }
    }
    return best;
  };
  const int nl = (int)c.L.size();
  B.bs_fine_gs = pick([&](int bs) { B.bs_fine_gs = bs; gs_b(c, B, 0, B.u32[0], B.f32[0], false); });
  if (nl > 1) B.bs_fine_res = pick([&](int bs) { B.bs_fine_res = bs; resid_b(c, B, 0, B.u32[0], B.f32[0], B.r32[0]); });
  for (int l = 1; l + 1 < nl; ++l) {
    B.bs_st_gs[l] = pick([&](int bs) { B.bs_st_gs[l] = bs; gs_b(c, B, l, B.u32[l], B.f32[l], false); });
    B.bs_st_res[l] = pick([&](int bs) { B.bs_st_res[l] = bs; resid_b(c, B, l, B.u32[l], B.f32[l], B.r32[l]); });
  }
  if (nl > 1)
    B.bs_transfer = pick([&](int bs) { B.bs_transfer = bs; prolong_b(c, B, 0, B.u32[1], B.r32[0]); });
  B.tuned = true;
}

int solve_batched(Ctx& c, const tos_solve_opts& o, tos_solve_stats* st) {
  BatchState& B = batch_prepare(c);
  Level& L0 = c.L[0];
  const int4* cl = L0.cl_lin;
  const int nch = L0.n_lin, K = B.K;
  const int64_t NP = L0.v.NP, n3k = 3 * NP * K;
  if (nch == 0) {  // no free node: u = 0 for every case
    for (int k = 0; k < K && st && k < 64; ++k) { st->iterations[k] = 0; st->rel_res[k] = 0; }
    return TOS_OK;
  }
  if (!B.tuned) batch_tune(c, B);
  const bool prof = c.prof_on;
  const unsigned nb = bgrid(nch, B.KP, kVecBS);
  const unsigned ni = (unsigned)(((int64_t)nch * 64 * K + 255) / 256);
  // F, U -> interleaved; r = f - A u; |f|^2, |r|^2
  hipLaunchKernelGGL(k_interleave, dim3(ni), dim3(256), 0, c.stream, cl, nch, NP, K, (const double*)c.F, B.F, 1);
  if (o.warm_start)
    hipLaunchKernelGGL(k_interleave, dim3(ni), dim3(256), 0, c.stream, cl, nch, NP, K, (const double*)c.U, B.U, 1);
  else
    hipLaunchKernelGGL(k_zero<double>, dim3(nblk(n3k)), dim3(256), 0, c.stream, B.U, n3k);
  hipLaunchKernelGGL((k_fine_apply_b<double, 1, kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, L0.v, cl, nch, K,
                     B.KP, (const double*)c.w64, (const double*)c.Kc64, (const double*)B.U, (const double*)B.F, B.R,
                     (double*)nullptr, B.f32[0]);
  hipLaunchKernelGGL((k_dotb<kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, cl, nch, NP, K, B.KP, (const double*)B.F,
                     (const double*)B.F, (const double*)B.R, (const double*)B.R, B.part);
  foldb(c, B, nb, 2, B_NF2, B_RN2, OPB_NONE);
  download(c.stream, B.h_sc, B.sc, (size_t)B_COUNT * K * sizeof(double));
  std::vector<double> nf(K), rt2(K, 0.0), act(K, 0.0);
  int n_act = 0;
  for (int k = 0; k < K; ++k) {
    nf[k] = std::sqrt(B.h_sc[B_NF2 * K + k]);
    rt2[k] = o.rtol * o.rtol * B.h_sc[B_NF2 * K + k];
    act[k] = (nf[k] > 0 && B.h_sc[B_RN2 * K + k] > rt2[k]) ? 1.0 : 0.0;
    n_act += act[k] != 0.0;
    if (!std::isfinite(B.h_sc[B_RN2 * K + k])) { c.err = "NaN/Inf in the initial residual"; return TOS_E_NAN; }
  }
  upload(c.stream, B.sc + B_RTOL2 * K, rt2.data(), K * sizeof(double));
  upload(c.stream, B.sc + B_ACTIVE * K, act.data(), K * sizeof(double));
  std::vector<int> its(K, 0);
  std::vector<double> rel(K, 0.0);
  for (int k = 0; k < K; ++k) rel[k] = nf[k] > 0 ? std::sqrt(B.h_sc[B_RN2 * K + k]) / nf[k] : 0.0;
  int it = 0;
  if (n_act) {
    vcycle_b(c, B, 0, B.f32[0], B.u32[0], o.pre, o.post);
    hipLaunchKernelGGL((k_pcgb_p<kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, cl, nch, NP, K, B.KP, B.P,
                       (const float*)B.u32[0], (const double*)B.sc, 1);
    hipLaunchKernelGGL((k_pcgb_zdots<kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, cl, nch, NP, K, B.KP,
                       (const float*)B.u32[0], (const double*)B.R, (const double*)nullptr, B.part);
    foldb(c, B, nb, 1, B_RZ, 0, OPB_NONE);
    LAUNCH_CHECK();
    if (B.exec && (B.graph_gal != c.gal_prec || B.graph_pre != o.pre || B.graph_post != o.post)) {
      (void)hipGraphExecDestroy(B.exec);
      (void)hipGraphDestroy(B.graph);
      B.exec = nullptr;
      B.graph = nullptr;
    }
    if (!B.exec && c.graph_replay) {
      B.graph_gal = c.gal_prec;
      B.graph_pre = o.pre;
      B.graph_post = o.post;
      c.prof_on = false;
      HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      batch_iteration(c, B, o.pre, o.post);
      HIPCHK(hipStreamEndCapture(c.stream, &B.graph));
      HIPCHK(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
      c.prof_on = prof;
    }
    while (it < o.maxit) {
      if (c.graph_replay)
        HIPCHK(hipGraphLaunch(B.exec, c.stream));
      else
        batch_iteration(c, B, o.pre, o.post);
      ++it;
      download(c.stream, B.h_sc, B.sc, (size_t)B_COUNT * K * sizeof(double));
      int still = 0;
      for (int k = 0; k < K; ++k) {
        const double rn2 = B.h_sc[B_RN2 * K + k];
        if (!std::isfinite(rn2)) { c.err = "NaN/Inf in batched FCG"; return TOS_E_NAN; }
        if (act[k] != 0.0) {
          its[k] = it;
          rel[k] = nf[k] > 0 ? std::sqrt(rn2) / nf[k] : 0.0;
        }
        act[k] = B.h_sc[B_ACTIVE * K + k];
        still += act[k] != 0.0;
      }
      if (!still) break;
    }
  }
  // true FP64 residual of every case, f - A u recomputed from the final u (the FCG loop
  // tracks the recursive one); convergence is judged on it, as in the one-case solve
  hipLaunchKernelGGL((k_fine_apply_b<double, 1, kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, L0.v, cl, nch, K,
                     B.KP, (const double*)c.w64, (const double*)c.Kc64, (const double*)B.U, (const double*)B.F, B.R,
                     (double*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL((k_dotb<kVecBS>), dim3(nb), dim3(kVecBS), 0, c.stream, cl, nch, NP, K, B.KP, (const double*)B.R,
                     (const double*)B.R, (const double*)nullptr, (const double*)nullptr, B.part);
  foldb(c, B, nb, 1, B_RN2, 0, OPB_NONE);
  hipLaunchKernelGGL(k_interleave, dim3(ni), dim3(256), 0, c.stream, cl, nch, NP, K, (const double*)B.U, c.U, 0);
  LAUNCH_CHECK();
  download(c.stream, B.h_sc, B.sc, (size_t)B_COUNT * K * sizeof(double));
  int rc = TOS_OK;
  for (int k = 0; k < K; ++k) {
    const double rn2 = B.h_sc[B_RN2 * K + k];
    if (!std::isfinite(rn2)) { c.err = "NaN/Inf in the batched true residual"; return TOS_E_NAN; }
    rel[k] = nf[k] > 0 ? std::sqrt(rn2) / nf[k] : 0.0;
    if (st && k < 64) { st->iterations[k] = its[k]; st->rel_res[k] = rel[k]; }
    if (rel[k] > o.rtol) rc = TOS_E_NOCONV;
  }
  return rc;
}

// ---------------------------------------------------------------------------------------
// Coarsest system (buildCoarsestSystem grid/Grid.cpp:3614-3669 assembles it on the host
// and solves with Eigen BDCSVD each V-cycle): dense assembly on the device, blocked
// symmetric Gauss-Jordan sweep leaving -A^{-1}, applied as a matvec in the V-cycle.
// ---------------------------------------------------------------------------------------
void coarse_alloc(Ctx& c) {
  Level& L = c.L.back();
  std::vector<int64_t> nodes;
  std::vector<int> xyz;
  std::vector<int> cmap(L.v.NP, -1);
  for (int z = 0; z < L.v.nz; ++z)
    for (int y = 0; y < L.v.ny; ++y)
      for (int x = 0; x < L.v.nx; ++x)
        if (L.h_free[x + (size_t)L.v.nx * (y + (size_t)L.v.ny * z)]) {
          cmap[L.v.pnode(x, y, z)] = (int)nodes.size();
          nodes.push_back(L.v.pnode(x, y, z));
          xyz.push_back(x); xyz.push_back(y); xyz.push_back(z);
        }
  c.cn = 3 * (int)nodes.size();
  c.cidx = dalloc<int64_t>(c, nodes.size());
  if (!nodes.empty()) upload(c.stream, c.cidx, nodes.data(), nodes.size() * sizeof(int64_t));
  c.cmap = dalloc<int>(c, cmap.size());
  upload(c.stream, c.cmap, cmap.data(), cmap.size() * sizeof(int));
  c.cinv = dalloc<double>(c, (size_t)c.cn * c.cn);
  c.cwork = dalloc<double>(c, (size_t)c.cn * kGJB * 2 + kGJB * kGJB + c.cn);
  c.cflag = dalloc<int>(c, 1);
  if (c.floating && c.cn) {
    // the rigid motions of the coarsest nodes (coordinates relative to their centroid; a
    // coarse level's nodes sit at 2^l times their indices on the fine lattice, the same
    // span), orthonormalised by modified Gram-Schmidt; dropped columns as the reference's
    const int n = c.cn, nq = n / 3;
    double o[3] = {0, 0, 0};
    for (int q = 0; q < nq; ++q)
      for (int k = 0; k < 3; ++k) o[k] += (double)xyz[3 * q + k] / nq;
    std::vector<double> R(6 * (size_t)n, 0.0);
    for (int q = 0; q < nq; ++q) {
      const double px = xyz[3 * q] - o[0], py = xyz[3 * q + 1] - o[1], pz = xyz[3 * q + 2] - o[2];
      R[0 * (size_t)n + 3 * q] = 1; R[1 * (size_t)n + 3 * q + 1] = 1; R[2 * (size_t)n + 3 * q + 2] = 1;
      R[3 * (size_t)n + 3 * q + 1] = pz; R[3 * (size_t)n + 3 * q + 2] = -py;
      R[4 * (size_t)n + 3 * q] = -pz; R[4 * (size_t)n + 3 * q + 2] = px;
      R[5 * (size_t)n + 3 * q] = py; R[5 * (size_t)n + 3 * q + 1] = -px;
    }
    int m = 0;
    for (int i = 0; i < 6; ++i) {
      double* ci = &R[(size_t)i * n];
      double old2 = 0;
      for (int t = 0; t < n; ++t) old2 += ci[t] * ci[t];
      for (int j = 0; j < m; ++j) {
        const double* cj = &R[(size_t)j * n];
        double d = 0;
        for (int t = 0; t < n; ++t) d += ci[t] * cj[t];
        for (int t = 0; t < n; ++t) ci[t] -= d * cj[t];
      }
      double r2 = 0;
      for (int t = 0; t < n; ++t) r2 += ci[t] * ci[t];
      if (!(old2 >= 1e-12 && r2 >= 1e-12 * old2)) continue;
      const double inv = 1.0 / std::sqrt(r2);
      double* dst = &R[(size_t)m * n];
      for (int t = 0; t < n; ++t) dst[t] = ci[t] * inv;
      ++m;
    }
    c.crig_m = m;
    c.crig = dalloc<double>(c, (size_t)std::max(m, 1) * n);
    if (m) upload(c.stream, c.crig, R.data(), (size_t)m * n * sizeof(double));
  }
}

// factor of the coarsest level's FP64 stencil into cinv, or (from32) of its FP32 stencil
// into cinv32
int coarse_factor(Ctx& c, bool from32) {
  if (c.cn == 0) return TOS_OK;
  Prof pr(c, P_COARSE_FACTOR);
  Level& L = c.L.back();
  const int n = c.cn, nnode = n / 3;
  if (from32 && !c.cinv32) c.cinv32 = dalloc<double>(c, (size_t)n * n);
  double* A = from32 ? c.cinv32 : c.cinv;
  hipLaunchKernelGGL(k_zero<double>, dim3(nblk((int64_t)n * n)), dim3(256), 0, c.stream, A, (int64_t)n * n);
  if (from32)
    hipLaunchKernelGGL(k_coarse_assemble<float>, dim3(nblk((int64_t)nnode * 27)), dim3(256), 0, c.stream, L.v,
                       (const int64_t*)c.cidx, nnode, (const int*)c.cmap, (const int*)L.slot,
                       (const float*)L.st32, L.NS, A, n);
  else
    hipLaunchKernelGGL(k_coarse_assemble<double>, dim3(nblk((int64_t)nnode * 27)), dim3(256), 0, c.stream, L.v,
                       (const int64_t*)c.cidx, nnode, (const int*)c.cmap, (const int*)L.slot,
                       (const double*)L.st64, L.NS, A, n);
  double* Dinv = c.cwork;
  double* Cp = Dinv + kGJB * kGJB;
  double* R = Cp + (size_t)n * kGJB;
  double* diag0 = R + (size_t)n * kGJB;
  hipLaunchKernelGGL(k_symmetrize, dim3(nblk((int64_t)n * n)), dim3(256), 0, c.stream, A, n, diag0);
  if (c.floating && c.crig_m) {
    // null space of a floating body: factor A + alpha R R^T (alpha = mean diagonal); for a
    // right-hand side orthogonal to R its solution is A's minimum-norm solution, the one
    // the reference's BDCSVD solve returns (grid/Grid.cpp:3654-3661, 4072)
    std::vector<double> d(n);
    download(c.stream, d.data(), diag0, (size_t)n * sizeof(double));
    double alpha = 0;
    for (double x : d) alpha += x / n;
    hipLaunchKernelGGL(k_add_lowrank, dim3(nblk((int64_t)n * n)), dim3(256), 0, c.stream, A, n,
                       (const double*)c.crig, c.crig_m, alpha);
  }
  HIPCHK(hipMemsetAsync(c.cflag, 0, sizeof(int), c.stream));
  for (int K0 = 0; K0 < n; K0 += kGJB) {
    const int b = std::min(kGJB, n - K0);
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(64), 0, c.stream, A, n, K0, b, Dinv, (const double*)diag0,
                       c.cflag);
    hipLaunchKernelGGL(k_gj_panel, dim3(nblk((int64_t)n * b)), dim3(256), 0, c.stream, (const double*)A, n, K0,
                       b, (const double*)Dinv, Cp, R);
    hipLaunchKernelGGL(k_gj_update, dim3(nblk((int64_t)n * n)), dim3(256), 0, c.stream, A, n, K0, b,
                       (const double*)Dinv, (const double*)Cp, (const double*)R);
  }
  LAUNCH_CHECK();
  int flag = 0;
  download(c.stream, &flag, c.cflag, sizeof(int));
  if (flag) {
    c.err = "coarsest system: non-positive pivot (no supports or a floating part?)";
    return TOS_E_SINGULAR;
  }
  return TOS_OK;
}

// ---------------------------------------------------------------------------------------
// Solve (flexible PCG or plain V-cycle iteration) for load case k
// ---------------------------------------------------------------------------------------
int solve_case(Ctx& c, int k, const tos_solve_opts& o, int* its, double* relres) {
  Level& L0 = c.L[0];
  const int64_t NP = L0.v.NP, n3 = 3 * NP;
  double* u = c.U + (size_t)k * n3;
  const double* f = c.F + (size_t)k * n3;
  const bool f32 = o.precision != 0;
  const unsigned nb = c.nb0;
  const int4* cl = L0.cl_lin;
  const int nch = L0.n_lin;
  if (!o.warm_start && nch)
    hipLaunchKernelGGL(k_zero_nodes<double>, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u);
  if (c.r_ready) c.r_ready = false;  // r = f - K u formed by the recycled start
  else apply_level<double, 1>(c, 0, u, f, c.r, nullptr, f32 ? L0.f32 : nullptr);  // r = f - K u
  dot_to(c, f, f, c.r, c.r, S_NF2, S_RN2, OP_NONE);
  const double* hs = read_scal(c);
  const double nf = std::sqrt(hs[S_NF2]);
  double rn = std::sqrt(hs[S_RN2]);
  *its = 0;
  if (nf == 0) {
    if (nch) hipLaunchKernelGGL(k_zero_nodes<double>, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u);
    *relres = 0;
    return TOS_OK;
  }
  if (!std::isfinite(rn)) {
    c.err = "NaN/Inf in the initial residual";
    return TOS_E_NAN;
  }
  int it = 0;
  // the learned correction pauses for the rest of this solve after 3 rejections in a row;
  // after a solve in which it was tried and never kept, it sits out the next 1, 3, 7, 15
  // solves (doubling while it keeps failing, reset when a correction is kept)
  LearnState* LS = (c.learn && c.learn->mode == 1) ? c.learn : nullptr;
  int nn_rej = 0;
  double nn_tried = hs[S_NN_TRIED], nn_kept = hs[S_NN_KEPT];
  const double nn_tried0 = nn_tried, nn_kept0 = nn_kept;
  if (LS) {
    LS->paused = LS->skip > 0;
    if (LS->skip > 0) --LS->skip;
  }
  const bool nn_off = !LS || LS->paused;  // no correction anywhere in this solve
  auto nn_watch = [&](const double* h) {
    if (!LS || LS->paused) return;
    if (h[S_NN_TRIED] > nn_tried) nn_rej = h[S_NN_KEPT] > nn_kept ? 0 : nn_rej + 1;
    nn_tried = h[S_NN_TRIED];
    nn_kept = h[S_NN_KEPT];
    if (nn_rej >= 3) LS->paused = true;
  };
  if (!o.use_pcg) {
    while (rn / nf > o.rtol && it < o.maxit) {
      precondition(c, o.precision, o.pre, o.post, f32, true);
      hipLaunchKernelGGL(k_add_nodes<double>, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u, (const double*)c.z);
      apply_level<double, 1>(c, 0, u, f, c.r, nullptr, f32 ? L0.f32 : nullptr);
      dot_to(c, c.r, c.r, nullptr, nullptr, S_RN2, 0, OP_NONE);
      rn = std::sqrt(read_scal(c)[S_RN2]);
      ++it;
      if (!std::isfinite(rn)) { c.err = "NaN/Inf in the V-cycle iteration"; return TOS_E_NAN; }
    }
  } else if (rn / nf > o.rtol) {
    // flexible CG, Polak-Ribiere beta = -alpha <z, q> / <z_old, r_old>.  With an FP32
    // V-cycle the search direction p is stored in FP32 (c.p32): q = K p, u += alpha p and
    // r -= alpha q stay FP64 operations on the stored p, so the recursive residual keeps
    // tracking f - K u; only p's rounding changes the direction.
    float* p32 = c.p32;
    precondition(c, o.precision, o.pre, o.post, f32, false);
    {
      Prof pr(c, P_VECTOR);
      if (f32) {
        hipLaunchKernelGGL((k_pcg_up<float, float>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u, p32,
                           (const float*)L0.u32, (const double*)c.scal, 1);
        hipLaunchKernelGGL((k_pcg_zdots<float>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, (const float*)L0.u32,
                           (const double*)c.r, (const double*)nullptr, c.partial);
      } else {
        hipLaunchKernelGGL((k_pcg_p<double, double>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, c.p,
                           (const double*)c.z, (const double*)c.scal, 1);
        hipLaunchKernelGGL((k_pcg_zdots<double>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, (const double*)c.z,
                           (const double*)c.r, (const double*)nullptr, c.partial);
      }
      finish0(c, 0, S_RZ, 0, OP_NONE);
    }
    // q = K p, alpha, u += alpha p, r -= alpha q, |r|^2
    auto tail = [&]() {
      if (f32)
        apply_fine_dir(c, p32, c.q, c.partial);
      else
        apply_level<double, 0>(c, 0, c.p, nullptr, c.q, c.partial);  // partials of p.q
      {
        Prof pr(c, P_VECTOR);
        finish0(c, 0, S_PQ, 0, OP_ALPHA);
        if (f32)  // (u += alpha p is deferred to the next direction update)
          hipLaunchKernelGGL(k_pcg_r, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, c.r, (const double*)c.q,
                             (const double*)c.scal, L0.f32, c.partial);
        else
          hipLaunchKernelGGL((k_pcg_ur<double>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u, c.r,
                             (const double*)c.p, (const double*)c.q, (const double*)c.scal, (float*)nullptr,
                             c.partial);
        finish0(c, 0, S_RN2, 0, OP_NONE);
      }
      LAUNCH_CHECK();
    };
    // z = B r, beta, p = z + beta p
    auto head = [&]() {
      precondition(c, o.precision, o.pre, o.post, f32, false);
      Prof pr(c, P_VECTOR);
      if (f32) {
        hipLaunchKernelGGL((k_pcg_zdots<float>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP,
                           (const float*)L0.u32, (const double*)c.r, (const double*)c.q, c.partial);
        finish0(c, 1, S_ZR, S_ZQ, OP_BETA);
        hipLaunchKernelGGL((k_pcg_up<float, float>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u, p32,
                           (const float*)L0.u32, (const double*)c.scal, 0);
      } else {
        hipLaunchKernelGGL((k_pcg_zdots<double>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP,
                           (const double*)c.z, (const double*)c.r, (const double*)c.q, c.partial);
        finish0(c, 1, S_ZR, S_ZQ, OP_BETA);
        hipLaunchKernelGGL((k_pcg_p<double, double>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, c.p,
                           (const double*)c.z, (const double*)c.scal, 0);
      }
      LAUNCH_CHECK();
    };
    // the iteration body [head, tail] is captured once into a graph and replayed (fixed
    // pointers; scalars stay on the device); not while kernels are timed with events or
    // while the learned correction takes host-side decisions
    const bool graphed =
        c.graph_replay && !c.prof_on && !(c.learn && (c.learn->ds || (c.learn->mode == 1 && !nn_off)));
    hipGraphExec_t exec = nullptr;
    if (graphed) {
      // a graph captured before any later allocation or release of context memory is
      // dropped (no captured pointer may outlive its buffer), then looked up by configuration
      bool stale = false;
      for (const Ctx::SolveGraph& g : c.graphs) stale |= g.epoch != c.alloc_epoch;
      if (stale) {
        for (const Ctx::SolveGraph& g : c.graphs) {
          (void)hipGraphExecDestroy(g.exec);
          (void)hipGraphDestroy(g.graph);
        }
        c.graphs.clear();
      }
      for (const Ctx::SolveGraph& g : c.graphs)
        if (g.precision == o.precision && g.pre == o.pre && g.post == o.post && g.sweep == c.sweep_mode &&
            g.gal == c.gal_prec && g.u == u)
          exec = g.exec;
    }
    bool first = true;
    while (it < o.maxit) {
      if (first) {
        tail();
        first = false;
      } else if (graphed) {
        if (!exec) {
          Ctx::SolveGraph g{o.precision, o.pre, o.post, c.sweep_mode, c.gal_prec, 0, u, 0, nullptr, nullptr};
          HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
          head();
          tail();
          HIPCHK(hipStreamEndCapture(c.stream, &g.graph));
          HIPCHK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
          g.epoch = c.alloc_epoch;
          c.graphs.push_back(g);
          exec = g.exec;
        }
        HIPCHK(hipGraphLaunch(exec, c.stream));
      } else {
        head();
        tail();
      }
      ++it;
      const double* h = read_scal(c);
      rn = std::sqrt(h[S_RN2]);
      nn_watch(h);
      if (!std::isfinite(rn)) { c.err = "NaN/Inf in PCG"; return TOS_E_NAN; }
      if (rn / nf <= o.rtol) break;
    }
    if (f32 && it > 0)  // the last iteration's deferred u += alpha p
      hipLaunchKernelGGL((k_pcg_ustep<float>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u, (const float*)p32,
                         (const double*)c.scal);
  }
  if (LS && !nn_off && nn_tried > nn_tried0) {
    if (nn_kept > nn_kept0) {
      LS->dry = 0;
    } else {
      LS->dry = std::min(LS->dry + 1, 4);
      LS->skip = (1 << LS->dry) - 1;
    }
  }
  // true FP64 residual
  apply_level<double, 1>(c, 0, u, f, c.r);
  dot_to(c, c.r, c.r, nullptr, nullptr, S_RN2, 0, OP_NONE);
  *relres = std::sqrt(read_scal(c)[S_RN2]) / nf;
  *its = it;
  return *relres <= o.rtol ? TOS_OK : TOS_E_NOCONV;
}

// Precision 3: the precision-2 solve with the FCG residual r and K p stored in FP32.  The
// operator keeps FP64 arithmetic on the FP32 direction (q = K p is exact before its
// rounding to FP32), u stays FP64, and the recursive residual r -= alpha q is kept in FP32
// (the V-cycle reads it directly); when it reports convergence, f - K u is formed in FP64
// and, if that misses the tolerance, the FCG restarts from it.  No reference counterpart.
int solve_case_r32(Ctx& c, int k, const tos_solve_opts& o, int* its, double* relres) {
  Level& L0 = c.L[0];
  const int64_t NP = L0.v.NP, n3 = 3 * NP;
  double* u = c.U + (size_t)k * n3;
  const double* f = c.F + (size_t)k * n3;
  const unsigned nb = c.nb0;
  const int4* cl = L0.cl_lin;
  const int nch = L0.n_lin;
  if (!c.q32) c.q32 = dalloc<float>(c, n3);
  float* r32 = L0.f32;  // the FCG residual is the V-cycle's right-hand side
  float* q32 = c.q32;
  float* p32 = c.p32;
  const float* z32 = L0.u32;
  if (!o.warm_start && nch)
    hipLaunchKernelGGL(k_zero_nodes<double>, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u);
  if (c.r_ready) c.r_ready = false;  // formed by the recycled start
  else apply_level<double, 1>(c, 0, u, f, c.r, nullptr, r32);  // FP64 residual and its FP32 copy
  dot_to(c, f, f, c.r, c.r, S_NF2, S_RN2, OP_NONE);
  const double* hs = read_scal(c);
  const double nf = std::sqrt(hs[S_NF2]);
  double rn = std::sqrt(hs[S_RN2]);
  *its = 0;
  if (nf == 0) {
    if (nch) hipLaunchKernelGGL(k_zero_nodes<double>, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u);
    *relres = 0;
    return TOS_OK;
  }
  if (!std::isfinite(rn)) {
    c.err = "NaN/Inf in the initial residual";
    return TOS_E_NAN;
  }
  // four direction buffers: direction j lives in P[j % 4] with its alpha in S_AH0 + j % 4,
  // and u takes four steps at a time (k_pcg_flush) instead of one read and write of the
  // FP64 u per iteration; the same steps in the same order as one at a time
  for (float*& b : c.p32s)
    if (!b) b = dalloc<float>(c, n3);
  {
    int64_t nd = 0;  // per-wave z.r / z.q sums of the last fine sweep (gs_sweep_pairw)
    for (int i = 0; i < 4; ++i) nd += L0.n_segw[i][1];
    if (nd && c.gsdots_n != nd) {
      c.gsdots = dalloc<double>(c, 2 * (size_t)nd);
      c.gsdots_n = nd;
    }
  }
  float* const P[4] = {p32, c.p32s[0], c.p32s[1], c.p32s[2]};
  auto flush = [&](int m) {  // u += alpha_k p_k for the directions in P[0 .. m-1]
    PcgDirs D{};
    for (int s = 0; s < m; ++s) {
      D.p[s] = P[s];
      D.aslot[s] = S_AH0 + s;
    }
    hipLaunchKernelGGL(k_pcg_flush, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u, D, m, (const double*)c.scal);
  };
  // q = K p_j (FP64 arithmetic, FP32 result), alpha_j, r -= alpha_j q, |r|^2
  auto tail = [&](int s) {
    {
      Prof pr(c, P_FINE_APPLY64);
      hipLaunchKernelGGL((k_fine_apply_dir<float>), dim3(cblk(L0.n_lin)), dim3(256), 0, c.stream, L0.v, cl, nch,
                         Prec<double>::w(c), Prec<double>::Ke(c), (const float*)P[s], c.partial, q32);
    }
    Prof pr(c, P_VECTOR);
    finish0(c, 0, S_PQ, S_AH0 + s, OP_ALPHA);
    hipLaunchKernelGGL(k_pcg_r32, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, r32, (const float*)q32,
                       (const double*)c.scal, c.partial);
    finish0(c, 0, S_RN2, 0, OP_NONE);
    LAUNCH_CHECK();
  };
  // z = B r, beta, (buffer 0: the four pending steps into u), p_j = z + beta p_{j-1}
  auto head = [&](int s) {
    c.dots_done = false;
    {
      struct DotsReq {  // z.r and z.q from the last fine sweep where it can (cleared on any exit)
        Ctx& x;
        DotsReq(Ctx& x_, const float* q) : x(x_) { x.dots_q = q; }
        ~DotsReq() { x.dots_q = nullptr; x.dots_last = false; }
      } req(c, q32);
      precondition(c, 2, o.pre, o.post, true, false);
    }
    Prof pr(c, P_VECTOR);
    if (c.dots_done) {
      finish_parts(c, c.gsdots, c.gsdots_n, S_ZR, S_ZQ, OP_BETA);
    } else {
      hipLaunchKernelGGL(k_pcg_zdots32, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, z32, (const float*)r32,
                         (const float*)q32, c.partial);
      finish0(c, 1, S_ZR, S_ZQ, OP_BETA);
    }
    c.dots_done = false;
    if (s == 0) flush(4);
    hipLaunchKernelGGL(k_pcg_pnew, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, P[s], z32,
                       (const float*)P[(s + 3) % 4], (const double*)c.scal);
    LAUNCH_CHECK();
  };
  const bool graphed = c.graph_replay && !c.prof_on;
  hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
  if (graphed) {
    bool stale = false;
    for (const Ctx::SolveGraph& g : c.graphs) stale |= g.epoch != c.alloc_epoch;
    if (stale) {
      for (const Ctx::SolveGraph& g : c.graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
      }
      c.graphs.clear();
    }
    for (const Ctx::SolveGraph& g : c.graphs)
      if (g.precision == 3 && g.pre == o.pre && g.post == o.post && g.sweep == c.sweep_mode && g.gal == c.gal_prec &&
          g.u == u)
        exec[g.slot] = g.exec;
  }
  int it = 0;
  while (rn / nf > o.rtol && it < o.maxit) {
    // FCG on r32 (from the true residual: at the start and after a restart); p_0 = z
    precondition(c, 2, o.pre, o.post, true, false);
    {
      Prof pr(c, P_VECTOR);
      hipLaunchKernelGGL((k_pcg_up<float, float>), dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, u, P[0], z32,
                         (const double*)c.scal, 1);
      hipLaunchKernelGGL(k_pcg_zdots32, dim3(nb), dim3(256), 0, c.stream, cl, nch, NP, z32, (const float*)r32,
                         (const float*)nullptr, c.partial);
      finish0(c, 0, S_RZ, 0, OP_NONE);
    }
    int j = -1;  // index of the last direction of this run
    while (it < o.maxit) {
      const int s = ++j % 4;
      if (j == 0) {
        tail(0);
      } else if (graphed) {
        if (!exec[s]) {
          Ctx::SolveGraph g{3, o.pre, o.post, c.sweep_mode, c.gal_prec, s, u, 0, nullptr, nullptr};
          HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
          head(s);
          tail(s);
          HIPCHK(hipStreamEndCapture(c.stream, &g.graph));
          HIPCHK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
          g.epoch = c.alloc_epoch;
          c.graphs.push_back(g);
          exec[s] = g.exec;
        }
        HIPCHK(hipGraphLaunch(exec[s], c.stream));
      } else {
        head(s);
        tail(s);
      }
      ++it;
      const double r_in = std::sqrt(read_scal(c)[S_RN2]);
      if (!std::isfinite(r_in)) { c.err = "NaN/Inf in PCG"; return TOS_E_NAN; }
      if (r_in / nf <= o.rtol) break;
    }
    flush(j % 4 + 1);  // the steps of the directions since the last flush
    // the true residual in FP64; its FP32 copy restarts the FCG when it misses the tolerance
    apply_level<double, 1>(c, 0, u, f, c.r, nullptr, r32);
    dot_to(c, c.r, c.r, nullptr, nullptr, S_RN2, 0, OP_NONE);
    rn = std::sqrt(read_scal(c)[S_RN2]);
    if (!std::isfinite(rn)) { c.err = "NaN/Inf in the true residual"; return TOS_E_NAN; }
  }
  *relres = rn / nf;
  *its = it;
  return *relres <= o.rtol ? TOS_OK : TOS_E_NOCONV;
}

float* sens_buf(Ctx& c, int which) { return which == 0 ? c.sens : which == 1 ? c.vsens : nullptr; }

double volume_of(Ctx& c) {
  const LevelView& v = c.L[0].v;
  hipLaunchKernelGGL(k_sum_active, dim3(kRedBlocks), dim3(256), 0, c.stream, v, (const uint8_t*)c.eact,
                     (const float*)c.rho, c.partial);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, c.stream, (const double*)c.partial, kRedBlocks, 0, S_VOL, 0,
                     OP_NONE, c.scal);
  LAUNCH_CHECK();
  return read_scal(c)[S_VOL] / (double)c.n_active;
}

// ---------------------------------------------------------------------------------------
// Spline tables (built on the host, once per tos_spline_init)
// ---------------------------------------------------------------------------------------
// Cox-de Boor recursion (The NURBS Book A2.2; SplineBasisX grid/Grid.cu:384-410): the
// `p+1` nonzero basis values at x on knot span `span`.
void cox_de_boor(int span, double x, int p, const std::vector<double>& U, double* N) {
  std::vector<double> left(p + 1), right(p + 1);
  N[0] = 1.0;
  for (int j = 1; j <= p; ++j) {
    left[j] = x - U[span + 1 - j];
    right[j] = U[span + j] - x;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      const double tmp = N[r] / (right[r + 1] + left[j - r]);
      N[r] = saved + right[r + 1] * tmp;
      saved = left[j - r] * tmp;
    }
    N[j] = saved;
  }
}

// uniform B-spline of order o, support [0, o), by the recursion on degree
double uniform_bspline(int o, double t) {
  if (t < 0 || t >= o) return 0.0;
  std::vector<double> b(o, 0.0);
  const int i0 = (int)std::floor(t);
  b[i0] = 1.0;  // order 1: indicator of [i0, i0+1)
  for (int k = 2; k <= o; ++k)
    for (int i = 0; i + k <= o; ++i)
      b[i] = ((t - i) * b[i] + (i + k - t) * b[i + 1]) / (k - 1);
  return b[0];
}

struct AxisTable {
  std::vector<int> idx;
  std::vector<float> wt;
  std::vector<uint8_t> ok;
  int nc = 0;
};

AxisTable spline_table(int mode, int order, int param, int ne, double h) {
  AxisTable A;
  A.idx.assign((size_t)ne * order, 0);
  A.wt.assign((size_t)ne * order, 0.f);
  A.ok.assign(ne, 1);
  if (mode == 0) {
    // clamped knots on the box enlarged by 2*box/(P+1) per side (Grid::set_spline_knot_series
    // grid/Grid.cpp:3253-3377), span index as in coeff2density (grid/Grid.cu:3393-3402)
    const int P = param;
    const double box = ne * h;
    const double lo = -2.0 * box / (P + 1), hi = box + 2.0 * box / (P + 1);
    const double step = (hi - lo) / (P + 1);
    std::vector<double> U(order, lo);
    double t = lo;
    for (int i = 0; i < P; ++i) { t += step; U.push_back(t); }
    for (int i = 0; i < order; ++i) U.push_back(hi);
    A.nc = order + P;
    std::vector<double> N(order);
    for (int e = 0; e < ne; ++e) {
      const double x = (e + 0.5) * h;
      const int i = (int)((x - lo) / step) + order;
      if (i < order || i > A.nc) { A.ok[e] = 0; continue; }
      cox_de_boor(i - 1, x, order - 1, U, N.data());
      for (int q = 0; q < order; ++q) {
        A.idx[(size_t)e * order + q] = i - order + q;
        A.wt[(size_t)e * order + q] = (float)N[q];
      }
    }
  } else {
    // uniform lattice: coefficient i at node param*i, replicated at the boundary
    const int s = param;
    A.nc = (ne + s - 1) / s + 1;
    for (int e = 0; e < ne; ++e) {
      const double t = (e + 0.5) / s;
      const int i0 = (int)std::floor(t - order / 2.0) + 1;
      for (int q = 0; q < order; ++q) {
        const int i = i0 + q;
        A.idx[(size_t)e * order + q] = std::min(std::max(i, 0), A.nc - 1);
        A.wt[(size_t)e * order + q] = (float)uniform_bspline(order, t - i + order / 2.0);
      }
    }
  }
  return A;
}

// ---------------------------------------------------------------------------------------
// MMA: host control of the primal-dual subproblem (mma_subproblem_t::solve mma_t.cu:330-499)
// over the fused device passes of mma.hip.
// ---------------------------------------------------------------------------------------
template <class T>
T* malloc_mma(tos_mma* M, size_t n) {
  return (T*)raw_alloc(M->allocs, M->stream, n * sizeof(T));
}

// run a pass kernel, fold its `cnt` columns, read them to M->h_red
template <class K, class... A>
const double* mma_pass(tos_mma* M, hipStream_t st, K kernel, int cnt, unsigned maxmask, A... args) {
  hipLaunchKernelGGL(kernel, dim3(M->nb), dim3(256), 0, st, args...);
  hipLaunchKernelGGL(k_mma_fold, dim3(cnt), dim3(64), 0, st, (const double*)M->partial, M->nb, cnt, maxmask, M->red);
  LAUNCH_CHECK();
  download(st, M->h_red, M->red, cnt * sizeof(double));
  return M->h_red;
}

// KKT residual (2-norm, inf-norm) from the folded n-length parts and the scalars
// (kkt_err mma_t.cu:574-678, same order of accumulation)
std::pair<double, double> mma_kkt_host(const tos_mma* M, const double* v, const double* y, const double* lam,
                                       const double* mu, const double* s, double z, double zeta, double damp) {
  const int m = M->V.m;
  double sum = v[0], mx = std::max(-10.0, v[3 + MMAX]);
  auto acc = [&](double r) { sum += r * r; mx = std::max(mx, std::fabs(r)); };
  for (int i = 0; i < m; ++i) acc(M->c[i] + M->d[i] * y[i] - lam[i] - mu[i]);
  double la = 0;
  for (int i = 0; i < m; ++i) la += lam[i] * M->a[i];
  acc(1.0 - zeta - la);
  for (int i = 0; i < m; ++i) acc((v[3 + i] - M->a[i] * z - y[i] + s[i] - M->b[i]) / damp);
  sum += v[1]; mx = std::max(mx, v[4 + MMAX]);
  sum += v[2]; mx = std::max(mx, v[5 + MMAX]);
  for (int i = 0; i < m; ++i) acc(mu[i] * y[i] - M->eps);
  acc(zeta * z - M->eps);
  for (int i = 0; i < m; ++i) acc(lam[i] * s[i] - M->eps);
  return {std::sqrt(sum), mx};
}

bool mma_update_dev(tos_mma* M, hipStream_t st, const double* df, const double* dg, const double* g,
                    std::function<void(int)> prof_mark = nullptr) {
  (void)prof_mark;
  MmaVecs& V = M->V;
  const int m = V.m;
  MmaSmall S{};
  S.move = M->move;
  constexpr unsigned KKT_MAX = (1u << (3 + MMAX)) | (1u << (4 + MMAX)) | (1u << (5 + MMAX));
  // p, q, b and the start point
  {
    const double* v = mma_pass(M, st, k_mma_pq, m, 0u, V, df, dg, M->partial);
    for (int i = 0; i < m; ++i) M->b[i] = v[i] - g[i];
  }
  for (int i = 0; i < m; ++i) {
    M->y[i] = 1; M->lam[i] = 1; M->s[i] = 1; M->mu[i] = std::max(M->c[i] / 2, 1.0);
  }
  M->z = 1; M->zeta = 1; M->eps = 1;
  for (int i = 0; i < m; ++i) S.lam[i] = M->lam[i];
  hipLaunchKernelGGL(k_mma_plam, dim3(M->nb), dim3(256), 0, st, V, S);
  LAUNCH_CHECK();
  double lasterr = 0;
  int nsub = 0;
  bool overflow = false;
  for (int lev = 0; lev < 6 && !overflow; ++lev) {
    double err = 1e30, damp = 1;
    int stays = 0;
    while (err > 0.9 * M->eps) {
      S.eps = M->eps;
      constexpr int CA = MMAX + MMAX * MMAX + MMAX;
      const double* v = mma_pass(M, st, k_mma_passA, CA, 0u, V, S, M->partial);
      // reduced (m+1) system (solveLinearSystem mma_t.cu:228-328), Gaussian elimination
      const int N1 = m + 1;
      double A[(MMAX + 1) * (MMAX + 1)] = {}, rhs[MMAX + 1], Dy[MMAX], dly[MMAX], dll[MMAX];
      for (int i = 0; i < m; ++i) {
        Dy[i] = M->d[i] + M->mu[i] / M->y[i];
        dly[i] = -(M->lam[i] + M->eps / M->y[i] - M->c[i] - M->d[i] * M->y[i]);
        dll[i] = -(M->z * M->a[i] + M->y[i] + M->b[i] - M->eps / M->lam[i] - v[i]);
      }
      double la = 0;
      for (int i = 0; i < m; ++i) la += M->lam[i] * M->a[i];
      const double dlz = -(la + M->eps / M->z - 1.0);
      for (int i = 0; i < m; ++i) {
        A[i * N1 + i] += M->s[i] / M->lam[i] + 1.0 / Dy[i];
        for (int k = 0; k < m; ++k) A[i * N1 + k] += v[MMAX + i * MMAX + k];
        rhs[i] = dll[i] + dly[i] / Dy[i] - v[MMAX + MMAX * MMAX + i];
        A[i * N1 + m] = M->a[i];
        A[m * N1 + i] = M->a[i];
      }
      A[m * N1 + m] = -M->zeta / M->z;
      rhs[m] = dlz;
      for (int k = 0; k < N1; ++k) {
        int piv = k;
        for (int i = k + 1; i < N1; ++i)
          if (std::fabs(A[i * N1 + k]) > std::fabs(A[piv * N1 + k])) piv = i;
        if (piv != k) {
          for (int j = 0; j < N1; ++j) std::swap(A[k * N1 + j], A[piv * N1 + j]);
          std::swap(rhs[k], rhs[piv]);
        }
        for (int i = k + 1; i < N1; ++i) {
          const double f = A[i * N1 + k] / A[k * N1 + k];
          for (int j = k; j < N1; ++j) A[i * N1 + j] -= f * A[k * N1 + j];
          rhs[i] -= f * rhs[k];
        }
      }
      for (int i = N1 - 1; i >= 0; --i) {
        double sacc = rhs[i];
        for (int j = i + 1; j < N1; ++j) sacc -= A[i * N1 + j] * rhs[j];
        rhs[i] = sacc / A[i * N1 + i];
      }
      double dlam[MMAX], dy[MMAX], dmu[MMAX], ds[MMAX];
      for (int i = 0; i < m; ++i) {
        dlam[i] = rhs[i];
        dy[i] = dlam[i] / Dy[i] - dly[i] / Dy[i];
        dmu[i] = M->eps / M->y[i] - dy[i] * M->mu[i] / M->y[i] - M->mu[i];
        ds[i] = M->eps / M->lam[i] - M->s[i] / M->lam[i] * dlam[i] - M->s[i];
        S.dlam[i] = dlam[i];
      }
      const double dz = rhs[m];
      const double dzeta = M->eps / M->z - M->zeta - (M->zeta / M->z) * dz;
      // step bound
      const double* tv = mma_pass(M, st, k_mma_passB, 3, 7u, V, S, M->partial);
      const double cf = 1.0 / 0.99;
      double t3 = tv[2];
      for (int i = 0; i < m; ++i) t3 = std::max(t3, -cf * dy[i] / M->y[i]);
      t3 = std::max(t3, -cf * dz / M->z);
      for (int i = 0; i < m; ++i) t3 = std::max(t3, -cf * dlam[i] / M->lam[i]);
      for (int i = 0; i < m; ++i) t3 = std::max(t3, -cf * dmu[i] / M->mu[i]);
      t3 = std::max(t3, -cf * dzeta / M->zeta);
      for (int i = 0; i < m; ++i) t3 = std::max(t3, -cf * ds[i] / M->s[i]);
      double step = 1.0 / std::max(std::max(tv[0], tv[1]), std::max(t3, 1.0));
      // backtracking on the KKT 2-norm (search_step mma_t.cu:501-571)
      for (int i = 0; i < m; ++i) S.lam[i] = M->lam[i];
      const double* kv = mma_pass(M, st, k_mma_kkt<false>, 3 + MMAX + 3, KKT_MAX, V, S, M->partial);
      const double old_err = mma_kkt_host(M, kv, M->y, M->lam, M->mu, M->s, M->z, M->zeta, 1.0).first;
      double new_err = 1e30, new_inf = 1e30;
      while (step > 1e-12 && new_err > old_err) {
        double ty[MMAX], tl[MMAX], tm[MMAX], ts[MMAX];
        for (int i = 0; i < m; ++i) {
          ty[i] = M->y[i] + dy[i] * step;
          tl[i] = M->lam[i] + dlam[i] * step;
          tm[i] = M->mu[i] + dmu[i] * step;
          ts[i] = M->s[i] + ds[i] * step;
          S.lam[i] = tl[i];
        }
        const double tz = M->z + dz * step, tzeta = M->zeta + dzeta * step;
        S.step = step;
        const double* kt = mma_pass(M, st, k_mma_kkt<true>, 3 + MMAX + 3, KKT_MAX, V, S, M->partial);
        const auto e2 = mma_kkt_host(M, kt, ty, tl, tm, ts, tz, tzeta, damp);
        new_err = e2.first;
        new_inf = e2.second;
        if (new_err < old_err) {
          std::swap(V.x, V.tx);
          std::swap(V.xi, V.txi);
          std::swap(V.eta, V.teta);
          for (int i = 0; i < m; ++i) { M->y[i] = ty[i]; M->lam[i] = tl[i]; M->mu[i] = tm[i]; M->s[i] = ts[i]; }
          M->z = tz;
          M->zeta = tzeta;
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
    M->eps *= 0.1;
  }
  M->last_newton = nsub;
  S.move = M->move;
  const double* av = mma_pass(M, st, k_mma_asym, 1, 0u, V, S, M->partial);
  if (std::sqrt(av[0]) < 1e-3) M->stop_counter++;
  return M->stop_counter > 10;
}

template <class F>
int guard(Ctx* c, F&& f) {
  if (!c) return TOS_E_ARG;
  try {
    HIPCHK(hipSetDevice(c->dev));
    return f();
  } catch (const HipError& e) {
    c->err = e.what();
    return TOS_E_HIP;
  } catch (const ArgError& e) {
    c->err = e.what();
    return TOS_E_ARG;
  } catch (const std::exception& e) {
    c->err = e.what();
    return TOS_E_ARG;
  }
}

#define REQUIRE(cond, msg) \
  if (!(cond)) throw ArgError(msg)

double* rig_partials(Ctx& c) {
  if (!c.rig_part) c.rig_part = dalloc<double>(c, 6 * (size_t)kRedBlocks);
  return c.rig_part;
}

// displacementProject (optimization/projection.cpp:374-395) on a fine-level vector of a
// floating context: u -= R M R^T u over the active nodes
void remove_rigid_dev(Ctx& c, double* u) {
  Level& L0 = c.L[0];
  double* part = rig_partials(c);
  hipLaunchKernelGGL(k_rigid_moments, dim3(kRedBlocks), dim3(256), 0, c.stream, (const int4*)L0.cl_lin, L0.n_lin,
                     L0.v.NP, (const double*)u, c.rig_all, part);
  hipLaunchKernelGGL(k_fold6, dim3(1), dim3(256), 0, c.stream, (const double*)part, (int)kRedBlocks, c.scal);
  hipLaunchKernelGGL(k_rigid_sub, dim3(c.nb0), dim3(256), 0, c.stream, (const int4*)L0.cl_lin, L0.n_lin, L0.v.NP,
                     u, c.rig_all, (const double*)c.scal);
  LAUNCH_CHECK();
}

// A load region on the device (wc_* buffers): padded indices, unit normals (when given),
// and on a floating context the coordinates relative to the region's rigid-motion
// projector, which is returned
struct Region {
  int n = 0;
  const double* nrm = nullptr;
  const double* pos = nullptr;
  RigidProj P{};
};
Region upload_region(Ctx& C, int n, const int64_t* nodes, const double* normals) {
  Level& L0 = C.L[0];
  const LevelView& v = L0.v;
  const int64_t nn = (int64_t)v.nx * v.ny * v.nz;
  std::vector<int64_t> pidx(n);
  std::vector<double> nrm(normals ? 3 * (size_t)n : 0), pos(3 * (size_t)n);
  std::vector<uint8_t> seen(nn, 0);
  for (int i = 0; i < n; ++i) {
    const int64_t id = nodes[i];
    REQUIRE(id >= 0 && id < nn, "load node out of range");
    REQUIRE(L0.h_free[id], "load node is not a free node");
    REQUIRE(!seen[id], "repeated load node");
    seen[id] = 1;
    const int x = (int)(id % v.nx), y = (int)((id / v.nx) % v.ny), z = (int)(id / ((int64_t)v.nx * v.ny));
    pidx[i] = v.pnode(x, y, z);
    pos[3 * (size_t)i] = x; pos[3 * (size_t)i + 1] = y; pos[3 * (size_t)i + 2] = z;
    if (normals) {
      const double* a = normals + 3 * (size_t)i;
      const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      REQUIRE(std::isfinite(len) && len > 0, "zero or non-finite normal");
      for (int k = 0; k < 3; ++k) nrm[3 * (size_t)i + k] = a[k] / len;
    }
  }
  if (C.wc_cap < n) {
    C.wc_idx = dalloc<int64_t>(C, n);
    C.wc_nrm = dalloc<double>(C, 3 * (size_t)n);
    C.wc_fs = dalloc<double>(C, 3 * (size_t)n);
    C.wc_src = dalloc<double>(C, 3 * (size_t)n);
    C.wc_pos = dalloc<double>(C, 3 * (size_t)n);
    C.wc_cap = n;
  }
  upload(C.stream, C.wc_idx, pidx.data(), sizeof(int64_t) * n);
  Region R;
  R.n = n;
  if (normals) {
    upload(C.stream, C.wc_nrm, nrm.data(), sizeof(double) * 3 * n);
    R.nrm = C.wc_nrm;
  }
  if (C.floating) {
    std::vector<double> rel;
    R.P = rigid_projector_region(n, pos.data(), normals ? nrm.data() : nullptr, rel);
    upload(C.stream, C.wc_pos, rel.data(), sizeof(double) * 3 * n);
    R.pos = C.wc_pos;
  }
  return R;
}

// scal[S_RIG0..] = R^T g over the region (g: the candidate from u or src_nm); no-op unless
// the context is floating
void region_moments(Ctx& C, const Region& R, const double* u, const double* src_nm) {
  if (!R.pos) return;
  const unsigned nbw = (unsigned)std::min<int64_t>(kRedBlocks, (R.n + 255) / 256);
  double* part = rig_partials(C);
  hipLaunchKernelGGL(k_worst_rig_moments, dim3(nbw), dim3(256), 0, C.stream, R.n, (const int64_t*)C.wc_idx, R.nrm,
                     R.pos, C.L[0].v.NP, u, src_nm, part);
  hipLaunchKernelGGL(k_fold6, dim3(1), dim3(256), 0, C.stream, (const double*)part, (int)nbw, C.scal);
  LAUNCH_CHECK();
}

}  // namespace

// =========================================================================================
// C ABI
// =========================================================================================
TOS_EXPORT int tos_ctx_create(const tos_grid_desc* d, tos_ctx** out) {
  if (!d || !out || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || !d->solid || !(d->h > 0)) return TOS_E_ARG;
  if ((int64_t)(d->nx + 4) * (d->ny + 4) * (d->nz + 4) > ((int64_t)1 << 31)) return TOS_E_ARG;
  *out = nullptr;
  tos_ctx* c = new tos_ctx();
  try {
    HIPCHK(hipGetDevice(&c->dev));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_gal, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev_pack, hipEventDisableTiming));
    c->nx = d->nx; c->ny = d->ny; c->nz = d->nz; c->h = d->h;
    const size_t ne = (size_t)d->nx * d->ny * d->nz;
    const size_t nn = (size_t)(d->nx + 1) * (d->ny + 1) * (d->nz + 1);
    c->h_solid.assign(d->solid, d->solid + ne);
    if (d->support) c->h_support.assign(d->support, d->support + nn);
    if (d->shell) c->h_shell.assign(d->shell, d->shell + ne);
    build_levels(*c, d->n_levels, d->coarsen_threshold > 0 ? d->coarsen_threshold : 400, d->skip_layer != 0);
    c->n_active = c->L[0].n_active;
    for (size_t l = 0; l < c->L.size(); ++l) {
      Level& L = c->L[l];
      const int64_t NP = L.v.NP;
      L.fre = dalloc<uint8_t>(*c, NP);
      auto pf = pad_nodes(L.v, L.h_free);
      upload(c->stream, L.fre, pf.data(), NP);
      build_chunks(*c, L, l > 0);
      if (l > 0) {
        L.st64 = dalloc<double>(*c, 243 * std::max<int64_t>(L.NS, 1));
        L.st32 = dalloc<float>(*c, 243 * std::max<int64_t>(L.NS, 1));
        L.u64 = dalloc<double>(*c, 3 * NP);
        L.f64 = dalloc<double>(*c, 3 * NP);
        L.u32 = dalloc<float>(*c, 3 * NP);
        L.f32 = dalloc<float>(*c, 3 * NP);
      }
      if (l + 1 < c->L.size()) {
        L.r64 = dalloc<double>(*c, 3 * NP);
        L.r32 = dalloc<float>(*c, 3 * NP);
      }
      if (l == 0) {
        L.u32 = dalloc<float>(*c, 3 * NP);
        L.f32 = dalloc<float>(*c, 3 * NP);
      }
    }
    const LevelView& v0 = c->L[0].v;
    c->eact = dalloc<uint8_t>(*c, v0.NE);
    auto pe = pad_elems(v0, c->h_solid.data());
    upload(c->stream, c->eact, pe.data(), v0.NE);
    {
      // active elements whose 8 nodes are all free (Galerkin fast path)
      std::vector<uint8_t> af(v0.NE, 0);
      const Level& L0h = c->L[0];
      for (int z = 0; z < v0.ez; ++z)
        for (int y = 0; y < v0.ey; ++y)
          for (int x = 0; x < v0.ex; ++x) {
            if (!pe[v0.pelem(x, y, z)]) continue;
            bool all = true;
            for (int j = 0; j < 8 && all; ++j)
              all = L0h.h_free[(x + (j & 1)) + (size_t)v0.nx * ((y + ((j >> 1) & 1)) + (size_t)v0.ny * (z + (j >> 2)))];
            af[v0.pelem(x, y, z)] = all;
          }
      c->allfree = dalloc<uint8_t>(*c, v0.NE);
      upload(c->stream, c->allfree, af.data(), v0.NE);
      if (c->skip) {
        // which nodes of each active fine element are supported (the penalty of the
        // skip-layer Galerkin operator)
        std::vector<uint8_t> fx(v0.NE, 0);
        if (!c->h_support.empty())
          for (int z = 0; z < v0.ez; ++z)
            for (int y = 0; y < v0.ey; ++y)
              for (int x = 0; x < v0.ex; ++x) {
                if (!pe[v0.pelem(x, y, z)]) continue;
                uint8_t m = 0;
                for (int j = 0; j < 8; ++j)
                  if (c->h_support[(x + (j & 1)) + (size_t)v0.nx * ((y + ((j >> 1) & 1)) + (size_t)v0.ny * (z + (j >> 2)))])
                    m |= (uint8_t)(1u << j);
                fx[v0.pelem(x, y, z)] = m;
              }
        c->efix = dalloc<uint8_t>(*c, v0.NE);
        upload(c->stream, c->efix, fx.data(), v0.NE);
        c->Mc4 = dalloc<double>(*c, 64 * 576);
      }
    }
    c->shell = dalloc<uint8_t>(*c, v0.NE);
    auto ps = pad_elems(v0, c->h_shell.empty() ? nullptr : c->h_shell.data());
    for (size_t i = 0; i < ps.size(); ++i) ps[i] = ps[i] && pe[i];
    upload(c->stream, c->shell, ps.data(), v0.NE);
    c->rho = dalloc<float>(*c, v0.NE);
    c->rho_pre = dalloc<float>(*c, v0.NE);
    c->sens = dalloc<float>(*c, v0.NE);
    c->vsens = dalloc<float>(*c, v0.NE);
    c->tmpf = dalloc<float>(*c, v0.NE);
    c->w64 = dalloc<double>(*c, v0.NE);
    c->w32 = dalloc<float>(*c, v0.NE);
    c->Ke64 = dalloc<double>(*c, 576);
    c->Ke32 = dalloc<float>(*c, 576);
    c->Kc64 = dalloc<double>(*c, 16);
    c->Kc32 = dalloc<float>(*c, 16);
    c->Mc64 = dalloc<double>(*c, 8 * 576);
    c->r = dalloc<double>(*c, 3 * v0.NP);
    c->z = dalloc<double>(*c, 3 * v0.NP);
    c->p = dalloc<double>(*c, 3 * v0.NP);
    c->p32 = dalloc<float>(*c, 3 * v0.NP);
    c->q = dalloc<double>(*c, 3 * v0.NP);
    c->nb0 = (int)cblk(c->L[0].n_lin);
    c->partial = dalloc<double>(*c, 2 * (size_t)std::max(kRedBlocks, c->nb0) + 2 * kStage);
    c->scal = dalloc<double>(*c, S_COUNT);
    HIPCHK(hipHostMalloc((void**)&c->h_scal, S_COUNT * sizeof(double)));
    c->K = 1;
    c->F = dalloc<double>(*c, 3 * v0.NP);
    c->U = dalloc<double>(*c, 3 * v0.NP);
    if (c->L.size() > 1) coarse_alloc(*c);
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "tos_ctx_create: %s\n", e.what());
    for (void* p : c->allocs) (void)hipFree(p);
    delete c;
    return TOS_E_HIP;
  }
  *out = c;
  return tos_set_material(c, c->E, c->nu, c->penal, c->rho_min);
}

TOS_EXPORT int tos_ctx_destroy(tos_ctx* c) {
  if (!c) return TOS_E_ARG;
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (const Ctx::SolveGraph& g : c->graphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  learn_release(*c);
  batch_release(*c);
  for (void* p : c->allocs) (void)hipFree(p);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stage) (void)hipFree(c->stage);
  if (c->h_scal) (void)hipHostFree(c->h_scal);
  if (c->ev_gal) (void)hipEventDestroy(c->ev_gal);
  if (c->ev_pack) (void)hipEventDestroy(c->ev_pack);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return TOS_OK;
}

TOS_EXPORT const char* tos_last_error(const tos_ctx* c) { return c ? c->err.c_str() : "null context"; }

TOS_EXPORT int tos_get_info(tos_ctx* c, tos_info* o) {
  if (!c || !o) return TOS_E_ARG;
  std::memset(o, 0, sizeof(*o));
  o->n_levels = (int)c->L.size();
  for (size_t l = 0; l < c->L.size() && l < 16; ++l) {
    o->level_dims[l][0] = c->L[l].v.ex;
    o->level_dims[l][1] = c->L[l].v.ey;
    o->level_dims[l][2] = c->L[l].v.ez;
    o->n_active_elements[l] = c->L[l].n_active;
    o->n_free_nodes[l] = c->L[l].n_free;
  }
  o->n_elements = (int64_t)c->nx * c->ny * c->nz;
  o->n_nodes = (int64_t)(c->nx + 1) * (c->ny + 1) * (c->nz + 1);
  o->coarse_dofs = c->cn;
  o->n_cases = c->K;
  o->n_coeffs = (int64_t)c->nc[0] * c->nc[1] * c->nc[2];
  o->has_support = c->floating ? 0 : 1;
  return TOS_OK;
}

TOS_EXPORT int tos_get_node_flags(tos_ctx* c, int level, uint8_t* out) {
  return guard(c, [&]() {
    REQUIRE(out && level >= 0 && level < (int)c->L.size(), "bad level");
    const Level& L = c->L[level];
    const LevelView& v = L.v;
    for (int z = 0; z < v.nz; ++z)
      for (int y = 0; y < v.ny; ++y)
        for (int x = 0; x < v.nx; ++x) {
          bool act = false;
          for (int a = 0; a < 8 && !act; ++a) {
            const int ex = x - 1 + (a & 1), ey = y - 1 + ((a >> 1) & 1), ez = z - 1 + (a >> 2);
            if (ex < 0 || ey < 0 || ez < 0 || ex >= v.ex || ey >= v.ey || ez >= v.ez) continue;
            act = L.h_eact[ex + (size_t)v.ex * (ey + (size_t)v.ey * ez)];
          }
          const size_t n = x + (size_t)v.nx * (y + (size_t)v.ny * z);
          const bool sup = level == 0 && !c->h_support.empty() && c->h_support[n];
          const int col = (x & 1) + 2 * (y & 1) + 4 * (z & 1);
          out[n] = (uint8_t)((act ? 1 : 0) | (sup ? 2 : 0) | (L.h_free[n] ? 4 : 0) | (col << 3));
        }
    return TOS_OK;
  });
}

TOS_EXPORT int tos_set_material(tos_ctx* c, double E, double nu, double p, double rmin) {
  return guard(c, [&]() {
    REQUIRE(E > 0 && nu > -1 && nu < 0.5 && p >= 1 && rmin > 0, "bad material");
    c->E = E; c->nu = nu; c->penal = p; c->rho_min = rmin;
    hex_stiffness(c->h, E, nu, c->h_Ke);
    child_products(c->h_Ke, c->h_Mc);
    float k32[576];
    for (int i = 0; i < 576; ++i) k32[i] = (float)c->h_Ke[i];
    upload(c->stream, c->Ke64, c->h_Ke, sizeof(c->h_Ke));
    upload(c->stream, c->Ke32, k32, sizeof(k32));
    // distinct values of the first block row; every entry of Ke must follow from them by
    // the reflection / axis symmetry the matrix-free kernels rely on
    double kc[16] = {0};
    double kmax = 0;
    for (int i = 0; i < 576; ++i) kmax = std::max(kmax, std::fabs(c->h_Ke[i]));
    for (int d = 0; d < 8; ++d)
      for (int r = 0; r < 3; ++r)
        for (int s2 = 0; s2 < 3; ++s2) kc[ke_key(d, r, s2)] = c->h_Ke[r * 24 + 3 * d + s2];
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b)
        for (int r = 0; r < 3; ++r)
          for (int s2 = 0; s2 < 3; ++s2) {
            const double sg = (((a >> r) & 1) ? -1.0 : 1.0) * (((a >> s2) & 1) ? -1.0 : 1.0);
            if (std::fabs(sg * kc[ke_key(a ^ b, r, s2)] - c->h_Ke[(3 * a + r) * 24 + 3 * b + s2]) > 1e-12 * kmax)
              throw std::runtime_error("element matrix lacks the cube symmetry of the hex element");
          }
    float kc32[16];
    for (int i = 0; i < 16; ++i) kc32[i] = (float)kc[i];
    upload(c->stream, c->Kc64, kc, sizeof(kc));
    upload(c->stream, c->Kc32, kc32, sizeof(kc32));
    upload(c->stream, c->Mc64, c->h_Mc, sizeof(c->h_Mc));
    if (c->skip) {
      std::vector<double> m4(64 * 576);
      child_products4(c->h_Ke, m4.data());
      upload(c->stream, c->Mc4, m4.data(), m4.size() * sizeof(double));
    }
    c->op_ready = false;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_get_template_matrix(tos_ctx* c, double* ke) {
  if (!c || !ke) return TOS_E_ARG;
  std::memcpy(ke, c->h_Ke, sizeof(c->h_Ke));
  return TOS_OK;
}

TOS_EXPORT int tos_set_loads(tos_ctx* c, int K, const double* f) {
  return guard(c, [&]() {
    REQUIRE(f && K > 0 && K <= 64, "bad loads");
    Level& L0 = c->L[0];
    const int64_t n = (int64_t)L0.v.nx * L0.v.ny * L0.v.nz, n3p = 3 * L0.v.NP;
    if (K != c->K) {
      // grow F, U (old buffers stay in the allocation list until destroy)
      c->F = dalloc<double>(*c, (size_t)K * n3p);
      c->U = dalloc<double>(*c, (size_t)K * n3p);
      c->K = K;
    }
    double* st = (double*)stage(*c, 3 * n * sizeof(double));
    for (int k = 0; k < K; ++k) {
      for (int64_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(f[(size_t)k * 3 * n + i])) throw ArgError("non-finite load");
      upload(c->stream, st, f + (size_t)k * 3 * n, 3 * n * sizeof(double));
      hipLaunchKernelGGL(k_pad_vec, dim3(nblk(n)), dim3(256), 0, c->stream, L0.v, (const double*)st,
                         c->F + (size_t)k * n3p, (const uint8_t*)L0.fre);
      hipLaunchKernelGGL(k_zero<double>, dim3(nblk(n3p)), dim3(256), 0, c->stream, c->U + (size_t)k * n3p, n3p);
    }
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

static int get_vec(tos_ctx* c, const double* dsrc, double* out) {
  Level& L0 = c->L[0];
  const int64_t n = (int64_t)L0.v.nx * L0.v.ny * L0.v.nz;
  double* st = (double*)stage(*c, 3 * n * sizeof(double));
  hipLaunchKernelGGL(k_unpad_vec, dim3(nblk(n)), dim3(256), 0, c->stream, L0.v, dsrc, st);
  LAUNCH_CHECK();
  download(c->stream, out, st, 3 * n * sizeof(double));
  return TOS_OK;
}

TOS_EXPORT int tos_get_loads(tos_ctx* c, int k, double* f) {
  return guard(c, [&]() {
    REQUIRE(f && k >= 0 && k < c->K, "bad case");
    return get_vec(c, c->F + (size_t)k * 3 * c->L[0].v.NP, f);
  });
}

TOS_EXPORT int tos_set_density(tos_ctx* c, const float* rho) {
  return guard(c, [&]() {
    REQUIRE(rho, "null density");
    const LevelView& v = c->L[0].v;
    const int64_t n = (int64_t)v.ex * v.ey * v.ez;
    for (int64_t i = 0; i < n; ++i)
      if (!std::isfinite(rho[i])) {
        c->err = "non-finite density";
        return TOS_E_NAN;
      }
    float* st = (float*)stage(*c, n * sizeof(float));
    upload(c->stream, st, rho, n * sizeof(float));
    hipLaunchKernelGGL(k_pad_elems, dim3(nblk(n)), dim3(256), 0, c->stream, v, (const float*)st, c->rho);
    hipLaunchKernelGGL(k_mask_inactive, dim3(nblk(v.NE)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact, c->rho);
    LAUNCH_CHECK();
    c->op_ready = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_get_density(tos_ctx* c, float* rho) {
  return guard(c, [&]() {
    REQUIRE(rho, "null density");
    const LevelView& v = c->L[0].v;
    const int64_t n = (int64_t)v.ex * v.ey * v.ez;
    float* st = (float*)stage(*c, n * sizeof(float));
    hipLaunchKernelGGL(k_unpad_elems, dim3(nblk(n)), dim3(256), 0, c->stream, v, (const float*)c->rho, st);
    LAUNCH_CHECK();
    download(c->stream, rho, st, n * sizeof(float));
    return TOS_OK;
  });
}

// update_stencil (grid/Grid.cpp:4167-4178): weights, Galerkin levels, coarsest factor
// coarse stencil of level C from level F (>= 1) in three axis passes through two boxes:
// (Xc, Yf, Zf) then (Xc, Yc, Zf), stored as TS (FP64: also the FP32 copy when copy32)
template <class TS>
void galerkin_by_axes(Ctx& c, Level& F, Level& C, bool copy32) {
  const GBox bf{F.v.nx, F.v.ny, F.v.nz};
  const GBox b1{C.v.nx, F.v.ny, F.v.nz}, b2{C.v.nx, C.v.ny, F.v.nz};
  const int64_t n1 = (int64_t)b1.X * b1.Y * b1.Z, n2 = (int64_t)b2.X * b2.Y * b2.Z;
  if (!C.gal_b1) {
    C.gal_b1 = dalloc<double>(c, 243 * n1);
    C.gal_b2 = dalloc<double>(c, 243 * n2);
    // b2 (X, Y, z) is read for coarse free nodes (X, Y, Z), |z - 2Z| <= 1; b1 (X, y, z) for
    // read b2 nodes (X, Y, z), |y - 2Y| <= 1
    std::vector<uint8_t> m2((size_t)n2, 0), m1((size_t)n1, 0);
    const LevelView& cv = C.v;
    for (int Z = 0; Z < cv.nz; ++Z)
      for (int Y = 0; Y < cv.ny; ++Y)
        for (int X = 0; X < cv.nx; ++X) {
          if (!C.h_free[X + (size_t)cv.nx * (Y + (size_t)cv.ny * Z)]) continue;
          for (int z = 2 * Z - 1; z <= 2 * Z + 1; ++z)
            if (z >= 0 && z < b2.Z) m2[X + (size_t)b2.X * (Y + (size_t)b2.Y * z)] = 1;
        }
    for (int z = 0; z < b2.Z; ++z)
      for (int Y = 0; Y < b2.Y; ++Y)
        for (int X = 0; X < b2.X; ++X) {
          if (!m2[X + (size_t)b2.X * (Y + (size_t)b2.Y * z)]) continue;
          for (int y = 2 * Y - 1; y <= 2 * Y + 1; ++y)
            if (y >= 0 && y < b1.Y) m1[X + (size_t)b1.X * (y + (size_t)b1.Y * z)] = 1;
        }
    C.gal_m1 = dalloc<uint8_t>(c, n1);
    C.gal_m2 = dalloc<uint8_t>(c, n2);
    upload(c.stream, C.gal_m1, m1.data(), n1);
    upload(c.stream, C.gal_m2, m2.data(), n2);
  }
  constexpr bool d64 = std::is_same<TS, double>::value;
  const TS* fin = d64 ? (const TS*)F.st64 : (const TS*)F.st32;
  TS* outp = d64 ? (TS*)C.st64 : (TS*)C.st32;
  TS* b1p = (TS*)C.gal_b1;  // (sized for FP64 entries)
  TS* b2p = (TS*)C.gal_b2;
  hipLaunchKernelGGL((k_gal_axis<0, true, false, TS>), dim3(nblk(n1), 27), dim3(256), 0, c.stream, F.v,
                     (const int*)F.slot, fin, F.NS, bf, b1, C.v, (const int4*)nullptr, 0, (const uint8_t*)nullptr,
                     b1p, (float*)nullptr, n1, (const uint8_t*)C.gal_m1);
  hipLaunchKernelGGL((k_gal_axis<1, false, false, TS>), dim3(nblk(n2), 27), dim3(256), 0, c.stream, F.v,
                     (const int*)nullptr, (const TS*)b1p, n1, b1, b2, C.v, (const int4*)nullptr, 0,
                     (const uint8_t*)nullptr, b2p, (float*)nullptr, n2, (const uint8_t*)C.gal_m2);
  hipLaunchKernelGGL((k_gal_axis<2, false, true, TS>), dim3(cblk(C.n_col), 27), dim3(256), 0, c.stream, F.v,
                     (const int*)nullptr, (const TS*)b2p, n2, b2, GBox{C.v.nx, C.v.ny, C.v.nz}, C.v,
                     (const int4*)C.cl_col, C.n_col, (const uint8_t*)C.fre, outp,
                     (d64 && copy32) ? C.st32 : (float*)nullptr, C.NS, (const uint8_t*)nullptr);
  LAUNCH_CHECK();
}

// Galerkin stencils of every level >= 1 stored as TS; an FP64 build also writes the FP32
// copies when copy32.  Level 1 comes from the element matrices (FP64 sums either way).
template <class TS>
void galerkin_levels(Ctx& c, bool copy32) {
  constexpr bool d64 = std::is_same<TS, double>::value;
  for (size_t l = 1; l < c.L.size(); ++l) {
    Level& F = c.L[l - 1];
    Level& C = c.L[l];
    if (C.n_col == 0) continue;
    if (l == 1 && c.skip)
      hipLaunchKernelGGL(k_galerkin_skip, dim3(cblk(C.n_col), 3), dim3(256), 0, c.stream, F.v, C.v,
                         (const int4*)C.cl_col, C.n_col, (const double*)c.w64, (const uint8_t*)c.efix,
                         (const double*)c.Ke64, (const double*)c.Mc4, d64 ? C.st64 : (double*)nullptr,
                         (!d64 || copy32) ? C.st32 : (float*)nullptr, C.NS);
    else if (l == 1)
      hipLaunchKernelGGL(k_galerkin_first_row, dim3(cblk(C.n_col), 9), dim3(256), 0, c.stream, F.v, C.v,
                         (const int4*)C.cl_col, C.n_col, (const uint8_t*)F.fre, (const uint8_t*)c.allfree,
                         (const uint8_t*)C.fre, (const double*)c.w64, (const double*)c.Ke64,
                         (const double*)c.Mc64, d64 ? C.st64 : (double*)nullptr,
                         (!d64 || copy32) ? C.st32 : (float*)nullptr, C.NS);
    else
      galerkin_by_axes<TS>(c, F, C, copy32);
    LAUNCH_CHECK();
  }
}

namespace {
// Galerkin precision 1: the FP64 stencils and their coarsest factor, built when first
// needed after an operator update (the FP32 stencils and their factor stay as built)
int ensure_fp64_operator(Ctx& c) {
  if (c.st64_valid) return TOS_OK;
  {
    Prof pr(c, P_GALERKIN);
    galerkin_levels<double>(c, false);
  }
  const int rc = coarse_factor(c, false);
  if (rc) return rc;
  c.st64_valid = true;
  return TOS_OK;
}
}  // namespace

TOS_EXPORT int tos_update_operator(tos_ctx* c) {
  return guard(c, [&]() {
    const LevelView& v0 = c->L[0].v;
    c->st16_valid = false;
    c->st64_valid = false;
    c->op_ready = false;
    {
      Prof pr(*c, P_GALERKIN);
      hipLaunchKernelGGL(k_weights, dim3(nblk(v0.NE)), dim3(256), 0, c->stream, v0, (const uint8_t*)c->eact,
                         (const float*)c->rho, c->penal, c->w64, c->w32);
      if (c->gal_prec == 0)
        galerkin_levels<double>(*c, true);
      else
        galerkin_levels<float>(*c, false);
    }
    // FP16 copies (read by the next solve when the last one read them) packed on the
    // second stream while the coarsest level is factored: they read the stencils the
    // factor reads and write nothing it touches
    const bool beside = pack_beside_ok(*c);
    if (beside) {
      HIPCHK(hipEventRecord(c->ev_gal, c->stream));
      HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_gal, 0));
      pack_half_stencils(*c, c->stream2);
      c->st16_valid = false;  // (until the library stream has waited for them)
      HIPCHK(hipEventRecord(c->ev_pack, c->stream2));
    }
    const int rc = coarse_factor(*c, c->gal_prec != 0);
    if (beside) {
      HIPCHK(hipStreamWaitEvent(c->stream, c->ev_pack, 0));
      c->st16_valid = true;
    }
    if (rc) return rc;
    c->st64_valid = c->gal_prec == 0;
    if (c->learn && c->learn->lrho) learn_logrho(*c);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->op_ready = true;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_get_stencil(tos_ctx* c, int level, double* out) {
  return guard(c, [&]() {
    REQUIRE(out && level >= 1 && level < (int)c->L.size(), "bad level");
    if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
    const int rc = ensure_fp64_operator(*c);
    if (rc) return rc;
    const Level& L = c->L[level];
    const LevelView& v = L.v;
    std::vector<double> h((size_t)243 * std::max<int64_t>(L.NS, 1));
    download(c->stream, h.data(), L.st64, h.size() * sizeof(double));
    const size_t nn = (size_t)v.nx * v.ny * v.nz;
    for (int z = 0; z < v.nz; ++z)
      for (int y = 0; y < v.ny; ++y)
        for (int x = 0; x < v.nx; ++x) {
          const size_t node = x + (size_t)v.nx * (y + (size_t)v.ny * z);
          const int sl = L.h_slot[v.pnode(x, y, z)];
          for (int e = 0; e < 243; ++e) out[e * nn + node] = sl < 0 ? 0.0 : h[(size_t)e * L.NS + sl];
        }
    return TOS_OK;
  });
}

TOS_EXPORT int tos_apply_operator(tos_ctx* c, const double* u, double* y) {
  return guard(c, [&]() {
    REQUIRE(u && y, "null vector");
    if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
    Level& L0 = c->L[0];
    const int64_t n = (int64_t)L0.v.nx * L0.v.ny * L0.v.nz;
    double* st = (double*)stage(*c, 3 * n * sizeof(double));
    upload(c->stream, st, u, 3 * n * sizeof(double));
    hipLaunchKernelGGL(k_pad_vec, dim3(nblk(n)), dim3(256), 0, c->stream, L0.v, (const double*)st, c->p,
                       (const uint8_t*)L0.fre);
    apply_level<double, 0>(*c, 0, c->p, nullptr, c->q);
    return get_vec(c, c->q, y);
  });
}

TOS_EXPORT int tos_vcycle(tos_ctx* c, const double* r, double* z, int precision, int pre, int post) {
  return guard(c, [&]() {
    REQUIRE(r && z && pre >= 0 && post >= 0 && precision >= 0 && precision <= 2, "bad arguments");
    if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
    const int rs = select_stencils(*c, precision);
    if (rs) return rs;
    Level& L0 = c->L[0];
    const int64_t n = (int64_t)L0.v.nx * L0.v.ny * L0.v.nz;
    double* st = (double*)stage(*c, 3 * n * sizeof(double));
    upload(c->stream, st, r, 3 * n * sizeof(double));
    hipLaunchKernelGGL(k_pad_vec, dim3(nblk(n)), dim3(256), 0, c->stream, L0.v, (const double*)st, c->r,
                       (const uint8_t*)L0.fre);
    precondition(*c, precision, pre, post, false, true);
    return get_vec(c, c->z, z);
  });
}

// ---- recycled start -----------------------------------------------------------------
// A sequence of solves whose operator changes a little per step (an optimisation loop)
// leaves earlier solutions that span most of the next one.  With depth m >= 1 the warm
// start u is replaced by the K-orthogonal (Galerkin) projection of the solution onto
// span{u, s_2, ..., s_m}, s_i the i-th newest stored solution of the case (the newest,
// s_1, is the warm start itself unless the displacement was set since): u0 = V c with
// (V^T K V) c = V^T f, i.e. the point of that span with the smallest energy error, never
// worse than u in the energy norm.  The Gram system is scaled to unit diagonal and solved
// through its eigenvalues, dropping those below 1e-10 of the largest (nearly dependent
// solutions).  No reference counterpart.
namespace {
bool rc_coefficients(int m, const double* A, const double* b, double* cf) {
  double d[kRcMax], S[kRcMax][kRcMax], Q[kRcMax][kRcMax], bs[kRcMax];
  for (int i = 0; i < m; ++i) d[i] = A[i * m + i] > 0 ? 1.0 / std::sqrt(A[i * m + i]) : 0.0;
  for (int i = 0; i < m; ++i) {
    bs[i] = b[i] * d[i];
    for (int j = 0; j < m; ++j) {
      S[i][j] = A[i * m + j] * d[i] * d[j];
      Q[i][j] = i == j ? 1.0 : 0.0;
    }
  }
  for (int sweep = 0; sweep < 60; ++sweep) {  // cyclic Jacobi
    double off = 0, dia = 0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) (i == j ? dia : off) += S[i][j] * S[i][j];
    if (off <= 1e-30 * dia) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        if (S[p][q] == 0) continue;
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
  double ch[kRcMax] = {0};
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

// ring storage for c.K cases (allocated on the first solve after the depth is set or the
// case count changed)
void rc_reserve(Ctx& c) {
  if (c.rc_ring && c.rc_K == c.K) return;
  const int64_t n3 = 3 * c.L[0].v.NP;
  c.rc_ring = dalloc<double>(c, (size_t)c.K * c.rc_depth * n3);
  c.rc_kv = dalloc<double>(c, (size_t)c.rc_depth * n3);
  constexpr int ncol = kRcMax * (kRcMax + 1) / 2 + kRcMax;
  c.rc_part = dalloc<double>(c, (size_t)ncol * kRedBlocks);
  c.rc_out = dalloc<double>(c, ncol);
  c.rc_K = c.K;
  c.rc_head.assign(c.K, 0);
  c.rc_cnt.assign(c.K, 0);
}

double* rc_slot(Ctx& c, int k, int i) {  // i-th newest stored solution of case k (i = 1 newest)
  const int64_t n3 = 3 * c.L[0].v.NP;
  const int s = (c.rc_head[k] - (i - 1) + c.rc_depth) % c.rc_depth;
  return c.rc_ring + ((size_t)k * c.rc_depth + s) * n3;
}

template <int M>
void rc_launch(Ctx& c, const RcVecs& V, const double* f, unsigned nbg) {
  Level& L0 = c.L[0];
  hipLaunchKernelGGL((k_rc_gram<M>), dim3(nbg), dim3(256), 0, c.stream, (const int4*)L0.cl_lin, L0.n_lin, L0.v.NP,
                     V, f, c.rc_part);
}

template <int M>
void rc_combine(Ctx& c, const RcVecs& V, const RcCoef& C, const double* f, double* u, float* r32) {
  Level& L0 = c.L[0];
  hipLaunchKernelGGL((k_rc_combine<M>), dim3(cblk(L0.n_lin)), dim3(256), 0, c.stream, (const int4*)L0.cl_lin,
                     L0.n_lin, L0.v.NP, V, C, f, u, c.r, r32);
  c.r_ready = true;
}

// u_k <- the projected start (before the solve of case k)
void rc_start(Ctx& c, int k, bool f32) {
  rc_reserve(c);
  float* r32 = f32 ? c.L[0].f32 : nullptr;
  Level& L0 = c.L[0];
  const int64_t n3 = 3 * L0.v.NP;
  double* u = c.U + (size_t)k * n3;
  const double* f = c.F + (size_t)k * n3;
  RcVecs V{};
  int m = 0;
  V.v[m++] = u;
  for (int i = 2; i <= c.rc_cnt[k] && m < c.rc_depth; ++i) V.v[m++] = rc_slot(c, k, i);
  for (int a = 0; a < m; ++a) {
    V.kv[a] = c.rc_kv + (size_t)a * n3;
    apply_level<double, 0>(c, 0, V.v[a], (const double*)nullptr, const_cast<double*>(V.kv[a]));
  }
  const int ncol = m * (m + 1) / 2 + m;
  {
    Prof pr(c, P_VECTOR);
    const unsigned nbg = (unsigned)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (L0.n_lin + 3) / 4));
    switch (m) {
      case 1: rc_launch<1>(c, V, f, nbg); break;
      case 2: rc_launch<2>(c, V, f, nbg); break;
      case 3: rc_launch<3>(c, V, f, nbg); break;
      case 4: rc_launch<4>(c, V, f, nbg); break;
      case 5: rc_launch<5>(c, V, f, nbg); break;
      case 6: rc_launch<6>(c, V, f, nbg); break;
      case 7: rc_launch<7>(c, V, f, nbg); break;
      default: rc_launch<8>(c, V, f, nbg); break;
    }
    hipLaunchKernelGGL(k_sf_fold, dim3(1), dim3(256), 0, c.stream, (const double*)c.rc_part, (int)nbg, ncol, 0,
                       c.rc_out);
    LAUNCH_CHECK();
  }
  double g[kRcMax * (kRcMax + 1) / 2 + kRcMax];
  download(c.stream, g, c.rc_out, ncol * sizeof(double));
  double A[kRcMax * kRcMax], b[kRcMax];
  int q = 0;
  for (int a = 0; a < m; ++a)
    for (int e = a; e < m; ++e) A[a * m + e] = A[e * m + a] = g[q++];
  for (int a = 0; a < m; ++a) b[a] = g[q++];
  RcCoef C{};
  if (!rc_coefficients(m, A, b, C.c)) {  // keep the warm start (u = 1 u exactly, r = f - K u)
    for (int a = 0; a < m; ++a) C.c[a] = a == 0 ? 1.0 : 0.0;
  }
  Prof pr(c, P_VECTOR);
  switch (m) {
    case 1: rc_combine<1>(c, V, C, f, u, r32); break;
    case 2: rc_combine<2>(c, V, C, f, u, r32); break;
    case 3: rc_combine<3>(c, V, C, f, u, r32); break;
    case 4: rc_combine<4>(c, V, C, f, u, r32); break;
    case 5: rc_combine<5>(c, V, C, f, u, r32); break;
    case 6: rc_combine<6>(c, V, C, f, u, r32); break;
    case 7: rc_combine<7>(c, V, C, f, u, r32); break;
    default: rc_combine<8>(c, V, C, f, u, r32); break;
  }
  LAUNCH_CHECK();
}

// store the solution of case k as its newest
void rc_push(Ctx& c, int k) {
  rc_reserve(c);
  Level& L0 = c.L[0];
  const int64_t n3 = 3 * L0.v.NP;
  c.rc_head[k] = (c.rc_head[k] + 1) % c.rc_depth;
  c.rc_cnt[k] = std::min(c.rc_cnt[k] + 1, c.rc_depth);
  if (L0.n_lin)
    hipLaunchKernelGGL(k_copy_nodes, dim3(cblk(L0.n_lin)), dim3(256), 0, c.stream, (const int4*)L0.cl_lin,
                       L0.n_lin, L0.v.NP, (const double*)(c.U + (size_t)k * n3), rc_slot(c, k, 1));
  LAUNCH_CHECK();
}
}  // namespace

TOS_EXPORT int tos_set_start_recycling(tos_ctx* c, int depth) {
  return guard(c, [&]() {
    REQUIRE(depth >= 0 && depth <= kRcMax, "bad recycling depth");
    if (depth != c->rc_depth) {  // a new ring (the old buffers stay with the context)
      c->rc_depth = depth;
      c->rc_ring = c->rc_kv = c->rc_part = c->rc_out = nullptr;
      c->rc_K = 0;
    }
    return TOS_OK;
  });
}

TOS_EXPORT int tos_solve(tos_ctx* c, const tos_solve_opts* o, tos_solve_stats* st) {
  return guard(c, [&]() {
    REQUIRE(o && o->rtol > 0 && o->maxit >= 0 && o->pre >= 0 && o->post >= 0 && o->precision >= 0 &&
                o->precision <= 3 && (o->precision < 3 || o->use_pcg),
            "bad options");
    if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
    REQUIRE(o->precision < 3 || !c->learn || (c->learn->mode == 0 && !c->learn->ds),
            "precision 3 takes no learned correction");
    const int rs = select_stencils(*c, o->precision == 3 ? 2 : o->precision);
    if (rs) return rs;
    const auto t0 = std::chrono::steady_clock::now();
    int rc_all = TOS_OK;
    for (int k = 0; k < c->K; ++k) {
      int its = 0;
      double rr = 0;
      c->r_ready = false;
      if (c->rc_depth > 0 && o->warm_start) rc_start(*c, k, o->precision != 0);
      const int rc = o->precision == 3 ? solve_case_r32(*c, k, *o, &its, &rr) : solve_case(*c, k, *o, &its, &rr);
      if (st && k < 64) { st->iterations[k] = its; st->rel_res[k] = rr; }
      if (rc == TOS_E_NAN) return rc;
      if (rc) rc_all = rc;
      if (c->rc_depth > 0 && rc == TOS_OK) rc_push(*c, k);
    }
    if (rc_all == TOS_E_NOCONV) c->err = "solver did not reach rtol";
    if (st) st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc_all;
  });
}

TOS_EXPORT int tos_solve_batched(tos_ctx* c, const tos_solve_opts* o, tos_solve_stats* st) {
  return guard(c, [&]() {
    REQUIRE(o && o->rtol > 0 && o->maxit >= 0 && o->pre >= 0 && o->post >= 0, "bad options");
    REQUIRE(o->precision == 1 && o->use_pcg, "the batched solve is FCG with an FP32 V-cycle");
    REQUIRE(!c->skip, "the batched solve has no skip-layer transfers");
    if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = solve_batched(*c, *o, st);
    if (rc == TOS_E_NOCONV) c->err = "solver did not reach rtol";
    if (st) st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
  });
}

TOS_EXPORT int tos_batch_tuning(tos_ctx* c, char* buf, int len) {
  return guard(c, [&]() {
    REQUIRE(buf && len > 0, "bad buffer");
    std::string s;
    if (c->batch && c->batch->tuned) {
      const BatchState& B = *c->batch;
      s = "K=" + std::to_string(B.K) + " fine_gs=" + std::to_string(B.bs_fine_gs) +
          " fine_res=" + std::to_string(B.bs_fine_res) + " transfer=" + std::to_string(B.bs_transfer);
      for (size_t l = 1; l + 1 < c->L.size(); ++l)
        s += " L" + std::to_string(l) + ":gs=" + std::to_string(B.bs_st_gs[l]) + ",res=" + std::to_string(B.bs_st_res[l]);
    }
    std::snprintf(buf, (size_t)len, "%s", s.c_str());
    return TOS_OK;
  });
}

TOS_EXPORT int tos_get_displacement(tos_ctx* c, int k, double* u) {
  return guard(c, [&]() {
    REQUIRE(u && k >= 0 && k < c->K, "bad case");
    return get_vec(c, c->U + (size_t)k * 3 * c->L[0].v.NP, u);
  });
}

TOS_EXPORT int tos_set_displacement(tos_ctx* c, int k, const double* u) {
  return guard(c, [&]() {
    REQUIRE(u && k >= 0 && k < c->K, "bad case");
    Level& L0 = c->L[0];
    const int64_t n = (int64_t)L0.v.nx * L0.v.ny * L0.v.nz;
    double* st = (double*)stage(*c, 3 * n * sizeof(double));
    upload(c->stream, st, u, 3 * n * sizeof(double));
    hipLaunchKernelGGL(k_pad_vec, dim3(nblk(n)), dim3(256), 0, c->stream, L0.v, (const double*)st,
                       c->U + (size_t)k * 3 * L0.v.NP, (const uint8_t*)L0.fre);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_compliance(tos_ctx* c, double* out) {
  return guard(c, [&]() {
    REQUIRE(out, "null output");
    const int64_t n3 = 3 * c->L[0].v.NP;
    for (int k = 0; k < c->K; ++k) {
      dot_to(*c, c->F + (size_t)k * n3, c->U + (size_t)k * n3, nullptr, nullptr, S_TMP0, 0, OP_NONE);
      out[k] = read_scal(*c)[S_TMP0];
    }
    return TOS_OK;
  });
}

TOS_EXPORT int tos_volume(tos_ctx* c, double* v) {
  return guard(c, [&]() {
    REQUIRE(v, "null output");
    *v = volume_of(*c);
    return TOS_OK;
  });
}

// splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014): the start force of the worst-case load search, mapped to [-1, 1)
static double splitmix_pm1(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

// Modified power method for the worst-case load (modifiedPM, optimization/optimization.cpp:
// 213-351).  Device-resident loop: per step one V-cycle, the residual r = f - K u with its
// squared norm, and the force update on the load nodes (two small passes, which also move
// r to the new force, so the next cycle needs no extra operator application); one
// readback of the scalars per step for the stop test, as the reference's host loop.
TOS_EXPORT int tos_worst_compliance(tos_ctx* c, const tos_worst_opts* o, tos_worst_stats* st) {
  return guard(c, [&]() {
    REQUIRE(o && o->n_load > 0 && o->nodes && o->max_iter >= 0 && o->pre >= 0 && o->post >= 0 &&
                o->precision >= 0 && o->precision <= 2 && o->fch_tol >= 0 && o->res_tol >= 0,
            "bad options");
    if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
    Ctx& C = *c;
    Level& L0 = C.L[0];
    const LevelView& v = L0.v;
    const int64_t NP = v.NP, n3p = 3 * NP;
    const int n = o->n_load;
    const Region RG = upload_region(C, n, o->nodes, o->normals);
    std::vector<double> f0(3 * (size_t)n);
    uint64_t seed = o->seed;
    for (double& x : f0) x = splitmix_pm1(seed);
    upload(C.stream, C.wc_src, f0.data(), sizeof(double) * 3 * n);
    const double* nrm_d = RG.nrm;
    const int rs = select_stencils(C, o->precision);
    if (rs) return rs;
    const auto t0 = std::chrono::steady_clock::now();
    const bool f32 = o->precision != 0;
    float* r32 = f32 ? L0.f32 : nullptr;
    double* u = C.U;  // load case 0
    double* f = C.F;
    const unsigned nbw = (unsigned)std::min<int64_t>(kRedBlocks, (n + 255) / 256);
    // scal[S_TMP1] = |g|^2, f = g / |g| on the load nodes, scal[S_TMP2] = |f_new - f_old|^2,
    // scal[S_TMP3] = |f_old|^2
    auto force_update = [&](const double* src_nm, int first) {
      Prof pr(C, P_VECTOR);
      region_moments(C, RG, u, src_nm);  // floating: R^T g of the candidate
      hipLaunchKernelGGL(k_worst_norm, dim3(nbw), dim3(256), 0, C.stream, n, (const int64_t*)C.wc_idx, nrm_d, NP,
                         (const double*)u, src_nm, RG.pos, RG.P, (const double*)C.scal, C.partial);
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, C.stream, (const double*)C.partial, (int)nbw, 0, S_TMP1,
                         0, OP_NONE, C.scal);
      hipLaunchKernelGGL(k_worst_apply, dim3(nbw), dim3(256), 0, C.stream, n, (const int64_t*)C.wc_idx, nrm_d, NP,
                         (const double*)u, src_nm, RG.pos, RG.P, (const double*)C.scal, (int)S_TMP1, first, f, C.r,
                         r32, C.wc_fs, C.partial);
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, C.stream, (const double*)C.partial, (int)nbw, 1, S_TMP2,
                         S_TMP3, OP_NONE, C.scal);
      LAUNCH_CHECK();
    };
    hipLaunchKernelGGL(k_zero<double>, dim3(nblk(n3p)), dim3(256), 0, C.stream, f, n3p);
    hipLaunchKernelGGL(k_zero<double>, dim3(nblk(n3p)), dim3(256), 0, C.stream, u, n3p);
    force_update(C.wc_src, 1);
    const double* h = read_scal(C);
    if (!(h[S_TMP1] > 0) || !std::isfinite(h[S_TMP1])) {
      C.err = "the start force is zero after the projection";
      return TOS_E_ARG;
    }
    apply_level<double, 1>(C, 0, u, f, C.r, nullptr, r32);  // r = f - K u
    int it = 0;
    double fch = 1, rel = 1;
    while (it < o->max_iter && (fch > o->fch_tol || rel > o->res_tol)) {
      // u += B r (an FP32 cycle's output added directly), r = f - K u with |r|^2 and |f|^2
      // (u stays FP64: an FP32 copy of u in the residual would add an error of the order
      // of 1e-7 |K| |u|, above the stop tolerance on high-contrast designs)
      precondition(C, o->precision, o->pre, o->post, f32, false);
      if (f32)
        hipLaunchKernelGGL(k_add_nodes<float>, dim3(C.nb0), dim3(256), 0, C.stream, (const int4*)L0.cl_lin,
                           L0.n_lin, NP, u, (const float*)L0.u32);
      else
        hipLaunchKernelGGL(k_add_nodes<double>, dim3(C.nb0), dim3(256), 0, C.stream, (const int4*)L0.cl_lin,
                           L0.n_lin, NP, u, (const double*)C.z);
      if (C.floating) remove_rigid_dev(C, u);  // optimization/optimization.cpp:282-285
      apply_level<double, 1>(C, 0, u, f, C.r, C.partial, r32);
      {
        Prof pr(C, P_VECTOR);
        finish0(C, 1, S_RN2, S_NF2, OP_NONE);
      }
      force_update(nullptr, 0);
      h = read_scal(C);
      ++it;
      rel = std::sqrt(h[S_RN2] / h[S_NF2]);
      if (!std::isfinite(rel)) { C.err = "NaN/Inf in the power iteration"; return TOS_E_NAN; }
      if (rel > 1e4) { C.err = "power iteration diverged (cycle residual above 1e4)"; return TOS_E_NOCONV; }
      if (!(h[S_TMP1] > 0) || !std::isfinite(h[S_TMP1])) {
        C.err = "the displacement vanished on the load nodes";
        return TOS_E_NAN;
      }
      fch = std::sqrt(h[S_TMP2]) / std::sqrt(h[S_TMP3]);
    }
    dot_to(C, f, u, nullptr, nullptr, S_TMP0, 0, OP_NONE);
    const double comp = read_scal(C)[S_TMP0];
    if (st) {
      st->iterations = it;
      st->compliance = comp;
      st->fch = fch;
      st->rel_res = rel;
      st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return TOS_OK;
  });
}

TOS_EXPORT int tos_balance_loads(tos_ctx* c, int k, int n, const int64_t* nodes, const double* normals) {
  return guard(c, [&]() {
    REQUIRE(k >= 0 && k < c->K && n > 0 && nodes, "bad arguments");
    Ctx& C = *c;
    const int64_t NP = C.L[0].v.NP;
    const Region RG = upload_region(C, n, nodes, normals);
    double* f = C.F + (size_t)k * 3 * NP;
    const unsigned nb = nblk(n);
    hipLaunchKernelGGL(k_gather_load, dim3(nb), dim3(256), 0, C.stream, n, (const int64_t*)C.wc_idx, NP,
                       (const double*)f, C.wc_src);
    region_moments(C, RG, nullptr, C.wc_src);
    hipLaunchKernelGGL(k_zero<double>, dim3(nblk(3 * NP)), dim3(256), 0, C.stream, f, 3 * NP);
    hipLaunchKernelGGL(k_balance_apply, dim3(nb), dim3(256), 0, C.stream, n, (const int64_t*)C.wc_idx, RG.nrm,
                       RG.pos, RG.P, (const double*)C.scal, NP, (const double*)C.wc_src, f);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(C.stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_remove_rigid(tos_ctx* c, int k) {
  return guard(c, [&]() {
    REQUIRE(k >= 0 && k < c->K, "bad case");
    if (!c->floating) { c->err = "the context has supports: no rigid motions to remove"; return TOS_E_STATE; }
    remove_rigid_dev(*c, c->U + (size_t)k * 3 * c->L[0].v.NP);
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_sensitivity(tos_ctx* c, int k) {
  return guard(c, [&]() {
    REQUIRE(k >= 0 && k < c->K, "bad case");
    const LevelView& v = c->L[0].v;
    const int64_t n = (int64_t)v.ex * v.ey * v.ez;
    Prof pr(*c, P_SENS_FILTER);
    hipLaunchKernelGGL((k_sensitivity<0>), dim3(nblk(n)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const float*)c->rho, (const double*)c->Ke64, (const double*)(c->U + (size_t)k * 3 * v.NP),
                       c->penal, c->sens, (double*)nullptr);
    LAUNCH_CHECK();
    return TOS_OK;
  });
}

TOS_EXPORT int tos_element_compliance(tos_ctx* c, int k, double* out) {
  return guard(c, [&]() {
    REQUIRE(out && k >= 0 && k < c->K, "bad case");
    const LevelView& v = c->L[0].v;
    const int64_t n = (int64_t)v.ex * v.ey * v.ez;
    double* st = (double*)stage(*c, n * sizeof(double));
    hipLaunchKernelGGL((k_sensitivity<1>), dim3(nblk(n)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const float*)c->rho, (const double*)c->Ke64, (const double*)(c->U + (size_t)k * 3 * v.NP),
                       c->penal, (float*)nullptr, st);
    LAUNCH_CHECK();
    download(c->stream, out, st, n * sizeof(double));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_get_sensitivity(tos_ctx* c, int which, float* out) {
  return guard(c, [&]() {
    float* b = sens_buf(*c, which);
    REQUIRE(b && out, "bad sensitivity request");
    const LevelView& v = c->L[0].v;
    const int64_t n = (int64_t)v.ex * v.ey * v.ez;
    float* st = (float*)stage(*c, n * sizeof(float));
    hipLaunchKernelGGL(k_unpad_elems, dim3(nblk(n)), dim3(256), 0, c->stream, v, (const float*)b, st);
    LAUNCH_CHECK();
    download(c->stream, out, st, n * sizeof(float));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_set_sensitivity(tos_ctx* c, int which, const float* in) {
  return guard(c, [&]() {
    float* b = sens_buf(*c, which);
    REQUIRE(b && in, "bad sensitivity request");
    const LevelView& v = c->L[0].v;
    const int64_t n = (int64_t)v.ex * v.ey * v.ez;
    float* st = (float*)stage(*c, n * sizeof(float));
    upload(c->stream, st, in, n * sizeof(float));
    hipLaunchKernelGGL(k_pad_elems, dim3(nblk(n)), dim3(256), 0, c->stream, v, (const float*)st, b);
    hipLaunchKernelGGL(k_mask_inactive, dim3(nblk(v.NE)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact, b);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_init_vol_sensitivity(tos_ctx* c) {
  return guard(c, [&]() {
    const LevelView& v = c->L[0].v;
    hipLaunchKernelGGL(k_fill_active, dim3(nblk(v.NE)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact, c->vsens,
                       (float)(1.0 / (double)c->n_active));
    LAUNCH_CHECK();
    return TOS_OK;
  });
}

TOS_EXPORT int tos_filter(tos_ctx* c, int which, double R) {
  return guard(c, [&]() {
    float* b = sens_buf(*c, which);
    REQUIRE(b && R > 0 && R < 64, "bad filter request");
    // ball table: offsets within floor(R + 1/2) with d^2 <= R^2, weight 1-6r^2+8r^3-3r^4
    const int Ri = (int)(R + 0.5);
    std::vector<int> off;
    std::vector<double> wt;
    for (int dz = -Ri; dz <= Ri; ++dz)
      for (int dy = -Ri; dy <= Ri; ++dy)
        for (int dx = -Ri; dx <= Ri; ++dx) {
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 > R * R) continue;
          const double r = std::sqrt(r2 / (R * R));
          off.push_back(dx); off.push_back(dy); off.push_back(dz);
          wt.push_back(1 - 6 * r * r + 8 * r * r * r - 3 * r * r * r * r);
        }
    const int nt = (int)wt.size();
    char* st = (char*)stage(*c, nt * (3 * sizeof(int) + sizeof(double)) + 64);
    int* doff = (int*)st;
    double* dwt = (double*)(st + ((nt * 3 * sizeof(int) + 15) / 16) * 16);
    upload(c->stream, doff, off.data(), off.size() * sizeof(int));
    upload(c->stream, dwt, wt.data(), wt.size() * sizeof(double));
    const LevelView& v = c->L[0].v;
    HIPCHK(hipMemcpyAsync(c->tmpf, b, v.NE * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    const int64_t n = (int64_t)v.ex * v.ey * v.ez;
    Prof pr(*c, P_SENS_FILTER);
    const dim3 tg((v.ex + kFzX - 1) / kFzX, (v.ey + kFzY - 1) / kFzY, (v.ez + kFzZ - 1) / kFzZ);
    const uint8_t* ea = (const uint8_t*)c->eact;
    const float* sb = (const float*)c->tmpf;
    if (Ri <= 1)
      hipLaunchKernelGGL(k_filter_zm<1>, tg, dim3(256), 0, c->stream, v, ea, sb, b, (const int*)doff,
                         (const double*)dwt, nt);
    else if (Ri == 2)
      hipLaunchKernelGGL(k_filter_zm<2>, tg, dim3(256), 0, c->stream, v, ea, sb, b, (const int*)doff,
                         (const double*)dwt, nt);
    else if (Ri == 3)
      hipLaunchKernelGGL(k_filter_zm<3>, tg, dim3(256), 0, c->stream, v, ea, sb, b, (const int*)doff,
                         (const double*)dwt, nt);
    else
      hipLaunchKernelGGL(k_filter, dim3(nblk(n)), dim3(256), 0, c->stream, v, ea, sb, b, (const int*)doff,
                         (const double*)dwt, nt);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->stream));  // the table lives in the staging buffer
    return TOS_OK;
  });
}

TOS_EXPORT int tos_oc_update(tos_ctx* c, double vgoal, double move, double damp, double* g_out, int* steps) {
  return guard(c, [&]() {
    REQUIRE(move > 0 && damp > 0, "bad OC parameters");
    const LevelView& v = c->L[0].v;
    Prof pr(*c, P_OPT);
    hipLaunchKernelGGL(k_absmax_partial, dim3(kRedBlocks), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const float*)c->sens, c->partial);
    hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(256), 0, c->stream, (const double*)c->partial, kRedBlocks, c->scal);
    for (int it = 0; it < 31; ++it) {
      hipLaunchKernelGGL(k_oc_trial, dim3(kRedBlocks), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                         (const uint8_t*)c->shell, (const float*)c->rho, (const float*)c->sens, move, damp,
                         c->rho_min, (const double*)c->scal, c->partial);
      hipLaunchKernelGGL(k_oc_decide, dim3(1), dim3(256), 0, c->stream, (const double*)c->partial, kRedBlocks,
                         (double)c->n_active, 0, vgoal, 31, -1.0, c->scal);
    }
    hipLaunchKernelGGL(k_oc_apply, dim3(nblk(v.NE)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const uint8_t*)c->shell, c->rho, (const float*)c->sens, move, damp, c->rho_min,
                       (const double*)c->scal);
    LAUNCH_CHECK();
    const double* hs = read_scal(*c);
    if (g_out) *g_out = hs[S_OC_G];
    if (steps) *steps = (int)hs[S_OC_STEPS];
    c->op_ready = false;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_oc_update_projected(tos_ctx* c, double vgoal, double move, double damp, double beta,
                                      double* g_out, int* steps, double* change) {
  return guard(c, [&]() {
    REQUIRE(move > 0 && damp > 0 && beta > 0, "bad OC parameters");
    const LevelView& v = c->L[0].v;
    Prof pr(*c, P_OPT);
    hipLaunchKernelGGL(k_absmax_partial, dim3(kRedBlocks), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const float*)c->sens, c->partial);
    hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(256), 0, c->stream, (const double*)c->partial, kRedBlocks, c->scal);
    for (int it = 0; it < 51; ++it) {
      hipLaunchKernelGGL(k_oc_trial_proj, dim3(kRedBlocks), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                         (const uint8_t*)c->shell, (const float*)c->rho_pre, (const float*)c->sens, move, damp,
                         c->rho_min, beta, (const double*)c->scal, c->partial);
      hipLaunchKernelGGL(k_oc_decide, dim3(1), dim3(256), 0, c->stream, (const double*)c->partial, kRedBlocks,
                         (double)c->n_active, 0, vgoal, 51, 1e-3, c->scal);
    }
    hipLaunchKernelGGL(k_oc_apply_proj, dim3(kRedBlocks), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const uint8_t*)c->shell, c->rho, c->rho_pre, (const float*)c->sens, move, damp, c->rho_min,
                       beta, (const double*)c->scal, c->partial);
    hipLaunchKernelGGL(k_change_fold, dim3(1), dim3(256), 0, c->stream, (const double*)c->partial, kRedBlocks,
                       c->scal);
    LAUNCH_CHECK();
    const double* hs = read_scal(*c);
    if (g_out) *g_out = hs[S_OC_G];
    if (steps) *steps = (int)hs[S_OC_STEPS];
    if (change) {
      change[0] = hs[S_TMP0];
      change[1] = hs[S_TMP1] / (double)c->n_active;
    }
    c->op_ready = false;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_project_density(tos_ctx* c, double beta) {
  return guard(c, [&]() {
    REQUIRE(beta > 0, "bad beta");
    const LevelView& v = c->L[0].v;
    hipLaunchKernelGGL(k_project, dim3(nblk(v.NE)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const uint8_t*)c->shell, c->rho, c->rho_pre, beta, c->rho_min);
    LAUNCH_CHECK();
    c->op_ready = false;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_projection_sensitivity(tos_ctx* c, int which, double beta) {
  return guard(c, [&]() {
    float* b = sens_buf(*c, which);
    REQUIRE(b && beta > 0, "bad projection request");
    const LevelView& v = c->L[0].v;
    hipLaunchKernelGGL(k_project_sens, dim3(nblk(v.NE)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const float*)c->rho_pre, b, beta);
    LAUNCH_CHECK();
    return TOS_OK;
  });
}

// ---- spline ------------------------------------------------------------------------------
TOS_EXPORT int tos_spline_init(tos_ctx* c, int mode, int order, int param) {
  return guard(c, [&]() {
    REQUIRE((mode == 0 || mode == 1) && order >= 2 && order <= 8 && param >= 1, "bad spline parameters");
    const int dims[3] = {c->nx, c->ny, c->nz};
    AxisTable T[3];
    for (int a = 0; a < 3; ++a) T[a] = spline_table(mode, order, param, dims[a], c->h);
    c->sp_mode = mode; c->sp_order = order; c->sp_param = param;
    // the surface-constraint state belongs to the previous spline (knots, coefficient count)
    for (double*& U : c->sf_knots) U = nullptr;
    c->sscsens = c->dripcsens = nullptr;
    c->sf_ncoef = 0;
    c->sf_valid = false;
    for (int a = 0; a < 3; ++a) {
      c->nc[a] = T[a].nc;
      c->sp_idx[a] = dalloc<int>(*c, T[a].idx.size());
      c->sp_wt[a] = dalloc<float>(*c, T[a].wt.size());
      c->sp_ok[a] = dalloc<uint8_t>(*c, T[a].ok.size());
      upload(c->stream, c->sp_idx[a], T[a].idx.data(), T[a].idx.size() * sizeof(int));
      upload(c->stream, c->sp_wt[a], T[a].wt.data(), T[a].wt.size() * sizeof(float));
      upload(c->stream, c->sp_ok[a], T[a].ok.data(), T[a].ok.size());
      // transpose (coefficient -> list of (element, weight)) for the deterministic gather
      std::vector<int> cnt(T[a].nc + 1, 0), ptr, elm;
      std::vector<float> cw;
      for (int e = 0; e < dims[a]; ++e)
        if (T[a].ok[e])
          for (int q = 0; q < order; ++q) cnt[T[a].idx[(size_t)e * order + q] + 1]++;
      ptr.assign(T[a].nc + 1, 0);
      for (int i = 0; i < T[a].nc; ++i) ptr[i + 1] = ptr[i] + cnt[i + 1];
      elm.assign(ptr.back(), 0);
      cw.assign(ptr.back(), 0.f);
      std::vector<int> fill(ptr.begin(), ptr.end() - 1);
      for (int e = 0; e < dims[a]; ++e)
        if (T[a].ok[e])
          for (int q = 0; q < order; ++q) {
            const int i = T[a].idx[(size_t)e * order + q];
            elm[fill[i]] = e;
            cw[fill[i]] = T[a].wt[(size_t)e * order + q];
            fill[i]++;
          }
      c->sp_nq[a] = ptr.back();
      c->sp_cptr[a] = dalloc<int>(*c, ptr.size());
      c->sp_celm[a] = dalloc<int>(*c, elm.size());
      c->sp_cw[a] = dalloc<float>(*c, cw.size());
      upload(c->stream, c->sp_cptr[a], ptr.data(), ptr.size() * sizeof(int));
      if (!elm.empty()) upload(c->stream, c->sp_celm[a], elm.data(), elm.size() * sizeof(int));
      if (!cw.empty()) upload(c->stream, c->sp_cw[a], cw.data(), cw.size() * sizeof(float));
    }
    const size_t n = (size_t)c->nc[0] * c->nc[1] * c->nc[2];
    c->coef = dalloc<float>(*c, n);
    c->coef_old = dalloc<float>(*c, n);
    c->csens = dalloc<float>(*c, n);
    c->cvsens = dalloc<float>(*c, n);
    c->sp_g = dalloc<double>(*c, n);  // last adjoint stage
    c->sp_t1 = dalloc<double>(*c, (size_t)c->nc[0] * c->ny * c->nz);
    c->sp_t2 = dalloc<double>(*c, std::max((size_t)c->nc[0] * c->nc[1] * c->nz, n));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_spline_set_coeffs(tos_ctx* c, const float* v) {
  return guard(c, [&]() {
    REQUIRE(v && c->sp_mode >= 0, "spline not initialised");
    upload(c->stream, c->coef, v, (size_t)c->nc[0] * c->nc[1] * c->nc[2] * sizeof(float));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_spline_get_coeffs(tos_ctx* c, float* v) {
  return guard(c, [&]() {
    REQUIRE(v && c->sp_mode >= 0, "spline not initialised");
    download(c->stream, v, c->coef, (size_t)c->nc[0] * c->nc[1] * c->nc[2] * sizeof(float));
    return TOS_OK;
  });
}

static void spline_eval_launch(tos_ctx* c) {
  const LevelView& v = c->L[0].v;
  const int64_t n = (int64_t)v.ex * v.ey * v.ez;
#define TOS_SPLINE_TABLES                                                                              \
  (const int*)c->sp_idx[0], (const float*)c->sp_wt[0], (const uint8_t*)c->sp_ok[0], (const int*)c->sp_idx[1], \
      (const float*)c->sp_wt[1], (const uint8_t*)c->sp_ok[1], (const int*)c->sp_idx[2],                        \
      (const float*)c->sp_wt[2], (const uint8_t*)c->sp_ok[2]
  if (c->sp_order == 4)
    hipLaunchKernelGGL(k_spline_eval_o<4>, dim3(nblk(n)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const uint8_t*)c->shell, (const float*)c->coef, c->nc[0], c->nc[1], TOS_SPLINE_TABLES,
                       c->rho_min, c->rho);
  else if (c->sp_order == 3)
    hipLaunchKernelGGL(k_spline_eval_o<3>, dim3(nblk(n)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const uint8_t*)c->shell, (const float*)c->coef, c->nc[0], c->nc[1], TOS_SPLINE_TABLES,
                       c->rho_min, c->rho);
  else
    hipLaunchKernelGGL(k_spline_eval, dim3(nblk(n)), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                       (const uint8_t*)c->shell, (const float*)c->coef, c->nc[0], c->nc[1], c->sp_order,
                       TOS_SPLINE_TABLES, c->rho_min, c->rho);
#undef TOS_SPLINE_TABLES
  LAUNCH_CHECK();
}

TOS_EXPORT int tos_spline_eval(tos_ctx* c) {
  return guard(c, [&]() {
    REQUIRE(c->sp_mode >= 0, "spline not initialised");
    Prof pr(*c, P_SPLINE);
    spline_eval_launch(c);
    c->op_ready = false;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_oc_update_coeffs(tos_ctx* c, double vgoal, double move, double damp, double cmin, double cmax,
                                    double* g_out, int* steps) {
  return guard(c, [&]() {
    REQUIRE(c->sp_mode >= 0 && move > 0 && damp > 0 && cmax > cmin, "bad coefficient OC request");
    Prof pr(*c, P_OPT);
    const LevelView& v = c->L[0].v;
    const int64_t n = (int64_t)c->nc[0] * c->nc[1] * c->nc[2];
    HIPCHK(hipMemcpyAsync(c->coef_old, c->coef, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_coef_absmax_partial, dim3(kRedBlocks), dim3(256), 0, c->stream, n, (const float*)c->csens,
                       c->partial);
    hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(256), 0, c->stream, (const double*)c->partial, kRedBlocks, c->scal);
    for (int it = 0; it < 31; ++it) {
      hipLaunchKernelGGL(k_coef_trial, dim3(nblk(n)), dim3(256), 0, c->stream, n, (const float*)c->coef_old,
                         (const float*)c->csens, move, damp, cmin, cmax, (const double*)c->scal, c->coef);
      spline_eval_launch(c);
      hipLaunchKernelGGL(k_sum_active, dim3(kRedBlocks), dim3(256), 0, c->stream, v, (const uint8_t*)c->eact,
                         (const float*)c->rho, c->partial);
      hipLaunchKernelGGL(k_oc_decide, dim3(1), dim3(256), 0, c->stream, (const double*)c->partial, kRedBlocks,
                         (double)c->n_active, 0, vgoal, 31, -1.0, c->scal);
    }
    LAUNCH_CHECK();
    const double* hs = read_scal(*c);
    if (g_out) *g_out = hs[S_OC_G];
    if (steps) *steps = (int)hs[S_OC_STEPS];
    c->op_ready = false;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_spline_adjoint(tos_ctx* c, int which) {
  return guard(c, [&]() {
    float* b = sens_buf(*c, which);
    REQUIRE(b && c->sp_mode >= 0, "spline not initialised");
    const LevelView& v = c->L[0].v;
    Prof pr(*c, P_SPLINE);
    const int n0 = c->nc[0], n1 = c->nc[1], n2 = c->nc[2];
    int64_t m1 = (int64_t)n0 * v.ey * v.ez, m2 = (int64_t)n0 * n1 * v.ez, m3 = (int64_t)n0 * n1 * n2;
    (void)m1;  // a workgroup per `rows` element rows, rows and lists staged in LDS (<= 60 KiB)
    const size_t tab0 = ((size_t)n0 + 1 + 2 * (size_t)c->sp_nq[0]) * 4;
    REQUIRE(tab0 + (size_t)v.ex * 4 <= 60 * 1024, "spline lists too long for the first adjoint stage");
    const int rows = (int)std::max<size_t>(1, std::min<size_t>(8, (60 * 1024 - tab0) / ((size_t)v.ex * 4)));
    const size_t lds0 = tab0 + (size_t)rows * v.ex * 4;
    hipLaunchKernelGGL(k_spline_stage0, dim3((unsigned)((int64_t)((v.ey + rows - 1) / rows) * v.ez)),
                       dim3(256), lds0, c->stream, v, (const uint8_t*)c->eact,
                       (const uint8_t*)c->shell, (const float*)b, (const uint8_t*)c->sp_ok[1],
                       (const uint8_t*)c->sp_ok[2], c->sp_t1, n0, (const int*)c->sp_cptr[0],
                       (const int*)c->sp_celm[0], (const float*)c->sp_cw[0], rows);
    hipLaunchKernelGGL(k_spline_stage, dim3(nblk(m2)), dim3(256), 0, c->stream, (const double*)c->sp_t1, c->sp_t2, 1,
                       n0, n1, v.ez, v.ey, (const int*)c->sp_cptr[1], (const int*)c->sp_celm[1],
                       (const float*)c->sp_cw[1]);
    hipLaunchKernelGGL(k_spline_stage, dim3(nblk(m3)), dim3(256), 0, c->stream, (const double*)c->sp_t2, c->sp_g, 2,
                       n0, n1, n2, v.ez, (const int*)c->sp_cptr[2], (const int*)c->sp_celm[2],
                       (const float*)c->sp_cw[2]);
    hipLaunchKernelGGL(k_d2f, dim3(nblk(m3)), dim3(256), 0, c->stream, (const double*)c->sp_g,
                       which == 0 ? c->csens : c->cvsens, m3);
    LAUNCH_CHECK();
    return TOS_OK;
  });
}

TOS_EXPORT int tos_spline_get_coeff_sens(tos_ctx* c, int which, float* out) {
  return guard(c, [&]() {
    REQUIRE(out && c->sp_mode >= 0 && which >= 0 && which <= 3, "bad request");
    const size_t n = (size_t)c->nc[0] * c->nc[1] * c->nc[2];
    const float* src = which == 0 ? c->csens : which == 1 ? c->cvsens : which == 2 ? c->sscsens : c->dripcsens;
    if (!src) {  // no surface constraints evaluated yet
      std::fill(out, out + n, 0.0f);
      return TOS_OK;
    }
    download(c->stream, out, src, n * sizeof(float));
    return TOS_OK;
  });
}

// ---- self-supporting and drip constraints on spline surface points -------------------
namespace {
// device description of the spline axes (knot vectors of mode 0 uploaded on first use)
SfAxes surface_axes(Ctx& c) {
  const int dims[3] = {c.nx, c.ny, c.nz};
  SfAxes S{};
  for (int a = 0; a < 3; ++a) {
    SfAxis& A = S.a[a];
    A.mode = c.sp_mode;
    A.k = c.sp_order;
    A.nc = c.nc[a];
    A.sh = c.sp_param * c.h;
    A.box = dims[a] * c.h;
    const int P = c.sp_param;
    const double box = dims[a] * c.h;
    A.lo = -2.0 * box / (P + 1);
    A.step = (box + 4.0 * box / (P + 1)) / (P + 1);
    A.U = nullptr;
    if (c.sp_mode == 0) {
      if (!c.sf_knots[a]) {  // clamped knots (Grid::set_spline_knot_series grid/Grid.cpp:3253-3377)
        const double hi = box + 2.0 * box / (P + 1);
        std::vector<double> U(c.sp_order, A.lo);
        double t = A.lo;
        for (int i = 0; i < P; ++i) { t += (hi - A.lo) / (P + 1); U.push_back(t); }
        for (int i = 0; i < c.sp_order; ++i) U.push_back(hi);
        c.sf_knots[a] = dalloc<double>(c, U.size());
        upload(c.stream, c.sf_knots[a], U.data(), U.size() * sizeof(double));
      }
      A.U = c.sf_knots[a];
    }
  }
  return S;
}
}  // namespace

TOS_EXPORT int tos_surface_constraints(tos_ctx* c, const tos_surface_opts* o, tos_surface_stats* st) {
  return guard(c, [&]() {
    REQUIRE(o && c->sp_mode >= 0 && o->cells >= 1 && o->cells <= 1000 && o->ss_mode >= 0 && o->ss_mode <= 5 &&
                o->drip_mode >= 0 && o->drip_mode <= 7 && o->beta > 0 && c->sp_order <= 8,
            "bad surface-constraint options");
    Ctx& C = *c;
    const int NB = o->cells + 2;
    const int64_t nl = (int64_t)NB * NB * NB, cap = 3 * nl;
    const size_t ncoef = (size_t)C.nc[0] * C.nc[1] * C.nc[2];
    if (C.sf_cells != o->cells) {  // lattice buffers for this lattice size
      C.sf_val = dalloc<double>(C, nl);
      C.sf_cnt = dalloc<int>(C, nl);
      C.sf_off = dalloc<int>(C, nl + 1);
      const int64_t nsb = (nl + kSfScanBlock - 1) / kSfScanBlock;
      C.sf_bsum = dalloc<int>(C, nsb);
      C.sf_boff = dalloc<int>(C, nsb + 1);
      C.sf_tab_i = dalloc<int>(C, 3 * 8 * (int64_t)NB);
      C.sf_tab_v = dalloc<double>(C, 3 * 8 * (int64_t)NB);
      C.sf_tab_ok = dalloc<uint8_t>(C, 3 * (int64_t)NB);
      C.sf_pts = dalloc<double>(C, 3 * cap);
      C.sf_grad = dalloc<double>(C, 3 * cap);
      C.sf_hess = dalloc<double>(C, 6 * cap);
      C.sf_nrm = dalloc<double>(C, 3 * cap);
      C.sf_wss = dalloc<double>(C, 3 * cap);
      C.sf_wdr = dalloc<double>(C, 6 * cap);
      C.sf_cells = o->cells;
    }
    if (C.sf_ncoef != (int64_t)ncoef) {  // per-coefficient (and per-knot-cell) buffers
      C.sscsens = dalloc<float>(C, ncoef);
      C.dripcsens = dalloc<float>(C, ncoef);
      C.sf_ccnt = dalloc<int>(C, ncoef);
      C.sf_cstart = dalloc<int>(C, ncoef + 1);
      const int64_t nsb = ((int64_t)ncoef + kSfScanBlock - 1) / kSfScanBlock;
      C.sf_cbsum = dalloc<int>(C, nsb);
      C.sf_cboff = dalloc<int>(C, nsb + 1);
      C.sf_ncoef = (int64_t)ncoef;
    }
    {  // all coefficients equal: no surface; the driver's constants (utils/test_utils.cpp:1106-1113)
      const unsigned nbq = (unsigned)std::min<int64_t>(kRedBlocks, ((int64_t)ncoef + 255) / 256);
      hipLaunchKernelGGL(k_sf_coef_differ, dim3(nbq), dim3(256), 0, C.stream, (const float*)C.coef,
                         (int64_t)ncoef, C.partial);
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, C.stream, (const double*)C.partial, (int)nbq, 0, S_TMP0, 0,
                         OP_NONE, C.scal);
      LAUNCH_CHECK();
      const bool equal = read_scal(C)[S_TMP0] == 0.0;
      if (equal) {
        hipLaunchKernelGGL(k_zero<float>, dim3(nblk((int64_t)ncoef)), dim3(256), 0, C.stream, C.sscsens,
                           (int64_t)ncoef);
        hipLaunchKernelGGL(k_zero<float>, dim3(nblk((int64_t)ncoef)), dim3(256), 0, C.stream, C.dripcsens,
                           (int64_t)ncoef);
        LAUNCH_CHECK();
        HIPCHK(hipStreamSynchronize(C.stream));
        C.sf_n = 0;
        C.ss_value = 1.0;
        C.drip_value = 0.0;
        C.sf_valid = true;
        if (st) *st = tos_surface_stats{0, 0, 1.0, 1.0, 0.0};
        return TOS_OK;
      }
    }
    const SfAxes S = surface_axes(C);
    const LevelView& v = C.L[0].v;
    SfLattice G{NB, {v.ex * C.h / o->cells, v.ey * C.h / o->cells, v.ez * C.h / o->cells}};
    Prof pr(C, P_SPLINE);
    hipLaunchKernelGGL(k_sf_axis_tab, dim3(nblk(3 * (int64_t)NB)), dim3(256), 0, C.stream, S, G, C.sf_tab_i,
                       C.sf_tab_v, C.sf_tab_ok);
    if (C.sp_order == 4)
      hipLaunchKernelGGL(k_sf_lattice<4>, dim3(nblk(nl)), dim3(256), 0, C.stream, G, C.sp_order, C.nc[0], C.nc[1],
                         (const int*)C.sf_tab_i, (const double*)C.sf_tab_v, (const uint8_t*)C.sf_tab_ok,
                         (const float*)C.coef, C.sf_val);
    else
      hipLaunchKernelGGL(k_sf_lattice<0>, dim3(nblk(nl)), dim3(256), 0, C.stream, G, C.sp_order, C.nc[0], C.nc[1],
                         (const int*)C.sf_tab_i, (const double*)C.sf_tab_v, (const uint8_t*)C.sf_tab_ok,
                         (const float*)C.coef, C.sf_val);
    hipLaunchKernelGGL(k_sf_count, dim3(nblk(nl)), dim3(256), 0, C.stream, G, v, (const uint8_t*)C.eact,
                       (const double*)C.sf_val, C.h, C.sf_cnt);
    {
      const unsigned nsb = (unsigned)((nl + kSfScanBlock - 1) / kSfScanBlock);
      hipLaunchKernelGGL(k_sf_scan_blocks, dim3(nsb), dim3(1024), 0, C.stream, (const int*)C.sf_cnt, nl, C.sf_bsum);
      hipLaunchKernelGGL(k_sf_scan, dim3(1), dim3(1024), 0, C.stream, (const int*)C.sf_bsum, (int64_t)nsb,
                         C.sf_boff);
      hipLaunchKernelGGL(k_sf_scan_apply, dim3(nsb), dim3(1024), 0, C.stream, (const int*)C.sf_cnt, nl,
                         (const int*)C.sf_boff, C.sf_off);
    }
    LAUNCH_CHECK();
    int total = 0;
    download(C.stream, &total, C.sf_off + nl, sizeof(int));
    const int64_t n = total;
    C.sf_n = n;
    hipLaunchKernelGGL(k_sf_points, dim3(nblk(nl)), dim3(256), 0, C.stream, G, v, (const uint8_t*)C.eact,
                       (const double*)C.sf_val, C.h, (const int*)C.sf_off, C.sf_pts);
    constexpr int kSfBlocks = 512;  // 3 columns of partials fit the context's partial buffer
    const unsigned nbp = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kSfBlocks, (n + 255) / 256));
    hipLaunchKernelGGL(k_sf_orient, dim3(nbp), dim3(256), 0, C.stream, S, (const float*)C.coef, n,
                       (const double*)C.sf_pts, o->beta, C.sf_grad, C.sf_hess, C.partial);
    hipLaunchKernelGGL(k_sf_fold, dim3(1), dim3(256), 0, C.stream, (const double*)C.partial, (int)nbp, 1, S_TMP3,
                       C.scal);
    SfParams P{std::cos(o->print_angle), std::cos(o->drip_angle), o->ss_mode, o->drip_mode, o->drip_version,
               o->p_norm, o->hfunction_c, o->sigmoid_c, o->drip_beta, o->drip_mu, o->drip_angle, o->drip_alpha,
               o->drip_beta_scale};
    hipLaunchKernelGGL(k_sf_values, dim3(nbp), dim3(256), 0, C.stream, P, n, (const double*)C.sf_grad,
                       (const double*)C.sf_hess, (const double*)C.scal, (int)S_TMP3, C.sf_nrm, C.sf_wss, C.sf_wdr,
                       C.partial);
    hipLaunchKernelGGL(k_sf_fold, dim3(1), dim3(256), 0, C.stream, (const double*)C.partial, (int)nbp, 3, S_TMP0,
                       C.scal);
    LAUNCH_CHECK();
    const double* hs = read_scal(C);
    const double dir = hs[S_TMP3] > 0 ? 1.0 : -1.0;
    const double ss_sum = hs[S_TMP0], dr_sum = hs[S_TMP1], n_down = hs[S_TMP2];
    // global values (global_selfsupp_constraint grid/Grid.cu:6101-6145, global_drip_constraint
    // :6485-6525) and the factors of their gradients
    const double cnt_ss = o->ss_mode % 2 == 0 ? n_down : (double)n;
    const double cnt_dr = o->drip_mode % 2 == 0 ? n_down : (double)n;
    double ss_val = 0, ss_fac = 0, dr_val = 0, dr_fac = 0;
    if (cnt_ss > 0) {
      const double mean = ss_sum / cnt_ss;
      if (o->ss_mode <= 1) {
        ss_val = std::pow(mean, 1.0 / o->p_norm) - 1;
        ss_fac = mean > 0 ? std::pow(mean, 1.0 / o->p_norm - 1) / o->p_norm / cnt_ss : 0.0;
      } else {
        ss_val = mean;
        ss_fac = 1.0 / cnt_ss;
      }
    }
    if (cnt_dr > 0) {
      dr_val = o->drip_mode <= 1 ? -1.0 : dr_sum / cnt_dr;
      dr_fac = o->drip_mode >= 6 ? 1.0 / cnt_dr : 0.0;
    }
    if (C.sf_bcap < std::max<int64_t>(n, 1)) {
      C.sf_bcap = std::max<int64_t>(n, 1024) * 2;
      C.sf_bidx = dalloc<int>(C, 24 * C.sf_bcap);
      C.sf_bN = dalloc<double>(C, 72 * C.sf_bcap);
      C.sf_key = dalloc<int>(C, C.sf_bcap);
      C.sf_list = dalloc<int>(C, C.sf_bcap);
    }
    // points grouped by knot cell (count, scan, fill, per-cell sort), then the gather
    const int64_t nc64 = (int64_t)ncoef;
    hipLaunchKernelGGL(k_zero<int>, dim3(nblk(nc64)), dim3(256), 0, C.stream, C.sf_ccnt, nc64);
    if (n) {
      hipLaunchKernelGGL(k_sf_basis, dim3(nblk(n)), dim3(256), 0, C.stream, S, n, (const double*)C.sf_pts,
                         C.sf_bidx, C.sf_bN);
      hipLaunchKernelGGL(k_sf_cell_count, dim3(nblk(n)), dim3(256), 0, C.stream, n, C.nc[0], C.nc[1],
                         (const int*)C.sf_bidx, C.sf_key, C.sf_ccnt);
    }
    {
      const unsigned nsb = (unsigned)((nc64 + kSfScanBlock - 1) / kSfScanBlock);
      hipLaunchKernelGGL(k_sf_scan_blocks, dim3(nsb), dim3(1024), 0, C.stream, (const int*)C.sf_ccnt, nc64,
                         C.sf_cbsum);
      hipLaunchKernelGGL(k_sf_scan, dim3(1), dim3(1024), 0, C.stream, (const int*)C.sf_cbsum, (int64_t)nsb,
                         C.sf_cboff);
      hipLaunchKernelGGL(k_sf_scan_apply, dim3(nsb), dim3(1024), 0, C.stream, (const int*)C.sf_ccnt, nc64,
                         (const int*)C.sf_cboff, C.sf_cstart);
    }
    hipLaunchKernelGGL(k_zero<int>, dim3(nblk(nc64)), dim3(256), 0, C.stream, C.sf_ccnt, nc64);
    if (n)
      hipLaunchKernelGGL(k_sf_cell_fill, dim3(nblk(n)), dim3(256), 0, C.stream, n, (const int*)C.sf_key,
                         (const int*)C.sf_cstart, C.sf_ccnt, C.sf_list);
    hipLaunchKernelGGL(k_sf_cell_sort, dim3(nblk(nc64)), dim3(256), 0, C.stream, nc64, (const int*)C.sf_cstart,
                       C.sf_list);
    hipLaunchKernelGGL(k_sf_grad, dim3(nblk(64 * nc64)), dim3(256), 0, C.stream, S, nc64, (const int*)C.sf_cstart,
                       (const int*)C.sf_list, (const int*)C.sf_bidx, (const double*)C.sf_bN,
                       (const double*)C.sf_wss, (const double*)C.sf_wdr, ss_fac, dr_fac, C.sscsens, C.dripcsens);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(C.stream));
    C.ss_value = ss_val;
    C.drip_value = dr_val;
    C.sf_valid = true;
    if (st) {
      st->n_points = n;
      st->n_down = (int64_t)n_down;
      st->direction = dir;
      st->ss_value = ss_val;
      st->drip_value = dr_val;
    }
    return TOS_OK;
  });
}

TOS_EXPORT int tos_surface_points(tos_ctx* c, int64_t cap, double* xyz, double* normals) {
  return guard(c, [&]() {
    REQUIRE(cap >= 0 && (cap == 0 || (xyz && normals)), "bad arguments");
    const int64_t n = std::min<int64_t>(cap, c->sf_valid ? c->sf_n : 0);
    if (n > 0) {
      download(c->stream, xyz, c->sf_pts, 3 * n * sizeof(double));
      download(c->stream, normals, c->sf_nrm, 3 * n * sizeof(double));
    }
    return TOS_OK;
  });
}

// ---- MMA -----------------------------------------------------------------------------------
TOS_EXPORT int tos_mma_create(int n, int m, const double* xmin, const double* xmax, double move, tos_mma** out) {
  if (n <= 0 || m <= 0 || m > MMAX || !xmin || !xmax || !out || !(move > 0)) return TOS_E_ARG;
  tos_mma* M = new tos_mma();
  try {
    HIPCHK(hipGetDevice(&M->dev));
    HIPCHK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
    MmaVecs& V = M->V;
    V.n = n; V.m = m;
    M->move = move;
    for (int i = 0; i < MMAX; ++i) { M->a[i] = 0; M->c[i] = 10; M->d[i] = 1; }  // mma_t.h:206-207
    double** vecs[] = {&V.x, &V.xmin, &V.xmax, &V.L, &V.U, &V.alpha, &V.beta, &V.lastdx, &V.xi, &V.eta,
                       &V.plam, &V.qlam, &V.curx, &V.Dx, &V.dlx, &V.dx, &V.dxi, &V.deta, &V.tx, &V.txi, &V.teta};
    for (double** v : vecs) *v = malloc_mma<double>(M, n);
    V.p = malloc_mma<double>(M, (size_t)(m + 1) * n);
    V.q = malloc_mma<double>(M, (size_t)(m + 1) * n);
    V.G = malloc_mma<double>(M, (size_t)m * n);
    M->nb = (int)std::min<int64_t>(512, std::max<int64_t>(1, (n + 255) / 256));
    M->partial = malloc_mma<double>(M, (size_t)MMA_CNT * M->nb);
    M->red = malloc_mma<double>(M, MMA_CNT);
    M->stage = malloc_mma<double>(M, (size_t)(m + 1) * n);
    HIPCHK(hipHostMalloc((void**)&M->h_red, MMA_CNT * sizeof(double)));
    // set_bound + clamp_asymptotes + init_subproblem_variable (mma_t.cu:747-855)
    std::vector<double> x(n), L(n), U(n), al(n), be(n);
    for (int j = 0; j < n; ++j) {
      const double st = xmax[j] - xmin[j];
      if (!(st > 0)) throw ArgError("xmax must exceed xmin");
      x[j] = 0.5 * (xmin[j] + xmax[j]);
      L[j] = std::min(std::max(xmin[j], x[j] - 10 * st), x[j] - 0.01 * st);
      U[j] = std::min(std::max(xmax[j], x[j] + 0.01 * st), x[j] + 10 * st);
      al[j] = std::max(xmin[j], std::max(L[j] + 0.1 * (x[j] - L[j]), x[j] - move * st));
      be[j] = std::min(xmax[j], std::min(U[j] - 0.1 * (U[j] - x[j]), x[j] + move * st));
      x[j] = 0.5 * (al[j] + be[j]);
    }
    upload(M->stream, V.x, x.data(), n * sizeof(double));
    upload(M->stream, V.xmin, xmin, n * sizeof(double));
    upload(M->stream, V.xmax, xmax, n * sizeof(double));
    upload(M->stream, V.L, L.data(), n * sizeof(double));
    upload(M->stream, V.U, U.data(), n * sizeof(double));
    upload(M->stream, V.alpha, al.data(), n * sizeof(double));
    upload(M->stream, V.beta, be.data(), n * sizeof(double));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "tos_mma_create: %s\n", e.what());
    for (void* p : M->allocs) (void)hipFree(p);
    delete M;
    return e.what()[0] == 'x' ? TOS_E_ARG : TOS_E_HIP;
  }
  *out = M;
  return TOS_OK;
}

TOS_EXPORT int tos_mma_destroy(tos_mma* M) {
  if (!M) return TOS_E_ARG;
  (void)hipSetDevice(M->dev);
  (void)hipStreamSynchronize(M->stream);
  for (void* p : M->allocs) (void)hipFree(p);
  if (M->h_red) (void)hipHostFree(M->h_red);
  (void)hipStreamDestroy(M->stream);
  delete M;
  return TOS_OK;
}

TOS_EXPORT int tos_mma_set_x(tos_mma* M, const double* x) {
  if (!M || !x) return TOS_E_ARG;
  try {
    HIPCHK(hipSetDevice(M->dev));
    const int n = M->V.n;
    std::vector<double> lo(n), hi(n), L(n), U(n), al(n), be(n);
    download(M->stream, lo.data(), M->V.xmin, n * sizeof(double));
    download(M->stream, hi.data(), M->V.xmax, n * sizeof(double));
    for (int j = 0; j < n; ++j) {
      const double st = hi[j] - lo[j];
      L[j] = std::min(std::max(x[j] - 0.5 * st, x[j] - 10 * st), x[j] - 0.01 * st);
      U[j] = std::min(std::max(x[j] + 0.5 * st, x[j] + 0.01 * st), x[j] + 10 * st);
      al[j] = std::max(lo[j], std::max(L[j] + 0.1 * (x[j] - L[j]), x[j] - M->move * st));
      be[j] = std::min(hi[j], std::min(U[j] - 0.1 * (U[j] - x[j]), x[j] + M->move * st));
    }
    upload(M->stream, M->V.x, x, n * sizeof(double));
    upload(M->stream, M->V.L, L.data(), n * sizeof(double));
    upload(M->stream, M->V.U, U.data(), n * sizeof(double));
    upload(M->stream, M->V.alpha, al.data(), n * sizeof(double));
    upload(M->stream, M->V.beta, be.data(), n * sizeof(double));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "tos_mma_set_x: %s\n", e.what());
    return TOS_E_HIP;
  }
  return TOS_OK;
}

TOS_EXPORT int tos_mma_get_x(tos_mma* M, double* x) {
  if (!M || !x) return TOS_E_ARG;
  try {
    HIPCHK(hipSetDevice(M->dev));
    download(M->stream, x, M->V.x, M->V.n * sizeof(double));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "tos_mma_get_x: %s\n", e.what());
    return TOS_E_HIP;
  }
  return TOS_OK;
}

TOS_EXPORT int tos_mma_update(tos_mma* M, const double* df, const double* dg, const double* g, int* converged) {
  if (!M || !df || !dg || !g) return TOS_E_ARG;
  try {
    HIPCHK(hipSetDevice(M->dev));
    const int n = M->V.n, m = M->V.m;
    double* s = M->stage;
    upload(M->stream, s, df, n * sizeof(double));
    upload(M->stream, s + n, dg, (size_t)m * n * sizeof(double));
    const bool cv = mma_update_dev(M, M->stream, s, s + n, g);
    if (converged) *converged = cv;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "tos_mma_update: %s\n", e.what());
    return TOS_E_HIP;
  }
  return TOS_OK;
}

__global__ void k_mma_inputs(int n, const float* __restrict__ cs, const float* __restrict__ cv, double ss, double vs,
                             double* __restrict__ df, double* __restrict__ dg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  df[i] = ss * (double)cs[i];
  dg[i] = vs * (double)cv[i];
}

TOS_EXPORT int tos_ctx_mma_step(tos_ctx* c, tos_mma* M, double sscale, double vscale, double vfrac, int* converged) {
  return guard(c, [&]() {
    REQUIRE(M && c->sp_mode >= 0 && M->V.m == 1 &&
                (int64_t)M->V.n == (int64_t)c->nc[0] * c->nc[1] * c->nc[2],
            "MMA size does not match the spline coefficients");
    const double vol = volume_of(*c);
    const int n = M->V.n;
    double* s = M->stage;
    const double g = vscale * (vol - vfrac);
    hipLaunchKernelGGL(k_mma_inputs, dim3(nblk(n)), dim3(256), 0, c->stream, n, (const float*)c->csens,
                       (const float*)c->cvsens, sscale, vscale, s, s + n);
    LAUNCH_CHECK();
    bool cv;
    {
      Prof pr(*c, P_OPT);
      cv = mma_update_dev(M, c->stream, s, s + n, &g);
    }
    if (converged) *converged = cv;
    hipLaunchKernelGGL((k_copy_cast<double, float>), dim3(nblk(n)), dim3(256), 0, c->stream, (const double*)M->V.x,
                       c->coef, (int64_t)n);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_ctx_mma_step_surface(tos_ctx* c, tos_mma* M, double sscale, double vscale, double vfrac,
                                       double ss_scale, double drip_scale, int* converged) {
  return guard(c, [&]() {
    REQUIRE(M && c->sp_mode >= 0 && M->V.m == 3 && (int64_t)M->V.n == (int64_t)c->nc[0] * c->nc[1] * c->nc[2],
            "MMA size does not match the spline coefficients and three constraints");
    if (!c->sf_valid) { c->err = "tos_surface_constraints first"; return TOS_E_STATE; }
    const double vol = volume_of(*c);
    const int n = M->V.n;
    double* s = M->stage;
    const double g[3] = {vscale * (vol - vfrac), ss_scale * (c->ss_value - 1.0), drip_scale * c->drip_value};
    hipLaunchKernelGGL(k_sf_mma_inputs, dim3(nblk(n)), dim3(256), 0, c->stream, (int64_t)n, (const float*)c->csens,
                       (const float*)c->cvsens, (const float*)c->sscsens, (const float*)c->dripcsens, sscale, vscale,
                       ss_scale, drip_scale, s);
    LAUNCH_CHECK();
    bool cv;
    {
      Prof pr(*c, P_OPT);
      cv = mma_update_dev(M, c->stream, s, s + n, g);
    }
    if (converged) *converged = cv;
    hipLaunchKernelGGL((k_copy_cast<double, float>), dim3(nblk(n)), dim3(256), 0, c->stream, (const double*)M->V.x,
                       c->coef, (int64_t)n);
    LAUNCH_CHECK();
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_set_graph_replay(tos_ctx* c, int enable) {
  return guard(c, [&]() {
    c->graph_replay = enable != 0;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_set_galerkin_precision(tos_ctx* c, int prec) {
  return guard(c, [&]() {
    REQUIRE(prec == 0 || prec == 1, "bad Galerkin precision");
    if (prec != c->gal_prec) {
      c->gal_prec = prec;
      c->op_ready = false;
    }
    return TOS_OK;
  });
}

TOS_EXPORT int tos_set_sweep_mode(tos_ctx* c, int mode) {
  return guard(c, [&]() {
    REQUIRE(mode >= 0 && mode <= 3, "bad sweep mode");
    c->sweep_mode = mode;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_profile(tos_ctx* c, int enable, tos_kernel_stats* out) {
  return guard(c, [&]() {
    if (enable) {
      HIPCHK(hipStreamSynchronize(c->stream));
      c->prof.clear();
      c->ev_used = 0;
      c->prof_on = true;
      return TOS_OK;
    }
    c->prof_on = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (out) {
      std::memset(out, 0, sizeof(*out));
      out->n_classes = P_COUNT;
      for (int i = 0; i < P_COUNT; ++i) std::snprintf(out->names[i], 32, "%s", kProfNames[i]);
      for (const auto& r : c->prof) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        out->total_ms[r.cls] += ms;
        out->launches[r.cls] += 1;
      }
    }
    c->prof.clear();
    c->ev_used = 0;
    return TOS_OK;
  });
}

// =========================================================================================
// Learned coarse-level correction: network, dataset, training, attachment to a context
// =========================================================================================
namespace {

template <class F>
int nguard(std::string* err, int dev, F&& f) {
  try {
    HIPCHK(hipSetDevice(dev));
    return f();
  } catch (const HipError& e) {
    if (err) *err = e.what();
    return TOS_E_HIP;
  } catch (const std::exception& e) {
    if (err) *err = e.what();
    return TOS_E_ARG;
  }
}

void unet_free(tos_unet* N) {
  unet_free_acts(N);
  for (void* p : N->allocs) (void)hipFree(p);
  if (N->h_scal) (void)hipHostFree(N->h_scal);
  if (N->stream) (void)hipStreamDestroy(N->stream);
  delete N;
}

// copy sample `i` of ds into slot b of the batch tensors of A
void load_sample(tos_unet* N, tos::NetActs* A, tos_dataset* ds, int b, int i) {
  const int64_t V = ds->V;
  HIPCHK(hipMemcpyAsync(A->feat + (int64_t)b * V * 8, ds->feat + (int64_t)i * V * 8, V * 8 * sizeof(uint16_t),
                        hipMemcpyDeviceToDevice, N->stream));
  HIPCHK(hipMemcpyAsync(A->tgt + (int64_t)b * V * 3, ds->tgt + (int64_t)i * V * 3, V * 3 * sizeof(float),
                        hipMemcpyDeviceToDevice, N->stream));
  HIPCHK(hipMemcpyAsync(A->mask + (int64_t)b * V, ds->mask + (int64_t)i * V, V, hipMemcpyDeviceToDevice, N->stream));
  HIPCHK(hipMemcpyAsync(A->inv_den + b, ds->inv_den + i, sizeof(float), hipMemcpyDeviceToDevice, N->stream));
}

double read_loss(tos_unet* N) {
  download(N->stream, N->h_scal, N->dscal, sizeof(double));
  return N->h_scal[0];
}

}  // namespace

TOS_EXPORT int tos_unet_create(int c0, int c1, int c2, uint64_t seed, tos_unet** out) {
  if (!out || c0 != 16 || c1 != 32 || c2 != 64) return TOS_E_ARG;
  *out = nullptr;
  tos_unet* N = new tos_unet();
  std::string err;
  const int rc = nguard(&err, 0, [&]() {
    HIPCHK(hipGetDevice(&N->dev));
    HIPCHK(hipStreamCreateWithFlags(&N->stream, hipStreamNonBlocking));
    N->ch[0] = c0; N->ch[1] = c1; N->ch[2] = c2;
    unet_layout(N);
    N->param = nalloc<float>(N->allocs, N->stream, N->nparam);
    N->grad = nalloc<float>(N->allocs, N->stream, N->nparam);
    N->am = nalloc<float>(N->allocs, N->stream, N->nparam);
    N->av = nalloc<float>(N->allocs, N->stream, N->nparam);
    size_t wp = 0;
    for (size_t li = 0; li < N->layers.size(); ++li) {
      tos::NetLayer& L = N->layers[li];
      L.wf = nalloc<uint16_t>(N->allocs, N->stream, (size_t)L.nout * L.kpf);
      if (li > 0) L.wb = nalloc<uint16_t>(N->allocs, N->stream, (size_t)L.cin * L.kpb);
      wp = std::max(wp, (size_t)kWgradGroups * 27 * ((L.cinp + 15) / 16 * 16) * L.nout);
    }
    N->wpart = nalloc<float>(N->allocs, N->stream, wp);
    N->wpart_n = wp;
    N->bpart = nalloc<float>(N->allocs, N->stream, (size_t)kBiasGroups * 64);
    N->dscal = nalloc<double>(N->allocs, N->stream, 8);
    HIPCHK(hipHostMalloc((void**)&N->h_scal, 8 * sizeof(double)));
    std::vector<float> h;
    unet_init_params(N, seed, h);
    upload(N->stream, N->param, h.data(), h.size() * sizeof(float));
    return TOS_OK;
  });
  if (rc) {
    std::fprintf(stderr, "tos_unet_create: %s\n", err.c_str());
    unet_free(N);
    return rc;
  }
  *out = N;
  return TOS_OK;
}

TOS_EXPORT int tos_unet_destroy(tos_unet* N) {
  if (!N) return TOS_E_ARG;
  (void)hipSetDevice(N->dev);
  if (N->stream) (void)hipStreamSynchronize(N->stream);
  unet_free(N);
  return TOS_OK;
}

TOS_EXPORT const char* tos_unet_last_error(const tos_unet* N) { return N ? N->err.c_str() : "null network"; }

TOS_EXPORT int tos_unet_num_params(const tos_unet* N, int64_t* n) {
  if (!N || !n) return TOS_E_ARG;
  *n = N->nparam;
  return TOS_OK;
}

TOS_EXPORT int tos_unet_get_params(tos_unet* N, float* out) {
  if (!N || !out) return TOS_E_ARG;
  return nguard(&N->err, N->dev, [&]() {
    download(N->stream, out, N->param, N->nparam * sizeof(float));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_unet_set_params(tos_unet* N, const float* in) {
  if (!N || !in) return TOS_E_ARG;
  return nguard(&N->err, N->dev, [&]() {
    upload(N->stream, N->param, in, N->nparam * sizeof(float));
    hipLaunchKernelGGL(k_zero<float>, dim3(nblk(N->nparam)), dim3(256), 0, N->stream, N->am, N->nparam);
    hipLaunchKernelGGL(k_zero<float>, dim3(nblk(N->nparam)), dim3(256), 0, N->stream, N->av, N->nparam);
    HIPCHK(hipStreamSynchronize(N->stream));
    N->adam_t = 0;
    N->dirty = true;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_unet_forward(tos_unet* N, int B, int X, int Y, int Z, const float* feat, float* out) {
  if (!N || !feat || !out || B <= 0 || X <= 0 || Y <= 0 || Z <= 0) return TOS_E_ARG;
  return nguard(&N->err, N->dev, [&]() {
    tos::NetActs* A = unet_acts(N, B, tos::NGrid{X, Y, Z}, false);
    const int64_t n = (int64_t)B * X * Y * Z;
    float* st = nalloc<float>(A->mem, N->stream, n * 4);
    upload(N->stream, st, feat, n * 4 * sizeof(float));
    hipLaunchKernelGGL(k_feat_pack, dim3(nblk(n)), dim3(256), 0, N->stream, (const float*)st, n, A->feat);
    unet_forward(N, A, N->stream);
    download(N->stream, out, A->y, n * 3 * sizeof(float));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_unet_loss_grad(tos_unet* N, int B, int X, int Y, int Z, const float* feat, const float* target,
                                  const uint8_t* mask, double* loss, float* grad) {
  if (!N || !feat || !target || !mask || B <= 0 || X <= 0 || Y <= 0 || Z <= 0) return TOS_E_ARG;
  return nguard(&N->err, N->dev, [&]() {
    tos::NetActs* A = unet_acts(N, B, tos::NGrid{X, Y, Z}, true);
    const int64_t n = (int64_t)B * X * Y * Z;
    float* st = nalloc<float>(A->mem, N->stream, n * 4);
    upload(N->stream, st, feat, n * 4 * sizeof(float));
    hipLaunchKernelGGL(k_feat_pack, dim3(nblk(n)), dim3(256), 0, N->stream, (const float*)st, n, A->feat);
    upload(N->stream, A->tgt, target, n * 3 * sizeof(float));
    upload(N->stream, A->mask, mask, n);
    batch_inv_den(N, A, N->stream);
    unet_forward(N, A, N->stream);
    unet_loss(N, A, N->stream);
    unet_backward(N, A, N->stream);
    if (loss) *loss = read_loss(N);
    if (grad) download(N->stream, grad, N->grad, N->nparam * sizeof(float));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_dataset_create(int X, int Y, int Z, int capacity, tos_dataset** out) {
  if (!out || X <= 0 || Y <= 0 || Z <= 0 || capacity <= 0) return TOS_E_ARG;
  *out = nullptr;
  tos_dataset* D = new tos_dataset();
  std::string err;
  const int rc = nguard(&err, 0, [&]() {
    HIPCHK(hipGetDevice(&D->dev));
    HIPCHK(hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking));
    D->g = tos::NGrid{X, Y, Z};
    D->V = (int64_t)X * Y * Z;
    D->cap = capacity;
    D->feat = nalloc<uint16_t>(D->allocs, D->stream, (size_t)capacity * D->V * 8);
    D->tgt = nalloc<float>(D->allocs, D->stream, (size_t)capacity * D->V * 3);
    D->mask = nalloc<uint8_t>(D->allocs, D->stream, (size_t)capacity * D->V);
    D->inv_den = nalloc<float>(D->allocs, D->stream, capacity);
    HIPCHK(hipStreamSynchronize(D->stream));
    return TOS_OK;
  });
  if (rc) {
    std::fprintf(stderr, "tos_dataset_create: %s\n", err.c_str());
    for (void* p : D->allocs) (void)hipFree(p);
    if (D->stream) (void)hipStreamDestroy(D->stream);
    delete D;
    return rc;
  }
  *out = D;
  return TOS_OK;
}

TOS_EXPORT int tos_dataset_destroy(tos_dataset* D) {
  if (!D) return TOS_E_ARG;
  (void)hipSetDevice(D->dev);
  for (void* p : D->allocs) (void)hipFree(p);
  if (D->stream) (void)hipStreamDestroy(D->stream);
  delete D;
  return TOS_OK;
}

TOS_EXPORT int tos_dataset_size(const tos_dataset* D, int* n) {
  if (!D || !n) return TOS_E_ARG;
  *n = D->n;
  return TOS_OK;
}

// sample i: features (4 channels as float), target (3), mask; any output may be null
TOS_EXPORT int tos_dataset_get(tos_dataset* D, int i, float* feat4, float* target, uint8_t* mask) {
  if (!D || i < 0 || i >= D->n) return TOS_E_ARG;
  return nguard(&D->err, D->dev, [&]() {
    const int64_t V = D->V;
    if (feat4) {
      std::vector<uint16_t> h(V * 8);
      download(D->stream, h.data(), D->feat + (int64_t)i * V * 8, V * 8 * sizeof(uint16_t));
      for (int64_t v = 0; v < V; ++v)
        for (int c = 0; c < 4; ++c) {
          const uint32_t u = (uint32_t)h[v * 8 + c] << 16;
          std::memcpy(&feat4[v * 4 + c], &u, 4);
        }
    }
    if (target) download(D->stream, target, D->tgt + (int64_t)i * V * 3, V * 3 * sizeof(float));
    if (mask) download(D->stream, mask, D->mask + (int64_t)i * V, V);
    return TOS_OK;
  });
}

// Adam on random batches of samples [first, first + count); loss_first = loss of the first
// batch, loss_last = mean loss over the last min(50, steps) batches (before their update)
TOS_EXPORT int tos_unet_train(tos_unet* N, tos_dataset* D, int first, int count, int steps, int batch, double lr,
                              uint64_t seed, double* loss_first, double* loss_last) {
  if (!N || !D || first < 0 || count <= 0 || first + count > D->n || steps < 0 || batch <= 0 || !(lr > 0))
    return TOS_E_ARG;
  return nguard(&N->err, N->dev, [&]() {
    HIPCHK(hipStreamSynchronize(D->stream));
    tos::NetActs* A = unet_acts(N, batch, D->g, true);
    uint64_t s = seed * 0x9e3779b97f4a7c15ULL + 1;
    auto next = [&]() {
      uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      return z ^ (z >> 31);
    };
    double lf = 0, ll = 0;
    int nl = 0;
    const int tail = std::min(50, steps);
    for (int it = 0; it < steps; ++it) {
      for (int b = 0; b < batch; ++b) load_sample(N, A, D, b, first + (int)(next() % (uint64_t)count));
      unet_forward(N, A, N->stream);
      unet_loss(N, A, N->stream);
      if (it == 0) lf = read_loss(N);
      if (it >= steps - tail) {
        ll += read_loss(N);
        ++nl;
      }
      unet_backward(N, A, N->stream);
      unet_adam(N, lr, N->stream);
    }
    HIPCHK(hipStreamSynchronize(N->stream));
    if (loss_first) *loss_first = lf;
    if (loss_last) *loss_last = nl ? ll / nl : 0.0;
    return TOS_OK;
  });
}

// mean per-sample loss of the network over samples [first, first + count)
TOS_EXPORT int tos_unet_eval(tos_unet* N, tos_dataset* D, int first, int count, double* mean_loss) {
  if (!N || !D || !mean_loss || first < 0 || count <= 0 || first + count > D->n) return TOS_E_ARG;
  return nguard(&N->err, N->dev, [&]() {
    HIPCHK(hipStreamSynchronize(D->stream));
    const int bs = std::min(count, 8);
    double tot = 0;
    for (int i0 = 0; i0 < count; i0 += bs) {
      const int b = std::min(bs, count - i0);
      tos::NetActs* A = unet_acts(N, b, D->g, true);
      for (int j = 0; j < b; ++j) load_sample(N, A, D, j, first + i0 + j);
      unet_forward(N, A, N->stream);
      unet_loss(N, A, N->stream);
      tot += read_loss(N) * b;
    }
    *mean_loss = tot / count;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_ctx_learn_attach(tos_ctx* c, tos_unet* net, int level, int mode) {
  return guard(c, [&]() {
    REQUIRE(mode == 0 || (net && mode == 1), "bad mode");
    if (mode == 0) {
      if (c->learn) { c->learn->mode = 0; c->learn->net = nullptr; }
      return TOS_OK;
    }
    REQUIRE(level >= 1 && level + 1 < (int)c->L.size(), "learned level must have a coarser level below it");
    REQUIRE(!c->skip, "the learned correction takes dyadic hierarchies only");
    learn_prepare(*c, level);
    LearnState& S = *c->learn;
    const Level& L = c->L[level];
    S.acts = unet_acts(net, 1, tos::NGrid{L.v.nx, L.v.ny, L.v.nz}, false);
    S.net = net;
    S.mode = 1;
    S.paused = false;
    S.dry = 0;
    S.skip = 0;
    hipLaunchKernelGGL(k_zero<double>, dim3(1), dim3(64), 0, c->stream, c->scal + S_NN_TRIED, (int64_t)3);
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOS_OK;
  });
}

TOS_EXPORT int tos_ctx_learn_log(tos_ctx* c, tos_dataset* ds, int level, int n_inner, int stride) {
  return guard(c, [&]() {
    if (!ds) {
      if (c->learn) c->learn->ds = nullptr;
      return TOS_OK;
    }
    REQUIRE(level >= 1 && level + 1 < (int)c->L.size(), "logged level must have a coarser level below it");
    REQUIRE(!c->skip, "the learned correction takes dyadic hierarchies only");
    REQUIRE(n_inner >= 1 && stride >= 1, "bad arguments");
    const Level& L = c->L[level];
    REQUIRE(ds->g.X == L.v.nx && ds->g.Y == L.v.ny && ds->g.Z == L.v.nz, "dataset grid differs from the level");
    REQUIRE(!c->learn || c->learn->level == level || c->learn->mode == 0, "network attached on another level");
    learn_prepare(*c, level);
    c->learn->ds = ds;
    c->learn->n_inner = n_inner;
    c->learn->stride = stride;
    c->learn->calls = 0;
    return TOS_OK;
  });
}

TOS_EXPORT int tos_ctx_learn_stats(tos_ctx* c, int* tried, int* accepted, double* mean_step) {
  return guard(c, [&]() {
    const double* hs = read_scal(*c);
    const bool on = c->learn != nullptr;
    if (tried) *tried = on ? (int)hs[S_NN_TRIED] : 0;
    if (accepted) *accepted = on ? (int)hs[S_NN_KEPT] : 0;
    if (mean_step) *mean_step = on && hs[S_NN_KEPT] > 0 ? hs[S_NN_STEP] / hs[S_NN_KEPT] : 0.0;
    return TOS_OK;
  });
}

// ---------------------------------------------------------------------------------------
// OC update without allocation: the bisection's block partials and scalars live in a
// file-scope device array, one copy per GPU, whose address is taken once per device.
// ---------------------------------------------------------------------------------------
namespace {
constexpr int kNoAllocChunk = 256 * 32;   // elements per block of the flat OC kernels
constexpr int kNoAllocParts = 8192;       // block partials the scratch holds
static_assert((int64_t)kNoAllocChunk * kNoAllocParts == TOS_NOALLOC_MAX_ELEMENTS, "capacity");
// [0, 2*kNoAllocParts): partials (max |s| and active counts, then trial volumes);
// [2*kNoAllocParts, +S_COUNT): the bisection scalars (S_OC_*, S_VOL = active count)
static __device__ __align__(256) double scratch[2 * kNoAllocParts + S_COUNT];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch (hipGetSymbolAddress once per device)
hipError_t scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}
}  // namespace

TOS_EXPORT int tos_oc_update_noalloc(int64_t n, const uint8_t* active, const uint8_t* shell, const float* rho,
                                     const float* sens, float* rho_out, double vgoal, double move, double damp,
                                     double rho_min, void* stream, double* g_out, int* steps_out) {
  if (n < 0 || !rho || !sens || !rho_out || !(move > 0) || !(damp > 0)) return TOS_E_ARG;
  if (n > TOS_NOALLOC_MAX_ELEMENTS) return TOS_E_TOO_LARGE;
  if (n == 0) {
    if (g_out) *g_out = 0.0;
    if (steps_out) *steps_out = 0;
    return TOS_OK;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return TOS_E_HIP;
  double* S = nullptr;
  if (scratch_address(dev, &S) != hipSuccess) return TOS_E_HIP;
  double* part = S;
  double* scal = S + 2 * kNoAllocParts;
  const hipStream_t st = (hipStream_t)stream;
  const int nb = (int)std::max<int64_t>(1, (n + kNoAllocChunk - 1) / kNoAllocChunk);
  // every scratch value a later kernel reads is written earlier in this call: the partials
  // by the pass before their fold, the scalars by k_oc_init_flat
  hipLaunchKernelGGL(k_absmax_count_flat, dim3(nb), dim3(256), 0, st, n, (int64_t)kNoAllocChunk, active, sens, nb,
                     part);
  hipLaunchKernelGGL(k_oc_init_flat, dim3(1), dim3(256), 0, st, (const double*)part, nb, scal);
  for (int it = 0; it < 31; ++it) {
    hipLaunchKernelGGL(k_oc_trial_flat, dim3(nb), dim3(256), 0, st, n, (int64_t)kNoAllocChunk, active, shell, rho,
                       sens, move, damp, rho_min, (const double*)scal, part);
    hipLaunchKernelGGL(k_oc_decide, dim3(1), dim3(256), 0, st, (const double*)part, nb, 0.0, 1, vgoal, 31, -1.0,
                       scal);
  }
  hipLaunchKernelGGL(k_oc_apply_flat, dim3(nblk(n)), dim3(256), 0, st, n, active, shell, rho, sens, rho_out, move,
                     damp, rho_min, (const double*)scal);
  if (hipGetLastError() != hipSuccess) return TOS_E_HIP;
  if (g_out || steps_out) {
    double h[S_COUNT];
    if (hipMemcpyAsync(h, scal, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess) return TOS_E_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return TOS_E_HIP;
    if (g_out) *g_out = h[S_OC_G];
    if (steps_out) *steps_out = (int)h[S_OC_STEPS];
  }
  return TOS_OK;
}

// ---------------------------------------------------------------------------------------
// Matrix-core element products against the matrix-free operator (a measurement; no
// counterpart in the reference).  On the fine level and a seeded u (free nodes), times
// `reps` launches each of y = K u in FP32 by the VALU kernel and by element products on
// v_mfma_f32_16x16x4f32 with the gathers and products only, then with the float-atomic
// scatter into the nodes; out = {VALU ms, MFMA products ms, MFMA + scatter ms, relative
// difference of the scattered result from the VALU one over the free nodes}.
TOS_EXPORT int tos_probe_element_mfma(tos_ctx* c, int reps, double* out) {
  return guard(c, [&]() {
    REQUIRE(out && reps >= 1, "bad arguments");
    if (!c->op_ready) { c->err = "update_operator first"; return TOS_E_STATE; }
    Ctx& C = *c;
    Level& L0 = C.L[0];
    const LevelView& v = L0.v;
    std::vector<int4> tiles;  // runs of up to 16 active elements along x
    for (int z = 0; z < v.ez; ++z)
      for (int y = 0; y < v.ey; ++y) {
        int x = 0;
        while (x < v.ex) {
          if (!L0.h_eact[x + (size_t)v.ex * (y + (size_t)v.ey * z)]) { ++x; continue; }
          int x1 = x;
          while (x1 < v.ex && x1 - x < 16 && L0.h_eact[x1 + (size_t)v.ex * (y + (size_t)v.ey * z)]) ++x1;
          tiles.push_back(make_int4(x, x1 - x, y, z));
          x = x1;
        }
      }
    const int nt = (int)tiles.size();
    const unsigned nbt = (unsigned)std::max(1, std::min(4 * 2048, (nt + 3) / 4));
    std::vector<void*> mem;
    int4* d_t = (int4*)raw_alloc(mem, C.stream, std::max<size_t>(1, tiles.size()) * sizeof(int4));
    float* uu = (float*)raw_alloc(mem, C.stream, 3 * v.NP * sizeof(float));
    float* y1 = (float*)raw_alloc(mem, C.stream, 3 * v.NP * sizeof(float));
    float* y2 = (float*)raw_alloc(mem, C.stream, 3 * v.NP * sizeof(float));
    float* chk = (float*)raw_alloc(mem, C.stream, (size_t)nbt * 256 * sizeof(float));
    auto release = [&]() {
      (void)hipStreamSynchronize(C.stream);
      for (void* p : mem) (void)hipFree(p);
    };
    try {
      if (nt) upload(C.stream, d_t, tiles.data(), tiles.size() * sizeof(int4));
      hipLaunchKernelGGL(k_probe_fill, dim3(cblk(L0.n_lin)), dim3(256), 0, C.stream, (const int4*)L0.cl_lin,
                         L0.n_lin, v.NP, uu, 12345u);
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      auto timed = [&](auto launch) {
        launch();  // warm-up
        HIPCHK(hipEventRecord(e0, C.stream));
        for (int i = 0; i < reps; ++i) launch();
        HIPCHK(hipEventRecord(e1, C.stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / reps;
      };
      out[0] = timed([&]() {
        hipLaunchKernelGGL((k_fine_apply<float, 0, float>), dim3(cblk(L0.n_lin)), dim3(256), 0, C.stream, v,
                           (const int4*)L0.cl_lin, L0.n_lin, (const float*)C.w32, (const float*)C.Kc32,
                           (const float*)uu, (const float*)nullptr, y1, (double*)nullptr, (float*)nullptr);
      });
      out[1] = timed([&]() {
        hipLaunchKernelGGL(k_probe_elem_mfma<0>, dim3(nbt), dim3(256), 0, C.stream, v, (const int4*)d_t, nt,
                           (const float*)C.w32, (const float*)C.Ke32, (const float*)uu, y2, chk);
      });
      out[2] = timed([&]() {
        hipLaunchKernelGGL(k_zero<float>, dim3(nblk(3 * v.NP)), dim3(256), 0, C.stream, y2, 3 * v.NP);
        hipLaunchKernelGGL(k_probe_elem_mfma<1>, dim3(nbt), dim3(256), 0, C.stream, v, (const int4*)d_t, nt,
                           (const float*)C.w32, (const float*)C.Ke32, (const float*)uu, y2, chk);
      });
      hipLaunchKernelGGL(k_probe_diff, dim3(C.nb0), dim3(256), 0, C.stream, (const int4*)L0.cl_lin, L0.n_lin, v.NP,
                         (const float*)y2, (const float*)y1, C.partial);
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, C.stream, (const double*)C.partial, (int)C.nb0, 1, S_TMP0,
                         S_TMP1, OP_NONE, C.scal);
      LAUNCH_CHECK();
      const double* h = read_scal(C);
      out[3] = h[S_TMP1] > 0 ? std::sqrt(h[S_TMP0] / h[S_TMP1]) : 0.0;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    } catch (...) {
      release();
      throw;
    }
    release();
    return TOS_OK;
  });
}
