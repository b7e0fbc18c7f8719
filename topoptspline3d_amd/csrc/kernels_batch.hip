// kernels_batch.hip -- multigrid and FCG kernels for K load cases solved together.
//
// Layout: case-interleaved vectors, v[(c*NP + p)*K + k] (component c, padded node p,
// case k), so the K values of one node and component are contiguous and every stencil,
// density and template-matrix read is shared by the K lanes of a node (one load serves
// K cases).  Thread mapping on a chunk list: 64*KP threads per chunk (KP = K rounded up
// to a power of two), thread i of a chunk -> node lane i / KP, case i % KP (idle when
// >= K).  Block size BS is a template parameter picked per kernel and level by the
// autotuner.  Per-case block partials are written as partial[block * K + k].
#include "tos_internal.hpp"

namespace tos {

template <int BS>
__device__ __forceinline__ bool bnode(const int4* __restrict__ cl, int nch, int s, int K, int KP, NodeRef& n,
                                      int& k) {
  const int64_t g = xcd_block() * BS + threadIdx.x;
  const int64_t tpc = 64LL * KP;
  const int64_t ch = g / tpc;
  if (ch >= nch) return false;
  const int i = (int)(g - ch * tpc);
  const int lane = i / KP;
  k = i - lane * KP;
  if (k >= K) return false;
  const int4 c = cl[ch];
  if (lane >= (c.y >> 16)) return false;
  n.x = (c.y & 0xffff) + s * lane;
  n.y = c.z;
  n.z = c.w;
  n.p = (int64_t)c.x + (int64_t)s * lane;
  n.slot = (int)ch * 64 + lane;
  return true;
}

// per-case block sums: partial[blockIdx.x * K + k]; every thread calls (v = 0 when idle)
template <int BS>
__device__ __forceinline__ void bpartial(double v, int K, int KP, double* __restrict__ partial) {
  __shared__ double sh[BS];
  sh[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x < K && threadIdx.x < KP) {
    double s = 0;
    for (int j = threadIdx.x; j < BS; j += KP) s += sh[j];
    partial[(int64_t)blockIdx.x * K + threadIdx.x] = s;
  }
  __syncthreads();
}

// neighbourhood of node p, case k; neighbours outside MASK read as 0 without a load
// (one per-lane base address; every neighbour and component offset is uniform, so each
// load costs one address add instead of a 64-bit multiply chain)
template <class T, unsigned MASK = 0x7ffffffu>
__device__ __forceinline__ void load_nbhd_b(const LevelView& L, const T* __restrict__ u, int64_t p, int K, int k,
                                            T (&un)[27][3]) {
  const T* __restrict__ b = u + p * K + k;
  const int64_t cs = L.NP * K;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int64_t o =
        ((int64_t)(q % 3 - 1) + (int64_t)L.NX * ((q / 3) % 3 - 1 + (int64_t)L.NY * (q / 9 - 1))) * K;
#pragma unroll
    for (int c = 0; c < 3; ++c) un[q][c] = ((MASK >> q) & 1u) ? b[c * cs + o] : T(0);
  }
}

// MODE 0: y = K u with per-case partials of u.y;  MODE 1: y = f - K u (+ float copy y32).
// TU: type of u (float: the FCG search direction, with FP64 arithmetic)
template <class T, int MODE, int BS, class TU = T>
__global__ __launch_bounds__(BS) void k_fine_apply_b(LevelView L, const int4* __restrict__ cl, int nch, int K, int KP,
                                                     const T* __restrict__ w, const T* __restrict__ Ke,
                                                     const TU* __restrict__ u, const T* __restrict__ f,
                                                     T* __restrict__ y, double* __restrict__ partial,
                                                     float* __restrict__ y32) {
  NodeRef n;
  int k = 0;
  const bool ok = bnode<BS>(cl, nch, 1, K, KP, n, k);
  double d = 0;
  if (ok) {
    TU un[27][3];
    T wa[8];
    load_nbhd_b(L, u, n.p, K, k, un);
    load_weights(L, w, n.x, n.y, n.z, wa);
    T acc[3] = {0, 0, 0}, D[9];
    fine_row<T, true, 0x7ffffffu, TU>(wa, Ke, un, acc, D);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = (c * L.NP + n.p) * K + k;
      if (MODE == 0) {
        y[i] = acc[c];
        d += (double)un[13][c] * (double)acc[c];
      } else {
        const T v = f[i] - acc[c];
        y[i] = v;
        if (y32) y32[i] = (float)v;
      }
    }
  }
  if (partial) bpartial<BS>(d, K, KP, partial);
}

// ZC >= 0: colour ZC of a first forward sweep from u = 0 (u is not cleared beforehand:
// only earlier colours' neighbours are read, the node's own start value is 0)
template <class T, bool BACKWARD, int BS, int ZC = -1>
__global__ __launch_bounds__(BS) void k_fine_gs_b(LevelView L, const int4* __restrict__ cl, int nch, int K, int KP,
                                                  const T* __restrict__ w, const T* __restrict__ Ke,
                                                  T* __restrict__ u, const T* __restrict__ f) {
  constexpr unsigned mask = nb_mask_of(ZC);
  NodeRef n;
  int k = 0;
  if (!bnode<BS>(cl, nch, 2, K, KP, n, k)) return;
  T un[27][3], wa[8];
  load_nbhd_b<T, ZC < 0 ? 0x7ffffffu : mask>(L, u, n.p, K, k, un);
  load_weights(L, w, n.x, n.y, n.z, wa);
  T s[3] = {0, 0, 0};
  T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  fine_row<T, false, mask>(wa, Ke, un, s, D);
  const int64_t i0 = n.p * K + k, cs = L.NP * K;
  T v0 = ZC < 0 ? un[13][0] : T(0), v1 = ZC < 0 ? un[13][1] : T(0), v2 = ZC < 0 ? un[13][2] : T(0);
  block_relax<T, BACKWARD>(f[i0] - s[0], f[cs + i0] - s[1], f[2 * cs + i0] - s[2], D, v0, v1, v2);
  u[i0] = v0;
  u[cs + i0] = v1;
  u[2 * cs + i0] = v2;
}

// residual after a first forward sweep from zero, nodes of colour ZC: later-colour
// neighbours and the node's own block only (as k_fine_resid_z, per case)
template <class T, int ZC, int BS>
__global__ __launch_bounds__(BS) void k_fine_resid_z_b(LevelView L, const int4* __restrict__ cl, int nch, int K,
                                                       int KP, const T* __restrict__ w, const T* __restrict__ Ke,
                                                       const T* __restrict__ u, T* __restrict__ r) {
  constexpr unsigned after = nb_mask_after(ZC);
  NodeRef n;
  int k = 0;
  if (!bnode<BS>(cl, nch, 2, K, KP, n, k)) return;
  T un[27][3], wa[8];
  load_nbhd_b<T, after | (1u << 13)>(L, u, n.p, K, k, un);
  load_weights(L, w, n.x, n.y, n.z, wa);
  T s[3] = {0, 0, 0};
  T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  fine_row<T, false, after>(wa, Ke, un, s, D);
  const T v1 = un[13][1], v2 = un[13][2];
  const int64_t i0 = n.p * K + k, cs = L.NP * K;
  r[i0] = -D[1] * v1 - D[2] * v2 - s[0];
  r[cs + i0] = -D[5] * v2 - s[1];
  r[2 * cs + i0] = -s[2];
}

template <class T, int MODE, int BS>
__global__ __launch_bounds__(BS) void k_st_apply_b(LevelView L, const int4* __restrict__ cl, int nch, int K, int KP,
                                                   const T* __restrict__ st, int64_t NS, const T* __restrict__ u,
                                                   const T* __restrict__ f, T* __restrict__ y) {
  NodeRef n;
  int k = 0;
  if (!bnode<BS>(cl, nch, 2, K, KP, n, k)) return;
  const int64_t cs = L.NP * K;
  const StFull<T> S{st, NS};
  T un[27][3];
  load_nbhd_b(L, u, n.p, K, k, un);
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const T u0 = un[q][0], u1 = un[q][1], u2 = un[q][2];
    const int64_t sl = n.slot;
    a0 = fma(S.get(sl, q, 0), u0, a0); a0 = fma(S.get(sl, q, 1), u1, a0); a0 = fma(S.get(sl, q, 2), u2, a0);
    a1 = fma(S.get(sl, q, 3), u0, a1); a1 = fma(S.get(sl, q, 4), u1, a1); a1 = fma(S.get(sl, q, 5), u2, a1);
    a2 = fma(S.get(sl, q, 6), u0, a2); a2 = fma(S.get(sl, q, 7), u1, a2); a2 = fma(S.get(sl, q, 8), u2, a2);
  }
  const int64_t i0 = n.p * K + k;
  if (MODE == 0) {
    y[i0] = a0; y[cs + i0] = a1; y[2 * cs + i0] = a2;
  } else {
    y[i0] = f[i0] - a0; y[cs + i0] = f[cs + i0] - a1; y[2 * cs + i0] = f[2 * cs + i0] - a2;
  }
}

template <class T, bool BACKWARD, int BS>
__global__ __launch_bounds__(BS) void k_st_gs_b(LevelView L, const int4* __restrict__ cl, int nch, int64_t slot0,
                                                int K, int KP, const T* __restrict__ st, int64_t NS,
                                                T* __restrict__ u, const T* __restrict__ f) {
  NodeRef n;
  int k = 0;
  if (!bnode<BS>(cl, nch, 2, K, KP, n, k)) return;
  const int64_t slot = slot0 + n.slot, cs = L.NP * K;
  const StFull<T> S{st, NS};
  T un[27][3];
  load_nbhd_b(L, u, n.p, K, k, un);
  T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    if (q == 13) continue;
    const T u0 = un[q][0], u1 = un[q][1], u2 = un[q][2];
    s0 = fma(S.get(slot, q, 0), u0, s0); s0 = fma(S.get(slot, q, 1), u1, s0); s0 = fma(S.get(slot, q, 2), u2, s0);
    s1 = fma(S.get(slot, q, 3), u0, s1); s1 = fma(S.get(slot, q, 4), u1, s1); s1 = fma(S.get(slot, q, 5), u2, s1);
    s2 = fma(S.get(slot, q, 6), u0, s2); s2 = fma(S.get(slot, q, 7), u1, s2); s2 = fma(S.get(slot, q, 8), u2, s2);
  }
  T D[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) D[e] = S.get(slot, 13, e);
  const int64_t i0 = n.p * K + k;
  T v0 = un[13][0], v1 = un[13][1], v2 = un[13][2];
  block_relax<T, BACKWARD>(f[i0] - s0, f[cs + i0] - s1, f[2 * cs + i0] - s2, D, v0, v1, v2);
  u[i0] = v0;
  u[cs + i0] = v1;
  u[2 * cs + i0] = v2;
}

template <class T, int BS>
__global__ __launch_bounds__(BS) void k_restrict_b(LevelView F, LevelView C, const int4* __restrict__ clc, int nch,
                                                   int K, int KP, const T* __restrict__ rf, T* __restrict__ fc) {
  NodeRef n;
  int k = 0;
  if (!bnode<BS>(clc, nch, 1, K, KP, n, k)) return;
  const int64_t fcs = F.NP * K, ccs = C.NP * K;
  T a0 = 0, a1 = 0, a2 = 0;
  const T* __restrict__ b = rf + F.pnode(2 * n.x, 2 * n.y, 2 * n.z) * K + k;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
    const T wgt = T((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0));
    const int64_t o = (dx + (int64_t)F.NX * (dy + (int64_t)F.NY * dz)) * K;  // (uniform)
    a0 += wgt * b[o]; a1 += wgt * b[fcs + o]; a2 += wgt * b[2 * fcs + o];
  }
  const int64_t i0 = n.p * K + k;
  fc[i0] = a0; fc[ccs + i0] = a1; fc[2 * ccs + i0] = a2;
}

template <class T, int BS>
__global__ __launch_bounds__(BS) void k_prolong_add_b(LevelView F, LevelView C, const int4* __restrict__ clf,
                                                      int nch, int K, int KP, const T* __restrict__ uc,
                                                      T* __restrict__ uf) {
  NodeRef n;
  int k = 0;
  if (!bnode<BS>(clf, nch, 1, K, KP, n, k)) return;
  // all 8 corners (inside the padding) with weight 0 where a fine coordinate is even: the
  // loads are issued together instead of in a divergent loop
  const int x = n.x, y = n.y, z = n.z;
  const T hx = (x & 1) ? T(0.5) : T(0), hy = (y & 1) ? T(0.5) : T(0), hz = (z & 1) ? T(0.5) : T(0);
  const T wx[2] = {T(1) - hx, hx}, wy[2] = {T(1) - hy, hy}, wz[2] = {T(1) - hz, hz};
  const int64_t ccs = C.NP * K, fcs = F.NP * K;
  const T* __restrict__ b = uc + C.pnode(x >> 1, y >> 1, z >> 1) * K + k;
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = q & 1, j = (q >> 1) & 1, l = q >> 2;
    const T wgt = wx[i] * wy[j] * wz[l];
    const int64_t o = (i + (int64_t)C.NX * (j + (int64_t)C.NY * l)) * K;  // (uniform)
    a0 = fma(wgt, b[o], a0); a1 = fma(wgt, b[ccs + o], a1); a2 = fma(wgt, b[2 * ccs + o], a2);
  }
  const int64_t i0 = n.p * K + k;
  uf[i0] += a0; uf[fcs + i0] += a1; uf[2 * fcs + i0] += a2;
}

template <class T, int BS>
__global__ __launch_bounds__(BS) void k_zero_nodes_b(const int4* __restrict__ cl, int nch, int64_t NP, int K, int KP,
                                                     T* __restrict__ u) {
  NodeRef n;
  int k = 0;
  if (!bnode<BS>(cl, nch, 1, K, KP, n, k)) return;
  const int64_t i0 = n.p * K + k, cs = NP * K;
  u[i0] = T(0); u[cs + i0] = T(0); u[2 * cs + i0] = T(0);
}

// coarsest level: u[:, k] = -cinv f[:, k] (cinv holds -A^{-1}), thread per (row, case)
// u_c = -M f_c for K cases: block = one row of M; thread t = j*K + k (j < J = 256/K) sums
// the columns j, j + J, ... for case k, the J sums are added in order j
template <class T>
__global__ __launch_bounds__(256) void k_coarse_solve_b(LevelView L, const int64_t* __restrict__ cidx, int n, int K,
                                                        const double* __restrict__ cinv, const T* __restrict__ f,
                                                        T* __restrict__ u) {
  __shared__ double sh[256];
  const int row = blockIdx.x;
  const int J = 256 / K, t = threadIdx.x, jl = t / K, k = t % K;
  const int64_t cs = L.NP * K;
  double s = 0;
  if (jl < J)
    for (int j = jl; j < n; j += J) s += cinv[(int64_t)row * n + j] * (double)f[(j % 3) * cs + cidx[j / 3] * K + k];
  sh[t] = s;
  __syncthreads();
  if (t < K) {
    double a = 0;
    for (int q = 0; q < J; ++q) a += sh[q * K + t];
    u[(row % 3) * cs + cidx[row / 3] * K + t] = (T)(-a);
  }
}

// ---- batched FCG (FP64 outer vectors, per-case scalars sc[slot * K + k]) ----------------
enum { B_RZ = 0, B_PQ, B_ALPHA, B_BETA, B_RN2, B_ZR, B_ZQ, B_NF2, B_ACTIVE, B_RTOL2, B_COUNT = 12 };

template <int BS>
__global__ __launch_bounds__(BS) void k_pcgb_ur(const int4* __restrict__ cl, int nch, int64_t NP, int K, int KP,
                                                double* __restrict__ u, double* __restrict__ r,
                                                const float* __restrict__ p, const double* __restrict__ q,
                                                const double* __restrict__ sc, float* __restrict__ r32,
                                                double* __restrict__ partial) {
  NodeRef n;
  int k = 0;
  const bool ok = bnode<BS>(cl, nch, 1, K, KP, n, k);
  double d = 0;
  if (ok) {
    const double a = sc[B_ALPHA * K + k];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = (c * NP + n.p) * K + k;
      u[i] += a * (double)p[i];
      const double v = r[i] - a * q[i];
      r[i] = v;
      r32[i] = (float)v;
      d += v * v;
    }
  }
  bpartial<BS>(d, K, KP, partial);
}

// partials of <z, r> (column 0) and <z, q> (column 1, when q != null): partial[2][nb][K]
template <int BS>
__global__ __launch_bounds__(BS) void k_pcgb_zdots(const int4* __restrict__ cl, int nch, int64_t NP, int K, int KP,
                                                   const float* __restrict__ z, const double* __restrict__ r,
                                                   const double* __restrict__ q, double* __restrict__ partial) {
  NodeRef n;
  int k = 0;
  const bool ok = bnode<BS>(cl, nch, 1, K, KP, n, k);
  double a = 0, b = 0;
  if (ok) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = (c * NP + n.p) * K + k;
      const double zi = (double)z[i];
      a += zi * r[i];
      if (q) b += zi * q[i];
    }
  }
  bpartial<BS>(a, K, KP, partial);
  if (q) bpartial<BS>(b, K, KP, partial + (int64_t)gridDim.x * K);
}

// p = z + beta p (first: p = z), formed in FP64 and stored in FP32
template <int BS>
__global__ __launch_bounds__(BS) void k_pcgb_p(const int4* __restrict__ cl, int nch, int64_t NP, int K, int KP,
                                               float* __restrict__ p, const float* __restrict__ z,
                                               const double* __restrict__ sc, int first) {
  NodeRef n;
  int k = 0;
  if (!bnode<BS>(cl, nch, 1, K, KP, n, k)) return;
  const double b = first ? 0.0 : sc[B_BETA * K + k];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t i = (c * NP + n.p) * K + k;
    p[i] = (float)((double)z[i] + b * (double)p[i]);
  }
}

// generic per-case partials of <a, b> (+ <a2, b2> into the second half)
template <int BS>
__global__ __launch_bounds__(BS) void k_dotb(const int4* __restrict__ cl, int nch, int64_t NP, int K, int KP,
                                             const double* __restrict__ a, const double* __restrict__ b,
                                             const double* __restrict__ a2, const double* __restrict__ b2,
                                             double* __restrict__ partial) {
  NodeRef n;
  int k = 0;
  const bool ok = bnode<BS>(cl, nch, 1, K, KP, n, k);
  double s0 = 0, s1 = 0;
  if (ok) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = (c * NP + n.p) * K + k;
      s0 += a[i] * b[i];
      if (a2) s1 += a2[i] * b2[i];
    }
  }
  bpartial<BS>(s0, K, KP, partial);
  if (a2) bpartial<BS>(s1, K, KP, partial + (int64_t)gridDim.x * K);
}

// stage 1: group g (grid) folds blocks [g*seg, (g+1)*seg) of `ncol` columns of per-case
// partials part[col][nb][K] into out[col][g][K].  Thread t = j*K + k (j < J = 256/K) sums
// blocks b0 + j, b0 + j + J, ... of case k (rows of K cases are contiguous, so a group of
// K threads reads one coalesced row); the J sums of a case are then added in order j.
__global__ __launch_bounds__(256) void k_foldb1(const double* __restrict__ part, int nb, int K, int ncol,
                                                double* __restrict__ out) {
  __shared__ double sh[256];
  const int G = gridDim.x, g = blockIdx.x;
  const int seg = (nb + G - 1) / G;
  const int b0 = g * seg, b1 = min(nb, b0 + seg);
  const int J = 256 / K, t = threadIdx.x, j = t / K, k = t % K;
  for (int col = 0; col < ncol; ++col) {
    double s = 0;
    if (j < J)
      for (int b = b0 + j; b < b1; b += J) s += part[((int64_t)col * nb + b) * K + k];
    sh[t] = s;
    __syncthreads();
    if (t < K) {
      double a = 0;
      for (int q = 0; q < J; ++q) a += sh[q * K + t];
      out[((int64_t)col * G + g) * K + t] = a;
    }
    __syncthreads();
  }
}

// stage 2: sc[slot_c * K + k] = sum_g out[c][g][K], then the FCG scalar update `op`
enum { OPB_NONE = 0, OPB_ALPHA, OPB_BETA, OPB_RN2, OPB_INIT };
// (thread t = j*K + k sums groups j, j + J, ... of case k; the J sums are added in order j)
__global__ __launch_bounds__(256) void k_foldb2(const double* __restrict__ st, int G, int K, int ncol, int slot0,
                                                int slot1, int op, double* __restrict__ sc) {
  __shared__ double sh[256];
  const int J = 256 / K, t = threadIdx.x, j = t / K, kk = t % K;
  for (int col = 0; col < ncol; ++col) {
    double s = 0;
    if (j < J)
      for (int g = j; g < G; g += J) s += st[((int64_t)col * G + g) * K + kk];
    sh[t] = s;
    __syncthreads();
    if (t < K) {
      double a = 0;
      for (int q = 0; q < J; ++q) a += sh[q * K + t];
      sc[(col == 0 ? slot0 : slot1) * K + t] = a;
    }
    __syncthreads();
  }
  const int k = t;
  if (k >= K) return;
  const bool act = sc[B_ACTIVE * K + k] != 0.0;
  if (op == OPB_ALPHA) sc[B_ALPHA * K + k] = act && sc[B_PQ * K + k] != 0.0 ? sc[B_RZ * K + k] / sc[B_PQ * K + k] : 0.0;
  if (op == OPB_RN2) sc[B_ACTIVE * K + k] = act && sc[B_RN2 * K + k] > sc[B_RTOL2 * K + k] ? 1.0 : 0.0;
  if (op == OPB_BETA) {
    sc[B_BETA * K + k] = sc[B_RZ * K + k] != 0.0 ? -sc[B_ALPHA * K + k] * sc[B_ZQ * K + k] / sc[B_RZ * K + k] : 0.0;
    sc[B_RZ * K + k] = sc[B_ZR * K + k];
  }
}

// interleave / de-interleave between case-major FP64 (K x 3 x NP) and the batched layout
__global__ void k_interleave(const int4* __restrict__ cl, int nch, int64_t NP, int K, const double* __restrict__ a,
                             double* __restrict__ b, int to_batched) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = 64LL * K;
  const int64_t ch = t / per;
  if (ch >= nch) return;
  const int i = (int)(t - ch * per), lane = i / K, k = i % K;
  const int4 c = cl[ch];
  if (lane >= (c.y >> 16)) return;
  const int64_t p = (int64_t)c.x + lane;
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    const int64_t ib = (cc * NP + p) * K + k, ic = (int64_t)k * 3 * NP + cc * NP + p;
    if (to_batched) b[ib] = a[ic];
    else b[ic] = a[ib];
  }
}

}  // namespace tos
