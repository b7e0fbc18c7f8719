// kernels_mg.hip -- multigrid kernels for gfx950 (wave64).
//
// Fine level (matrix-free, "on-the-fly assembly"): each thread owns one node, keeps the
// 27 neighbouring displacements in VGPRs and sums rho_e^p * Ke[loc_i][loc_j] u_j over the
// 8 incident elements (update_residual_OTFA_WS_kernel grid/Grid.cu:2277-2333,
// gs_relax_OTFA_WS_kernel :1700-1842).  Ke is read through a uniform pointer, so the
// compiler serves it from the scalar cache into SGPR operands of the FMAs.
// Coarse levels: stored 27x9 Galerkin stencil, one thread per node (gs_relax_kernel
// :1455-1566, update_residual_kernel_1 :2127-2224).
#include "tos_internal.hpp"

namespace tos {

__device__ __forceinline__ int64_t thread_id() {
  return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}

// neighbour offset of stencil slot k in the padded node array
__device__ __forceinline__ int64_t noff(const LevelView& L, int k) {
  return (int64_t)(k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
}

// ---------------------------------------------------------------------------------------
// Fine level: y = K(w) u  (MODE 0)  or  y = f - K(w) u  (MODE 1), rows of non-free nodes 0.
// ---------------------------------------------------------------------------------------
template <class T>
struct FineNode {
  T un[27][3];
};

template <class T>
__device__ __forceinline__ void load_nbhd(const LevelView& L, const T* __restrict__ u, int64_t p,
                                          T (&un)[27][3]) {
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    int64_t q = p + (int64_t)(k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) un[k][c] = u[c * L.NP + q];
  }
}

// acc += sum_a w_a sum_j Ke[li][j] u_j ; with SELF = false the node itself (j = li) is
// skipped and its diagonal block accumulated into D instead.
template <class T, bool SELF>
__device__ __forceinline__ void fine_row(const LevelView& L, const T* __restrict__ w,
                                         const T* __restrict__ Ke, int x, int y, int z,
                                         const T (&un)[27][3], T (&acc)[3], T (&D)[9]) {
  const int64_t e0 = L.pelem(x - 1, y - 1, z - 1);
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ax = a & 1, ay = (a >> 1) & 1, az = a >> 2;
    const T we = w[e0 + ax + (int64_t)L.EX * (ay + (int64_t)L.EY * az)];
    const int li = 7 - a;
    T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int nb = (ax + (j & 1)) + 3 * (ay + ((j >> 1) & 1)) + 9 * (az + (j >> 2));
      if (!SELF && j == li) continue;
      const T* kr = Ke + (3 * li) * 24 + 3 * j;
      t0 += kr[0] * un[nb][0] + kr[1] * un[nb][1] + kr[2] * un[nb][2];
      t1 += kr[24] * un[nb][0] + kr[25] * un[nb][1] + kr[26] * un[nb][2];
      t2 += kr[48] * un[nb][0] + kr[49] * un[nb][1] + kr[50] * un[nb][2];
    }
    acc[0] += we * t0;
    acc[1] += we * t1;
    acc[2] += we * t2;
    if (!SELF) {
      const T* kd = Ke + (3 * li) * 24 + 3 * li;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) D[r * 3 + c] += we * kd[r * 24 + c];
    }
  }
}

template <class T, int MODE>
__global__ __launch_bounds__(256) void k_fine_apply(LevelView L, const uint8_t* __restrict__ fre,
                                                    const T* __restrict__ w, const T* __restrict__ Ke,
                                                    const T* __restrict__ u, const T* __restrict__ f,
                                                    T* __restrict__ y) {
  const int64_t t = thread_id();
  const int64_t n = (int64_t)L.nx * L.ny * L.nz;
  if (t >= n) return;
  const int x = t % L.nx, yy = (t / L.nx) % L.ny, z = t / ((int64_t)L.nx * L.ny);
  const int64_t p = L.pnode(x, yy, z);
  T acc[3] = {0, 0, 0};
  if (fre[p]) {
    T un[27][3];
    load_nbhd(L, u, p, un);
    T D[9];
    fine_row<T, true>(L, w, Ke, x, yy, z, un, acc, D);
  }
  if (MODE == 0) {
    y[p] = acc[0]; y[L.NP + p] = acc[1]; y[2 * L.NP + p] = acc[2];
  } else {
    const bool fr = fre[p];
    y[p] = fr ? f[p] - acc[0] : T(0);
    y[L.NP + p] = fr ? f[L.NP + p] - acc[1] : T(0);
    y[2 * L.NP + p] = fr ? f[2 * L.NP + p] - acc[2] : T(0);
  }
}

// y = K p fused with the partial sums of p.y (for PCG): partial[block]
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) s += sh[i];
  }
  __syncthreads();
  return s;
}

// Gauss-Seidel colour pass on the fine level.  BACKWARD relaxes components z, y, x.
template <class T, bool BACKWARD>
__global__ __launch_bounds__(256) void k_fine_gs(LevelView L, const uint8_t* __restrict__ fre,
                                                 const T* __restrict__ w, const T* __restrict__ Ke,
                                                 T* __restrict__ u, const T* __restrict__ f, int col) {
  const int px = col & 1, py = (col >> 1) & 1, pz = col >> 2;
  const int cx = (L.nx - px + 1) >> 1, cy = (L.ny - py + 1) >> 1, cz = (L.nz - pz + 1) >> 1;
  const int64_t t = thread_id();
  if (t >= (int64_t)cx * cy * cz) return;
  const int x = 2 * (t % cx) + px, yy = 2 * ((t / cx) % cy) + py, z = 2 * (t / ((int64_t)cx * cy)) + pz;
  const int64_t p = L.pnode(x, yy, z);
  if (!fre[p]) return;
  T un[27][3];
  load_nbhd(L, u, p, un);
  T s[3] = {0, 0, 0};
  T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  fine_row<T, false>(L, w, Ke, x, yy, z, un, s, D);
  T b0 = f[p] - s[0], b1 = f[L.NP + p] - s[1], b2 = f[2 * L.NP + p] - s[2];
  T v0 = un[13][0], v1 = un[13][1], v2 = un[13][2];
  if (!BACKWARD) {
    v0 = (b0 - D[1] * v1 - D[2] * v2) / D[0];
    v1 = (b1 - D[3] * v0 - D[5] * v2) / D[4];
    v2 = (b2 - D[6] * v0 - D[7] * v1) / D[8];
  } else {
    v2 = (b2 - D[6] * v0 - D[7] * v1) / D[8];
    v1 = (b1 - D[3] * v0 - D[5] * v2) / D[4];
    v0 = (b0 - D[1] * v1 - D[2] * v2) / D[0];
  }
  u[p] = v0; u[L.NP + p] = v1; u[2 * L.NP + p] = v2;
}

// ---------------------------------------------------------------------------------------
// Stored-stencil levels
// ---------------------------------------------------------------------------------------
template <class T, int MODE>
__global__ __launch_bounds__(256) void k_st_apply(LevelView L, const uint8_t* __restrict__ fre,
                                                  const T* __restrict__ st, const T* __restrict__ u,
                                                  const T* __restrict__ f, T* __restrict__ y) {
  const int64_t t = thread_id();
  const int64_t n = (int64_t)L.nx * L.ny * L.nz;
  if (t >= n) return;
  const int x = t % L.nx, yy = (t / L.nx) % L.ny, z = t / ((int64_t)L.nx * L.ny);
  const int64_t p = L.pnode(x, yy, z);
  T acc[3] = {0, 0, 0};
  const bool fr = fre[p];
  if (fr) {
#pragma unroll
    for (int k = 0; k < 27; ++k) {
      const int64_t q = p + (int64_t)(k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
      const T u0 = u[q], u1 = u[L.NP + q], u2 = u[2 * L.NP + q];
      const T* s = st + (int64_t)k * 9 * L.NP + p;
      acc[0] += s[0] * u0 + s[L.NP] * u1 + s[2 * L.NP] * u2;
      acc[1] += s[3 * L.NP] * u0 + s[4 * L.NP] * u1 + s[5 * L.NP] * u2;
      acc[2] += s[6 * L.NP] * u0 + s[7 * L.NP] * u1 + s[8 * L.NP] * u2;
    }
  }
  if (MODE == 0) {
    y[p] = acc[0]; y[L.NP + p] = acc[1]; y[2 * L.NP + p] = acc[2];
  } else {
    y[p] = fr ? f[p] - acc[0] : T(0);
    y[L.NP + p] = fr ? f[L.NP + p] - acc[1] : T(0);
    y[2 * L.NP + p] = fr ? f[2 * L.NP + p] - acc[2] : T(0);
  }
}

template <class T, bool BACKWARD>
__global__ __launch_bounds__(256) void k_st_gs(LevelView L, const uint8_t* __restrict__ fre,
                                               const T* __restrict__ st, T* __restrict__ u,
                                               const T* __restrict__ f, int col) {
  const int px = col & 1, py = (col >> 1) & 1, pz = col >> 2;
  const int cx = (L.nx - px + 1) >> 1, cy = (L.ny - py + 1) >> 1, cz = (L.nz - pz + 1) >> 1;
  const int64_t t = thread_id();
  if (t >= (int64_t)cx * cy * cz) return;
  const int x = 2 * (t % cx) + px, yy = 2 * ((t / cx) % cy) + py, z = 2 * (t / ((int64_t)cx * cy)) + pz;
  const int64_t p = L.pnode(x, yy, z);
  if (!fre[p]) return;
  T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    if (k == 13) continue;
    const int64_t q = p + (int64_t)(k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
    const T u0 = u[q], u1 = u[L.NP + q], u2 = u[2 * L.NP + q];
    const T* s = st + (int64_t)k * 9 * L.NP + p;
    s0 += s[0] * u0 + s[L.NP] * u1 + s[2 * L.NP] * u2;
    s1 += s[3 * L.NP] * u0 + s[4 * L.NP] * u1 + s[5 * L.NP] * u2;
    s2 += s[6 * L.NP] * u0 + s[7 * L.NP] * u1 + s[8 * L.NP] * u2;
  }
  const T* d = st + (int64_t)13 * 9 * L.NP + p;
  T D[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) D[e] = d[e * L.NP];
  T b0 = f[p] - s0, b1 = f[L.NP + p] - s1, b2 = f[2 * L.NP + p] - s2;
  T v0 = u[p], v1 = u[L.NP + p], v2 = u[2 * L.NP + p];
  if (!BACKWARD) {
    v0 = (b0 - D[1] * v1 - D[2] * v2) / D[0];
    v1 = (b1 - D[3] * v0 - D[5] * v2) / D[4];
    v2 = (b2 - D[6] * v0 - D[7] * v1) / D[8];
  } else {
    v2 = (b2 - D[6] * v0 - D[7] * v1) / D[8];
    v1 = (b1 - D[3] * v0 - D[5] * v2) / D[4];
    v0 = (b0 - D[1] * v1 - D[2] * v2) / D[0];
  }
  u[p] = v0; u[L.NP + p] = v1; u[2 * L.NP + p] = v2;
}

// ---------------------------------------------------------------------------------------
// Transfers: trilinear P (weights 1, 1/2, 1/4, 1/8), restriction = P^T, both masked.
// (restrict_residual_kernel grid/Grid.cu:2456-2507, prolongate_correction_kernel :2594-2632)
// ---------------------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_restrict(LevelView F, LevelView C, const uint8_t* __restrict__ cfre,
                                                  const T* __restrict__ rf, T* __restrict__ fc) {
  const int64_t t = thread_id();
  if (t >= (int64_t)C.nx * C.ny * C.nz) return;
  const int x = t % C.nx, y = (t / C.nx) % C.ny, z = t / ((int64_t)C.nx * C.ny);
  const int64_t I = C.pnode(x, y, z);
  T a0 = 0, a1 = 0, a2 = 0;
  if (cfre[I]) {
    const int64_t c0 = F.pnode(2 * x, 2 * y, 2 * z);
#pragma unroll
    for (int k = 0; k < 27; ++k) {
      const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
      const T wgt = T((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0));
      const int64_t a = c0 + dx + (int64_t)F.NX * (dy + (int64_t)F.NY * dz);
      a0 += wgt * rf[a]; a1 += wgt * rf[F.NP + a]; a2 += wgt * rf[2 * F.NP + a];
    }
  }
  fc[I] = a0; fc[C.NP + I] = a1; fc[2 * C.NP + I] = a2;
}

template <class T>
__global__ __launch_bounds__(256) void k_prolong_add(LevelView F, LevelView C, const uint8_t* __restrict__ ffre,
                                                     const T* __restrict__ uc, T* __restrict__ uf) {
  const int64_t t = thread_id();
  if (t >= (int64_t)F.nx * F.ny * F.nz) return;
  const int x = t % F.nx, y = (t / F.nx) % F.ny, z = t / ((int64_t)F.nx * F.ny);
  const int64_t a = F.pnode(x, y, z);
  if (!ffre[a]) return;
  const int nxp = (x & 1) + 1, nyp = (y & 1) + 1, nzp = (z & 1) + 1;
  const T wx = (x & 1) ? T(0.5) : T(1), wy = (y & 1) ? T(0.5) : T(1), wz = (z & 1) ? T(0.5) : T(1);
  const int64_t I0 = C.pnode(x >> 1, y >> 1, z >> 1);
  T a0 = 0, a1 = 0, a2 = 0;
  for (int k = 0; k < nzp; ++k)
    for (int j = 0; j < nyp; ++j)
      for (int i = 0; i < nxp; ++i) {
        const int64_t I = I0 + i + (int64_t)C.NX * (j + (int64_t)C.NY * k);
        const T wgt = wx * wy * wz;
        a0 += wgt * uc[I]; a1 += wgt * uc[C.NP + I]; a2 += wgt * uc[2 * C.NP + I];
      }
  uf[a] += a0; uf[F.NP + a] += a1; uf[2 * F.NP + a] += a2;
}

// ---------------------------------------------------------------------------------------
// Galerkin coarse operators.  Thread = (coarse node I, neighbour slot k = blockIdx.y).
// First coarse level: the coarse element matrices are sum over active children of
// rho^p P_c^T Ke P_c (8 precomputed 24x24 products Mc), evaluated with the eliminated
// rows/columns when a child touches a non-free fine node
// (restrict_stencil_nondyadic_OTFA_WS_kernel grid/Grid.cu:1112-1225 computes the same
// triple product per coarse node directly).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double pw(int c, int i, int m) {
  double w = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int t = ((c >> a) & 1) + ((i >> a) & 1);
    w *= ((m >> a) & 1) ? t * 0.5 : (2 - t) * 0.5;
  }
  return w;
}

__global__ __launch_bounds__(256) void k_galerkin_first(LevelView F, LevelView C, const uint8_t* __restrict__ ffre,
                                                        const uint8_t* __restrict__ cfre,
                                                        const double* __restrict__ w, const double* __restrict__ Ke,
                                                        const double* __restrict__ Mc, double* __restrict__ st) {
  __shared__ double sMc[8 * 576];
  for (int i = threadIdx.x; i < 8 * 576; i += blockDim.x) sMc[i] = Mc[i];
  __syncthreads();
  const int k = blockIdx.y;
  const int64_t t = thread_id();
  if (t >= (int64_t)C.nx * C.ny * C.nz) return;
  const int X = t % C.nx, Y = (t / C.nx) % C.ny, Z = t / ((int64_t)C.nx * C.ny);
  const int64_t I = C.pnode(X, Y, Z);
  const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t J = I + dx + (int64_t)C.NX * (dy + (int64_t)C.NY * dz);
  if (cfre[I] && cfre[J]) {
    // coarse elements E = I - 1 + bits(a) that also contain J
    for (int a = 0; a < 8; ++a) {
      const int ax = a & 1, ay = (a >> 1) & 1, az = a >> 2;
      const int nx_ = 1 + dx - ax, ny_ = 1 + dy - ay, nz_ = 1 + dz - az;  // J - E: J's local coords
      if (nx_ < 0 || nx_ > 1 || ny_ < 0 || ny_ > 1 || nz_ < 0 || nz_ > 1) continue;
      const int Ex = X - 1 + ax, Ey = Y - 1 + ay, Ez = Z - 1 + az;
      if (Ex < 0 || Ey < 0 || Ez < 0 || Ex >= C.ex || Ey >= C.ey || Ez >= C.ez) continue;
      const int mI = 7 - a, n = nx_ + 2 * ny_ + 4 * nz_;
      for (int ch = 0; ch < 8; ++ch) {
        const int fx = 2 * Ex + (ch & 1), fy = 2 * Ey + ((ch >> 1) & 1), fz = 2 * Ez + (ch >> 2);
        if (fx >= F.ex || fy >= F.ey || fz >= F.ez) continue;
        const double we = w[F.pelem(fx, fy, fz)];
        if (we == 0) continue;
        const int64_t fb = F.pnode(fx, fy, fz);
        bool fr[8];
        bool all = true;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          fr[i] = ffre[fb + (i & 1) + (int64_t)F.NX * (((i >> 1) & 1) + (int64_t)F.NY * (i >> 2))];
          all = all && fr[i];
        }
        if (all) {
          const double* m = sMc + ch * 576 + (3 * mI) * 24 + 3 * n;
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[r * 3 + c] += we * m[r * 24 + c];
        } else {
          for (int i = 0; i < 8; ++i) {
            if (!fr[i]) continue;
            const double pi = pw(ch, i, mI);
            if (pi == 0) continue;
            for (int j = 0; j < 8; ++j) {
              if (!fr[j]) continue;
              const double pj = pw(ch, j, n);
              if (pj == 0) continue;
              const double s = we * pi * pj;
#pragma unroll
              for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[r * 3 + c] += s * Ke[(3 * i + r) * 24 + 3 * j + c];
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) st[(int64_t)(k * 9 + e) * C.NP + I] = acc[e];
}

// Deeper levels: K_c[I, I+d] = sum_{a near 2I} P[a,I] sum_{b} K_f[a,b] P[b,I+d], per axis
// the admissible (fine offset of a, fine offset of b) pairs.
__global__ __launch_bounds__(256) void k_galerkin_next(LevelView F, LevelView C, const uint8_t* __restrict__ cfre,
                                                       const double* __restrict__ stf, double* __restrict__ stc) {
  const int k = blockIdx.y;
  const int64_t t = thread_id();
  if (t >= (int64_t)C.nx * C.ny * C.nz) return;
  const int X = t % C.nx, Y = (t / C.nx) % C.ny, Z = t / ((int64_t)C.nx * C.ny);
  const int64_t I = C.pnode(X, Y, Z);
  const int d[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t J = I + d[0] + (int64_t)C.NX * (d[1] + (int64_t)C.NY * d[2]);
  if (cfre[I] && cfre[J]) {
    // per-axis pair lists: a = 2I + da, b = 2J + db, |b - a| <= 1
    int pa[3][7], pk[3][7], np_[3];
    double pwt[3][7];
    const int Ic[3] = {X, Y, Z};
    const int Fn[3] = {F.nx, F.ny, F.nz};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      int cnt = 0;
      for (int da = -1; da <= 1; ++da)
        for (int db = -1; db <= 1; ++db) {
          const int kin = 2 * d[ax] + db - da;
          if (kin < -1 || kin > 1) continue;
          const int ai = 2 * Ic[ax] + da, bi = ai + kin;
          if (ai < 0 || ai >= Fn[ax] || bi < 0 || bi >= Fn[ax]) continue;
          pa[ax][cnt] = da;
          pk[ax][cnt] = kin;
          pwt[ax][cnt] = (da ? 0.5 : 1.0) * (db ? 0.5 : 1.0);
          ++cnt;
        }
      np_[ax] = cnt;
    }
    const int64_t A0 = F.pnode(2 * X, 2 * Y, 2 * Z);
    for (int iz = 0; iz < np_[2]; ++iz)
      for (int iy = 0; iy < np_[1]; ++iy)
        for (int ix = 0; ix < np_[0]; ++ix) {
          const int64_t A = A0 + pa[0][ix] + (int64_t)F.NX * (pa[1][iy] + (int64_t)F.NY * pa[2][iz]);
          const int kin = (pk[0][ix] + 1) + 3 * (pk[1][iy] + 1) + 9 * (pk[2][iz] + 1);
          const double wgt = pwt[0][ix] * pwt[1][iy] * pwt[2][iz];
          const double* s = stf + (int64_t)kin * 9 * F.NP + A;
#pragma unroll
          for (int e = 0; e < 9; ++e) acc[e] += wgt * s[e * F.NP];
        }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) stc[(int64_t)(k * 9 + e) * C.NP + I] = acc[e];
}

__global__ void k_cast_d2f(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t t = thread_id();
  if (t < n) b[t] = (float)a[t];
}

// ---------------------------------------------------------------------------------------
// Coarsest level: dense matrix over the free nodes, then an in-place blocked symmetric
// sweep (Gauss-Jordan) that leaves -A^{-1}; the V-cycle applies it as a dense matvec.
// ---------------------------------------------------------------------------------------
__global__ void k_coarse_assemble(LevelView L, const int64_t* __restrict__ cidx, int nnode,
                                  const int* __restrict__ cmap_pad, const double* __restrict__ st,
                                  double* __restrict__ A, int n) {
  const int64_t t = thread_id();
  if (t >= (int64_t)nnode * 27) return;
  const int pnode = t / 27, k = t % 27;
  const int64_t i = cidx[pnode];
  const int64_t j = i + (k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
  const int q = cmap_pad[j];
  if (q < 0) return;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      A[(int64_t)(3 * pnode + r) * n + 3 * q + c] = st[(int64_t)(k * 9 + r * 3 + c) * L.NP + i];
}

__global__ void k_symmetrize(double* A, int n, double* __restrict__ diag0) {
  const int64_t t = thread_id();
  if (t >= (int64_t)n * n) return;
  const int i = t / n, j = t % n;
  if (i == j) diag0[i] = A[t];
  if (j >= i) return;
  const double v = 0.5 * (A[(int64_t)i * n + j] + A[(int64_t)j * n + i]);
  A[(int64_t)i * n + j] = v;
  A[(int64_t)j * n + i] = v;
}

constexpr int kGJB = 32;

// Invert the diagonal block K (size b) by a scalar sweep in LDS; copy row panel R.
// A pivot (Schur complement) below 1e-12 of the unswept diagonal marks the system as
// singular -- a floating part or a structure without supports.
__global__ __launch_bounds__(256) void k_gj_pivot(double* __restrict__ A, int n, int K0, int b,
                                                  double* __restrict__ Dinv, double* __restrict__ R,
                                                  const double* __restrict__ diag0, int* __restrict__ flag) {
  __shared__ double S[kGJB][kGJB + 1];
  __shared__ double T[kGJB][kGJB + 1];
  for (int t = threadIdx.x; t < b * b; t += blockDim.x) S[t / b][t % b] = A[(int64_t)(K0 + t / b) * n + K0 + t % b];
  __syncthreads();
  for (int k = 0; k < b; ++k) {
    const double piv = S[k][k];  // same LDS word for every thread: the branch is uniform
    if (!(piv > 1e-12 * diag0[K0 + k])) {
      if (threadIdx.x == 0) *flag = 1;
      return;
    }
    const double d = 1.0 / piv;
    for (int t = threadIdx.x; t < b * b; t += blockDim.x) {
      const int i = t / b, j = t % b;
      double v = S[i][j];
      if (i != k && j != k) v = v - S[i][k] * S[k][j] * d;
      else if (i == k && j == k) v = -d;
      else v = v * d;
      T[i][j] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < b * b; t += blockDim.x) S[t / b][t % b] = T[t / b][t % b];
    __syncthreads();
  }
  for (int t = threadIdx.x; t < b * b; t += blockDim.x) Dinv[t] = -S[t / b][t % b];  // +A_KK^{-1}
  __syncthreads();
  for (int64_t t = threadIdx.x; t < (int64_t)b * n; t += blockDim.x) R[t] = A[(int64_t)(K0 + t / n) * n + t % n];
}

// C = A[:, K] * Dinv  (n x b)
__global__ void k_gj_panel(const double* __restrict__ A, int n, int K0, int b, const double* __restrict__ Dinv,
                           double* __restrict__ Cp) {
  const int64_t t = thread_id();
  if (t >= (int64_t)n * b) return;
  const int i = t / b, j = t % b;
  double s = 0;
  for (int q = 0; q < b; ++q) s += A[(int64_t)i * n + K0 + q] * Dinv[q * b + j];
  Cp[t] = s;
}

__global__ void k_gj_update(double* __restrict__ A, int n, int K0, int b, const double* __restrict__ Dinv,
                            const double* __restrict__ Cp, const double* __restrict__ R) {
  const int64_t t = thread_id();
  if (t >= (int64_t)n * n) return;
  const int i = t / n, j = t % n;
  const bool iK = i >= K0 && i < K0 + b, jK = j >= K0 && j < K0 + b;
  double v;
  if (iK && jK) v = -Dinv[(i - K0) * b + (j - K0)];
  else if (jK) v = Cp[(int64_t)i * b + (j - K0)];
  else if (iK) v = Cp[(int64_t)j * b + (i - K0)];
  else {
    v = A[t];
    for (int q = 0; q < b; ++q) v -= Cp[(int64_t)i * b + q] * R[(int64_t)q * n + j];
  }
  A[t] = v;
}

// u_c = -M f_c over the compact coarse unknowns; one wave per row.
template <class T>
__global__ __launch_bounds__(256) void k_coarse_solve(LevelView L, const int64_t* __restrict__ cidx, int n,
                                                      const double* __restrict__ M, const T* __restrict__ f,
                                                      T* __restrict__ u) {
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0;
  for (int j = lane; j < n; j += 64) {
    const int64_t node = cidx[j / 3];
    s += M[(int64_t)row * n + j] * (double)f[(j % 3) * L.NP + node];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (lane == 0) {
    const int64_t node = cidx[row / 3];
    u[(row % 3) * L.NP + node] = (T)(-s);
  }
}

// ---------------------------------------------------------------------------------------
// Vector kernels (FP64 PCG and casts)
// ---------------------------------------------------------------------------------------
template <class A, class B>
__global__ void k_copy_cast(const A* __restrict__ a, B* __restrict__ b, int64_t n) {
  const int64_t t = thread_id();
  if (t < n) b[t] = (B)a[t];
}

template <class T>
__global__ void k_zero(T* __restrict__ a, int64_t n) {
  const int64_t t = thread_id();
  if (t < n) a[t] = T(0);
}

}  // namespace tos
