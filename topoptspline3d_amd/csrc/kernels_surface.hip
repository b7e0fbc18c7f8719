// kernels_surface.hip -- self-supporting and drip constraints on the surface points of the
// spline density (deal_surface_points, optimization/optimization.cpp:1103-1128): the
// marching-cubes vertices of {spline = 1/2} on a background lattice, their normals and
// orientation (grid/Grid.cu:5529-5763), the per-point constraint values (:6015-6100,
// 6744-6925) and the coefficient gradients as a deterministic per-coefficient gather over
// the points in the coefficient's support (the reference scatters with unsynchronised
// `+=`, :6290, 7133).
#include "tos_internal.hpp"

namespace tos {

// one axis of the spline: mode 1 = uniform lattice of spacing sh with replicated boundary
// coefficients, mode 0 = clamped knots U (Grid::set_spline_knot_series grid/Grid.cpp:3253-3377)
struct SfAxis {
  int mode, k, nc;
  double sh;        // mode 1: param * h
  double lo, step;  // mode 0: first knot and interior knot step
  double box;       // grid extent n * h
  const double* U;  // mode 0: knot vector
};

struct SfAxes {
  SfAxis a[3];
};

// cardinal B-spline of order o on the knots 0..o (Cox-de Boor on uniform knots)
__device__ __forceinline__ double sf_cardinal(int o, double t) {
  if (o <= 0 || t < 0.0 || t >= (double)o) return 0.0;
  double N[9];
  for (int i = 0; i < o; ++i) N[i] = (t >= i && t < i + 1) ? 1.0 : 0.0;
  for (int p = 1; p < o; ++p)
    for (int i = 0; i + p < o; ++i) N[i] = ((t - i) * N[i] + (i + p + 1 - t) * N[i + 1]) / p;
  return N[0];
}

// basis values and first / second derivatives of the p+1 functions nonzero on knot span
// `span` (The NURBS Book algorithm A2.3; SplineBasisDeriX grid/Grid.cu:469-530)
__device__ void sf_ders_basis(int span, double x, int p, const double* U, double ders[3][8]) {
  double ndu[8][8], left[8], right[8], a[2][8];
  ndu[0][0] = 1.0;
  for (int j = 1; j <= p; ++j) {
    left[j] = x - U[span + 1 - j];
    right[j] = U[span + j] - x;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      ndu[j][r] = right[r + 1] + left[j - r];
      const double temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    ndu[j][j] = saved;
  }
  for (int j = 0; j <= p; ++j) {
    ders[0][j] = ndu[j][p];
    ders[1][j] = 0.0;
    ders[2][j] = 0.0;
  }
  for (int r = 0; r <= p; ++r) {
    int s1 = 0, s2 = 1;
    a[0][0] = 1.0;
    for (int k = 1; k <= 2 && k <= p; ++k) {
      double d = 0.0;
      const int rk = r - k, pk = p - k;
      if (r >= k) {
        a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
        d = a[s2][0] * ndu[rk][pk];
      }
      const int j1 = rk >= -1 ? 1 : -rk, j2 = (r - 1 <= pk) ? k - 1 : p - r;
      for (int j = j1; j <= j2; ++j) {
        a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
        d += a[s2][j] * ndu[rk + j][pk];
      }
      if (r <= pk) {
        a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
        d += a[s2][k] * ndu[r][pk];
      }
      ders[k][r] = d;
      const int t = s1;
      s1 = s2;
      s2 = t;
    }
  }
  int r = p;
  for (int k = 1; k <= 2; ++k) {
    for (int j = 0; j <= p; ++j) ders[k][j] *= r;
    r *= (p - k);
  }
}

// coefficient index and value / d / d2 of the k terms along one axis at x; false outside
// the spline's range (mode 0: the knot range)
__device__ bool sf_axis(const SfAxis& A, double x, int idx[8], double N[3][8]) {
  const int k = A.k;
  if (A.mode == 1) {
    const double t = x / A.sh;
    const int i0 = (int)floor(t - k / 2.0) + 1;
    for (int q = 0; q < k; ++q) {
      const int i = i0 + q;
      const double u = t - i + k / 2.0;
      idx[q] = min(max(i, 0), A.nc - 1);
      N[0][q] = sf_cardinal(k, u);
      N[1][q] = (sf_cardinal(k - 1, u) - sf_cardinal(k - 1, u - 1)) / A.sh;
      N[2][q] = (sf_cardinal(k - 2, u) - 2.0 * sf_cardinal(k - 2, u - 1) + sf_cardinal(k - 2, u - 2)) / (A.sh * A.sh);
    }
    return true;
  }
  const int span1 = (int)((x - A.lo) / A.step) + k;  // calc_node grid/Grid.cu:5550-5552
  if (span1 < k || span1 > A.nc) return false;
  double d[3][8];
  sf_ders_basis(span1 - 1, x, k - 1, A.U, d);
  for (int q = 0; q < k; ++q) {
    idx[q] = span1 - k + q;
    N[0][q] = d[0][q];
    N[1][q] = d[1][q];
    N[2][q] = d[2][q];
  }
  return true;
}

// spline value, gradient and Hessian (xx, yy, zz, xy, xz, yz) at p
__device__ bool sf_point(const SfAxes& S, const float* __restrict__ coef, const double p[3], double& phi,
                         double g[3], double H[6]) {
  int ix[8], iy[8], iz[8];
  double X[3][8], Y[3][8], Z[3][8];
  if (!sf_axis(S.a[0], p[0], ix, X) || !sf_axis(S.a[1], p[1], iy, Y) || !sf_axis(S.a[2], p[2], iz, Z)) return false;
  const int k = S.a[0].k, ncx = S.a[0].nc, ncy = S.a[1].nc;
  phi = 0;
  for (int i = 0; i < 3; ++i) g[i] = 0;
  for (int i = 0; i < 6; ++i) H[i] = 0;
  for (int qz = 0; qz < k; ++qz)
    for (int qy = 0; qy < k; ++qy)
      for (int qx = 0; qx < k; ++qx) {
        const double cv = coef[ix[qx] + (int64_t)ncx * (iy[qy] + (int64_t)ncy * iz[qz])];
        phi += cv * X[0][qx] * Y[0][qy] * Z[0][qz];
        g[0] += cv * X[1][qx] * Y[0][qy] * Z[0][qz];
        g[1] += cv * X[0][qx] * Y[1][qy] * Z[0][qz];
        g[2] += cv * X[0][qx] * Y[0][qy] * Z[1][qz];
        H[0] += cv * X[2][qx] * Y[0][qy] * Z[0][qz];
        H[1] += cv * X[0][qx] * Y[2][qy] * Z[0][qz];
        H[2] += cv * X[0][qx] * Y[0][qy] * Z[2][qz];
        H[3] += cv * X[1][qx] * Y[1][qy] * Z[0][qz];
        H[4] += cv * X[1][qx] * Y[0][qy] * Z[1][qz];
        H[5] += cv * X[0][qx] * Y[1][qy] * Z[1][qz];
      }
  return true;
}

// background lattice of NB = cells + 2 nodes per axis at (i + 1/2) eh - eh
struct SfLattice {
  int NB;
  double eh[3];
  __device__ double pos(int a, int i) const { return -eh[a] + (i + 0.5) * eh[a]; }
};

// The lattice is separable: per axis and lattice index r = a * NB + i, the k coefficient
// indices and basis values of the spline at the lattice position (sf_axis),
// tab_i / tab_v[r * 8 + q], and tab_ok[r]: bit 0 inside the box, bit 1 inside the spline's
// range.
__global__ void k_sf_axis_tab(SfAxes S, SfLattice G, int* __restrict__ tab_i, double* __restrict__ tab_v,
                              uint8_t* __restrict__ tab_ok) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 3 * G.NB) return;
  const int a = r / G.NB, i = r % G.NB;
  const double p = G.pos(a, i);
  int idx[8];
  double N[3][8];
  const bool ok = sf_axis(S.a[a], p, idx, N);
  for (int q = 0; q < 8; ++q) {
    const bool use = ok && q < S.a[a].k;
    tab_i[r * 8 + q] = use ? idx[q] : 0;
    tab_v[r * 8 + q] = use ? N[0][q] : 0.0;
  }
  tab_ok[r] = (uint8_t)((p >= 0.0 && p <= S.a[a].box ? 1 : 0) | (ok ? 2 : 0));
}

// node values: spline - 1/2 inside the box (-0.2 - 1/2 outside the spline range), -1/2
// outside the box (compute_background_mcPoints_value grid/Grid.cu:3986-4080), from the
// per-axis tables; K = the spline order when fixed at compile time (registers), else 0
template <int K>
__global__ __launch_bounds__(256) void k_sf_lattice(SfLattice G, int k, int ncx, int ncy, const int* __restrict__ tab_i,
                                                    const double* __restrict__ tab_v,
                                                    const uint8_t* __restrict__ tab_ok, const float* __restrict__ coef,
                                                    double* __restrict__ val) {
  const int64_t n = (int64_t)G.NB * G.NB * G.NB;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % G.NB), j = (int)((t / G.NB) % G.NB), kk = (int)(t / ((int64_t)G.NB * G.NB));
  const int rx = i, ry = G.NB + j, rz = 2 * G.NB + kk;
  const int ok = tab_ok[rx] & tab_ok[ry] & tab_ok[rz];
  double phi = 0;
  if ((ok & 1) && !(ok & 2)) phi = -0.2;
  if (ok == 3 && K > 0) {
    int xi[K > 0 ? K : 1];
    double xv[K > 0 ? K : 1];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      xi[q] = tab_i[rx * 8 + q];
      xv[q] = tab_v[rx * 8 + q];
    }
    for (int qz = 0; qz < K; ++qz) {
      const int iz = tab_i[rz * 8 + qz];
      const double z = tab_v[rz * 8 + qz];
      for (int qy = 0; qy < K; ++qy) {
        const int iy = tab_i[ry * 8 + qy];
        const double y = tab_v[ry * 8 + qy];
        const float* row = coef + (int64_t)ncx * (iy + (int64_t)ncy * iz);
#pragma unroll
        for (int qx = 0; qx < K; ++qx) phi += (double)row[xi[qx]] * xv[qx] * y * z;
      }
    }
  } else if (ok == 3) {
    for (int qz = 0; qz < k; ++qz) {
      const int iz = tab_i[rz * 8 + qz];
      const double z = tab_v[rz * 8 + qz];
      for (int qy = 0; qy < k; ++qy) {
        const int iy = tab_i[ry * 8 + qy];
        const double y = tab_v[ry * 8 + qy];
        const float* row = coef + (int64_t)ncx * (iy + (int64_t)ncy * iz);
        for (int qx = 0; qx < k; ++qx) phi += (double)row[tab_i[rx * 8 + qx]] * tab_v[rx * 8 + qx] * y * z;
      }
    }
  }
  val[t] = ((ok & 1) ? phi : 0.0) - 0.5;
}

// the point on the lattice edge from node t along axis d, when the edge changes sign and
// the point lies in an active element
__device__ __forceinline__ bool sf_edge_point(const SfLattice& G, LevelView L, const uint8_t* __restrict__ eact,
                                              const double* __restrict__ val, double h, int i, int j, int kk, int d,
                                              double p[3]) {
  int q[3] = {i, j, kk};
  if (++q[d] >= G.NB) return false;
  const double va = val[i + (int64_t)G.NB * (j + (int64_t)G.NB * kk)];
  const double vb = val[q[0] + (int64_t)G.NB * (q[1] + (int64_t)G.NB * q[2])];
  if ((va < 0) == (vb < 0)) return false;
  const double t = (0.0 - va) / (vb - va);
  const int ijk[3] = {i, j, kk};
  const int dims[3] = {L.ex, L.ey, L.ez};
  int e[3];
  for (int a = 0; a < 3; ++a) {
    p[a] = G.pos(a, ijk[a]) + (a == d ? t * G.eh[a] : 0.0);
    e[a] = (int)floor(p[a] / h);
    if (e[a] < 0 || e[a] >= dims[a]) return false;
  }
  return eact[L.pelem(e[0], e[1], e[2])] != 0;
}

__global__ void k_sf_count(SfLattice G, LevelView L, const uint8_t* __restrict__ eact, const double* __restrict__ val,
                           double h, int* __restrict__ cnt) {
  const int64_t n = (int64_t)G.NB * G.NB * G.NB;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % G.NB), j = (int)((t / G.NB) % G.NB), kk = (int)(t / ((int64_t)G.NB * G.NB));
  int c = 0;
  double p[3];
  for (int d = 0; d < 3; ++d) c += sf_edge_point(G, L, eact, val, h, i, j, kk, d, p);
  cnt[t] = c;
}

// exclusive prefix sum of n counts in one workgroup (1024 threads, contiguous chunks):
// off[i] = sum of cnt[0..i), off[n] = total
__global__ __launch_bounds__(1024) void k_sf_scan(const int* __restrict__ cnt, int64_t n, int* __restrict__ off) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024, b = t * per, e = min(n, b + per);
  int s = 0;
  for (int64_t i = b; i < e; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive scan of the chunk sums
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int64_t i = b; i < e; ++i) {
    off[i] = run;
    run += cnt[i];
  }
  if (t == 1023) off[n] = part[1023];
}

// Exclusive prefix sum over many workgroups: k_sf_scan_blocks sums each 2048-count block,
// k_sf_scan (one workgroup) scans those sums, k_sf_scan_apply scans each block locally
// and adds its block offset; off[n] = total.
constexpr int kSfScanBlock = 2048;
__global__ __launch_bounds__(1024) void k_sf_scan_blocks(const int* __restrict__ cnt, int64_t n,
                                                         int* __restrict__ bsum) {
  __shared__ int sh[1024];
  const int64_t b0 = (int64_t)blockIdx.x * kSfScanBlock;
  const int t = threadIdx.x;
  int s = 0;
  for (int k = 0; k < 2; ++k) {
    const int64_t i = b0 + 2 * t + k;
    if (i < n) s += cnt[i];
  }
  sh[t] = s;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  if (t == 0) bsum[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(1024) void k_sf_scan_apply(const int* __restrict__ cnt, int64_t n,
                                                        const int* __restrict__ boff, int* __restrict__ off) {
  __shared__ int sh[1024];
  const int64_t b0 = (int64_t)blockIdx.x * kSfScanBlock;
  const int t = threadIdx.x;
  const int64_t i0 = b0 + 2 * t;
  const int c0 = i0 < n ? cnt[i0] : 0, c1 = i0 + 1 < n ? cnt[i0 + 1] : 0;
  sh[t] = c0 + c1;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive scan of the thread pairs
    const int v = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += v;
    __syncthreads();
  }
  const int base = boff[blockIdx.x] + sh[t] - (c0 + c1);
  if (i0 < n) off[i0] = base;
  if (i0 + 1 < n) off[i0 + 1] = base + c0;
  if (i0 + 1 == n - 1 || i0 == n - 1) off[n] = base + c0 + c1;  // the last count's block writes the total
}

// 1 when some coefficient differs from the first (partials summed)
__global__ __launch_bounds__(256) void k_sf_coef_differ(const float* __restrict__ coef, int64_t n,
                                                        double* __restrict__ partial) {
  double s = 0;
  const float c0 = coef[0];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += coef[i] != c0 ? 1.0 : 0.0;
  block_partial(s, partial, 0);
}

__global__ void k_sf_points(SfLattice G, LevelView L, const uint8_t* __restrict__ eact,
                            const double* __restrict__ val, double h, const int* __restrict__ off,
                            double* __restrict__ pts) {
  const int64_t n = (int64_t)G.NB * G.NB * G.NB;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % G.NB), j = (int)((t / G.NB) % G.NB), kk = (int)(t / ((int64_t)G.NB * G.NB));
  int w = off[t];
  double p[3];
  for (int d = 0; d < 3; ++d)
    if (sf_edge_point(G, L, eact, val, h, i, j, kk, d, p)) {
      for (int a = 0; a < 3; ++a) pts[3 * (int64_t)w + a] = p[a];
      ++w;
    }
}

// gradient and Hessian at every point, and the orientation vote (+1 when a step of
// 1e-3 grad spline ends in void, Heaviside < 1/2; correct_spline_surface_point_normal_direction
// grid/Grid.cu:5620-5763); per-block vote sums in part[b]
__global__ __launch_bounds__(256) void k_sf_orient(SfAxes S, const float* __restrict__ coef, int64_t n,
                                                   const double* __restrict__ pts, double beta,
                                                   double* __restrict__ grad, double* __restrict__ hess,
                                                   double* __restrict__ part) {
  double vote = 0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const double p[3] = {pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]};
    double phi, g[3], H[6];
    if (!sf_point(S, coef, p, phi, g, H)) {
      for (double& v : g) v = 0;
      for (double& v : H) v = 0;
    }
    for (int a = 0; a < 3; ++a) grad[3 * t + a] = g[a];
    for (int a = 0; a < 6; ++a) hess[6 * t + a] = H[a];
    double pd[3], v2 = -0.1, g2[3], H2[6];
    for (int a = 0; a < 3; ++a) pd[a] = p[a] + 1e-3 * g[a];
    if (!sf_point(S, coef, pd, v2, g2, H2)) v2 = -0.1;
    const double eta = 0.5, den = tanh(beta * eta) + tanh(beta * (1 - eta));
    const double hs = (tanh(beta * eta) + tanh(beta * (v2 - eta))) / den;
    vote += hs < 0.5 ? 1.0 : -1.0;
  }
  vote = block_sum256(vote);
  if (threadIdx.x == 0) part[blockIdx.x] = vote;
}

struct SfParams {
  double cs, cd;  // cos(print_angle), cos(drip_angle)
  int ss_mode, drip_mode, drip_version;
  double p_norm, hfunction_c, sigmoid_c, drip_beta, drip_mu, drip_angle, drip_alpha, drip_beta_scale;
};

__device__ __forceinline__ double sf_sigm(double x) { return 1.0 / (1.0 + exp(-x)); }

// per-point values and gradient weights with the orientation dir = scal[slot]; partials
// [3][nblk]: self-support sum, drip sum, points with n_z < 0
__global__ __launch_bounds__(256) void k_sf_values(SfParams P, int64_t n, const double* __restrict__ grad,
                                                   const double* __restrict__ hess, const double* __restrict__ scal,
                                                   int dir_slot, double* __restrict__ nrm, double* __restrict__ wss,
                                                   double* __restrict__ wdr, double* __restrict__ part) {
  const double dir = scal[dir_slot] > 0 ? 1.0 : -1.0;
  double s_ss = 0, s_dr = 0, s_dn = 0;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    double nv[3];
    for (int a = 0; a < 3; ++a) nrm[3 * t + a] = nv[a] = dir * grad[3 * t + a];
    for (int a = 0; a < 3; ++a) wss[3 * t + a] = 0;
    for (int a = 0; a < 6; ++a) wdr[6 * t + a] = 0;
    const double nn = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    const bool down = nv[2] < 0;
    s_dn += down ? 1.0 : 0.0;
    if (nn == 0) continue;
    // d inner / d c_j = dinner . grad N_j / cos, dinner = dir (e_z |n| - n_z n / |n|) / |n|^2
    double di[3];
    for (int a = 0; a < 3; ++a) di[a] = dir * ((a == 2 ? nn : 0.0) - nv[2] * nv[a] / nn) / (nn * nn);
    // self-support (compute_spline_selfsupp_constraint grid/Grid.cu:6015-6100, h/oh :117-160)
    const double inner = nv[2] / nn / P.cs;
    double v = 0, dv = 0;
    if (!(P.ss_mode % 2 == 0 && !down)) {
      if (P.ss_mode <= 1) {
        v = pow(inner, P.p_norm);
        dv = P.p_norm * pow(inner, P.p_norm - 1);
      } else if (P.ss_mode <= 3) {
        v = inner >= 1 ? P.hfunction_c * (inner - 1) * (inner - 1) : 0.0;
        dv = inner >= 1 ? 2 * P.hfunction_c * (inner - 1) : 0.0;
      } else {
        const double s = sf_sigm((inner - 1.5) / P.sigmoid_c);
        v = s + 1;
        dv = s * (1 - s) / P.sigmoid_c;
      }
    }
    s_ss += v;
    for (int a = 0; a < 3; ++a) wss[3 * t + a] = dv * di[a] / P.cs;
    // drip (compute_spline_drip_constraint_test grid/Grid.cu:6744-6925), modes 6 / 7
    if (P.drip_mode >= 6 && (P.drip_mode == 7 || down)) {
      const double* H = &hess[6 * t];  // xx, yy, zz, xy, xz, yz
      const double hin = nv[2] / nn / P.cd;
      double u, du[3];
      if (P.drip_version == 1) {
        const double px = nv[0] / nv[2], py = nv[1] / nv[2];
        u = exp(-P.drip_alpha * (px * px + py * py));
        for (int a = 0; a < 3; ++a) {
          const double dpx = ((a == 0 ? nv[2] : 0.0) - (a == 2 ? nv[0] : 0.0)) / (nv[2] * nv[2]);
          const double dpy = ((a == 1 ? nv[2] : 0.0) - (a == 2 ? nv[1] : 0.0)) / (nv[2] * nv[2]);
          du[a] = -2 * P.drip_alpha * (px * dpx + py * dpy) * u * dir;
        }
      } else {
        double s, ds;
        if (P.drip_angle > 0.5 * M_PI) {
          s = sf_sigm((hin - 1.2) * P.drip_mu);
          ds = s * (1 - s) * P.drip_mu;
        } else {
          s = sf_sigm(-(hin + 1.2) * P.drip_mu);
          ds = -s * (1 - s) * P.drip_mu;
        }
        u = s;
        for (int a = 0; a < 3; ++a) du[a] = ds * di[a] / P.cd;
      }
      const double x1 = H[0], x2 = H[0] * H[1] - H[3] * H[3];
      const double e1 = exp(P.drip_beta * x1), e2 = exp(-P.drip_beta * x2);
      const double q1 = P.drip_beta_scale * x1 * x1 / (1 + e1);
      const double q2 = P.drip_beta_scale * x2 * x2 / (1 + e2);
      const double dq1 = P.drip_beta_scale * (2 * x1 / (1 + e1) - x1 * x1 * P.drip_beta * e1 / ((1 + e1) * (1 + e1)));
      const double dq2 = P.drip_beta_scale * (2 * x2 / (1 + e2) + x2 * x2 * P.drip_beta * e2 / ((1 + e2) * (1 + e2)));
      s_dr += u * q1 * q2;
      for (int a = 0; a < 3; ++a) wdr[6 * t + a] = du[a] * q1 * q2;
      wdr[6 * t + 3] = u * (dq1 * q2 + q1 * dq2 * H[1]);  // with N_xx
      wdr[6 * t + 4] = u * q1 * dq2 * H[0];                // with N_yy
      wdr[6 * t + 5] = u * q1 * dq2 * (-2 * H[3]);         // with N_xy
    }
  }
  s_ss = block_sum256(s_ss);
  s_dr = block_sum256(s_dr);
  s_dn = block_sum256(s_dn);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s_ss;
    part[gridDim.x + blockIdx.x] = s_dr;
    part[2 * gridDim.x + blockIdx.x] = s_dn;
  }
}

// folds `ncol` columns of nb block partials into scal[slot0 + col]
__global__ __launch_bounds__(256) void k_sf_fold(const double* __restrict__ part, int nb, int ncol, int slot0,
                                                 double* __restrict__ scal) {
  for (int col = 0; col < ncol; ++col) {
    double s = 0;
    for (int i = threadIdx.x; i < nb; i += 256) s += part[(int64_t)col * nb + i];
    s = block_sum256(s);
    if (threadIdx.x == 0) scal[slot0 + col] = s;
  }
}

// per point and axis, the k coefficient indices and basis values / d / d2 (sf_axis), once:
// bidx[(3 t + a) * 8 + q] (-1 past k or outside the spline's range),
// bN[((3 t + a) * 3 + d) * 8 + q]
__global__ void k_sf_basis(SfAxes S, int64_t n, const double* __restrict__ pts, int* __restrict__ bidx,
                           double* __restrict__ bN) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  for (int a = 0; a < 3; ++a) {
    int idx[8];
    double N[3][8];
    const bool ok = sf_axis(S.a[a], pts[3 * t + a], idx, N);
    for (int q = 0; q < 8; ++q) {
      const bool use = ok && q < S.a[a].k;
      bidx[(3 * t + a) * 8 + q] = use ? idx[q] : -1;
      for (int d = 0; d < 3; ++d) bN[((3 * t + a) * 3 + d) * 8 + q] = use ? N[d][q] : 0.0;
    }
  }
}

// the contributions d/dc of coefficient index ci along axis a at point t (sf_basis_at from
// the per-point table): sum over the terms q whose (possibly clamped) index equals ci
__device__ __forceinline__ void sf_basis_tab(const int* __restrict__ bidx, const double* __restrict__ bN, int64_t t,
                                             int a, int ci, double out[3]) {
  out[0] = out[1] = out[2] = 0;
  const int* ix = bidx + (3 * t + a) * 8;
  const double* nn = bN + (3 * t + a) * 24;
  for (int q = 0; q < 8; ++q)
    if (ix[q] == ci)
      for (int d = 0; d < 3; ++d) out[d] += nn[d * 8 + q];
}

// The coefficient gradients gather over points grouped by knot cell: a point's cell is the
// first coefficient index of its span along each axis (clamped into the coefficient range,
// sf_axis), and every coefficient the point touches lies in the k^3 block that starts at
// that cell, so coefficient c gathers from the cells c - [0, k)^3.  key[t] = the cell
// (-1 outside the spline's range), cnt[cell] += 1.
__global__ void k_sf_cell_count(int64_t n, int ncx, int ncy, const int* __restrict__ bidx, int* __restrict__ key,
                                int* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int bx = bidx[(3 * t + 0) * 8], by = bidx[(3 * t + 1) * 8], bz = bidx[(3 * t + 2) * 8];
  int c = -1;
  if (bx >= 0 && by >= 0 && bz >= 0) {
    c = bx + ncx * (by + ncy * bz);
    atomicAdd(&cnt[c], 1);
  }
  key[t] = c;
}

// list[start[cell] ..) = the cell's points; cur (zeroed) hands out the slots in arrival
// order, k_sf_cell_sort puts each cell's points back in ascending order
__global__ void k_sf_cell_fill(int64_t n, const int* __restrict__ key, const int* __restrict__ start,
                               int* __restrict__ cur, int* __restrict__ list) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || key[t] < 0) return;
  list[start[key[t]] + atomicAdd(&cur[key[t]], 1)] = (int)t;
}

__global__ void k_sf_cell_sort(int64_t ncell, const int* __restrict__ start, int* __restrict__ list) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int b = start[c], e = start[c + 1];
  for (int i = b + 1; i < e; ++i) {  // insertion sort: a cell holds a few points
    const int v = list[i];
    int p = i - 1;
    while (p >= b && list[p] > v) {
      list[p + 1] = list[p];
      --p;
    }
    list[p + 1] = v;
  }
}

// gss/gdr[c] = factor * sum over the points of the cells c - [0, k)^3 of w . (basis
// derivatives of c).  One wavefront per coefficient: lane l takes the window cells l,
// l + 64, ... (cell order, then point order), and lane 0 sums the 64 lanes in a fixed
// butterfly order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_sf_grad(SfAxes S, int64_t ncoef, const int* __restrict__ start,
                                                 const int* __restrict__ list, const int* __restrict__ bidx,
                                                 const double* __restrict__ bN, const double* __restrict__ wss,
                                                 const double* __restrict__ wdr, double fss, double fdr,
                                                 float* __restrict__ gss, float* __restrict__ gdr) {
  const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (j >= ncoef) return;  // whole wavefronts
  const int k = S.a[0].k, ncx = S.a[0].nc, ncy = S.a[1].nc;
  const int ci[3] = {(int)(j % ncx), (int)((j / ncx) % ncy), (int)(j / ((int64_t)ncx * ncy))};
  double a_ss = 0, a_dr = 0;
  int seen = 0;
  for (int w = lane; w < k * k * k; w += 64) {
    const int bx = ci[0] - w % k, by = ci[1] - (w / k) % k, bz = ci[2] - w / (k * k);
    if (bx < 0 || by < 0 || bz < 0) continue;
    const int64_t cell = bx + (int64_t)ncx * (by + (int64_t)ncy * bz);
    const int e = start[cell + 1];
    for (int p = start[cell]; p < e; ++p) {
      const int64_t t = list[p];
      ++seen;
      double X[3], Y[3], Z[3];
      sf_basis_tab(bidx, bN, t, 0, ci[0], X);
      if (X[0] == 0 && X[1] == 0 && X[2] == 0) continue;
      sf_basis_tab(bidx, bN, t, 1, ci[1], Y);
      sf_basis_tab(bidx, bN, t, 2, ci[2], Z);
      const double gx = X[1] * Y[0] * Z[0], gy = X[0] * Y[1] * Z[0], gz = X[0] * Y[0] * Z[1];
      const double* ws = &wss[3 * t];
      const double* wd = &wdr[6 * t];
      a_ss += ws[0] * gx + ws[1] * gy + ws[2] * gz;
      a_dr += wd[0] * gx + wd[1] * gy + wd[2] * gz + wd[3] * X[2] * Y[0] * Z[0] + wd[4] * X[0] * Y[2] * Z[0] +
              wd[5] * X[1] * Y[1] * Z[0];
    }
  }
  if (__ballot(seen > 0) != 0) {
    for (int o = 1; o < 64; o <<= 1) {
      a_ss += __shfl_xor(a_ss, o, 64);
      a_dr += __shfl_xor(a_dr, o, 64);
    }
  }
  if (lane == 0) {
    gss[j] = (float)(fss * a_ss);
    gdr[j] = (float)(fdr * a_dr);
  }
}

}  // namespace tos

namespace tos {
// df and the three constraint gradient rows of the self-supporting MMA step
__global__ void k_sf_mma_inputs(int64_t n, const float* __restrict__ cs, const float* __restrict__ cv,
                                const float* __restrict__ css, const float* __restrict__ cdr, double sscale,
                                double vscale, double ss_scale, double drip_scale, double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  out[j] = sscale * (double)cs[j];
  out[n + j] = vscale * (double)cv[j];
  out[2 * n + j] = ss_scale * (double)css[j];
  out[3 * n + j] = drip_scale * (double)cdr[j];
}
}  // namespace tos
