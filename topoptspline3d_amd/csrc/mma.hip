// mma.hip -- method of moving asymptotes (MMA::mma_t, optimization/mma_t.h:159-266,
// mma_t.cpp:18-51, mma_t.cu:15-807), FP64.
//
// The reference evaluates every vector expression of the primal-dual Newton iteration as
// its own kernel (dozens per step) and solves the (m+1)x(m+1) reduced system on the host.
// Here each n-length sweep of a Newton step is one fused grid-stride kernel over the whole
// GPU that also produces the step's reductions as fixed-order per-block partials; one
// block folds them and the host reads the few scalars that steer the iteration (the
// small system, the step bound, the KKT line search), exactly in the reference's order.
#include "tos_internal.hpp"

namespace tos {

constexpr int MMAX = 4;       // constraints supported
constexpr int MMA_CNT = 24;   // widest reduction (m + m*m + m for m = 4)

struct MmaVecs {
  int n, m;
  double *x, *xmin, *xmax, *L, *U, *alpha, *beta, *lastdx, *xi, *eta;
  double *p, *q;  // (m+1) x n
  double *plam, *qlam, *curx, *Dx, *G, *dlx, *dx, *dxi, *deta, *tx, *txi, *teta;
};

struct MmaSmall {  // scalars passed by value
  double lam[MMAX], dlam[MMAX];
  double eps, step, move;
};

__device__ __forceinline__ int64_t mtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int64_t mstride() { return (int64_t)gridDim.x * blockDim.x; }

template <int CNT>
__device__ __forceinline__ void block_partials(double (&v)[CNT], const bool (&ismax)[CNT], double* partial) {
  __shared__ double sh[4][CNT];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < CNT; ++i) {
    double a = v[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double b = __shfl_down(a, o, 64);
      a = ismax[i] ? fmax(a, b) : a + b;
    }
    if (lane == 0) sh[wid][i] = a;
  }
  __syncthreads();
  if (threadIdx.x < CNT) {
    const int i = threadIdx.x;
    const double r = ismax[i] ? fmax(fmax(sh[0][i], sh[1][i]), fmax(sh[2][i], sh[3][i]))
                              : (sh[0][i] + sh[1][i]) + (sh[2][i] + sh[3][i]);
    partial[(int64_t)i * gridDim.x + blockIdx.x] = r;
  }
}

// p_i, q_i at the current x, the constraint proxies there (for b), then the subproblem
// start point (mma_subproblem_t::init mma_t.cu:15-118 and reset :702-713).
__global__ __launch_bounds__(256) void k_mma_pq(MmaVecs V, const double* __restrict__ df,
                                                const double* __restrict__ dg, double* partial) {
  const int n = V.n, m = V.m;
  double v[MMAX] = {0, 0, 0, 0};
  const bool mx[MMAX] = {false, false, false, false};
  for (int64_t j = mtid(); j < n; j += mstride()) {
    const double x = V.x[j], ux = V.U[j] - x, xl = x - V.L[j];
    const double r = 1e-5 / (V.xmax[j] - V.xmin[j]);
    for (int i = 0; i <= m; ++i) {
      const double d = i == 0 ? df[j] : dg[(int64_t)(i - 1) * n + j];
      const double dp = fmax(d, 0.0), dm = fmax(-d, 0.0);
      const double pv = ux * ux * (1.001 * dp + 0.001 * dm + r);
      const double qv = xl * xl * (0.001 * dp + 1.001 * dm + r);
      V.p[(int64_t)i * n + j] = pv;
      V.q[(int64_t)i * n + j] = qv;
      if (i > 0) v[i - 1] += pv / ux + qv / xl;
    }
    V.curx[j] = x;
    const double xn = 0.5 * (V.alpha[j] + V.beta[j]);
    V.x[j] = xn;
    V.xi[j] = fmax(1.0 / (xn - V.alpha[j]), 1.0);
    V.eta[j] = fmax(1.0 / (V.beta[j] - xn), 1.0);
  }
  block_partials<MMAX>(v, mx, partial);
}

// Dx, delta_x and G at the current point; sums for the reduced system (mma_t.cu:358-395,
// solveLinearSystem :228-328): proxies g_i, A11 = G Dx^-1 G^T, b1 part G (delta_x / Dx).
__global__ __launch_bounds__(256) void k_mma_passA(MmaVecs V, MmaSmall S, double* partial) {
  const int n = V.n, m = V.m;
  constexpr int CA = MMAX + MMAX * MMAX + MMAX;
  double v[CA];
  bool mx[CA];
#pragma unroll
  for (int i = 0; i < CA; ++i) { v[i] = 0; mx[i] = false; }
  for (int64_t j = mtid(); j < n; j += mstride()) {
    const double xj = V.x[j], ux = V.U[j] - xj, xl = xj - V.L[j];
    const double xa = xj - V.alpha[j], bx = V.beta[j] - xj;
    const double psi = 2 * V.plam[j] / (ux * ux * ux) + 2 * V.qlam[j] / (xl * xl * xl);
    const double Dx = psi + V.xi[j] / xa + V.eta[j] / bx;
    const double phi = V.plam[j] / (ux * ux) - V.qlam[j] / (xl * xl);
    const double dlx = -(S.eps / xa - S.eps / bx - phi);
    V.Dx[j] = Dx;
    V.dlx[j] = dlx;
    double Gi[MMAX];
    for (int i = 0; i < m; ++i) {
      const double pi = V.p[(int64_t)(i + 1) * n + j], qi = V.q[(int64_t)(i + 1) * n + j];
      Gi[i] = pi / (ux * ux) - qi / (xl * xl);
      V.G[(int64_t)i * n + j] = Gi[i];
      v[i] += pi / ux + qi / xl;
    }
    for (int i = 0; i < m; ++i) {
      for (int k = 0; k < m; ++k) v[MMAX + i * MMAX + k] += Gi[i] * (Gi[k] / Dx);
      v[MMAX + MMAX * MMAX + i] += Gi[i] * (dlx / Dx);
    }
  }
  block_partials<CA>(v, mx, partial);
}

// Newton direction in x, xi, eta and the fraction-to-boundary bounds (mma_t.cu:302-317,
// :396-402)
__global__ __launch_bounds__(256) void k_mma_passB(MmaVecs V, MmaSmall S, double* partial) {
  const int n = V.n, m = V.m;
  double t[3] = {-1e300, -1e300, -1e300};
  const bool mx[3] = {true, true, true};
  const double cf = 1.0 / 0.99;
  for (int64_t j = mtid(); j < n; j += mstride()) {
    double dxv = -V.dlx[j] / V.Dx[j];
    for (int i = 0; i < m; ++i) dxv -= S.dlam[i] * V.G[(int64_t)i * n + j] / V.Dx[j];
    const double xa = V.x[j] - V.alpha[j], bx = V.beta[j] - V.x[j];
    const double dxi = S.eps / xa - dxv / xa * V.xi[j] - V.xi[j];
    const double deta = S.eps / bx + V.eta[j] / bx * dxv - V.eta[j];
    V.dx[j] = dxv;
    V.dxi[j] = dxi;
    V.deta[j] = deta;
    t[0] = fmax(t[0], -cf * (dxv / xa));
    t[1] = fmax(t[1], cf * (dxv / bx));
    t[2] = fmax(t[2], fmax(-cf * dxi / V.xi[j], -cf * deta / V.eta[j]));
  }
  block_partials<3>(t, mx, partial);
}

// plam/qlam for lambda, then the n-length parts of the KKT residual (kkt_err
// mma_t.cu:574-678) at the current point (TRIAL = false) or at the trial point
// x + step dx (TRIAL = true, written to tx/txi/teta).
template <bool TRIAL>
__global__ __launch_bounds__(256) void k_mma_kkt(MmaVecs V, MmaSmall S, double* partial) {
  const int n = V.n, m = V.m;
  constexpr int CK = 3 + MMAX + 3;
  double v[CK];
  bool mx[CK];
#pragma unroll
  for (int i = 0; i < CK; ++i) { v[i] = 0; mx[i] = i >= 3 + MMAX; }
  for (int64_t j = mtid(); j < n; j += mstride()) {
    double X = V.x[j], XI = V.xi[j], ETA = V.eta[j];
    if (TRIAL) {
      X = X + V.dx[j] * S.step;
      XI = XI + V.dxi[j] * S.step;
      ETA = ETA + V.deta[j] * S.step;
      V.tx[j] = X; V.txi[j] = XI; V.teta[j] = ETA;
    }
    double pl = V.p[j], ql = V.q[j];
    for (int i = 0; i < m; ++i) {
      pl += S.lam[i] * V.p[(int64_t)(i + 1) * n + j];
      ql += S.lam[i] * V.q[(int64_t)(i + 1) * n + j];
    }
    V.plam[j] = pl;
    V.qlam[j] = ql;
    const double ux = V.U[j] - X, xl = X - V.L[j];
    const double rx = pl / (ux * ux) - ql / (xl * xl) - XI + ETA;
    const double r1 = XI * (X - V.alpha[j]) - S.eps;
    const double r2 = ETA * (V.beta[j] - X) - S.eps;
    v[0] += rx * rx; v[1] += r1 * r1; v[2] += r2 * r2;
    for (int i = 0; i < m; ++i) v[3 + i] += V.p[(int64_t)(i + 1) * n + j] / ux + V.q[(int64_t)(i + 1) * n + j] / xl;
    v[3 + MMAX] = fmax(v[3 + MMAX], fabs(rx));
    v[4 + MMAX] = fmax(v[4 + MMAX], fabs(r1));
    v[5 + MMAX] = fmax(v[5 + MMAX], fabs(r2));
  }
  block_partials<CK>(v, mx, partial);
}

// plam/qlam only (after the reset, with lambda = 1)
__global__ __launch_bounds__(256) void k_mma_plam(MmaVecs V, MmaSmall S) {
  const int n = V.n, m = V.m;
  for (int64_t j = mtid(); j < n; j += mstride()) {
    double pl = V.p[j], ql = V.q[j];
    for (int i = 0; i < m; ++i) {
      pl += S.lam[i] * V.p[(int64_t)(i + 1) * n + j];
      ql += S.lam[i] * V.q[(int64_t)(i + 1) * n + j];
    }
    V.plam[j] = pl;
    V.qlam[j] = ql;
  }
}

// asymptotes and move limits for the next iteration (adjust_asym mma_t.cu:164-226)
__global__ __launch_bounds__(256) void k_mma_asym(MmaVecs V, MmaSmall S, double* partial) {
  const int n = V.n;
  double v[1] = {0};
  const bool mx[1] = {false};
  for (int64_t j = mtid(); j < n; j += mstride()) {
    const double xv = V.x[j], d1 = xv - V.curx[j], d2 = V.lastdx[j];
    double gam = 1;
    if (d1 * d2 > 0) gam = 1.2;
    if (d1 * d2 < 0) gam = 0.7;
    const double st = V.xmax[j] - V.xmin[j];
    double al = xv - gam * (xv - d1 - V.L[j]);
    double au = xv + gam * (V.U[j] - xv + d1);
    al = fmin(fmax(al, xv - 10 * st), xv - 0.01 * st);
    au = fmin(fmax(au, xv + 0.01 * st), xv + 10 * st);
    V.L[j] = al;
    V.U[j] = au;
    double na = V.xmin[j];
    na = fmax(na, al + 0.1 * (xv - al));
    na = fmax(na, xv - S.move * st);
    double nb = V.xmax[j];
    nb = fmin(nb, au - 0.1 * (au - xv));
    nb = fmin(nb, xv + S.move * st);
    V.alpha[j] = na;
    V.beta[j] = nb;
    V.lastdx[j] = d1;
    v[0] += d1 * d1;
  }
  block_partials<1>(v, mx, partial);
}

// fold `cnt` columns of `nb` per-block partials in a fixed order: one 64-lane wave per
// column (grid = cnt), lane-strided partial sums, then a fixed butterfly
__global__ __launch_bounds__(64) void k_mma_fold(const double* __restrict__ partial, int nb, int cnt,
                                                 unsigned maxmask, double* __restrict__ out) {
  const int c = blockIdx.x;
  if (c >= cnt) return;
  const bool mx = (maxmask >> c) & 1u;
  const int lane = threadIdx.x;
  const double* col = partial + (int64_t)c * nb;
  double s = lane < nb ? col[lane] : (mx ? -HUGE_VAL : 0.0);
  for (int b = lane + 64; b < nb; b += 64) {
    const double v = col[b];
    s = mx ? fmax(s, v) : s + v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(s, o, 64);
    s = mx ? fmax(s, v) : s + v;
  }
  if (lane == 0) out[c] = s;
}

}  // namespace tos
