// mma.hip -- method of moving asymptotes (MMA::mma_t, optimization/mma_t.h:159-266,
// mma_t.cpp:18-51, mma_t.cu:15-807) as one persistent workgroup.
//
// The reference runs every vector expression of the primal-dual Newton iteration as a
// separate kernel and solves the (m+1)x(m+1) reduced system on the host, so each Newton
// step costs dozens of launches and several device->host copies.  Here one 1024-thread
// workgroup runs the whole update -- p/q, the epsilon continuation, Newton steps, the
// KKT line search and the asymptote update -- with block reductions in LDS and the small
// system solved by thread 0, in FP64.  Control flow only depends on values every thread
// reads from LDS after a barrier, so all barriers are reached uniformly.
#include "tos_internal.hpp"

namespace tos {

constexpr int MMAX = 4;      // constraints supported
constexpr int MT = 1024;     // threads
constexpr int NW = MT / 64;  // waves

struct MmaDev {
  int n, m;
  double move;
  double *x, *xmin, *xmax, *L, *U, *alpha, *beta, *lastdx, *xi, *eta;
  double *p, *q;          // (m+1) x n
  double *plam, *qlam, *curx, *Dx, *G, *dlx, *dx, *dxi, *deta, *tx, *txi, *teta;
  const double *df, *dg, *g;
  double* sv;             // persistent small state (see SV_*)
  double* out;            // out[0] = converged flag, out[1] = Newton steps, out[2] = final err
};

enum { SV_Y = 0, SV_LAM = SV_Y + MMAX, SV_MU = SV_LAM + MMAX, SV_S = SV_MU + MMAX, SV_Z = SV_S + MMAX,
       SV_ZETA, SV_STOP, SV_COUNT };

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Reduce cnt values per thread (sum or max) across the block; every thread gets the
// totals.  Fixed order -> deterministic.
template <int CNT>
__device__ void bred(double (&v)[CNT], const bool (&ismax)[CNT], double* sh /* NW*CNT + CNT */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < CNT; ++i) v[i] = ismax[i] ? wmax(v[i]) : wsum(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < CNT; ++i) sh[wid * CNT + i] = v[i];
  __syncthreads();
  if (threadIdx.x < CNT) {
    const int i = threadIdx.x;
    double s = sh[i];
    for (int w = 1; w < NW; ++w) s = ismax[i] ? fmax(s, sh[w * CNT + i]) : s + sh[w * CNT + i];
    sh[NW * CNT + i] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < CNT; ++i) v[i] = sh[NW * CNT + i];
  __syncthreads();
}

struct Small {  // w's small part and the scalars of one Newton step (in LDS)
  double y[MMAX], lam[MMAX], mu[MMAX], s[MMAX], z, zeta;
  double a[MMAX], c[MMAX], d[MMAX], b[MMAX];
  double dy[MMAX], dlam[MMAX], dmu[MMAX], ds[MMAX], dz, dzeta;
  double ty[MMAX], tlam[MMAX], tmu[MMAX], ts[MMAX], tz, tzeta;
  double eps, step, old_err, new_err, new_inf, err, lasterr, damp;
  int stays, nsub, overflow, accept, go;
};

// plam/qlam from lambda, then the KKT residual (2-norm, inf-norm) of the point
// (X, XI, ETA, small trial or current) -- kkt_err, mma_t.cu:574-678.
__device__ void kkt(const MmaDev& M, Small& S, const double* X, const double* XI, const double* ETA,
                    const double* y, const double* lam, const double* mu, const double* s, double z, double zeta,
                    double damp, double* sh, double* out2, double* outinf) {
  const int n = M.n, m = M.m;
  double v[3 + MMAX + 3];
  constexpr int CNT = 3 + MMAX + 3;
  bool ismax[CNT];
#pragma unroll
  for (int i = 0; i < CNT; ++i) { v[i] = 0; ismax[i] = i >= 3 + MMAX; }
  for (int j = threadIdx.x; j < n; j += MT) {
    double pl = M.p[j], ql = M.q[j];
    for (int i = 0; i < m; ++i) { pl += lam[i] * M.p[(int64_t)(i + 1) * n + j]; ql += lam[i] * M.q[(int64_t)(i + 1) * n + j]; }
    M.plam[j] = pl; M.qlam[j] = ql;
    const double ux = M.U[j] - X[j], xl = X[j] - M.L[j];
    const double rx = pl / (ux * ux) - ql / (xl * xl) - XI[j] + ETA[j];
    const double r1 = XI[j] * (X[j] - M.alpha[j]) - S.eps;
    const double r2 = ETA[j] * (M.beta[j] - X[j]) - S.eps;
    v[0] += rx * rx; v[1] += r1 * r1; v[2] += r2 * r2;
    for (int i = 0; i < m; ++i) v[3 + i] += M.p[(int64_t)(i + 1) * n + j] / ux + M.q[(int64_t)(i + 1) * n + j] / xl;
    v[3 + MMAX] = fmax(v[3 + MMAX], fabs(rx));
    v[4 + MMAX] = fmax(v[4 + MMAX], fabs(r1));
    v[5 + MMAX] = fmax(v[5 + MMAX], fabs(r2));
  }
  bred<CNT>(v, ismax, sh);
  if (threadIdx.x == 0) {
    double sum = v[0], mx = fmax(-10.0, v[3 + MMAX]);
    auto acc = [&](double r) { sum += r * r; mx = fmax(mx, fabs(r)); };
    for (int i = 0; i < m; ++i) acc(S.c[i] + S.d[i] * y[i] - lam[i] - mu[i]);
    double la = 0;
    for (int i = 0; i < m; ++i) la += lam[i] * S.a[i];
    acc(1.0 - zeta - la);  // a0 = 1
    for (int i = 0; i < m; ++i) acc((v[3 + i] - S.a[i] * z - y[i] + s[i] - S.b[i]) / damp);
    sum += v[1]; mx = fmax(mx, v[4 + MMAX]);
    sum += v[2]; mx = fmax(mx, v[5 + MMAX]);
    for (int i = 0; i < m; ++i) acc(mu[i] * y[i] - S.eps);
    acc(zeta * z - S.eps);
    for (int i = 0; i < m; ++i) acc(lam[i] * s[i] - S.eps);
    *out2 = sqrt(sum);
    *outinf = mx;
  }
  __syncthreads();
}

__global__ __launch_bounds__(MT) void k_mma_update(MmaDev M) {
  __shared__ Small S;
  __shared__ double sh[NW * 32 + 32];
  const int n = M.n, m = M.m;
  const int tid = threadIdx.x;

  if (tid == 0) {
    for (int i = 0; i < m; ++i) { S.a[i] = 0; S.c[i] = 10; S.d[i] = 1; }
  }
  __syncthreads();

  // ---- p, q at the current x (mma_subproblem_t::init, mma_t.cu:15-118) ----
  {
    double v[MMAX];
    bool ismax[MMAX];
#pragma unroll
    for (int i = 0; i < MMAX; ++i) { v[i] = 0; ismax[i] = false; }
    for (int j = tid; j < n; j += MT) {
      const double ux = M.U[j] - M.x[j], xl = M.x[j] - M.L[j];
      const double r = 1e-5 / (M.xmax[j] - M.xmin[j]);
      for (int i = 0; i <= m; ++i) {
        const double d = i == 0 ? M.df[j] : M.dg[(int64_t)(i - 1) * n + j];
        const double dp = fmax(d, 0.0), dm = fmax(-d, 0.0);
        const double pv = ux * ux * (1.001 * dp + 0.001 * dm + r);
        const double qv = xl * xl * (0.001 * dp + 1.001 * dm + r);
        M.p[(int64_t)i * n + j] = pv;
        M.q[(int64_t)i * n + j] = qv;
        if (i > 0) v[i - 1] += pv / ux + qv / xl;
      }
      // reset (mma_t.cu:702-713)
      M.curx[j] = M.x[j];
      const double xn = 0.5 * (M.alpha[j] + M.beta[j]);
      M.x[j] = xn;
      M.xi[j] = fmax(1.0 / (xn - M.alpha[j]), 1.0);
      M.eta[j] = fmax(1.0 / (M.beta[j] - xn), 1.0);
    }
    bred<MMAX>(v, ismax, sh);
    if (tid == 0) {
      for (int i = 0; i < m; ++i) {
        S.b[i] = v[i] - M.g[i];
        S.y[i] = 1; S.lam[i] = 1; S.s[i] = 1; S.mu[i] = fmax(S.c[i] / 2, 1.0);
      }
      S.z = 1; S.zeta = 1; S.eps = 1;
      S.lasterr = 0; S.nsub = 0; S.overflow = 0;
    }
    __syncthreads();
  }
  // plam/qlam with lambda = 1
  for (int j = tid; j < n; j += MT) {
    double pl = M.p[j], ql = M.q[j];
    for (int i = 0; i < m; ++i) { pl += S.lam[i] * M.p[(int64_t)(i + 1) * n + j]; ql += S.lam[i] * M.q[(int64_t)(i + 1) * n + j]; }
    M.plam[j] = pl; M.qlam[j] = ql;
  }
  __syncthreads();

  for (int lev = 0; lev < 6; ++lev) {
    if (S.overflow) break;
    if (tid == 0) { S.err = 1e30; S.damp = 1; S.stays = 0; }
    __syncthreads();
    while (S.err > 0.9 * S.eps) {
      // ---- pass A: Dx, G, delta_x; reduced system sums ----
      constexpr int CA = MMAX + MMAX * MMAX + MMAX;
      double v[CA];
      bool ismax[CA];
#pragma unroll
      for (int i = 0; i < CA; ++i) { v[i] = 0; ismax[i] = false; }
      for (int j = tid; j < n; j += MT) {
        const double xj = M.x[j], ux = M.U[j] - xj, xl = xj - M.L[j];
        const double xa = xj - M.alpha[j], bx = M.beta[j] - xj;
        const double psi = 2 * M.plam[j] / (ux * ux * ux) + 2 * M.qlam[j] / (xl * xl * xl);
        const double Dx = psi + M.xi[j] / xa + M.eta[j] / bx;
        const double phi = M.plam[j] / (ux * ux) - M.qlam[j] / (xl * xl);
        const double dlx = -(S.eps / xa - S.eps / bx - phi);
        M.Dx[j] = Dx; M.dlx[j] = dlx;
        double Gi[MMAX];
        for (int i = 0; i < m; ++i) {
          Gi[i] = M.p[(int64_t)(i + 1) * n + j] / (ux * ux) - M.q[(int64_t)(i + 1) * n + j] / (xl * xl);
          M.G[(int64_t)i * n + j] = Gi[i];
          v[i] += M.p[(int64_t)(i + 1) * n + j] / ux + M.q[(int64_t)(i + 1) * n + j] / xl;  // gproxy
        }
        for (int i = 0; i < m; ++i) {
          for (int k = 0; k < m; ++k) v[MMAX + i * MMAX + k] += Gi[i] * (Gi[k] / Dx);
          v[MMAX + MMAX * MMAX + i] += Gi[i] * (dlx / Dx);
        }
      }
      bred<CA>(v, ismax, sh);
      if (tid == 0) {
        // reduced (m+1) system (solveLinearSystem mma_t.cu:228-328), Gaussian elimination
        const int N1 = m + 1;
        double A[(MMAX + 1) * (MMAX + 1)], rhs[MMAX + 1], Dy[MMAX], dly[MMAX], dll[MMAX];
        for (int i = 0; i < N1 * N1; ++i) A[i] = 0;
        for (int i = 0; i < m; ++i) {
          Dy[i] = S.d[i] + S.mu[i] / S.y[i];
          dly[i] = -(S.lam[i] + S.eps / S.y[i] - S.c[i] - S.d[i] * S.y[i]);
          dll[i] = -(S.z * S.a[i] + S.y[i] + S.b[i] - S.eps / S.lam[i] - v[i]);
        }
        double la = 0;
        for (int i = 0; i < m; ++i) la += S.lam[i] * S.a[i];
        const double dlz = -(la + S.eps / S.z - 1.0);
        for (int i = 0; i < m; ++i) {
          A[i * N1 + i] += S.s[i] / S.lam[i] + 1.0 / Dy[i];
          for (int k = 0; k < m; ++k) A[i * N1 + k] += v[MMAX + i * MMAX + k];
          rhs[i] = dll[i] + dly[i] / Dy[i] - v[MMAX + MMAX * MMAX + i];
          A[i * N1 + m] = S.a[i];
          A[m * N1 + i] = S.a[i];
        }
        A[m * N1 + m] = -S.zeta / S.z;
        rhs[m] = dlz;
        for (int k = 0; k < N1; ++k) {
          int piv = k;
          for (int i = k + 1; i < N1; ++i)
            if (fabs(A[i * N1 + k]) > fabs(A[piv * N1 + k])) piv = i;
          if (piv != k) {
            for (int j = 0; j < N1; ++j) { const double t = A[k * N1 + j]; A[k * N1 + j] = A[piv * N1 + j]; A[piv * N1 + j] = t; }
            const double t = rhs[k]; rhs[k] = rhs[piv]; rhs[piv] = t;
          }
          for (int i = k + 1; i < N1; ++i) {
            const double f = A[i * N1 + k] / A[k * N1 + k];
            for (int j = k; j < N1; ++j) A[i * N1 + j] -= f * A[k * N1 + j];
            rhs[i] -= f * rhs[k];
          }
        }
        for (int i = N1 - 1; i >= 0; --i) {
          double s = rhs[i];
          for (int j = i + 1; j < N1; ++j) s -= A[i * N1 + j] * rhs[j];
          rhs[i] = s / A[i * N1 + i];
        }
        for (int i = 0; i < m; ++i) {
          S.dlam[i] = rhs[i];
          S.dy[i] = rhs[i] / Dy[i] - dly[i] / Dy[i];
          S.dmu[i] = S.eps / S.y[i] - S.dy[i] * S.mu[i] / S.y[i] - S.mu[i];
          S.ds[i] = S.eps / S.lam[i] - S.s[i] / S.lam[i] * S.dlam[i] - S.s[i];
        }
        S.dz = rhs[m];
        S.dzeta = S.eps / S.z - S.zeta - (S.zeta / S.z) * S.dz;
      }
      __syncthreads();
      // ---- pass B: dx, dxi, deta and the step bound (fraction 0.99 to the boundary) ----
      {
        constexpr int CB = 3;
        double t[CB] = {-1e300, -1e300, -1e300};
        const bool mx[CB] = {true, true, true};
        const double cf = 1.0 / 0.99;
        for (int j = tid; j < n; j += MT) {
          double dxv = -M.dlx[j] / M.Dx[j];
          for (int i = 0; i < m; ++i) dxv -= S.dlam[i] * M.G[(int64_t)i * n + j] / M.Dx[j];
          const double xa = M.x[j] - M.alpha[j], bx = M.beta[j] - M.x[j];
          const double dxi = S.eps / xa - dxv / xa * M.xi[j] - M.xi[j];
          const double deta = S.eps / bx + M.eta[j] / bx * dxv - M.eta[j];
          M.dx[j] = dxv; M.dxi[j] = dxi; M.deta[j] = deta;
          t[0] = fmax(t[0], -cf * (dxv / xa));
          t[1] = fmax(t[1], cf * (dxv / bx));
          t[2] = fmax(t[2], fmax(-cf * dxi / M.xi[j], -cf * deta / M.eta[j]));
        }
        bred<CB>(t, mx, sh);
        if (tid == 0) {
          double t3 = t[2];
          for (int i = 0; i < m; ++i) t3 = fmax(t3, -cf * S.dy[i] / S.y[i]);
          t3 = fmax(t3, -cf * S.dz / S.z);
          for (int i = 0; i < m; ++i) t3 = fmax(t3, -cf * S.dlam[i] / S.lam[i]);
          for (int i = 0; i < m; ++i) t3 = fmax(t3, -cf * S.dmu[i] / S.mu[i]);
          t3 = fmax(t3, -cf * S.dzeta / S.zeta);
          for (int i = 0; i < m; ++i) t3 = fmax(t3, -cf * S.ds[i] / S.s[i]);
          S.step = 1.0 / fmax(fmax(t[0], t[1]), fmax(t3, 1.0));
        }
        __syncthreads();
      }
      // ---- line search on the KKT 2-norm (search_step, mma_t.cu:501-571) ----
      {
        double e2, einf;
        kkt(M, S, M.x, M.xi, M.eta, S.y, S.lam, S.mu, S.s, S.z, S.zeta, 1.0, sh, &e2, &einf);
        if (tid == 0) { S.old_err = e2; S.new_err = 1e30; S.new_inf = 1e30; S.accept = 0; }
        __syncthreads();
        while (S.step > 1e-12 && S.new_err > S.old_err) {
          const double st = S.step;
          for (int j = tid; j < n; j += MT) {
            M.tx[j] = M.x[j] + M.dx[j] * st;
            M.txi[j] = M.xi[j] + M.dxi[j] * st;
            M.teta[j] = M.eta[j] + M.deta[j] * st;
          }
          if (tid == 0) {
            for (int i = 0; i < m; ++i) {
              S.ty[i] = S.y[i] + S.dy[i] * st;
              S.tlam[i] = S.lam[i] + S.dlam[i] * st;
              S.tmu[i] = S.mu[i] + S.dmu[i] * st;
              S.ts[i] = S.s[i] + S.ds[i] * st;
            }
            S.tz = S.z + S.dz * st;
            S.tzeta = S.zeta + S.dzeta * st;
          }
          __syncthreads();
          kkt(M, S, M.tx, M.txi, M.teta, S.ty, S.tlam, S.tmu, S.ts, S.tz, S.tzeta, S.damp, sh, &e2, &einf);
          if (tid == 0) {
            S.new_err = e2;
            S.new_inf = einf;
            S.accept = e2 < S.old_err;
            if (!S.accept) S.step = st / 2;
          }
          __syncthreads();
          if (S.accept) {
            for (int j = tid; j < n; j += MT) { M.x[j] = M.tx[j]; M.xi[j] = M.txi[j]; M.eta[j] = M.teta[j]; }
            if (tid == 0) {
              for (int i = 0; i < m; ++i) { S.y[i] = S.ty[i]; S.lam[i] = S.tlam[i]; S.mu[i] = S.tmu[i]; S.s[i] = S.ts[i]; }
              S.z = S.tz; S.zeta = S.tzeta;
            }
            __syncthreads();
            break;
          }
        }
        if (tid == 0) {
          S.err = S.new_inf;
          if ((S.err - S.lasterr) / S.err < 1e-5) S.stays++; else S.stays = 0;
          S.lasterr = S.err;
          if (S.stays > 6) { S.damp *= 10; S.stays = 0; }
          if (S.nsub++ > 80) S.overflow = 1;
        }
        __syncthreads();
        if (S.overflow) break;
      }
    }
    if (tid == 0) S.eps *= 0.1;
    __syncthreads();
  }

  // ---- asymptotes and move limits (adjust_asym, mma_t.cu:164-226) ----
  {
    double v[1] = {0};
    const bool mx[1] = {false};
    for (int j = tid; j < n; j += MT) {
      const double xv = M.x[j], d1 = xv - M.curx[j], d2 = M.lastdx[j];
      double gam = 1;
      if (d1 * d2 > 0) gam = 1.2;
      if (d1 * d2 < 0) gam = 0.7;
      const double st = M.xmax[j] - M.xmin[j];
      double al = xv - gam * (xv - d1 - M.L[j]);
      double au = xv + gam * (M.U[j] - xv + d1);
      al = fmin(fmax(al, xv - 10 * st), xv - 0.01 * st);
      au = fmin(fmax(au, xv + 0.01 * st), xv + 10 * st);
      M.L[j] = al; M.U[j] = au;
      double na = M.xmin[j];
      na = fmax(na, al + 0.1 * (xv - al));
      na = fmax(na, xv - M.move * st);
      double nb = M.xmax[j];
      nb = fmin(nb, au - 0.1 * (au - xv));
      nb = fmin(nb, xv + M.move * st);
      M.alpha[j] = na; M.beta[j] = nb;
      M.lastdx[j] = d1;
      v[0] += d1 * d1;
    }
    bred<1>(v, mx, sh);
    if (tid == 0) {
      if (sqrt(v[0]) < 1e-3) M.sv[SV_STOP] += 1;
      M.out[0] = M.sv[SV_STOP] > 10 ? 1.0 : 0.0;
      M.out[1] = S.nsub;
      M.out[2] = S.err;
      for (int i = 0; i < m; ++i) {
        M.sv[SV_Y + i] = S.y[i]; M.sv[SV_LAM + i] = S.lam[i]; M.sv[SV_MU + i] = S.mu[i]; M.sv[SV_S + i] = S.s[i];
      }
      M.sv[SV_Z] = S.z; M.sv[SV_ZETA] = S.zeta;
    }
  }
}

}  // namespace tos
