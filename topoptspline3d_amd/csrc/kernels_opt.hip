// kernels_opt.hip -- element-wise and reduction kernels of the optimisation step:
// sensitivities, filter, OC bisection, Heaviside projection, spline evaluation and its
// adjoint, deterministic FP64 reductions and the PCG vector updates.
#include "tos_internal.hpp"

namespace tos {

__device__ __forceinline__ int64_t tid64() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// ---------------------------------------------------------------------------------------
// Deterministic reductions: every block writes one partial (fixed block -> element
// mapping), one block sums the partials in a fixed order.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// returns the block total in thread 0 (blockDim = 256)
__device__ __forceinline__ double block_sum256(double v) {
  __shared__ double sh[4];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return s;
}
__device__ __forceinline__ double block_max256(double v) {
  __shared__ double sh[4];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0) s = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
  __syncthreads();
  return s;
}

// Sum `np` partials (two interleaved sets when two != 0) into scal[slot0], scal[slot1],
// then apply a small scalar program `op`.
enum { OP_NONE = 0, OP_ALPHA, OP_BETA, OP_RN2 };
// first stage of a two-stage reduction of np (x2) block partials: block b sums a fixed
// contiguous segment into out[b] (and out[kStage + b]); deterministic order.
constexpr int kStage = 256;
__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ partial, int np, int two,
                                                         double* __restrict__ out) {
  const int seg = (np + kStage - 1) / kStage;
  const int i0 = blockIdx.x * seg, i1 = min(np, i0 + seg);
  double s0 = 0, s1 = 0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
    s0 += partial[i];
    if (two) s1 += partial[np + i];
  }
  s0 = block_sum256(s0);
  s1 = block_sum256(s1);
  if (threadIdx.x == 0) {
    out[blockIdx.x] = s0;
    out[kStage + blockIdx.x] = s1;
  }
}

__global__ __launch_bounds__(256) void k_finish(const double* __restrict__ partial, int np, int two, int slot0,
                                                int slot1, int op, double* __restrict__ scal) {
  double s0 = 0, s1 = 0;
  for (int i = threadIdx.x; i < np; i += 256) {
    s0 += partial[i];
    if (two) s1 += partial[np + i];
  }
  s0 = block_sum256(s0);
  s1 = block_sum256(s1);
  if (threadIdx.x == 0) {
    scal[slot0] = s0;
    if (two) scal[slot1] = s1;
    if (op == OP_ALPHA) {
      scal[S_ALPHA] = scal[S_RZ] / scal[S_PQ];
      if (!two && slot1 > 0) scal[slot1] = scal[S_ALPHA];  // (the direction's alpha kept)
    }
    if (op == OP_BETA) {  // flexible (Polak-Ribiere) beta = -alpha <z,q> / rz_old
      scal[S_BETA] = -scal[S_ALPHA] * scal[S_ZQ] / scal[S_RZ];
      scal[S_RZ] = scal[S_ZR];
    }
  }
}

// ---------------------------------------------------------------------------------------
// Element weights w = rho^p (FP64 and FP32 copies), 0 for inactive / ghost elements.
// ---------------------------------------------------------------------------------------
__global__ void k_weights(LevelView L, const uint8_t* __restrict__ eact, const float* __restrict__ rho,
                          double penal, double* __restrict__ w64, float* __restrict__ w32) {
  const int64_t i = tid64();
  if (i >= L.NE) return;
  const double v = eact[i] ? pow((double)rho[i], penal) : 0.0;
  w64[i] = v;
  w32[i] = (float)v;
}

// ---------------------------------------------------------------------------------------
// Sensitivity s_e = -p rho^(p-1) u_e^T Ke u_e (computeSensitivity_kernel
// optimization/optimization.cu:96-159), one thread per element, no atomics; MODE 1
// writes the element compliance rho^p u_e^T Ke u_e (elementCompliance_kernel
// grid/Grid.cu:5389-5430) as FP64 into out64.
// ---------------------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256) void k_sensitivity(LevelView L, const uint8_t* __restrict__ eact,
                                                     const float* __restrict__ rho, const double* __restrict__ Ke,
                                                     const double* __restrict__ u, double penal,
                                                     float* __restrict__ sens, double* __restrict__ out64) {
  const int64_t t = tid64();
  if (t >= (int64_t)L.ex * L.ey * L.ez) return;
  const int x = t % L.ex, y = (t / L.ex) % L.ey, z = t / ((int64_t)L.ex * L.ey);
  const int64_t e = L.pelem(x, y, z);
  if (!eact[e]) {
    if (MODE == 0) sens[e] = 0.f;
    else out64[t] = 0.0;
    return;
  }
  const int64_t n0 = L.pnode(x, y, z);
  double ue[24];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t q = n0 + (j & 1) + (int64_t)L.NX * (((j >> 1) & 1) + (int64_t)L.NY * (j >> 2));
#pragma unroll
    for (int c = 0; c < 3; ++c) ue[3 * j + c] = u[c * L.NP + q];
  }
  // u^T Ke u over the upper triangle (Ke is symmetric): sum_i u_i (Ke_ii u_i + 2 sum_{j>i} Ke_ij u_j)
  double en = 0;
#pragma unroll
  for (int i = 0; i < 24; ++i) {
    double s = 0;
#pragma unroll
    for (int j = i + 1; j < 24; ++j) s += Ke[i * 24 + j] * ue[j];
    en += ue[i] * (Ke[i * 25] * ue[i] + 2.0 * s);
  }
  const double r = rho[e];
  if (MODE == 0) sens[e] = (float)(-penal * pow(r, penal - 1.0) * en);
  else out64[t] = pow(r, penal) * en;
}

// ---------------------------------------------------------------------------------------
// Filter (filterSensitivity_kernel grid/Grid.cu:4131-4216): weighted mean over the active
// elements of a ball; offsets and weights come from a host-built table.
// k_filter_zm<H> (table offsets within H per axis): a workgroup owns a 64 x 8 column of
// elements and marches it along z over kFzZ planes, keeping the 2H + 1 element planes of the
// current output plane, each with an H-element halo in x and y, in an LDS ring as
// (value, 1) for active elements and (0, 0) otherwise; every thread forms two outputs per
// plane from LDS.  Each element is loaded (64 + 2H)(8 + 2H)/512 * (kFzZ + 2H)/kFzZ times
// (1.8 for H = 2) instead of 4.25 times with cube tiles of 64 x 4 x 4.  Both kernels walk
// the table in its order with FP64 sums, so they give the same result (an inactive or
// outside tap adds exactly zero to both sums).  k_filter serves larger radii.
// ---------------------------------------------------------------------------------------
constexpr int kFzX = 64, kFzY = 8, kFzZ = 32;

template <int H>
__global__ __launch_bounds__(256) void k_filter_zm(LevelView L, const uint8_t* __restrict__ eact,
                                                   const float* __restrict__ src, float* __restrict__ dst,
                                                   const int* __restrict__ off, const double* __restrict__ wt,
                                                   int nt) {
  constexpr int LX = kFzX + 2 * H, LY = kFzY + 2 * H, NZ = 2 * H + 1, PL = LX * LY;
  constexpr int NT = (2 * H + 1) * (2 * H + 1) * (2 * H + 1);  // table size bound
  __shared__ float2 ring[NZ * PL];
  __shared__ int tdz[NT], tin[NT];  // tap plane (dz + H) and offset within a plane
  __shared__ int toffz[NT];         // LDS offset of each tap for the current output plane
  __shared__ double twt[NT];
  const int x0 = blockIdx.x * kFzX, y0 = blockIdx.y * kFzY, z0 = blockIdx.z * kFzZ;
  for (int i = threadIdx.x; i < nt; i += 256) {
    tdz[i] = off[3 * i + 2] + H;
    tin[i] = off[3 * i] + LX * off[3 * i + 1];
    twt[i] = wt[i];
  }
  const int tx = threadIdx.x % kFzX, ty = threadIdx.x / kFzX;  // rows ty and ty + 4
  const int nplanes = min(kFzZ, L.ez - z0) + 2 * H;
  for (int zi = 0; zi < nplanes; ++zi) {
    const int gz = z0 - H + zi;  // plane loaded in this step, into slot zi % NZ
    float2* pl = ring + (zi % NZ) * PL;
    for (int i = threadIdx.x; i < PL; i += 256) {
      const int gx = x0 + i % LX - H, gy = y0 + i / LX - H;
      const bool in = gx >= 0 && gy >= 0 && gz >= 0 && gx < L.ex && gy < L.ey && gz < L.ez;
      const int64_t q = L.pelem(in ? gx : 0, in ? gy : 0, in ? gz : 0);
      const uint8_t a = eact[q];  // both loads issue before either is used
      const float v = src[q];
      pl[i] = (in && a) ? make_float2(v, 1.f) : make_float2(0.f, 0.f);
    }
    // output plane gz - H reads the planes loaded in steps zi - 2H .. zi
    if (zi >= 2 * H)
      for (int i = threadIdx.x; i < nt; i += 256) toffz[i] = ((zi - 2 * H + tdz[i]) % NZ) * PL + tin[i];
    __syncthreads();
    if (zi >= 2 * H) {
      const float2* cen = ring + ((zi - H) % NZ) * PL;
      for (int r = 0; r < 2; ++r) {
        const int yy = ty + 4 * r, x = x0 + tx, y = y0 + yy;
        if (x >= L.ex || y >= L.ey) continue;
        const int64_t e = L.pelem(x, y, gz - H);
        const int c = (yy + H) * LX + tx + H;
        if (cen[c].y == 0.f) {
          dst[e] = 0.f;
          continue;
        }
        double gs = 0, ws = 0;
        for (int i = 0; i < nt; ++i) {
          const float2 sv = ring[toffz[i] + c];
          gs += twt[i] * (double)sv.x;
          ws += twt[i] * (double)sv.y;
        }
        dst[e] = (float)(gs / ws);
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_filter(LevelView L, const uint8_t* __restrict__ eact,
                                                const float* __restrict__ src, float* __restrict__ dst,
                                                const int* __restrict__ off, const double* __restrict__ wt, int nt) {
  const int64_t t = tid64();
  if (t >= (int64_t)L.ex * L.ey * L.ez) return;
  const int x = t % L.ex, y = (t / L.ex) % L.ey, z = t / ((int64_t)L.ex * L.ey);
  const int64_t e = L.pelem(x, y, z);
  if (!eact[e]) { dst[e] = 0.f; return; }
  double gs = 0, ws = 0;
  for (int i = 0; i < nt; ++i) {
    const int dx = off[3 * i], dy = off[3 * i + 1], dz = off[3 * i + 2];
    const int qx = x + dx, qy = y + dy, qz = z + dz;
    if (qx < 0 || qy < 0 || qz < 0 || qx >= L.ex || qy >= L.ey || qz >= L.ez) continue;
    const int64_t q = e + dx + (int64_t)L.EX * (dy + (int64_t)L.EY * dz);
    if (!eact[q]) continue;
    gs += wt[i] * (double)src[q];
    ws += wt[i];
  }
  dst[e] = (float)(gs / ws);
}

// ---------------------------------------------------------------------------------------
// OC bisection on the device (updateDensities / trySensMultiplier_kernel
// optimization/optimization.cu:355-460).  Each trial: partial sums of the trial density;
// the finisher updates lo/hi, the next midpoint and the done flag, so the host only
// waits once at the end.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float oc_rho(float r0, float s, double g, double move, double damp, double rmin,
                                        bool shell) {
  const double gg = s > 0 ? 0.0 : -(double)s;
  double rn = (double)r0 * pow(gg / g, damp);
  rn = fmin(fmax(rn, (double)r0 - move), (double)r0 + move);
  rn = fmin(fmax(rn, rmin), 1.0);
  if (shell) rn = 1.0;
  return (float)rn;
}

__global__ __launch_bounds__(256) void k_absmax_partial(LevelView L, const uint8_t* __restrict__ eact,
                                                        const float* __restrict__ s, double* __restrict__ partial) {
  double m = 0;
  for (int64_t i = tid64(); i < L.NE; i += (int64_t)gridDim.x * blockDim.x)
    if (eact[i]) m = fmax(m, fabs((double)s[i]));
  m = block_max256(m);
  if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

__global__ __launch_bounds__(256) void k_oc_init(const double* __restrict__ partial, int np, double* __restrict__ scal) {
  double m = 0;
  for (int i = threadIdx.x; i < np; i += 256) m = fmax(m, partial[i]);
  m = block_max256(m);
  if (threadIdx.x == 0) {
    scal[S_OC_GMAX] = m;
    scal[S_OC_LO] = 0.0;
    scal[S_OC_HI] = m;
    scal[S_OC_G] = 0.5 * m;
    scal[S_OC_DONE] = 0.0;
    scal[S_OC_STEPS] = 0.0;
  }
}

__global__ __launch_bounds__(256) void k_oc_trial(LevelView L, const uint8_t* __restrict__ eact,
                                                  const uint8_t* __restrict__ shell, const float* __restrict__ rho,
                                                  const float* __restrict__ s, double move, double damp, double rmin,
                                                  const double* __restrict__ scal, double* __restrict__ partial) {
  if (scal[S_OC_DONE] != 0.0) return;
  const double g = scal[S_OC_G];
  double v = 0;
  for (int64_t i = tid64(); i < L.NE; i += (int64_t)gridDim.x * blockDim.x)
    if (eact[i]) v += (double)oc_rho(rho[i], s[i], g, move, damp, rmin, shell[i] != 0);
  v = block_sum256(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// nact: active element count, read from scal[S_VOL] when nact_dev is set.  Stop rules:
// |V - vgoal| <= 1e-4 or max_trials trials (updateDensities: 31, optimization.cu:446), and
// with rel_bracket > 0 also (hi - lo)/(lo + hi) <= rel_bracket (updateDensities2: 1e-3 and
// 51 trials, optimization.cu:534)
__global__ __launch_bounds__(256) void k_oc_decide(const double* __restrict__ partial, int np, double nact,
                                                   int nact_dev, double vgoal, int max_trials, double rel_bracket,
                                                   double* __restrict__ scal) {
  if (scal[S_OC_DONE] != 0.0) return;
  double v = 0;
  for (int i = threadIdx.x; i < np; i += 256) v += partial[i];
  v = block_sum256(v);
  if (threadIdx.x == 0) {
    const double V = v / (nact_dev ? scal[S_VOL] : nact);
    const double g = scal[S_OC_G];
    scal[S_OC_V] = V;
    scal[S_OC_STEPS] += 1.0;
    if (V > vgoal) scal[S_OC_LO] = g;
    else if (V < vgoal) scal[S_OC_HI] = g;
    const double lo = scal[S_OC_LO], hi = scal[S_OC_HI];
    const bool bracket = rel_bracket > 0 && !((hi - lo) / (lo + hi) > rel_bracket);
    if (fabs(V - vgoal) <= 1e-4 || bracket || scal[S_OC_STEPS] >= (double)max_trials) scal[S_OC_DONE] = 1.0;
    else scal[S_OC_G] = 0.5 * (lo + hi);
  }
}

__global__ void k_oc_apply(LevelView L, const uint8_t* __restrict__ eact, const uint8_t* __restrict__ shell,
                           float* __restrict__ rho, const float* __restrict__ s, double move, double damp,
                           double rmin, const double* __restrict__ scal) {
  const int64_t i = tid64();
  if (i >= L.NE || !eact[i]) return;
  rho[i] = oc_rho(rho[i], s[i], scal[S_OC_G], move, damp, rmin, shell[i] != 0);
}

// Heaviside projection of one design value (computeProjection_kernel
// optimization/optimization.cu:313-337): eta = 0.5, clamp [rmin, 1], shell -> 1
__device__ __forceinline__ float heaviside_rho(double x, double beta, double rmin, bool shell) {
  const double eta = 0.5, den = tanh(beta * eta) + tanh(beta * (1 - eta));
  double r = (tanh(beta * eta) + tanh(beta * (x - eta))) / den;
  r = fmin(fmax(r, rmin), 1.0);
  return shell ? 1.0f : (float)r;
}

// projected OC trial (updateDensities2 optimization.cu:490-520): the design variable x
// (rho_pre) gets the OC update, its projection enters the volume
__global__ __launch_bounds__(256) void k_oc_trial_proj(LevelView L, const uint8_t* __restrict__ eact,
                                                       const uint8_t* __restrict__ shell,
                                                       const float* __restrict__ x, const float* __restrict__ s,
                                                       double move, double damp, double rmin, double beta,
                                                       const double* __restrict__ scal, double* __restrict__ partial) {
  if (scal[S_OC_DONE] != 0.0) return;
  const double g = scal[S_OC_G];
  double v = 0;
  for (int64_t i = tid64(); i < L.NE; i += (int64_t)gridDim.x * blockDim.x)
    if (eact[i]) {
      const bool sh = shell[i] != 0;
      v += (double)heaviside_rho((double)oc_rho(x[i], s[i], g, move, damp, rmin, sh), beta, rmin, sh);
    }
  v = block_sum256(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// final trial as the new design variable, its projection as the density, and per-block
// max / sum of |x_new - x_old| (partial[b], partial[gridDim.x + b])
__global__ __launch_bounds__(256) void k_oc_apply_proj(LevelView L, const uint8_t* __restrict__ eact,
                                                       const uint8_t* __restrict__ shell, float* __restrict__ rho,
                                                       float* __restrict__ x, const float* __restrict__ s, double move,
                                                       double damp, double rmin, double beta,
                                                       const double* __restrict__ scal, double* __restrict__ partial) {
  const double g = scal[S_OC_G];
  double m = 0, sum = 0;
  for (int64_t i = tid64(); i < L.NE; i += (int64_t)gridDim.x * blockDim.x)
    if (eact[i]) {
      const bool sh = shell[i] != 0;
      const float x0 = x[i];
      const float xn = oc_rho(x0, s[i], g, move, damp, rmin, sh);
      const double d = fabs((double)xn - (double)x0);
      m = fmax(m, d);
      sum += d;
      x[i] = xn;
      rho[i] = heaviside_rho((double)xn, beta, rmin, sh);
    }
  m = block_max256(m);
  sum = block_sum256(sum);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = m;
    partial[gridDim.x + blockIdx.x] = sum;
  }
}

// folds k_oc_apply_proj's partials: scal[S_TMP0] = max, scal[S_TMP1] = sum
__global__ __launch_bounds__(256) void k_change_fold(const double* __restrict__ partial, int np,
                                                     double* __restrict__ scal) {
  double m = 0, sum = 0;
  for (int i = threadIdx.x; i < np; i += 256) {
    m = fmax(m, partial[i]);
    sum += partial[np + i];
  }
  m = block_max256(m);
  sum = block_sum256(sum);
  if (threadIdx.x == 0) {
    scal[S_TMP0] = m;
    scal[S_TMP1] = sum;
  }
}

__global__ __launch_bounds__(256) void k_sum_active(LevelView L, const uint8_t* __restrict__ eact,
                                                    const float* __restrict__ v, double* __restrict__ partial) {
  double s = 0;
  for (int64_t i = tid64(); i < L.NE; i += (int64_t)gridDim.x * blockDim.x)
    if (eact[i]) s += (double)v[i];
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------------------
// OC over a dense element array (the no-allocation entry point): block b covers the
// elements [b*chunk, (b+1)*chunk) with stride 256 per thread, so the partial count is
// ceil(n / chunk); masks may be null (every element active / no shell).  Same trial
// density (oc_rho) and decision kernel as the context's OC.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_absmax_count_flat(int64_t n, int64_t chunk, const uint8_t* __restrict__ act,
                                                           const float* __restrict__ s, int nb,
                                                           double* __restrict__ partial) {
  const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  double m = 0, cnt = 0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256)
    if (!act || act[i]) {
      m = fmax(m, fabs((double)s[i]));
      cnt += 1.0;
    }
  m = block_max256(m);
  cnt = block_sum256(cnt);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = m;
    partial[nb + blockIdx.x] = cnt;
  }
}

// folds the k_absmax_count_flat partials: bisection bounds as k_oc_init, active count into
// scal[S_VOL]
__global__ __launch_bounds__(256) void k_oc_init_flat(const double* __restrict__ partial, int nb,
                                                      double* __restrict__ scal) {
  double m = 0, cnt = 0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    m = fmax(m, partial[i]);
    cnt += partial[nb + i];
  }
  m = block_max256(m);
  cnt = block_sum256(cnt);
  if (threadIdx.x == 0) {
    scal[S_OC_GMAX] = m;
    scal[S_OC_LO] = 0.0;
    scal[S_OC_HI] = m;
    scal[S_OC_G] = 0.5 * m;
    scal[S_OC_DONE] = 0.0;
    scal[S_OC_STEPS] = 0.0;
    scal[S_VOL] = cnt;
  }
}

__global__ __launch_bounds__(256) void k_oc_trial_flat(int64_t n, int64_t chunk, const uint8_t* __restrict__ act,
                                                       const uint8_t* __restrict__ shell, const float* __restrict__ rho,
                                                       const float* __restrict__ s, double move, double damp,
                                                       double rmin, const double* __restrict__ scal,
                                                       double* __restrict__ partial) {
  if (scal[S_OC_DONE] != 0.0) return;
  const double g = scal[S_OC_G];
  const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  double v = 0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256)
    if (!act || act[i]) v += (double)oc_rho(rho[i], s[i], g, move, damp, rmin, shell && shell[i]);
  v = block_sum256(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// active elements get the OC density of the final threshold, the others keep their input
__global__ void k_oc_apply_flat(int64_t n, const uint8_t* __restrict__ act, const uint8_t* __restrict__ shell,
                                const float* rho, const float* __restrict__ s, float* rho_out, double move,
                                double damp, double rmin, const double* __restrict__ scal) {
  const int64_t i = tid64();
  if (i >= n) return;
  const float r0 = rho[i];
  rho_out[i] = (!act || act[i]) ? oc_rho(r0, s[i], scal[S_OC_G], move, damp, rmin, shell && shell[i]) : r0;
}

// ---------------------------------------------------------------------------------------
// Heaviside projection (computeProjection_kernel / computeProjectionSensitivity_kernel
// optimization/optimization.cu:161-183, 313-337)
// ---------------------------------------------------------------------------------------
__global__ void k_project(LevelView L, const uint8_t* __restrict__ eact, const uint8_t* __restrict__ shell,
                          float* __restrict__ rho, float* __restrict__ rho_pre, double beta, double rmin) {
  const int64_t i = tid64();
  if (i >= L.NE || !eact[i]) return;
  const float r0 = rho[i];
  rho_pre[i] = r0;
  rho[i] = heaviside_rho((double)r0, beta, rmin, shell[i] != 0);
}

__global__ void k_project_sens(LevelView L, const uint8_t* __restrict__ eact, const float* __restrict__ rho_pre,
                               float* __restrict__ s, double beta) {
  const int64_t i = tid64();
  if (i >= L.NE || !eact[i]) return;
  const double eta = 0.5, den = tanh(beta * eta) + tanh(beta * (1 - eta));
  const double t = tanh(beta * ((double)rho_pre[i] - eta));
  s[i] = (float)((double)s[i] * (beta * (1 - t * t) / den));
}

__global__ void k_mask_inactive(LevelView L, const uint8_t* __restrict__ eact, float* __restrict__ v) {
  const int64_t i = tid64();
  if (i < L.NE && !eact[i]) v[i] = 0.f;
}

__global__ void k_fill_active(LevelView L, const uint8_t* __restrict__ eact, float* __restrict__ v, float val) {
  const int64_t i = tid64();
  if (i < L.NE) v[i] = eact[i] ? val : 0.f;
}

// ---------------------------------------------------------------------------------------
// Spline (coeff2density grid/Grid.cu:3325-3456; adjoint :3643-3859).  Per-axis tables
// give, for each element index along an axis, `order` (coefficient index, basis value)
// pairs; the 3-D value is their tensor product.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_spline_eval(LevelView L, const uint8_t* __restrict__ eact,
                                                     const uint8_t* __restrict__ shell, const float* __restrict__ coef,
                                                     int ncx, int ncy, int order,
                                                     const int* __restrict__ ix, const float* __restrict__ wx,
                                                     const uint8_t* __restrict__ okx,
                                                     const int* __restrict__ iy, const float* __restrict__ wy,
                                                     const uint8_t* __restrict__ oky,
                                                     const int* __restrict__ iz, const float* __restrict__ wz,
                                                     const uint8_t* __restrict__ okz,
                                                     double rmin, float* __restrict__ rho) {
  const int64_t t = tid64();
  if (t >= (int64_t)L.ex * L.ey * L.ez) return;
  const int x = t % L.ex, y = (t / L.ex) % L.ey, z = t / ((int64_t)L.ex * L.ey);
  const int64_t e = L.pelem(x, y, z);
  if (!eact[e]) { rho[e] = 0.f; return; }
  float v;
  if (!okx[x] || !oky[y] || !okz[z]) {
    v = -0.2f;
  } else {
    v = 0.f;
    for (int k = 0; k < order; ++k) {
      const float wk = wz[z * order + k];
      const int64_t bk = (int64_t)iz[z * order + k] * ncy;
      for (int j = 0; j < order; ++j) {
        const float wjk = wy[y * order + j] * wk;
        const int64_t bj = (bk + iy[y * order + j]) * ncx;
        float acc = 0.f;
        for (int i = 0; i < order; ++i) acc += coef[bj + ix[x * order + i]] * wx[x * order + i];
        v += acc * wjk;
      }
    }
  }
  double r = fmin(fmax((double)v, rmin), 1.0);
  if (shell[e]) r = 1.0;
  rho[e] = (float)r;
}

// k_spline_eval with the order known at compile time: the basis tables and all ORDER^3
// coefficients of an element are loaded together instead of one dependent load at a time.
// Same products in the same order as the loops above.
template <int ORDER>
__global__ __launch_bounds__(256) void k_spline_eval_o(LevelView L, const uint8_t* __restrict__ eact,
                                                       const uint8_t* __restrict__ shell,
                                                       const float* __restrict__ coef, int ncx, int ncy,
                                                       const int* __restrict__ ix, const float* __restrict__ wx,
                                                       const uint8_t* __restrict__ okx,
                                                       const int* __restrict__ iy, const float* __restrict__ wy,
                                                       const uint8_t* __restrict__ oky,
                                                       const int* __restrict__ iz, const float* __restrict__ wz,
                                                       const uint8_t* __restrict__ okz, double rmin,
                                                       float* __restrict__ rho) {
  const int64_t t = tid64();
  if (t >= (int64_t)L.ex * L.ey * L.ez) return;
  const int x = t % L.ex, y = (t / L.ex) % L.ey, z = t / ((int64_t)L.ex * L.ey);
  const int64_t e = L.pelem(x, y, z);
  if (!eact[e]) { rho[e] = 0.f; return; }
  float v;
  if (!okx[x] || !oky[y] || !okz[z]) {
    v = -0.2f;
  } else {
    int jx[ORDER], jy[ORDER], jz[ORDER];
    float ax[ORDER], ay[ORDER], az[ORDER];
#pragma unroll
    for (int i = 0; i < ORDER; ++i) {
      jx[i] = ix[x * ORDER + i]; ax[i] = wx[x * ORDER + i];
      jy[i] = iy[y * ORDER + i]; ay[i] = wy[y * ORDER + i];
      jz[i] = iz[z * ORDER + i]; az[i] = wz[z * ORDER + i];
    }
    float cv[ORDER][ORDER][ORDER];
#pragma unroll
    for (int k = 0; k < ORDER; ++k)
#pragma unroll
      for (int j = 0; j < ORDER; ++j)
#pragma unroll
        for (int i = 0; i < ORDER; ++i)
          cv[k][j][i] = coef[((int64_t)jz[k] * ncy + jy[j]) * ncx + jx[i]];
    v = 0.f;
#pragma unroll
    for (int k = 0; k < ORDER; ++k) {
#pragma unroll
      for (int j = 0; j < ORDER; ++j) {
        const float wjk = ay[j] * az[k];
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < ORDER; ++i) acc += cv[k][j][i] * ax[i];
        v += acc * wjk;
      }
    }
  }
  double r = fmin(fmax((double)v, rmin), 1.0);
  if (shell[e]) r = 1.0;
  rho[e] = (float)r;
}

// First adjoint stage (x) read straight from the padded element field g.  A workgroup
// (256 threads) takes `rows` element rows (y0 .. y0 + rows - 1, z): it stages the
// coefficient lists of axis x (cptr, celm, cw: the same for every row) and the rows' masked
// values (active, outside the shell, zero when the y or z row lies outside the spline
// support; x is masked by the lists, which hold only supported elements) in LDS with
// coalesced loads, then forms out[c, y, z] = the FP64 sum over the list of coefficient c
// of w * row[x] in list order -- the same sum as a masked copy followed by k_spline_stage
// axis 0.  Dynamic LDS: rows * ex floats, n0 + 1 + 2 * nq words (nq = cptr[n0]).
__global__ __launch_bounds__(256) void k_spline_stage0(LevelView L, const uint8_t* __restrict__ eact,
                                                       const uint8_t* __restrict__ shell, const float* __restrict__ g,
                                                       const uint8_t* __restrict__ oky, const uint8_t* __restrict__ okz,
                                                       double* __restrict__ out, int n0, const int* __restrict__ cptr,
                                                       const int* __restrict__ celm, const float* __restrict__ cw,
                                                       int rows) {
  extern __shared__ float sp_lds[];
  const int nq = cptr[n0];
  float* srow = sp_lds;                           // rows x ex
  int* tcp = (int*)(sp_lds + rows * L.ex);        // n0 + 1
  int* tce = tcp + n0 + 1;                        // nq
  float* tcw = (float*)(tce + nq);                // nq
  const unsigned nyb = (unsigned)((L.ey + rows - 1) / rows);
  const int y0 = (int)(blockIdx.x % nyb) * rows, z = (int)(blockIdx.x / nyb);
  for (int i = threadIdx.x; i <= n0; i += 256) tcp[i] = cptr[i];
  for (int i = threadIdx.x; i < nq; i += 256) {
    tce[i] = celm[i];
    tcw[i] = cw[i];
  }
  const int nr = min(rows, L.ey - y0);
  const bool okzz = okz[z] != 0;
  for (int r = 0; r < nr; ++r) {
    const bool ok = okzz && oky[y0 + r];
    const int64_t row = L.pelem(0, y0 + r, z);
    for (int x = threadIdx.x; x < L.ex; x += 256) {
      const int64_t e = row + x;
      srow[r * L.ex + x] = (ok && eact[e] && !shell[e]) ? g[e] : 0.f;
    }
  }
  __syncthreads();
  // rows fastest across the lanes: the lanes' row reads sit ex words apart (spread over
  // the LDS banks) instead of the coefficient spacing apart (a few banks)
  for (int i = threadIdx.x; i < nr * n0; i += 256) {
    const int r = i % nr, c = i / nr;
    const float* sr = srow + r * L.ex;
    double acc = 0;
    for (int q = tcp[c]; q < tcp[c + 1]; ++q) acc += (double)tcw[q] * (double)sr[tce[q]];
    out[c + (int64_t)n0 * (y0 + r + (int64_t)L.ey * z)] = acc;
  }
}

// One separable adjoint stage along axis a of a 3-D array with dims (d0, d1, d2) (x
// fastest): out[.. c ..] = sum over the CSR list of coefficient c of w * in[.. e ..],
// where the contracted axis has length nin in `in` and nout in `out`.
__global__ void k_spline_stage(const double* __restrict__ in, double* __restrict__ out, int axis,
                               int d0, int d1, int d2,  // dims of `out`
                               int nin, const int* __restrict__ cptr, const int* __restrict__ celm,
                               const float* __restrict__ cw) {
  const int64_t t = tid64();
  if (t >= (int64_t)d0 * d1 * d2) return;
  const int i0 = t % d0, i1 = (t / d0) % d1, i2 = t / ((int64_t)d0 * d1);
  int c, s0, s1, s2;  // contracted index and strides of `in`
  int64_t base, stride;
  if (axis == 0) { c = i0; base = (int64_t)nin * (i1 + (int64_t)d1 * i2); stride = 1; }
  else if (axis == 1) { c = i1; base = i0 + (int64_t)d0 * nin * i2; stride = d0; }
  else { c = i2; base = i0 + (int64_t)d0 * i1; stride = (int64_t)d0 * d1; }
  (void)s0; (void)s1; (void)s2;
  double acc = 0;
  for (int q = cptr[c]; q < cptr[c + 1]; ++q) acc += (double)cw[q] * in[base + stride * celm[q]];
  out[t] = acc;
}

// OC on the coefficients (tryCSensMultiplier_kernel optimization/optimization.cu:559-587)
__global__ __launch_bounds__(256) void k_coef_absmax_partial(int64_t n, const float* __restrict__ s,
                                                             double* __restrict__ partial) {
  double m = 0;
  for (int64_t i = tid64(); i < n; i += (int64_t)gridDim.x * blockDim.x) m = fmax(m, fabs((double)s[i]));
  m = block_max256(m);
  if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

__global__ void k_coef_trial(int64_t n, const float* __restrict__ old, const float* __restrict__ s, double move,
                             double damp, double cmin, double cmax, const double* __restrict__ scal,
                             float* __restrict__ coef) {
  const int64_t i = tid64();
  if (i >= n || scal[S_OC_DONE] != 0.0) return;
  const double g = scal[S_OC_G];
  const double gg = s[i] > 0 ? 0.0 : -(double)s[i];
  const double c0 = old[i];
  double cn = c0 * pow(gg / g, damp);
  cn = fmin(fmax(cn, c0 - move), c0 + move);
  cn = fmin(fmax(cn, cmin), cmax);
  coef[i] = (float)cn;
}

__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t i = tid64();
  if (i < n) b[i] = (float)a[i];
}

// element array: unpadded host layout <-> padded device layout
__global__ void k_pad_elems(LevelView L, const float* __restrict__ src, float* __restrict__ dst) {
  const int64_t t = tid64();
  if (t >= (int64_t)L.ex * L.ey * L.ez) return;
  const int x = t % L.ex, y = (t / L.ex) % L.ey, z = t / ((int64_t)L.ex * L.ey);
  dst[L.pelem(x, y, z)] = src[t];
}
__global__ void k_unpad_elems(LevelView L, const float* __restrict__ src, float* __restrict__ dst) {
  const int64_t t = tid64();
  if (t >= (int64_t)L.ex * L.ey * L.ez) return;
  const int x = t % L.ex, y = (t / L.ex) % L.ey, z = t / ((int64_t)L.ex * L.ey);
  dst[t] = src[L.pelem(x, y, z)];
}
// node vectors (3 SoA components): unpadded <-> padded, optional free mask on the way in
__global__ void k_pad_vec(LevelView L, const double* __restrict__ src, double* __restrict__ dst,
                          const uint8_t* __restrict__ fre) {
  const int64_t t = tid64();
  const int64_t n = (int64_t)L.nx * L.ny * L.nz;
  if (t >= n) return;
  const int x = t % L.nx, y = (t / L.nx) % L.ny, z = t / ((int64_t)L.nx * L.ny);
  const int64_t p = L.pnode(x, y, z);
  const bool ok = fre == nullptr || fre[p];
  for (int c = 0; c < 3; ++c) dst[c * L.NP + p] = ok ? src[c * n + t] : 0.0;
}
__global__ void k_unpad_vec(LevelView L, const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t t = tid64();
  const int64_t n = (int64_t)L.nx * L.ny * L.nz;
  if (t >= n) return;
  const int x = t % L.nx, y = (t / L.nx) % L.ny, z = t / ((int64_t)L.nx * L.ny);
  const int64_t p = L.pnode(x, y, z);
  for (int c = 0; c < 3; ++c) dst[c * n + t] = src[c * L.NP + p];
}

// ---------------------------------------------------------------------------------------
// Worst-case load, modified power method (optimization/optimization.cpp:289-300 with
// forceProject, optimization/projection.cpp:289-350).  Load node i sits at padded index
// pidx[i]; its force direction is free (nrm == nullptr) or along the unit normal
// nrm[3i..3i+2].  The candidate force is g_i = u_i, or (n_i . u_i) n_i; a node-major
// source (src_nm, 3 values per node) replaces u at the start.  The reference removes the
// two tangential components (f -= N^T N f with an orthonormal tangent pair), which
// leaves the same normal component.  Grid-stride loops over a fixed grid: every block
// writes one partial, so the sums do not depend on timing.
// ---------------------------------------------------------------------------------------
// R^T v at one node (b += ...) and the rigid field R w at one node, for the columns
// t_x, t_y, t_z, (0, z, -y), (-z, 0, x), (y, -x, 0) at coordinates (px, py, pz)
__device__ __forceinline__ void rigid_dot(double px, double py, double pz, const double v[3], double b[6]) {
  b[0] += v[0];
  b[1] += v[1];
  b[2] += v[2];
  b[3] += pz * v[1] - py * v[2];
  b[4] += px * v[2] - pz * v[0];
  b[5] += py * v[0] - px * v[1];
}
__device__ __forceinline__ void rigid_field(double px, double py, double pz, const double w[6], double c[3]) {
  c[0] = w[0] - pz * w[4] + py * w[5];
  c[1] = w[1] + pz * w[3] - px * w[5];
  c[2] = w[2] - py * w[3] + px * w[4];
}
// w = M b with b = scal[S_RIG0 .. S_RIG0 + 5]
__device__ __forceinline__ void rigid_weights(const RigidProj& P, const double* __restrict__ scal, double w[6]) {
  double b[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) b[k] = scal[S_RIG0 + k];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double a = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) a = fma(P.M[r * 6 + k], b[k], a);
    w[r] = a;
  }
}

// the load-node candidate g_i: u (or src_nm, node-major) at the node, along its normal when
// normals are given; with `pos` (floating contexts) minus its rigid-motion part, whose
// weights come from R^T g in scal (k_worst_rig_moments), the part taken along the normal
// as well
__device__ __forceinline__ void worst_candidate(int i, const int64_t* __restrict__ pidx,
                                                const double* __restrict__ nrm, int64_t NP,
                                                const double* __restrict__ u, const double* __restrict__ src_nm,
                                                double g[3]) {
  if (src_nm) {
    g[0] = src_nm[3 * i]; g[1] = src_nm[3 * i + 1]; g[2] = src_nm[3 * i + 2];
  } else {
    const int64_t p = pidx[i];
    g[0] = u[p]; g[1] = u[NP + p]; g[2] = u[2 * NP + p];
  }
  if (nrm) {
    const double nx = nrm[3 * i], ny = nrm[3 * i + 1], nz = nrm[3 * i + 2];
    const double d = fma(nz, g[2], fma(ny, g[1], nx * g[0]));
    g[0] = d * nx; g[1] = d * ny; g[2] = d * nz;
  }
}
__device__ __forceinline__ void worst_candidate_bal(int i, const int64_t* __restrict__ pidx,
                                                    const double* __restrict__ nrm, int64_t NP,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ src_nm,
                                                    const double* __restrict__ pos, const RigidProj& P,
                                                    const double* __restrict__ scal, double g[3]) {
  worst_candidate(i, pidx, nrm, NP, u, src_nm, g);
  if (!pos) return;
  double w[6], c[3];
  rigid_weights(P, scal, w);
  rigid_field(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], w, c);
  if (nrm) {
    const double nx = nrm[3 * i], ny = nrm[3 * i + 1], nz = nrm[3 * i + 2];
    const double d = fma(nz, c[2], fma(ny, c[1], nx * c[0]));
    c[0] = d * nx; c[1] = d * ny; c[2] = d * nz;
  }
  g[0] -= c[0]; g[1] -= c[1]; g[2] -= c[2];
}

// partials (6 sets) of R^T g over the load nodes, g the candidate before the rigid part
// is removed (positions relative to the projector's origin)
__global__ __launch_bounds__(256) void k_worst_rig_moments(int n, const int64_t* __restrict__ pidx,
                                                           const double* __restrict__ nrm,
                                                           const double* __restrict__ pos, int64_t NP,
                                                           const double* __restrict__ u,
                                                           const double* __restrict__ src_nm,
                                                           double* __restrict__ partial) {
  double b[6] = {0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double g[3];
    worst_candidate(i, pidx, nrm, NP, u, src_nm, g);
    rigid_dot(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], g, b);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) block_partial(b[k], partial, k);
}

// partials (6 sets) of R^T u over the free nodes of a chunk list (fixed grid, chunks
// strided over the blocks' waves)
__global__ __launch_bounds__(256) void k_rigid_moments(const int4* __restrict__ cl, int nch, int64_t NP,
                                                       const double* __restrict__ u, RigidProj P,
                                                       double* __restrict__ partial) {
  double b[6] = {0, 0, 0, 0, 0, 0};
  const int lane = threadIdx.x & 63;
  for (int ch = blockIdx.x * 4 + (threadIdx.x >> 6); ch < nch; ch += gridDim.x * 4) {
    NodeRef nd;
    if (!chunk_node_at(cl, nch, ch, lane, 1, nd)) continue;
    const double v[3] = {u[nd.p], u[NP + nd.p], u[2 * NP + nd.p]};
    rigid_dot(nd.x - P.o[0], nd.y - P.o[1], nd.z - P.o[2], v, b);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) block_partial(b[k], partial, k);
}

// scal[S_RIG0 + k] = sum of partial set k (np partials per set)
__global__ __launch_bounds__(256) void k_fold6(const double* __restrict__ partial, int np,
                                               double* __restrict__ scal) {
  for (int k = 0; k < 6; ++k) {
    double s = 0;
    for (int i = threadIdx.x; i < np; i += 256) s += partial[(int64_t)k * np + i];
    s = block_sum256(s);
    if (threadIdx.x == 0) scal[S_RIG0 + k] = s;
  }
}

// u -= R (M b) over the free nodes of a chunk list (displacementProject)
__global__ __launch_bounds__(256) void k_rigid_sub(const int4* __restrict__ cl, int nch, int64_t NP,
                                                   double* __restrict__ u, RigidProj P,
                                                   const double* __restrict__ scal) {
  NodeRef nd;
  if (!chunk_node(cl, nch, 1, nd)) return;
  double w[6], c[3];
  rigid_weights(P, scal, w);
  rigid_field(nd.x - P.o[0], nd.y - P.o[1], nd.z - P.o[2], w, c);
  u[nd.p] -= c[0];
  u[NP + nd.p] -= c[1];
  u[2 * NP + nd.p] -= c[2];
}

// gather the 3-vectors of the load nodes (node-major)
__global__ __launch_bounds__(256) void k_gather_load(int n, const int64_t* __restrict__ pidx, int64_t NP,
                                                     const double* __restrict__ v, double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int64_t p = pidx[i];
  out[3 * i] = v[p];
  out[3 * i + 1] = v[NP + p];
  out[3 * i + 2] = v[2 * NP + p];
}

// forceProject's write-back: f = projected candidate on the load nodes (f is zero elsewhere)
__global__ __launch_bounds__(256) void k_balance_apply(int n, const int64_t* __restrict__ pidx,
                                                       const double* __restrict__ nrm,
                                                       const double* __restrict__ pos, RigidProj P,
                                                       const double* __restrict__ scal, int64_t NP,
                                                       const double* __restrict__ src_nm, double* __restrict__ f) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double g[3];
  // (u is not read when src_nm is given; a null literal there crashed the device compiler)
  worst_candidate_bal(i, pidx, nrm, NP, f, src_nm, pos, P, scal, g);
  const int64_t p = pidx[i];
  f[p] = g[0];
  f[NP + p] = g[1];
  f[2 * NP + p] = g[2];
}

// coarsest system of a floating context: A += alpha R R^T (R: m orthonormal columns of
// length n, column k at R[k * n])
__global__ void k_add_lowrank(double* __restrict__ A, int n, const double* __restrict__ R, int m, double alpha) {
  const int64_t t = thread_id();
  if (t >= (int64_t)n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  double s = 0;
  for (int k = 0; k < m; ++k) s = fma(R[(int64_t)k * n + i], R[(int64_t)k * n + j], s);
  A[t] = fma(alpha, s, A[t]);
}

// partial[b] = sum of |g_i|^2 over the block's load nodes
__global__ __launch_bounds__(256) void k_worst_norm(int n, const int64_t* __restrict__ pidx,
                                                    const double* __restrict__ nrm, int64_t NP,
                                                    const double* __restrict__ u, const double* __restrict__ src_nm,
                                                    const double* __restrict__ pos, RigidProj P,
                                                    const double* __restrict__ scal, double* __restrict__ partial) {
  double s = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double g[3];
    worst_candidate_bal(i, pidx, nrm, NP, u, src_nm, pos, P, scal, g);
    s = fma(g[2], g[2], fma(g[1], g[1], fma(g[0], g[0], s)));
  }
  block_partial(s, partial, 0);
}

// f_i = g_i / sqrt(scal[g2_slot]) on the load nodes; r (and its FP32 copy r32, when
// given) += f_new - f_old there, so r stays f - K u; fs (node-major) = f_new.
// partials: set 0 |f_new - f_old|^2, set 1 |f_old|^2.  `first`: f_old = 0, r untouched.
__global__ __launch_bounds__(256) void k_worst_apply(int n, const int64_t* __restrict__ pidx,
                                                     const double* __restrict__ nrm, int64_t NP,
                                                     const double* __restrict__ u,
                                                     const double* __restrict__ src_nm,
                                                     const double* __restrict__ pos, RigidProj P,
                                                     const double* __restrict__ scal, int g2_slot, int first,
                                                     double* __restrict__ f, double* __restrict__ r,
                                                     float* __restrict__ r32, double* __restrict__ fs,
                                                     double* __restrict__ partial) {
  const double g2 = scal[g2_slot];
  const double inv = g2 > 0 ? 1.0 / sqrt(g2) : 0.0;  // zero candidate: the host reports it
  double sd = 0, so = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double g[3];
    worst_candidate_bal(i, pidx, nrm, NP, u, src_nm, pos, P, scal, g);
    const int64_t p = pidx[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double fn = g[c] * inv;
      const double fo = first ? 0.0 : fs[3 * i + c];
      const double d = fn - fo;
      sd = fma(d, d, sd);
      so = fma(fo, fo, so);
      f[c * NP + p] = fn;
      fs[3 * i + c] = fn;
      if (!first) {
        const double rn = r[c * NP + p] + d;
        r[c * NP + p] = rn;
        if (r32) r32[c * NP + p] = (float)rn;
      }
    }
  }
  block_partial(sd, partial, 0);
  block_partial(so, partial, 1);
}

}  // namespace tos
