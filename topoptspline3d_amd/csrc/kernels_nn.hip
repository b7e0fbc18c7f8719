// kernels_nn.hip -- kernels of the learned coarse-level correction: a 3-D convolutional
// U-Net over a multigrid level's node box (no counterpart in the reference).
//
// Tensors are channel-last BF16 ("x[v * cs + co + c]", v = x + X*(y + Y*z) + b*V) with a
// channel stride `cs` and offset `co`, so a decoder's skip-concatenation buffer is written
// in place by the encoder.  Convolutions are implicit GEMMs on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation): one wave computes 16 consecutive output
// voxels of a row (A rows) x all output channels (B columns, NOUT/16 tiles); the K
// dimension runs over (tap, input channel) with tap = (dx+1) + 3(dy+1) + 9(dz+1), so one
// lane's 8 consecutive K values are 8 channels of one tap (one 16-byte load).
//   MODE 0 forward, stride 1:   y[o] = sum_t W_t x[o + d_t]
//   MODE 1 forward, stride 2:   y[o] = sum_t W_t x[2o + d_t]
//   MODE 2 data gradient, s 1:  dx[u] = sum_t W_t^T dy[u - d_t]
//   MODE 3 data gradient, s 2:  dx[u] = sum_{t: u - d_t even} W_t^T dy[(u - d_t) / 2]
// Zero padding outside the box.  Weight gradients: per tap a [Cin x Cout] GEMM over
// voxels, staged transposed through LDS; per-group partials folded in a fixed order.
#include "tos_internal.hpp"

namespace tos {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct NGrid {
  int X, Y, Z;
};

__device__ __forceinline__ bool ngrid_in(const NGrid& g, int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < g.X && y < g.Y && z < g.Z;
}

struct ConvArgs {
  const uint16_t* x;  // input tensor (channel stride xcs, offset xco)
  int xcs, xco;
  NGrid gi;
  const uint16_t* w;  // [NOUT][KP] BF16, K index = tap * CINP + channel
  const float* bias;  // NOUT or null
  uint16_t* y;        // BF16 output (stride ycs, offset yco) ...
  int ycs, yco;
  float* y32;         // ... or FP32 output of the first `nstore` channels (stride nstore)
  int nstore;
  NGrid go;
  const uint16_t* mask;  // optional: zero where mask <= 0 (ReLU derivative)
  int mcs, mco;
  const uint16_t* add;   // optional addend
  int acs, aco;
  int relu;
  int B;
};

template <int CINP, int NOUT, int MODE>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  constexpr int NT = NOUT / 16;
  constexpr int KP = ((27 * CINP + 31) / 32) * 32;
  const int lane = threadIdx.x & 63, g = lane >> 4, i16 = lane & 15;
  const int tpr = (a.go.X + 15) >> 4;
  const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t ntile = (int64_t)a.B * a.go.Z * a.go.Y * tpr;
  if (tile >= ntile) return;  // uniform per wave
  int64_t t = tile;
  const int tx = (int)(t % tpr);
  t /= tpr;
  const int oy = (int)(t % a.go.Y);
  t /= a.go.Y;
  const int oz = (int)(t % a.go.Z);
  const int b = (int)(t / a.go.Z);
  const int ox = tx * 16 + i16;
  const bool rowok = ox < a.go.X;
  const int64_t Vi = (int64_t)a.gi.X * a.gi.Y * a.gi.Z;
  const uint16_t* xb = a.x + (int64_t)b * Vi * a.xcs + a.xco;
  f32x4_t acc[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) acc[n] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < KP; kc += 32) {
    const int k0 = kc + 8 * g;
    const int tap = k0 / CINP, ci = k0 % CINP;
    uint4 av = make_uint4(0, 0, 0, 0);
    if (tap < 27 && rowok) {
      const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
      int ix, iy, iz;
      bool ok = true;
      if (MODE == 0) {
        ix = ox + dx; iy = oy + dy; iz = oz + dz;
      } else if (MODE == 1) {
        ix = 2 * ox + dx; iy = 2 * oy + dy; iz = 2 * oz + dz;
      } else if (MODE == 2) {
        ix = ox - dx; iy = oy - dy; iz = oz - dz;
      } else {
        const int sx = ox - dx, sy = oy - dy, sz = oz - dz;
        ok = ((sx | sy | sz) & 1) == 0;
        ix = sx >> 1; iy = sy >> 1; iz = sz >> 1;
      }
      if (ok && ngrid_in(a.gi, ix, iy, iz))
        av = *(const uint4*)(xb + (((int64_t)iz * a.gi.Y + iy) * a.gi.X + ix) * a.xcs + ci);
    }
    const bf16x8_t A = __builtin_bit_cast(bf16x8_t, av);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const uint4 bv = *(const uint4*)(a.w + (int64_t)(n * 16 + i16) * KP + k0);
      acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, __builtin_bit_cast(bf16x8_t, bv), acc[n], 0, 0, 0);
    }
  }
  // lane holds D[row 4g + r][col i16]: voxel tx*16 + 4g + r, channel n*16 + i16
  const int64_t Vo = (int64_t)a.go.X * a.go.Y * a.go.Z;
  const int64_t vrow = (int64_t)b * Vo + ((int64_t)oz * a.go.Y + oy) * a.go.X;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int vx = tx * 16 + 4 * g + r;
    if (vx >= a.go.X) continue;
    const int64_t v = vrow + vx;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int co = n * 16 + i16;
      float val = acc[n][r];
      if (a.bias) val += a.bias[co];
      if (a.add) val += bf2f(a.add[v * a.acs + a.aco + co]);
      if (a.relu) val = fmaxf(val, 0.f);
      if (a.mask && !(bf2f(a.mask[v * a.mcs + a.mco + co]) > 0.f)) val = 0.f;
      if (a.y32) {
        if (co < a.nstore) a.y32[v * a.nstore + co] = val;
      } else {
        a.y[v * a.ycs + a.yco + co] = f2bf(val);
      }
    }
  }
}

// weight gradient partials part[grp][tap][ci][co] (ci < CI16) over voxel tiles grp, grp+G, ...
// x: forward input (stride xcs, offset xco) on grid gi; dy: gradient of the (pre-activation)
// forward output, contiguous with NOUT channels, on grid go.
template <int CINP, int NOUT, int STRIDE>
__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t* __restrict__ x, int xcs, int xco, NGrid gi,
                                                    const uint16_t* __restrict__ dy, NGrid go, int B,
                                                    float* __restrict__ part) {
  constexpr int CI16 = ((CINP + 15) / 16) * 16;
  constexpr int MT = CI16 / 16, NT = NOUT / 16, NTILE = MT * NT, PER = (NTILE + 3) / 4;
  constexpr int LD = 64 + 8;  // row pitch (elements): 144 bytes, 16-byte aligned rows
  __shared__ __attribute__((aligned(16))) uint16_t sx[CI16 * LD];
  __shared__ __attribute__((aligned(16))) uint16_t sd[NOUT * LD];
  const int tap = blockIdx.x, grp = blockIdx.y, G = gridDim.y;
  const int dx = tap % 3 - 1, dy_ = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const int lane = threadIdx.x & 63, g = lane >> 4, i16 = lane & 15, wv = threadIdx.x >> 6;
  const int64_t Vo = (int64_t)go.X * go.Y * go.Z, Vi = (int64_t)gi.X * gi.Y * gi.Z;
  const int64_t NV = (int64_t)B * Vo, ntl = (NV + 63) / 64;
  for (int q = threadIdx.x; q < CI16 * LD; q += 256) sx[q] = 0;
  f32x4_t acc[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) acc[p] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  __syncthreads();
  for (int64_t tl = grp; tl < ntl; tl += G) {
    const int64_t v0 = tl * 64;
    for (int q = threadIdx.x; q < 64 * (NOUT / 8); q += 256) {
      const int vv = q / (NOUT / 8), c8 = (q % (NOUT / 8)) * 8;
      uint4 val = make_uint4(0, 0, 0, 0);
      if (v0 + vv < NV) val = *(const uint4*)(dy + (v0 + vv) * NOUT + c8);
      const uint16_t* h = (const uint16_t*)&val;
#pragma unroll
      for (int j = 0; j < 8; ++j) sd[(c8 + j) * LD + vv] = h[j];
    }
    for (int q = threadIdx.x; q < 64 * (CINP / 8); q += 256) {
      const int vv = q / (CINP / 8), c8 = (q % (CINP / 8)) * 8;
      uint4 val = make_uint4(0, 0, 0, 0);
      const int64_t v = v0 + vv;
      if (v < NV) {
        const int b = (int)(v / Vo);
        int64_t r = v - (int64_t)b * Vo;
        const int ox = (int)(r % go.X);
        r /= go.X;
        const int oy = (int)(r % go.Y), oz = (int)(r / go.Y);
        const int ix = STRIDE * ox + dx, iy = STRIDE * oy + dy_, iz = STRIDE * oz + dz;
        if (ngrid_in(gi, ix, iy, iz))
          val = *(const uint4*)(x + ((int64_t)b * Vi + ((int64_t)iz * gi.Y + iy) * gi.X + ix) * xcs + xco + c8);
      }
      const uint16_t* h = (const uint16_t*)&val;
#pragma unroll
      for (int j = 0; j < 8; ++j) sx[(c8 + j) * LD + vv] = h[j];
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int tid = wv + 4 * p;
      if (tid < NTILE) {
        const int mt = tid / NT, nt = tid % NT;
#pragma unroll
        for (int ks = 0; ks < 64; ks += 32) {
          const uint4 av = *(const uint4*)(sx + (mt * 16 + i16) * LD + ks + 8 * g);
          const uint4 bv = *(const uint4*)(sd + (nt * 16 + i16) * LD + ks + 8 * g);
          acc[p] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, av),
                                                           __builtin_bit_cast(bf16x8_t, bv), acc[p], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int tid = wv + 4 * p;
    if (tid < NTILE) {
      const int mt = tid / NT, nt = tid % NT;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ci = mt * 16 + 4 * g + r, co = nt * 16 + i16;
        part[(((int64_t)grp * 27 + tap) * CI16 + ci) * NOUT + co] = acc[p][r];
      }
    }
  }
}

// grad[woff + (tap*cin + ci)*cout + co] = sum_grp part[grp][tap][ci][co]  (fixed order)
__global__ void k_wgrad_fold(const float* __restrict__ part, int G, int ci16, int nout, int cin, int cout,
                             float* __restrict__ grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = 27LL * cin * cout;
  if (i >= n) return;
  const int co = (int)(i % cout), ci = (int)((i / cout) % cin), tap = (int)(i / ((int64_t)cout * cin));
  float s = 0.f;
  for (int gq = 0; gq < G; ++gq) s += part[(((int64_t)gq * 27 + tap) * ci16 + ci) * nout + co];
  grad[i] = s;
}

// bias gradient: part[grp][co] = sum over voxels grp, grp+G, ... of dy[v][co]
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* __restrict__ dy, int64_t NV, int nout,
                                                   float* __restrict__ part) {
  __shared__ float sh[256];
  const int per = 256 / nout, co = threadIdx.x % nout, sub = threadIdx.x / nout;
  float s = 0.f;
  if (sub < per)
    for (int64_t v = (int64_t)blockIdx.x * per + sub; v < NV; v += (int64_t)gridDim.x * per)
      s += bf2f(dy[v * nout + co]);
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < nout) {
    float t = 0.f;
    for (int j = 0; j < per; ++j) t += sh[j * nout + threadIdx.x];
    part[(int64_t)blockIdx.x * nout + threadIdx.x] = t;
  }
}
__global__ void k_bias_fold(const float* __restrict__ part, int G, int nout, int cout, float* __restrict__ grad) {
  const int co = threadIdx.x;
  if (co >= cout) return;
  float s = 0.f;
  for (int gq = 0; gq < G; ++gq) s += part[(int64_t)gq * nout + co];
  grad[co] = s;
}

// BF16 weight copies from the FP32 master W[tap][ci][co]:
//   wf[co][tap*cinp + ci] (forward, NOUT rows, KPF columns),
//   wb[ci][tap*nout + co] (data gradient, cinb rows, KPB columns)
__global__ void k_prep_weights(const float* __restrict__ W, int cin, int cout, int cinp, int nout, int kpf,
                               uint16_t* __restrict__ wf, int cinb, int kpb, uint16_t* __restrict__ wb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nf = (int64_t)nout * kpf, nb = wb ? (int64_t)cinb * kpb : 0;
  if (i < nf) {
    const int co = (int)(i / kpf), k = (int)(i % kpf), tap = k / cinp, ci = k % cinp;
    float v = 0.f;
    if (tap < 27 && ci < cin && co < cout) v = W[((int64_t)tap * cin + ci) * cout + co];
    wf[i] = f2bf(v);
  } else if (i < nf + nb) {
    const int64_t j = i - nf;
    const int ci = (int)(j / kpb), k = (int)(j % kpb), tap = k / nout, co = k % nout;
    float v = 0.f;
    if (tap < 27 && ci < cin && co < cout) v = W[((int64_t)tap * cin + ci) * cout + co];
    wb[j] = f2bf(v);
  }
}

// node-aligned trilinear upsampling by 2 (fine u = 2I exact, odd u halfway), 8 channels per
// thread: y[u][yco + c] = sum w x[I][c], x contiguous with C channels
__device__ __forceinline__ int up_taps(int u, int n, int* I, float* w) {
  if ((u & 1) == 0) {
    I[0] = u >> 1; w[0] = 1.f;
    return 1;
  }
  I[0] = u >> 1; w[0] = 0.5f;
  if ((u >> 1) + 1 < n) {
    I[1] = (u >> 1) + 1; w[1] = 0.5f;
    return 2;
  }
  return 1;
}
__global__ void k_upsample(const uint16_t* __restrict__ x, int C, NGrid gc, uint16_t* __restrict__ y, int ycs,
                           int yco, NGrid gf, int B) {
  const int c8n = C / 8;
  const int64_t Vf = (int64_t)gf.X * gf.Y * gf.Z, Vc = (int64_t)gc.X * gc.Y * gc.Z;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * Vf * c8n) return;
  const int c8 = (int)(i % c8n) * 8;
  int64_t v = i / c8n;
  const int b = (int)(v / Vf);
  int64_t r = v - (int64_t)b * Vf;
  const int ux = (int)(r % gf.X);
  r /= gf.X;
  const int uy = (int)(r % gf.Y), uz = (int)(r / gf.Y);
  int Ix[2], Iy[2], Iz[2];
  float wx[2], wy[2], wz[2];
  const int nx = up_taps(ux, gc.X, Ix, wx), ny = up_taps(uy, gc.Y, Iy, wy), nz = up_taps(uz, gc.Z, Iz, wz);
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int kz = 0; kz < nz; ++kz)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx) {
        const float w = wx[kx] * wy[ky] * wz[kz];
        const uint4 val =
            *(const uint4*)(x + ((int64_t)b * Vc + ((int64_t)Iz[kz] * gc.Y + Iy[ky]) * gc.X + Ix[kx]) * C + c8);
        const uint16_t* h = (const uint16_t*)&val;
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += w * bf2f(h[j]);
      }
  uint16_t o[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = f2bf(acc[j]);
  *(uint4*)(y + v * ycs + yco + c8) = *(const uint4*)o;
}

// adjoint of k_upsample: x[I][c] = (sum_{u: I in taps(u)} w dy[u][dco + c]) (* ReLU mask)
__global__ void k_upsample_adj(const uint16_t* __restrict__ dy, int dcs, int dco, NGrid gf, uint16_t* __restrict__ x,
                               int C, NGrid gc, const uint16_t* __restrict__ mask, int B) {
  const int c8n = C / 8;
  const int64_t Vf = (int64_t)gf.X * gf.Y * gf.Z, Vc = (int64_t)gc.X * gc.Y * gc.Z;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * Vc * c8n) return;
  const int c8 = (int)(i % c8n) * 8;
  int64_t v = i / c8n;
  const int b = (int)(v / Vc);
  int64_t r = v - (int64_t)b * Vc;
  const int Ix = (int)(r % gc.X);
  r /= gc.X;
  const int Iy = (int)(r % gc.Y), Iz = (int)(r / gc.Y);
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int dz = -1; dz <= 1; ++dz)
    for (int dyy = -1; dyy <= 1; ++dyy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int ux = 2 * Ix + dx, uy = 2 * Iy + dyy, uz = 2 * Iz + dz;
        if (!ngrid_in(gf, ux, uy, uz)) continue;
        const float w = (dx ? 0.5f : 1.f) * (dyy ? 0.5f : 1.f) * (dz ? 0.5f : 1.f);
        const uint4 val = *(const uint4*)(dy + ((int64_t)b * Vf + ((int64_t)uz * gf.Y + uy) * gf.X + ux) * dcs + dco + c8);
        const uint16_t* h = (const uint16_t*)&val;
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += w * bf2f(h[j]);
      }
  uint16_t o[8];
  const uint16_t* m = mask ? mask + v * C + c8 : nullptr;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = f2bf((m && !(bf2f(m[j]) > 0.f)) ? 0.f : acc[j]);
  *(uint4*)(x + v * C + c8) = *(const uint4*)o;
}

// normalised MSE over the masked voxels: loss = 1/B sum_b inv_den[b] sum_v m |y - t|^2,
// gy (BF16, 16 channels, zero beyond 3) = d loss / d y; block partials of the loss
__global__ __launch_bounds__(256) void k_nn_loss(const float* __restrict__ y, const float* __restrict__ tgt,
                                                 const uint8_t* __restrict__ mask, const float* __restrict__ inv_den,
                                                 int64_t V, int B, uint16_t* __restrict__ gy,
                                                 double* __restrict__ partial) {
  __shared__ double sh[4];
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double s = 0;
  if (v < (int64_t)B * V) {
    const int b = (int)(v / V);
    const float sc = inv_den[b] / (float)B;
    const float m = mask[v] ? 1.f : 0.f;
    uint16_t o[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) o[c] = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float d = m * (y[v * 3 + c] - tgt[v * 3 + c]);
      s += (double)(sc * d * d);
      o[c] = f2bf(2.f * sc * d);
    }
    uint4* dst = (uint4*)(gy + v * 16);
    dst[0] = ((const uint4*)o)[0];
    dst[1] = ((const uint4*)o)[1];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Adam on the FP32 master parameters
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// host float features (4 channels) -> BF16 8-channel input tensor
__global__ void k_feat_pack(const float* __restrict__ f4, int64_t n, uint16_t* __restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint16_t o[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = c < 4 ? f2bf(f4[v * 4 + c]) : (uint16_t)0;
  *(uint4*)(out + v * 8) = *(const uint4*)o;
}

// ---------------------------------------------------------------------------------------
// Level-side kernels: scaled residual r~ = D^{-1/2} r (D = diagonal of the level's
// Galerkin centre block), features, correction and training targets.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float diag_of(const float* __restrict__ st, int64_t NS, int s, int c) {
  return st[(int64_t)(13 * 9 + 4 * c) * NS + s];
}

// block partials of sum_c r_c^2 / D_c over the free nodes (chunk list `cl`, x stride 1)
__global__ __launch_bounds__(256) void k_nn_norm(LevelView L, const int4* __restrict__ cl, int nch,
                                                 const float* __restrict__ r, const float* __restrict__ st, int64_t NS,
                                                 const int* __restrict__ slot, double* __restrict__ partial) {
  __shared__ double sh[4];
  double s = 0;
  const int ch = (int)xcd_block() * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (ch < nch) {
    const int4 c = cl[ch];
    if (lane < (c.y >> 16)) {
      const int64_t p = (int64_t)c.x + lane;
      const int sl = slot[p];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double v = r[k * L.NP + p];
        s += v * v / (double)diag_of(st, NS, sl, k);
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// one value: out = sum of np partials (fixed order)
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, int np,
                                                      double* __restrict__ out) {
  __shared__ double sh[4];
  double s = 0;
  for (int i = threadIdx.x; i < np; i += 256) s += partial[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// input tensor over the level's node box: [r~ / s (3), log-density (1), 0 x 4] with the
// RMS scale s = sqrt(|r~|^2 / n_free) (n2 holds |r~|^2)
__global__ void k_nn_features(LevelView L, const float* __restrict__ r, const float* __restrict__ st, int64_t NS,
                              const int* __restrict__ slot, const float* __restrict__ lrho,
                              const double* __restrict__ n2, double nfree, uint16_t* __restrict__ feat) {
  const int64_t V = (int64_t)L.nx * L.ny * L.nz;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int x = (int)(v % L.nx), y = (int)((v / L.nx) % L.ny), z = (int)(v / ((int64_t)L.nx * L.ny));
  const int64_t p = L.pnode(x, y, z);
  const int sl = slot[p];
  const double nn = *n2;
  const float inv = nn > 0 ? (float)sqrt(nfree / nn) : 0.f;
  uint16_t o[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = 0;
  if (sl >= 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = f2bf(r[c * L.NP + p] * rsqrtf(diag_of(st, NS, sl, c)) * inv);
  }
  o[3] = f2bf(lrho[v]);
  *(uint4*)(feat + v * 8) = *(const uint4*)o;
}

// correction e = D^{-1/2} y s on the free nodes, 0 on the other nodes of the box
__global__ void k_nn_correction(LevelView L, const float* __restrict__ y, const float* __restrict__ st, int64_t NS,
                                const int* __restrict__ slot, const double* __restrict__ n2, double nfree,
                                float* __restrict__ e) {
  const int64_t V = (int64_t)L.nx * L.ny * L.nz;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int x = (int)(v % L.nx), yy = (int)((v / L.nx) % L.ny), z = (int)(v / ((int64_t)L.nx * L.ny));
  const int64_t p = L.pnode(x, yy, z);
  const int sl = slot[p];
  const float nrm = (float)sqrt(*n2 / nfree);
#pragma unroll
  for (int c = 0; c < 3; ++c) e[c * L.NP + p] = sl >= 0 ? y[v * 3 + c] * rsqrtf(diag_of(st, NS, sl, c)) * nrm : 0.f;
}

// training target y* = D^{1/2} e / s and the free mask over the node box
__global__ void k_nn_target(LevelView L, const float* __restrict__ e, const float* __restrict__ st, int64_t NS,
                            const int* __restrict__ slot, const double* __restrict__ n2, double nfree,
                            float* __restrict__ tgt, uint8_t* __restrict__ mask) {
  const int64_t V = (int64_t)L.nx * L.ny * L.nz;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int x = (int)(v % L.nx), yy = (int)((v / L.nx) % L.ny), z = (int)(v / ((int64_t)L.nx * L.ny));
  const int64_t p = L.pnode(x, yy, z);
  const int sl = slot[p];
  const double nn = *n2;
  const float inv = nn > 0 ? (float)sqrt(nfree / nn) : 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) tgt[v * 3 + c] = sl >= 0 ? e[c * L.NP + p] * sqrtf(diag_of(st, NS, sl, c)) * inv : 0.f;
  mask[v] = sl >= 0 ? 1 : 0;
}

// block partials of sum m |t|^2 over V voxels of one sample
__global__ __launch_bounds__(256) void k_target_norm(const float* __restrict__ tgt, const uint8_t* __restrict__ mask,
                                                     int64_t V, double* __restrict__ partial) {
  __shared__ double sh[4];
  double s = 0;
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < V; v += (int64_t)gridDim.x * 256)
    if (mask[v])
      s += (double)tgt[v * 3] * tgt[v * 3] + (double)tgt[v * 3 + 1] * tgt[v * 3 + 1] +
           (double)tgt[v * 3 + 2] * tgt[v * 3 + 2];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// inv_den[i] = 1 / max(den, tiny) (den from the partials in `partial`, fixed order)
__global__ __launch_bounds__(256) void k_inv_den(const double* __restrict__ partial, int np, float* __restrict__ out) {
  __shared__ double sh[4];
  double s = 0;
  for (int i = threadIdx.x; i < np; i += 256) s += partial[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    *out = t > 1e-30 ? (float)(1.0 / t) : 0.f;
  }
}

// nodal log-density of a level whose elements are `f` x `f` x `f` fine elements:
// mean of rho over the fine elements in the node's (2f)^3 neighbourhood inside the box
// (inactive elements count as rho_min), mapped to log10(rho) / 3
__global__ void k_node_logrho(LevelView F, const uint8_t* __restrict__ eact, const float* __restrict__ rho, int fex,
                              int fey, int fez, int f, float rmin, int nx, int ny, int nz, float* __restrict__ out) {
  const int64_t V = (int64_t)nx * ny * nz;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((int64_t)nx * ny));
  const int x0 = max(0, f * (x - 1)), x1 = min(fex, f * (x + 1));
  const int y0 = max(0, f * (y - 1)), y1 = min(fey, f * (y + 1));
  const int z0 = max(0, f * (z - 1)), z1 = min(fez, f * (z + 1));
  double s = 0;
  int n = 0;
  for (int k = z0; k < z1; ++k)
    for (int j = y0; j < y1; ++j)
      for (int i = x0; i < x1; ++i) {
        const int64_t e = F.pelem(i, j, k);
        s += eact[e] ? (double)rho[e] : (double)rmin;
        ++n;
      }
  const double m = n ? fmax(s / n, (double)rmin) : (double)rmin;
  out[v] = (float)(log10(m) / 3.0);
}

// block partials of e.r (column 0) and e.t (column 1) over the free nodes
__global__ __launch_bounds__(256) void k_nn_dots(LevelView L, const int4* __restrict__ cl, int nch,
                                                 const float* __restrict__ e, const float* __restrict__ r,
                                                 const float* __restrict__ t, double* __restrict__ partial) {
  __shared__ double sh[2][4];
  double a = 0, b = 0;
  const int ch = (int)xcd_block() * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (ch < nch) {
    const int4 c = cl[ch];
    if (lane < (c.y >> 16)) {
      const int64_t p = (int64_t)c.x + lane;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double ek = e[k * L.NP + p];
        a += ek * r[k * L.NP + p];
        b += ek * t[k * L.NP + p];
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    partial[gridDim.x + blockIdx.x] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
  }
}

// out[0] = sum partial[0..np), out[1] = sum partial[np..2np)  (fixed order)
__global__ __launch_bounds__(256) void k_sum2_partials(const double* __restrict__ partial, int np,
                                                       double* __restrict__ out) {
  __shared__ double sh[2][4];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < np; i += 256) {
    a += partial[i];
    b += partial[np + i];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    out[1] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
  }
}

// energy-optimal use of the proposal e for A x = r: with er = e.r and eAe = e.r - e.t
// (t = r - A e) the step alpha = er / eAe minimises the A-norm error along e.  The
// correction is kept (u += alpha e) only when alpha lies in [amin, amax]: alpha > 0 means
// it lowers the energy error, and a step far from 1 means the proposal is mostly a wrong
// direction or scale, so plain multigrid is left alone.  ab = {e.r, e.t}; block 0 /
// thread 0 keeps the statistics st = {tries, kept, sum of alpha}.
__global__ __launch_bounds__(256) void k_nn_accept(LevelView L, const int4* __restrict__ cl, int nch,
                                                   const double* __restrict__ ab, double amin, double amax,
                                                   const float* __restrict__ e, float* __restrict__ u,
                                                   double* __restrict__ st) {
  const double er = ab[0], eAe = ab[0] - ab[1];
  double alpha = (er > 0 && eAe > 0) ? er / eAe : 0.0;
  if (alpha < amin || alpha > amax) alpha = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[0] += 1;
    if (alpha > 0) {
      st[1] += 1;
      st[2] += alpha;
    }
  }
  if (!(alpha > 0)) return;
  const int ch = (int)xcd_block() * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (ch >= nch) return;
  const int4 c = cl[ch];
  if (lane >= (c.y >> 16)) return;
  const int64_t p = (int64_t)c.x + lane;
  const float a = (float)alpha;
#pragma unroll
  for (int k = 0; k < 3; ++k) u[k * L.NP + p] += a * e[k * L.NP + p];
}

// y += a on the free nodes (3 components)
__global__ __launch_bounds__(256) void k_add_nodes_f(const int4* __restrict__ cl, int nch, int64_t NP,
                                                     float* __restrict__ y, const float* __restrict__ a) {
  const int ch = (int)xcd_block() * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (ch >= nch) return;
  const int4 c = cl[ch];
  if (lane >= (c.y >> 16)) return;
  const int64_t p = (int64_t)c.x + lane;
#pragma unroll
  for (int k = 0; k < 3; ++k) y[k * NP + p] += a[k * NP + p];
}

}  // namespace tos
