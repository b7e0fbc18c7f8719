// kernels_probe.hip -- matrix-core element products, a measurement of the alternative to
// the matrix-free VALU operator of the fine level (not used by the solver).
//
// Per wave, up to 16 consecutive active elements of one row are the columns of
// B (24 x 16: each element's rho^p-weighted node values, k = 3 * local node + component);
// Ke (24 x 24, rows padded to 32) is A, and D = A B runs on v_mfma_f32_16x16x4f32 as two
// row tiles by six k-steps.  Lane layout of that instruction: A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i / 4), register i % 4.
// MODE 0 keeps a per-lane checksum of D (the gathers and products only, a lower bound on
// any node-assembling variant); MODE 1 adds every row of D into its node with a float
// atomic (non-deterministic order).
#pragma once

namespace tos {

typedef float probe_f4 __attribute__((ext_vector_type(4)));

template <int MODE>
__global__ __launch_bounds__(256) void k_probe_elem_mfma(LevelView L, const int4* __restrict__ tiles, int ntiles,
                                                         const float* __restrict__ w, const float* __restrict__ Ke,
                                                         const float* __restrict__ u, float* __restrict__ y,
                                                         float* __restrict__ check) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lg = lane >> 4;
  float a[2][6];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int row = 16 * m + li;
      a[m][s] = row < 24 ? Ke[row * 24 + 4 * s + lg] : 0.0f;
    }
  float cs = 0.0f;
  for (int t = blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles; t += gridDim.x * 4) {
    const int4 tl = tiles[t];  // {x of the first element, count, y, z}
    const int x = tl.x + li;
    const bool ok = li < tl.y;
    const float we = ok ? w[L.pelem(x, tl.z, tl.w)] : 0.0f;
    float b[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int k = 4 * s + lg, n = k / 3, c = k % 3;
      const int64_t p = L.pnode(x + (n & 1), tl.z + ((n >> 1) & 1), tl.w + (n >> 2));
      b[s] = ok ? we * u[c * L.NP + p] : 0.0f;
    }
    probe_f4 d0 = {0.0f, 0.0f, 0.0f, 0.0f}, d1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][s], b[s], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1][s], b[s], d1, 0, 0, 0);
    }
    if (MODE == 0) {
      cs += (d0[0] + d0[1]) + (d0[2] + d0[3]) + (d1[0] + d1[1]) + (d1[2] + d1[3]);
    } else if (ok) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i0 = 4 * lg + r, n0 = i0 / 3, c0 = i0 % 3;
        atomicAdd(&y[c0 * L.NP + L.pnode(x + (n0 & 1), tl.z + ((n0 >> 1) & 1), tl.w + (n0 >> 2))], d0[r]);
        const int i1 = 16 + 4 * lg + r;
        if (i1 < 24) {
          const int n1 = i1 / 3, c1 = i1 % 3;
          atomicAdd(&y[c1 * L.NP + L.pnode(x + (n1 & 1), tl.z + ((n1 >> 1) & 1), tl.w + (n1 >> 2))], d1[r]);
        }
      }
    }
  }
  if (MODE == 0) check[(int64_t)blockIdx.x * 256 + threadIdx.x] = cs;
}

// seeded values in [-1, 1) on the free nodes of a chunk list
__global__ __launch_bounds__(256) void k_probe_fill(const int4* __restrict__ cl, int nch, int64_t NP,
                                                    float* __restrict__ u, uint32_t seed) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uint32_t h = (uint32_t)(n.p * 3 + c) * 2654435761u ^ seed;
    h ^= h >> 16; h *= 2246822519u; h ^= h >> 13; h *= 3266489917u; h ^= h >> 16;
    u[c * NP + n.p] = (float)(h >> 8) * (1.0f / 8388608.0f) - 1.0f;
  }
}

// partials over the free nodes: set 0 |a - b|^2, set 1 |b|^2
__global__ __launch_bounds__(256) void k_probe_diff(const int4* __restrict__ cl, int nch, int64_t NP,
                                                    const float* __restrict__ a, const float* __restrict__ b,
                                                    double* __restrict__ partial) {
  NodeRef n;
  double s0 = 0, s1 = 0;
  if (chunk_node(cl, nch, 1, n))
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double d = (double)a[c * NP + n.p] - (double)b[c * NP + n.p];
      s0 += d * d;
      s1 += (double)b[c * NP + n.p] * (double)b[c * NP + n.p];
    }
  block_partial(s0, partial, 0);
  block_partial(s1, partial, 1);
}

}  // namespace tos
