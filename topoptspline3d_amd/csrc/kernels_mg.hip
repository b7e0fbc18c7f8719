// kernels_mg.hip -- multigrid kernels for gfx950 (wave64).
//
// Launch geometry: every node kernel walks a chunk list.  A chunk is a run of at most 64
// free nodes of one grid row (x-contiguous for the plain list, every second x for the
// per-colour lists), so one wave covers one chunk with unit-stride (or stride-2) lanes
// and void / eliminated parts of the box cost nothing.  int4 chunk = {padded index of the
// first node, x0 | count << 16, y, z}.
//
// Fine level (matrix-free, "on-the-fly assembly"): each thread owns one node, keeps the
// 27 neighbouring displacements in VGPRs and sums rho_e^p * Ke[loc_i][loc_j] u_j over the
// 8 incident elements (update_residual_OTFA_WS_kernel grid/Grid.cu:2277-2333,
// gs_relax_OTFA_WS_kernel :1700-1842).  Ke is read through a uniform pointer, so the
// compiler serves it from the scalar cache into SGPR operands of the FMAs; only the 14
// distinct values of its first block row are used (reflection and axis symmetry).
// Coarse levels: Galerkin 27x9 stencil stored colour-compact: slot = 64 * (index of the
// chunk in the colour-ordered list) + lane, entry (k*9 + r*3 + c)*NS + slot, so a GS
// colour pass reads its stencil rows fully coalesced (gs_relax_kernel :1455-1566,
// update_residual_kernel_1 :2127-2224 use the same 27x9 layout in colour order).
#include "tos_internal.hpp"

namespace tos {

__device__ __forceinline__ int64_t thread_id() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

struct NodeRef {
  int64_t p;
  int x, y, z, slot;
};

// node of lane `lane` of chunk `ch` of list `cl` (stride s); false for idle lanes
__device__ __forceinline__ bool chunk_node_at(const int4* __restrict__ cl, int nch, int ch, int lane, int s,
                                              NodeRef& n) {
  if (ch >= nch) return false;
  const int4 c = cl[ch];
  if (lane >= (c.y >> 16)) return false;
  n.x = (c.y & 0xffff) + s * lane;
  n.y = c.z;
  n.z = c.w;
  n.p = (int64_t)c.x + (int64_t)s * lane;
  n.slot = ch * 64 + lane;
  return true;
}

// node of this lane in chunk list `cl` (stride s); false for idle lanes
__device__ __forceinline__ bool chunk_node(const int4* __restrict__ cl, int nch, int s, NodeRef& n) {
  return chunk_node_at(cl, nch, (int)xcd_block() * 4 + (threadIdx.x >> 6), threadIdx.x & 63, s, n);
}

// block (256 threads) sum written to partial[blockIdx.x + k*gridDim.x]; all threads call
__device__ __forceinline__ void block_partial(double v, double* __restrict__ partial, int k) {
  __shared__ double sh[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[(int64_t)k * gridDim.x + blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
}

// ---------------------------------------------------------------------------------------
// Fine level
// ---------------------------------------------------------------------------------------
// Neighbours read by a colour pass.  ZC = -1: all 27.  ZC = c in 0..7: the first sweep
// from a zero initial guess at colour c, which reads only the neighbours of colours
// relaxed before c in this sweep (the others are still 0); the neighbour at offset d has
// colour c ^ (d_x&1 | (d_y&1)<<1 | (d_z&1)<<2), the node itself (still 0) is skipped too.
__host__ __device__ constexpr unsigned nb_mask_of(int zc) {
  unsigned m = 0;
  for (int k = 0; k < 27; ++k) {
    const int par = ((k % 3 - 1) & 1) | ((((k / 3) % 3 - 1) & 1) << 1) | (((k / 9 - 1) & 1) << 2);
    if (zc < 0 || (zc ^ par) < zc) m |= 1u << k;
  }
  return m;
}
// neighbours of colours after c (relaxed later in the sweep)
__host__ __device__ constexpr unsigned nb_mask_after(int c) {
  unsigned m = 0;
  for (int k = 0; k < 27; ++k) {
    const int par = ((k % 3 - 1) & 1) | ((((k / 3) % 3 - 1) & 1) << 1) | (((k / 9 - 1) & 1) << 2);
    if ((c ^ par) > c) m |= 1u << k;
  }
  return m;
}

template <class T, int ZC = -1, unsigned MASK = nb_mask_of(ZC)>
__device__ __forceinline__ void load_nbhd(const LevelView& L, const T* __restrict__ u, int64_t p,
                                          T (&un)[27][3]) {
  constexpr unsigned mask = MASK;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int64_t q = p + (int64_t)(k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      un[k][c] = ((mask >> k) & 1u) ? u[c * L.NP + q] : T(0);
    }
  }
}

// acc += sum_a w_a sum_j Ke[li][j] u_j ; with SELF = false the node itself (j = li) is
// skipped and its diagonal block accumulated into D instead.
// Only the first block row of Ke is needed: reflecting the cube along the axes set in li
// maps local node li to 0 and flips those displacement components, so
// Ke[li][j](r,s) = sg(li,r) sg(li,s) Ke[0][li^j](r,s) with sg = -1 on reflected axes, and
// that row holds 14 distinct values Kc[ke_key(d, r, s)] (cube symmetry of the isotropic
// element).  They stay in scalar registers (all 576 entries would spill).
template <class T>
__device__ __forceinline__ void load_weights(const LevelView& L, const T* __restrict__ w, int x, int y, int z,
                                             T (&wa)[8]) {
  const int64_t e0 = L.pelem(x - 1, y - 1, z - 1);
#pragma unroll
  for (int a = 0; a < 8; ++a) wa[a] = w[e0 + (a & 1) + (int64_t)L.EX * (((a >> 1) & 1) + (int64_t)L.EY * (a >> 2))];
}

// every operand of a node's update is loaded before any arithmetic: one memory round trip
// per wave instead of one per neighbour row (the scheduler would otherwise interleave
// loads with FMAs to save registers and serialise the latencies)
#define TOS_LOADS_DONE() __builtin_amdgcn_sched_barrier(0)

template <class T, bool SELF, unsigned MASK = 0x7ffffffu, class TU = T>
__device__ __forceinline__ void fine_row(const T (&wa)[8], const T* __restrict__ Kc, const TU (&un)[27][3],
                                         T (&acc)[3], T (&D)[9]) {
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ax = a & 1, ay = (a >> 1) & 1, az = a >> 2;
    const T we = wa[a];
    const int li = 7 - a;
    const T sx = (li & 1) ? T(-1) : T(1), sy = (li & 2) ? T(-1) : T(1), sz = (li & 4) ? T(-1) : T(1);
    T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int nb = (ax + (j & 1)) + 3 * (ay + ((j >> 1) & 1)) + 9 * (az + (j >> 2));
      if (!SELF && j == li) continue;
      if (!((MASK >> nb) & 1u)) continue;
      const int d = li ^ j;
      // neighbour components reflected into the frame of local node 0
      const T v0 = sx * T(un[nb][0]), v1 = sy * T(un[nb][1]), v2 = sz * T(un[nb][2]);
      t0 = fma(Kc[ke_key(d, 0, 0)], v0, t0); t0 = fma(Kc[ke_key(d, 0, 1)], v1, t0);
      t0 = fma(Kc[ke_key(d, 0, 2)], v2, t0);
      t1 = fma(Kc[ke_key(d, 1, 0)], v0, t1); t1 = fma(Kc[ke_key(d, 1, 1)], v1, t1);
      t1 = fma(Kc[ke_key(d, 1, 2)], v2, t1);
      t2 = fma(Kc[ke_key(d, 2, 0)], v0, t2); t2 = fma(Kc[ke_key(d, 2, 1)], v1, t2);
      t2 = fma(Kc[ke_key(d, 2, 2)], v2, t2);
    }
    acc[0] = fma(we * sx, t0, acc[0]);
    acc[1] = fma(we * sy, t1, acc[1]);
    acc[2] = fma(we * sz, t2, acc[2]);
    if (!SELF) {
      const T sg[3] = {sx, sy, sz};
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) D[r * 3 + c] = fma(we * (sg[r] * sg[c]), Kc[ke_key(0, r, c)], D[r * 3 + c]);
    }
  }
}

// MODE 0: y = K u (into y32 instead when it is set), optional partial sums of u.y;
// MODE 1: y = f - K u, optional float copy y32, optional partial sums of y.y (set 0) and
// f.f (set 1).  Only free nodes are
// touched (the others stay 0 in every vector).
// TU: type of u (float u with FP64 arithmetic: the FCG search direction of the mixed
// precision solves).  FP32 arithmetic: all 81 neighbour values live at once, register
// budget sized for 4 waves per SIMD.  FP64 keeps the compiler's interleaved schedule
// (twice the registers would halve the occupancy, which measured slower).
#define TOS_FINE_BOUNDS(T) __launch_bounds__(256, sizeof(T) == 4 ? 4 : 1)

template <class T, int MODE, class TU = T>
__global__ TOS_FINE_BOUNDS(T) void k_fine_apply(LevelView L, const int4* __restrict__ cl, int nch,
                                                 const T* __restrict__ w, const T* __restrict__ Ke,
                                                 const TU* __restrict__ u, const T* __restrict__ f,
                                                 T* __restrict__ y, double* __restrict__ partial,
                                                 float* __restrict__ y32) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double d = 0, d2 = 0;
  if (ok) {
    TU un[27][3];
    T wa[8], fv[3] = {0, 0, 0};
    load_nbhd(L, u, n.p, un);
    load_weights(L, w, n.x, n.y, n.z, wa);
    if (MODE == 1)
#pragma unroll
      for (int c = 0; c < 3; ++c) fv[c] = f[c * L.NP + n.p];
    if constexpr (sizeof(T) == 4) TOS_LOADS_DONE();
    T acc[3] = {0, 0, 0}, D[9];
    fine_row<T, true, 0x7ffffffu, TU>(wa, Ke, un, acc, D);
    if (MODE == 0) {  // (y32 set: the result stored in FP32 instead of y)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (y32) y32[c * L.NP + n.p] = (float)acc[c];
        else y[c * L.NP + n.p] = acc[c];
        d += (double)un[13][c] * (double)acc[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T v = fv[c] - acc[c];
        y[c * L.NP + n.p] = v;
        if (y32) y32[c * L.NP + n.p] = (float)v;
        d += (double)v * (double)v;
        d2 += (double)fv[c] * (double)fv[c];
      }
    }
  }
  if (partial) {
    block_partial(d, partial, 0);
    if (MODE == 1) block_partial(d2, partial, 1);
  }
}

// The elements of half AZ of a node's cube (az = AZ: a = 4 AZ .. 4 AZ + 3), the node's
// row of each, accumulated as in fine_row<T, true> (same FMAs in the same order) from two
// node planes: lower = the plane at dz = AZ - 1, upper = dz = AZ, index (dx+1) + 3 (dy+1).
template <int AZ, class T, class TU>
__device__ __forceinline__ void fine_row_half(const T (&wa)[8], const T* __restrict__ Kc, const TU (&lower)[9][3],
                                              const TU (&upper)[9][3], T (&acc)[3]) {
#pragma unroll
  for (int a = 4 * AZ; a < 4 * AZ + 4; ++a) {
    const int ax = a & 1, ay = (a >> 1) & 1;
    const T we = wa[a];
    const int li = 7 - a;
    const T sx = (li & 1) ? T(-1) : T(1), sy = (li & 2) ? T(-1) : T(1), sz = (li & 4) ? T(-1) : T(1);
    T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int nb = (ax + (j & 1)) + 3 * (ay + ((j >> 1) & 1));
      const TU(&pl)[9][3] = (j >> 2) ? upper : lower;
      const int d = li ^ j;
      const T v0 = sx * T(pl[nb][0]), v1 = sy * T(pl[nb][1]), v2 = sz * T(pl[nb][2]);
      t0 = fma(Kc[ke_key(d, 0, 0)], v0, t0); t0 = fma(Kc[ke_key(d, 0, 1)], v1, t0);
      t0 = fma(Kc[ke_key(d, 0, 2)], v2, t0);
      t1 = fma(Kc[ke_key(d, 1, 0)], v0, t1); t1 = fma(Kc[ke_key(d, 1, 1)], v1, t1);
      t1 = fma(Kc[ke_key(d, 1, 2)], v2, t1);
      t2 = fma(Kc[ke_key(d, 2, 0)], v0, t2); t2 = fma(Kc[ke_key(d, 2, 1)], v1, t2);
      t2 = fma(Kc[ke_key(d, 2, 2)], v2, t2);
    }
    acc[0] = fma(we * sx, t0, acc[0]);
    acc[1] = fma(we * sy, t1, acc[1]);
    acc[2] = fma(we * sz, t2, acc[2]);
  }
}

// q = K p for an FP32 search direction p with FP64 arithmetic (the FCG operator of the
// mixed-precision solves), q stored as TQ (FP64, or FP32 at solve precision 3), with the
// partial sums of p.q.  The elements below the node (az = 0) read node planes z-1 and z,
// those above planes z and z+1; plane z+1 is loaded into plane z-1's registers after the
// first half, so 54 instead of 81 neighbour values are live: 116 VGPRs, 4 waves per SIMD
// (k_fine_apply<double, 0, float> holds all 81: 250 VGPRs, 2 waves).  Same sums in the
// same order as fine_row<double, true> with float neighbours (it visits the elements
// a = 0 .. 7 in order), so the same results bit for bit.
template <class TQ>
__global__ __launch_bounds__(256) void k_fine_apply_dir(LevelView L, const int4* __restrict__ cl, int nch,
                                                        const double* __restrict__ w, const double* __restrict__ Kc,
                                                        const float* __restrict__ p, double* __restrict__ partial,
                                                        TQ* __restrict__ q) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double d = 0;
  if (ok) {
    float lo[9][3], mid[9][3];
    const int64_t NXY = (int64_t)L.NX * L.NY;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int64_t a = n.p + (k % 3 - 1) + (int64_t)L.NX * (k / 3 - 1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        lo[k][c] = p[c * L.NP + a - NXY];
        mid[k][c] = p[c * L.NP + a];
      }
    }
    double wa[8];
    load_weights(L, w, n.x, n.y, n.z, wa);
    double acc[3] = {0, 0, 0};
    fine_row_half<0>(wa, Kc, lo, mid, acc);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int64_t a = n.p + (k % 3 - 1) + (int64_t)L.NX * (k / 3 - 1) + NXY;
#pragma unroll
      for (int c = 0; c < 3; ++c) lo[k][c] = p[c * L.NP + a];
    }
    fine_row_half<1>(wa, Kc, mid, lo, acc);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      q[c * L.NP + n.p] = (TQ)acc[c];
      d += (double)mid[4][c] * acc[c];
    }
  }
  if (partial) block_partial(d, partial, 0);
}

// y = K u for an FP64 u in the same two halves (54 FP64 neighbour values live; 102 VGPRs,
// 4 waves per SIMD), with the partial sums of u.y.  Same results bit for bit as
// k_fine_apply<double, 0> (fine_row's element order and FMAs).
__global__ __launch_bounds__(256) void k_fine_apply_u64(LevelView L, const int4* __restrict__ cl, int nch,
                                                        const double* __restrict__ w, const double* __restrict__ Kc,
                                                        const double* __restrict__ u, double* __restrict__ y,
                                                        double* __restrict__ partial) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double d = 0;
  if (ok) {
    double lo[9][3], mid[9][3];
    const int64_t NXY = (int64_t)L.NX * L.NY;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int64_t a = n.p + (k % 3 - 1) + (int64_t)L.NX * (k / 3 - 1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        lo[k][c] = u[c * L.NP + a - NXY];
        mid[k][c] = u[c * L.NP + a];
      }
    }
    double wa[8];
    load_weights(L, w, n.x, n.y, n.z, wa);
    double acc[3] = {0, 0, 0};
    fine_row_half<0>(wa, Kc, lo, mid, acc);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int64_t a = n.p + (k % 3 - 1) + (int64_t)L.NX * (k / 3 - 1) + NXY;
#pragma unroll
      for (int c = 0; c < 3; ++c) lo[k][c] = u[c * L.NP + a];
    }
    fine_row_half<1>(wa, Kc, mid, lo, acc);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      y[c * L.NP + n.p] = acc[c];
      d += mid[4][c] * acc[c];
    }
  }
  if (partial) block_partial(d, partial, 0);
}

// Gauss-Seidel colour pass on the fine level.  BACKWARD relaxes components z, y, x.
// ZC: see load_nbhd (-1 = every neighbour, c = first forward sweep from u = 0 at colour c)
// point Gauss-Seidel on the node's 3x3 diagonal block: components x, y, z (BACKWARD:
// z, y, x), each with the others' latest values; b = f - off-diagonal sum
template <class T, bool BACKWARD>
// (explicit fmas: every GS kernel rounds the same way, so the schedules stay bit-identical
// whatever the compiler would contract in each of them)
__device__ __forceinline__ void block_relax(T b0, T b1, T b2, const T (&D)[9], T& v0, T& v1, T& v2) {
  if (!BACKWARD) {
    v0 = fma(-D[2], v2, fma(-D[1], v1, b0)) / D[0];
    v1 = fma(-D[5], v2, fma(-D[3], v0, b1)) / D[4];
    v2 = fma(-D[7], v1, fma(-D[6], v0, b2)) / D[8];
  } else {
    v2 = fma(-D[7], v1, fma(-D[6], v0, b2)) / D[8];
    v1 = fma(-D[5], v2, fma(-D[3], v0, b1)) / D[4];
    v0 = fma(-D[2], v2, fma(-D[1], v1, b0)) / D[0];
  }
}

// relax node n (one colour pass body)
template <class T, bool BACKWARD, int ZC>
__device__ __forceinline__ void relax_fine(const LevelView& L, const T* __restrict__ w, const T* __restrict__ Ke,
                                           T* __restrict__ u, const T* __restrict__ f, const NodeRef& n) {
  T un[27][3], wa[8];
  load_nbhd<T, ZC>(L, u, n.p, un);
  load_weights(L, w, n.x, n.y, n.z, wa);
  const int64_t p = n.p;
  const T f0 = f[p], f1 = f[L.NP + p], f2 = f[2 * L.NP + p];
  if constexpr (sizeof(T) == 4) TOS_LOADS_DONE();
  T s[3] = {0, 0, 0};
  T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  fine_row<T, false>(wa, Ke, un, s, D);
  T v0 = un[13][0], v1 = un[13][1], v2 = un[13][2];
  block_relax<T, BACKWARD>(f0 - s[0], f1 - s[1], f2 - s[2], D, v0, v1, v2);
  u[p] = v0;
  u[L.NP + p] = v1;
  u[2 * L.NP + p] = v2;
}

// Two colours of one sweep in one launch.  The colours c and c^1 (same y, z parity,
// opposite x parity) depend on each other only along x within a grid row, so one
// workgroup takes a whole row: first its nodes of colour CA, then - after a workgroup
// barrier - those of colour CB (CA relaxed first in the sweep order).  Same updates as
// two colour passes; each neighbourhood is streamed from HBM once for both.
// rows[r] = {first chunk of colour CA in the colour-ordered list, count, same for CB}.
// ZA / ZB: -1, or the colour for the first sweep from zero (see load_nbhd).
template <class T, bool BACKWARD, int ZA, int ZB>
__global__ TOS_FINE_BOUNDS(T) void k_fine_gs_pair(LevelView L, const int4* __restrict__ cl,
                                                  const int4* __restrict__ rows, int nrows,
                                                  const T* __restrict__ w, const T* __restrict__ Ke,
                                                  T* __restrict__ u, const T* __restrict__ f) {
  const int r = (int)xcd_block();
  if (r >= nrows) return;  // whole workgroup
  const int4 R = rows[r];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll 1
  for (int ph = 0; ph < 2; ++ph) {
    if (ph) __syncthreads();
    const int beg = ph ? R.z : R.x, cnt = ph ? R.w : R.y;
    for (int i = wv; i < cnt; i += 4) {
      const int4 c = cl[beg + i];
      if (lane < (c.y >> 16)) {
        const NodeRef n{(int64_t)c.x + 2 * lane, (c.y & 0xffff) + 2 * lane, c.z, c.w, 0};
        if constexpr (ZA == ZB) {
          relax_fine<T, BACKWARD, ZA>(L, w, Ke, u, f, n);
        } else {
          if (ph)
            relax_fine<T, BACKWARD, ZB>(L, w, Ke, u, f, n);
          else
            relax_fine<T, BACKWARD, ZA>(L, w, Ke, u, f, n);
        }
      }
    }
  }
}

// Two colours of one sweep in one pass, paired inside a wave.  Colours c1 (first in the
// sweep order) and c2 = c1 ^ 1 share the y, z parities, so a c2 node depends on c1 only
// through its two x-neighbours in the same grid row.  A wave takes a row segment: lane l
// relaxes the c1 node x1 = xs + 2l and then the c2 node x1 + 1 (l < 63), whose c1
// neighbours are lanes l and l + 1 -- their new values, and lane l+1's neighbour column,
// arrive by lane shuffles instead of a second pass over memory.  Segments of a row start
// 126 apart, so the c1 node of lane 63 is also lane 0 of the next segment (computed twice
// from the same inputs, written once).  Those two computations read the old values of
// that c1 node itself (its start value) and of the c2 nodes of lanes 0 and 62, and the
// two waves run in no fixed order, so the new values of these three nodes go to `ub` and
// are copied into u by k_fine_pairw_copy after the pass.  Same updates as two colour
// passes (bit-identical); each neighbourhood is streamed once for both colours.
// seg = {padded index of (xs, y, z), (xs + 1) | (xe + 1) << 15 | cont << 30, y, z}: xe the
// row's last free node, cont = 1 when lane 0 repeats the previous segment's lane 63.
// ZA / ZB: -1, or the colours of a first forward sweep from zero (see load_nbhd).
// DOTS (the last sweep of a V-cycle's post-smoothing, whose new values are final): each
// wave also writes its sums of u.f and u.q over the nodes it relaxed to dpart[si] and
// dpart[dT + si], so the FCG's z.r and z.q need no pass of their own.
template <class T, bool BACKWARD, int ZA, int ZB, bool DOTS = false>
__global__ TOS_FINE_BOUNDS(T) void k_fine_gs_pairw(LevelView L, const int4* __restrict__ seg, int nseg,
                                                   const uint8_t* __restrict__ fre, const T* __restrict__ w,
                                                   const T* __restrict__ Kc, T* __restrict__ u,
                                                   const T* __restrict__ f, T* __restrict__ ub,
                                                   const T* __restrict__ q, double* __restrict__ dpart,
                                                   int64_t dT) {
  constexpr unsigned MA = nb_mask_of(ZA), MB = nb_mask_of(ZB);
  const int si = (int)xcd_block() * 4 + (threadIdx.x >> 6);
  if (si >= nseg) return;  // whole wave
  const int lane = threadIdx.x & 63;
  const int4 S = seg[si];
  const int xs = (S.y & 0x7fff) - 1, xe = ((S.y >> 15) & 0x7fff) - 1;
  const bool cont = (S.y >> 30) & 1;
  const int x1 = xs + 2 * lane, y = S.z, z = S.w;
  const int64_t p1 = (int64_t)S.x + 2 * lane;
  const bool live = x1 <= xe + 1;
  // loads are unconditional, from addresses clamped into the row (lanes past the row read
  // its end, the x - 1 column of x1 = -1 reads x = -1): a value read at a clamped address
  // is never used, and no load needs a branch of its own
  const int xa = live ? x1 : xe + 1;
  const int64_t pa = (int64_t)S.x + (xa - xs);
  const bool f1 = live && fre[pa] != 0;  // (x1 = -1 is padding: never free)
  const bool f2 = live && lane < 63 && x1 + 1 <= xe && fre[pa + 1] != 0;
  // c1's neighbourhood (k = (dx+1) + 3 (dy+1) + 9 (dz+1)); column dx = -1 serves c1 only,
  // columns 0 and +1 are also c2's columns -1 and 0, and column 0 is lane l-1's column +1
  T un[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k % 3 - 1;
    const bool need = dx < 0 ? ((MA >> k) & 1u) != 0
                              : (((MA >> k) & 1u) || ((MB >> (k - 1)) & 1u) ||
                                 (dx == 0 && ((MB >> (k + 1)) & 1u)));
    const int64_t q = (dx < 0 ? (xa >= 0 ? pa - 1 : pa) : pa + dx) +
                      (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) un[k][c] = need ? u[c * L.NP + q] : T(0);
  }
  T wa[8];  // c1's elements (x1 - 1 + ax, y - 1 + ay, z - 1 + az), a = ax + 2 ay + 4 az
  {
    const int64_t e0 = L.pelem(xa >= 0 ? xa - 1 : -1, y - 1, z - 1);
    const int64_t e1 = L.pelem(xa <= L.ex ? xa : L.ex, y - 1, z - 1);
#pragma unroll
    for (int a = 0; a < 8; ++a)
      wa[a] = w[((a & 1) ? e1 : e0) + (int64_t)L.EX * (((a >> 1) & 1) + (int64_t)L.EY * (a >> 2))];
  }
  T fa[3], fb[3], qa[3], qb[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    fa[c] = f[c * L.NP + pa];
    fb[c] = f[c * L.NP + pa + 1];
    if constexpr (DOTS) {
      qa[c] = q[c * L.NP + pa];
      qb[c] = q[c * L.NP + pa + 1];
    }
  }
  if constexpr (sizeof(T) == 4) TOS_LOADS_DONE();
  // c1 (from zero: its own start value is 0; un[13] may hold the old value, loaded for c2)
  T v0 = ZA < 0 ? un[13][0] : T(0), v1 = ZA < 0 ? un[13][1] : T(0), v2 = ZA < 0 ? un[13][2] : T(0);
  if (f1) {
    T s[3] = {0, 0, 0};
    T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    fine_row<T, false, MA>(wa, Kc, un, s, D);
    block_relax<T, BACKWARD>(fa[0] - s[0], fa[1] - s[1], fa[2] - s[2], D, v0, v1, v2);
  }
  // c2's neighbourhood: own columns 0, +1 and lane l+1's column 0, with the new c1 values
  T vn[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k % 3 - 1;
    if (dx < 0) {  // dx = -1: own column 0 (centre row: the new c1 value, set below)
#pragma unroll
      for (int c = 0; c < 3; ++c) vn[k][c] = un[k + 1][c];
    } else if (dx == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) vn[k][c] = un[k + 1][c];
    }
  }
  vn[12][0] = v0; vn[12][1] = v1; vn[12][2] = v2;
#pragma unroll
  for (int k = 2; k < 27; k += 3) {  // dx = +1: lane l+1's column 0
    const int k0 = k - 1;             // its index in the neighbour lane's c1 frame
    const bool need = ((MB >> k) & 1u) != 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T mine = k0 == 13 ? (c == 0 ? v0 : (c == 1 ? v1 : v2)) : un[k0][c];
      vn[k][c] = need ? __shfl_down(mine, 1, 64) : T(0);
    }
  }
  T wb[8];  // c2's elements: column x1 (own ax = 1) and x1 + 1 (lane l+1's ax = 0)
#pragma unroll
  for (int a = 0; a < 8; a += 2) {
    wb[a] = wa[a + 1];
    wb[a + 1] = __shfl_down(wa[a], 1, 64);
  }
  T t0 = ZB < 0 ? vn[13][0] : T(0), t1 = ZB < 0 ? vn[13][1] : T(0), t2 = ZB < 0 ? vn[13][2] : T(0);
  if (f2) {
    T s[3] = {0, 0, 0};
    T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    fine_row<T, false, MB>(wb, Kc, vn, s, D);
    block_relax<T, BACKWARD>(fb[0] - s[0], fb[1] - s[1], fb[2] - s[2], D, t0, t1, t2);
  }
  if (f1 && !(cont && lane == 0)) {
    T* dst = lane == 63 ? ub : u;  // (the next segment's lane 0 reads its start value)
    dst[p1] = v0; dst[L.NP + p1] = v1; dst[2 * L.NP + p1] = v2;
  }
  if (f2) {
    T* dst = (lane == 0 || lane == 62) ? ub : u;
    dst[p1 + 1] = t0; dst[L.NP + p1 + 1] = t1; dst[2 * L.NP + p1 + 1] = t2;
  }
  if constexpr (DOTS) {  // each node once: where it is written (to u or to ub)
    double zr = 0, zq = 0;
    if (f1 && !(cont && lane == 0)) {
      zr = (double)v0 * (double)fa[0] + (double)v1 * (double)fa[1] + (double)v2 * (double)fa[2];
      zq = (double)v0 * (double)qa[0] + (double)v1 * (double)qa[1] + (double)v2 * (double)qa[2];
    }
    if (f2) {
      zr += (double)t0 * (double)fb[0] + (double)t1 * (double)fb[1] + (double)t2 * (double)fb[2];
      zq += (double)t0 * (double)qb[0] + (double)t1 * (double)qb[1] + (double)t2 * (double)qb[2];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      zr += __shfl_xor(zr, o, 64);
      zq += __shfl_xor(zq, o, 64);
    }
    if (lane == 0) {
      dpart[si] = zr;
      dpart[dT + si] = zq;
    }
  }
}

// k_fine_resid_z for the two colours c1 = ZA, c2 = ZB = ZA + 1 of a class in one pass, on
// the forward segments of k_fine_gs_pairw: lane l forms the residual of the c1 node x1 and
// of the c2 node x1 + 1 from one load of x1's neighbourhood plus lane l+1's neighbour
// column.  u is only read, so the repeated c1 node of lane 0 is simply written twice with
// the same value.
template <class T, int ZA, int ZB>
__global__ TOS_FINE_BOUNDS(T) void k_fine_resid_pairw(LevelView L, const int4* __restrict__ seg, int nseg,
                                                      const uint8_t* __restrict__ fre, const T* __restrict__ w,
                                                      const T* __restrict__ Kc, const T* __restrict__ u,
                                                      T* __restrict__ r) {
  constexpr unsigned MA = nb_mask_after(ZA), MB = nb_mask_after(ZB);
  constexpr unsigned LA = MA | (1u << 13), LB = MB | (1u << 13);  // (own values too)
  const int si = (int)xcd_block() * 4 + (threadIdx.x >> 6);
  if (si >= nseg) return;  // whole wave
  const int lane = threadIdx.x & 63;
  const int4 S = seg[si];
  const int xs = (S.y & 0x7fff) - 1, xe = ((S.y >> 15) & 0x7fff) - 1;
  const int x1 = xs + 2 * lane, y = S.z, z = S.w;
  const bool live = x1 <= xe + 1;
  const int xa = live ? x1 : xe + 1;  // clamped as in k_fine_gs_pairw
  const int64_t pa = (int64_t)S.x + (xa - xs);
  const bool f1 = live && fre[pa] != 0;
  const bool f2 = live && lane < 63 && x1 + 1 <= xe && fre[pa + 1] != 0;
  T un[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k % 3 - 1;
    const bool need = dx < 0 ? ((LA >> k) & 1u) != 0
                              : (((LA >> k) & 1u) || ((LB >> (k - 1)) & 1u) ||
                                 (dx == 0 && ((LB >> (k + 1)) & 1u)));
    const int64_t q = (dx < 0 ? (xa >= 0 ? pa - 1 : pa) : pa + dx) +
                      (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) un[k][c] = need ? u[c * L.NP + q] : T(0);
  }
  T wa[8];
  {
    const int64_t e0 = L.pelem(xa >= 0 ? xa - 1 : -1, y - 1, z - 1);
    const int64_t e1 = L.pelem(xa <= L.ex ? xa : L.ex, y - 1, z - 1);
#pragma unroll
    for (int a = 0; a < 8; ++a)
      wa[a] = w[((a & 1) ? e1 : e0) + (int64_t)L.EX * (((a >> 1) & 1) + (int64_t)L.EY * (a >> 2))];
  }
  if constexpr (sizeof(T) == 4) TOS_LOADS_DONE();
  if (f1) {
    T s[3] = {0, 0, 0};
    T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    fine_row<T, false, MA>(wa, Kc, un, s, D);
    const T v1 = un[13][1], v2 = un[13][2];
    r[pa] = -D[1] * v1 - D[2] * v2 - s[0];
    r[L.NP + pa] = -D[5] * v2 - s[1];
    r[2 * L.NP + pa] = -s[2];
  }
  T vn[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k % 3 - 1;
    if (dx <= 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) vn[k][c] = un[k + 1][c];
    } else {
      const bool need = ((LB >> k) & 1u) != 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) vn[k][c] = need ? __shfl_down(un[k - 1][c], 1, 64) : T(0);
    }
  }
  T wb[8];
#pragma unroll
  for (int a = 0; a < 8; a += 2) {
    wb[a] = wa[a + 1];
    wb[a + 1] = __shfl_down(wa[a], 1, 64);
  }
  if (f2) {
    T s[3] = {0, 0, 0};
    T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    fine_row<T, false, MB>(wb, Kc, vn, s, D);
    const T v1 = vn[13][1], v2 = vn[13][2];
    r[pa + 1] = -D[1] * v1 - D[2] * v2 - s[0];
    r[L.NP + pa + 1] = -D[5] * v2 - s[1];
    r[2 * L.NP + pa + 1] = -s[2];
  }
}

// the deferred nodes of k_fine_gs_pairw / k_st_gs_pairw (the c2 nodes of lanes 0 and 62
// and the c1 node of lane 63 of every segment, x = xs + 1, xs + 125, xs + 126): ub -> u.
// Each was written in the pass exactly when it is free and at most xe.
template <class T>
__global__ __launch_bounds__(256) void k_fine_pairw_copy(LevelView L, const int4* __restrict__ seg, int nseg,
                                                         const uint8_t* __restrict__ fre, const T* __restrict__ ub,
                                                         T* __restrict__ u) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * (int64_t)nseg) return;
  const int4 S = seg[t / 3];
  const int xs = (S.y & 0x7fff) - 1, xe = ((S.y >> 15) & 0x7fff) - 1;
  const int j = (int)(t % 3);
  const int dxn = j == 0 ? 1 : (j == 1 ? 125 : 126);
  if (xs + dxn > xe) return;
  const int64_t p = (int64_t)S.x + dxn;
  if (!fre[p]) return;
  u[p] = ub[p]; u[L.NP + p] = ub[L.NP + p]; u[2 * L.NP + p] = ub[2 * L.NP + p];
}

template <class T, bool BACKWARD, int ZC = -1>
__global__ TOS_FINE_BOUNDS(T) void k_fine_gs(LevelView L, const int4* __restrict__ cl, int nch,
                                             const T* __restrict__ w, const T* __restrict__ Ke,
                                             T* __restrict__ u, const T* __restrict__ f) {
  NodeRef n;
  if (!chunk_node(cl, nch, 2, n)) return;
  relax_fine<T, BACKWARD, ZC>(L, w, Ke, u, f, n);
}

// Residual after ONE forward sweep that started from u = 0, for the nodes of colour ZC.
// When node i was relaxed, the neighbours of earlier colours already had their final
// values and those of later colours were 0, and the point GS inside the 3x3 block left
// the block equations with residual (-D01 v1 - D02 v2, -D12 v2, 0).  Hence
//   r_i = (-D01 v1 - D02 v2, -D12 v2, 0) - sum_{j of a later colour} A_ij u_j :
// only the later-colour half of the neighbourhood (and no f) is read.
template <class T, int ZC>
__global__ TOS_FINE_BOUNDS(T) void k_fine_resid_z(LevelView L, const int4* __restrict__ cl, int nch,
                                                  const T* __restrict__ w, const T* __restrict__ Ke,
                                                  const T* __restrict__ u, T* __restrict__ r) {
  constexpr unsigned after = nb_mask_after(ZC);
  NodeRef n;
  if (!chunk_node(cl, nch, 2, n)) return;
  T un[27][3], wa[8];
  load_nbhd<T, ZC, after | (1u << 13)>(L, u, n.p, un);
  load_weights(L, w, n.x, n.y, n.z, wa);
  if constexpr (sizeof(T) == 4) TOS_LOADS_DONE();
  T s[3] = {0, 0, 0};
  T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  fine_row<T, false, after>(wa, Ke, un, s, D);
  const T v1 = un[13][1], v2 = un[13][2];
  const int64_t p = n.p;
  r[p] = -D[1] * v1 - D[2] * v2 - s[0];
  r[L.NP + p] = -D[5] * v2 - s[1];
  r[2 * L.NP + p] = -s[2];
}

// The same pass over several chunk ranges at once (one wavefront step): block b belongs
// to the range g with tab.blk[g] <= b < tab.blk[g+1].
__device__ __forceinline__ bool seg_node(const int4* __restrict__ cl, const SegTab& tab, NodeRef& n) {
  const int b = (int)xcd_block();
  int g = 0;
#pragma unroll
  for (int i = 1; i < 8; ++i)
    if (i < tab.n && b >= tab.blk[i]) g = i;
  const int loc = (b - tab.blk[g]) * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (loc >= tab.cnt[g]) return false;
  const int4 c = cl[tab.beg[g] + loc];
  if (lane >= (c.y >> 16)) return false;
  n.x = (c.y & 0xffff) + 2 * lane;
  n.y = c.z;
  n.z = c.w;
  n.p = (int64_t)c.x + 2 * lane;
  n.slot = 0;
  return true;
}

template <class T, bool BACKWARD>
__global__ TOS_FINE_BOUNDS(T) void k_fine_gs_wf(LevelView L, const int4* __restrict__ cl, SegTab tab,
                                                const T* __restrict__ w, const T* __restrict__ Ke,
                                                T* __restrict__ u, const T* __restrict__ f) {
  NodeRef n;
  if (!seg_node(cl, tab, n)) return;
  relax_fine<T, BACKWARD, -1>(L, w, Ke, u, f, n);
}

// ---------------------------------------------------------------------------------------
// Stored-stencil levels (colour-compact stencil, NS slots per entry)
// ---------------------------------------------------------------------------------------
// Readers of the compact stencil: entry e (3x3 row-major) of block k of the node in slot s.
// row[i] of a stored-stencil row through a buffer resource: the row base is uniform
// (kernel arguments and unrolled constants), so the address needs no per-lane 64-bit
// arithmetic (scalar base, 32-bit lane byte offset: a level >= 1 has at most 2^28 slots
// under the 2^31 padded-node limit of a context)
template <class T>
__device__ __forceinline__ T ld_row(const T* row, uint32_t i) {
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(row), (short)0, -1,
                                                                      0x00020000);
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8u, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, i * 4u, 0, 0));
}

template <class T>
struct StFull {  // T entries at (9k + e) NS + s
  const T* __restrict__ p;
  int64_t NS;
  static constexpr bool kScaled = false;
  __device__ __forceinline__ T get(int64_t s, int k, int e) const {
    return ld_row(p + (int64_t)(k * 9 + e) * NS, (uint32_t)s);
  }
  __device__ __forceinline__ T scale(int64_t) const { return T(1); }
};

__device__ __forceinline__ float half_bits(uint32_t h) {
  return (float)__builtin_bit_cast(_Float16, (unsigned short)h);
}
__device__ __forceinline__ uint32_t to_half_bits(float x) {
  return (uint32_t)__builtin_bit_cast(unsigned short, (_Float16)x);
}

// FP16 entries of the node's stencil divided by its largest magnitude m = msc[s] (every
// stored value in [-1, 1]), two per 32-bit word: entries 2w, 2w+1 of block k in word
// 4k + w (w < 4); entry 8 of block k in half k&1 of word 108 + k/2.  122 words per node
// instead of 243 floats.
constexpr int kHalfWords = 122;
struct StHalf {
  const uint32_t* __restrict__ p;
  const float* __restrict__ msc;
  int64_t NS;
  static constexpr bool kScaled = true;
  __device__ __forceinline__ float get(int64_t s, int k, int e) const {
    const int w = e < 8 ? 4 * k + (e >> 1) : 108 + (k >> 1);
    const int hi = e < 8 ? (e & 1) : (k & 1);
    const uint32_t v = ld_row(p + (int64_t)w * NS, (uint32_t)s);
    return half_bits(hi ? (v >> 16) : (v & 0xffffu));
  }
  __device__ __forceinline__ float scale(int64_t s) const { return msc[s]; }
};

// FP16 copy (StHalf layout) of a compact FP32 stencil; thread per slot of the colour list.
// The node's 243 entries are read once into registers (all loads in flight together).
__global__ __launch_bounds__(256) void k_st_pack16(const int4* __restrict__ cl, int nch,
                                                   const float* __restrict__ st, int64_t NS,
                                                   uint32_t* __restrict__ h, float* __restrict__ msc) {
  NodeRef n;
  if (!chunk_node(cl, nch, 2, n)) return;
  const int64_t s = n.slot;
  float v[243];
#pragma unroll
  for (int q = 0; q < 243; ++q) v[q] = st[(int64_t)q * NS + s];
  float mx = 0;
#pragma unroll
  for (int q = 0; q < 243; ++q) mx = fmaxf(mx, fabsf(v[q]));
  if (!(mx > 0)) mx = 1;
  const float inv = 1.0f / mx;
  msc[s] = mx;
#pragma unroll
  for (int k = 0; k < 27; ++k)
#pragma unroll
    for (int w = 0; w < 4; ++w)
      h[(int64_t)(4 * k + w) * NS + s] =
          to_half_bits(v[k * 9 + 2 * w] * inv) | (to_half_bits(v[k * 9 + 2 * w + 1] * inv) << 16);
#pragma unroll
  for (int t = 0; t < 14; ++t)
    h[(int64_t)(108 + t) * NS + s] =
        to_half_bits(v[2 * t * 9 + 8] * inv) | (2 * t + 1 < 27 ? to_half_bits(v[(2 * t + 1) * 9 + 8] * inv) << 16 : 0u);
}

// y = K u (MODE 0) or f - K u (MODE 1) at node p with stencil slot s
template <class T, int MODE, class ST>
__device__ __forceinline__ void st_apply_node(const LevelView& L, int64_t p, int64_t s, const ST& st,
                                              const T* __restrict__ u, const T* __restrict__ f,
                                              T* __restrict__ y) {
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int64_t q = p + (int64_t)(k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
    const T u0 = u[q], u1 = u[L.NP + q], u2 = u[2 * L.NP + q];
    a0 = fma(st.get(s, k, 0), u0, a0); a0 = fma(st.get(s, k, 1), u1, a0); a0 = fma(st.get(s, k, 2), u2, a0);
    a1 = fma(st.get(s, k, 3), u0, a1); a1 = fma(st.get(s, k, 4), u1, a1); a1 = fma(st.get(s, k, 5), u2, a1);
    a2 = fma(st.get(s, k, 6), u0, a2); a2 = fma(st.get(s, k, 7), u1, a2); a2 = fma(st.get(s, k, 8), u2, a2);
  }
  if constexpr (ST::kScaled) {
    const T m = st.scale(s);
    a0 *= m; a1 *= m; a2 *= m;
  }
  if (MODE == 0) {
    y[p] = a0; y[L.NP + p] = a1; y[2 * L.NP + p] = a2;
  } else {
    y[p] = f[p] - a0; y[L.NP + p] = f[L.NP + p] - a1; y[2 * L.NP + p] = f[2 * L.NP + p] - a2;
  }
}

// y = K u (MODE 0) or f - K u (MODE 1) over the whole colour-ordered list
template <class T, int MODE, class ST>
__global__ __launch_bounds__(256) void k_st_apply(LevelView L, const int4* __restrict__ cl, int nch, ST st,
                                                  const T* __restrict__ u, const T* __restrict__ f,
                                                  T* __restrict__ y) {
  NodeRef n;
  if (!chunk_node(cl, nch, 2, n)) return;
  st_apply_node<T, MODE>(L, n.p, n.slot, st, u, f, y);
}

// Relax one node from its loaded neighbourhood un (k = (dx+1) + 3 (dy+1) + 9 (dz+1)) and
// its stencil row at `slot`; MASK: the neighbours read (the others hold 0 and are skipped);
// v: the node's value (its start value on entry).  With a scaled stencil the block
// equations are solved in the scaled form (K~ = K / m, f / m), which gives the same update.
template <class T, bool BACKWARD, unsigned MASK, class ST>
__device__ __forceinline__ void st_relax(const ST& st, int64_t slot, const T (&un)[27][3], T f0, T f1, T f2,
                                         T& v0, T& v1, T& v2) {
  T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    if (k == 13 || !((MASK >> k) & 1u)) continue;  // skipped neighbours hold 0: no stencil read
    const T u0 = un[k][0], u1 = un[k][1], u2 = un[k][2];
    s0 = fma(st.get(slot, k, 0), u0, s0); s0 = fma(st.get(slot, k, 1), u1, s0); s0 = fma(st.get(slot, k, 2), u2, s0);
    s1 = fma(st.get(slot, k, 3), u0, s1); s1 = fma(st.get(slot, k, 4), u1, s1); s1 = fma(st.get(slot, k, 5), u2, s1);
    s2 = fma(st.get(slot, k, 6), u0, s2); s2 = fma(st.get(slot, k, 7), u1, s2); s2 = fma(st.get(slot, k, 8), u2, s2);
  }
  T D[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) D[e] = st.get(slot, 13, e);
  if constexpr (ST::kScaled) {
    const T inv = T(1) / st.scale(slot);
    f0 *= inv; f1 *= inv; f2 *= inv;
  }
  block_relax<T, BACKWARD>(f0 - s0, f1 - s1, f2 - s2, D, v0, v1, v2);
}

// relax node p (stencil slot `slot`) of colour ZC (-1: any colour, all neighbours read)
template <class T, bool BACKWARD, int ZC, class ST>
__device__ __forceinline__ void st_gs_node(const LevelView& L, int64_t p, int64_t slot, const ST& st,
                                           T* __restrict__ u, const T* __restrict__ f) {
  constexpr unsigned mask = nb_mask_of(ZC);
  T un[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const bool ld = k != 13 && ((mask >> k) & 1u);
    const int64_t q = p + (int64_t)(k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) un[k][c] = ld ? u[c * L.NP + q] : T(0);
  }
  T v0 = 0, v1 = 0, v2 = 0;
  if constexpr (ZC < 0) {
    v0 = u[p]; v1 = u[L.NP + p]; v2 = u[2 * L.NP + p];
  }
  st_relax<T, BACKWARD, mask>(st, slot, un, f[p], f[L.NP + p], f[2 * L.NP + p], v0, v1, v2);
  u[p] = v0;
  u[L.NP + p] = v1;
  u[2 * L.NP + p] = v2;
}

// one colour: cl points at the colour's chunks, slot0 = 64 * (their offset in the list)
template <class T, bool BACKWARD, int ZC, class ST>
__global__ __launch_bounds__(256) void k_st_gs(LevelView L, const int4* __restrict__ cl, int nch, int64_t slot0,
                                               ST st, T* __restrict__ u, const T* __restrict__ f) {
  NodeRef n;
  if (!chunk_node(cl, nch, 2, n)) return;
  st_gs_node<T, BACKWARD, ZC>(L, n.p, slot0 + n.slot, st, u, f);
}

// Stored-stencil counterpart of k_fine_gs_pairw (same segments, the same repeated c1 node
// at segment boundaries and the same two deferred c2 nodes per segment, written to `ub`
// and copied by k_fine_pairw_copy; levels whose rows are single segments pass ub = null
// and need no copy).  Stencil rows are found through the level's slot map.
template <class T, bool BACKWARD, int ZA, int ZB, class ST>
__global__ __launch_bounds__(256) void k_st_gs_pairw(LevelView L, const int4* __restrict__ seg, int nseg,
                                                     const uint8_t* __restrict__ fre,
                                                     const int* __restrict__ slotmap, ST st,
                                                     T* __restrict__ u, const T* __restrict__ f, T* __restrict__ ub) {
  constexpr unsigned MA = nb_mask_of(ZA), MB = nb_mask_of(ZB);
  const int si = (int)xcd_block() * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);  // (1 or 4 waves per block)
  if (si >= nseg) return;  // whole wave
  const int lane = threadIdx.x & 63;
  const int4 S = seg[si];
  const int xs = (S.y & 0x7fff) - 1, xe = ((S.y >> 15) & 0x7fff) - 1;
  const bool cont = (S.y >> 30) & 1;
  const int x1 = xs + 2 * lane;
  const bool live = x1 <= xe + 1;
  const int xa = live ? x1 : xe + 1;  // loads from addresses clamped into the row
  const int64_t pa = (int64_t)S.x + (xa - xs);
  const bool f1 = live && fre[pa] != 0;
  const bool f2 = live && lane < 63 && x1 + 1 <= xe && fre[pa + 1] != 0;
  T un[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k % 3 - 1;
    const bool need = dx < 0 ? ((MA >> k) & 1u) != 0
                              : (((MA >> k) & 1u) || ((MB >> (k - 1)) & 1u) ||
                                 (dx == 0 && ((MB >> (k + 1)) & 1u)));
    const int64_t q = (dx < 0 ? (xa >= 0 ? pa - 1 : pa) : pa + dx) +
                      (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) un[k][c] = need ? u[c * L.NP + q] : T(0);
  }
  const int sl1 = slotmap[pa], sl2 = slotmap[pa + 1];
  T fa[3], fb[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    fa[c] = f[c * L.NP + pa];
    fb[c] = f[c * L.NP + pa + 1];
  }
  // c1 (from zero: its own start value is 0)
  T v0 = ZA < 0 ? un[13][0] : T(0), v1 = ZA < 0 ? un[13][1] : T(0), v2 = ZA < 0 ? un[13][2] : T(0);
  if (f1) st_relax<T, BACKWARD, MA>(st, sl1, un, fa[0], fa[1], fa[2], v0, v1, v2);
  // c2's neighbourhood: own columns 0, +1 and lane l+1's column 0, with the new c1 values
  T vn[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    if (k % 3 - 1 <= 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) vn[k][c] = un[k + 1][c];
    }
  }
  vn[12][0] = v0; vn[12][1] = v1; vn[12][2] = v2;
#pragma unroll
  for (int k = 2; k < 27; k += 3) {
    const int k0 = k - 1;
    const bool need = ((MB >> k) & 1u) != 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T mine = k0 == 13 ? (c == 0 ? v0 : (c == 1 ? v1 : v2)) : un[k0][c];
      vn[k][c] = need ? __shfl_down(mine, 1, 64) : T(0);
    }
  }
  T t0 = ZB < 0 ? vn[13][0] : T(0), t1 = ZB < 0 ? vn[13][1] : T(0), t2 = ZB < 0 ? vn[13][2] : T(0);
  if (f2) st_relax<T, BACKWARD, MB>(st, sl2, vn, fb[0], fb[1], fb[2], t0, t1, t2);
  if (f1 && !(cont && lane == 0)) {
    T* dst = (ub && lane == 63) ? ub : u;  // (the next segment's lane 0 reads its start value)
    dst[pa] = v0; dst[L.NP + pa] = v1; dst[2 * L.NP + pa] = v2;
  }
  if (f2) {
    T* dst = (ub && (lane == 0 || lane == 62)) ? ub : u;
    dst[pa + 1] = t0; dst[L.NP + pa + 1] = t1; dst[2 * L.NP + pa + 1] = t2;
  }
}

// residual after a first forward sweep from zero of a node with stencil slot `slot` from
// its loaded neighbourhood un: later-colour blocks (MASK = nb_mask_after of its colour)
// and its own block (un[13] holds its value)
template <class T, unsigned MASK, class ST>
__device__ __forceinline__ void st_resid_after(const ST& st, int64_t slot, const T (&un)[27][3], T& r0, T& r1,
                                               T& r2) {
  T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    if (!((MASK >> k) & 1u)) continue;
    const T u0 = un[k][0], u1 = un[k][1], u2 = un[k][2];
    s0 = fma(st.get(slot, k, 0), u0, s0); s0 = fma(st.get(slot, k, 1), u1, s0); s0 = fma(st.get(slot, k, 2), u2, s0);
    s1 = fma(st.get(slot, k, 3), u0, s1); s1 = fma(st.get(slot, k, 4), u1, s1); s1 = fma(st.get(slot, k, 5), u2, s1);
    s2 = fma(st.get(slot, k, 6), u0, s2); s2 = fma(st.get(slot, k, 7), u1, s2); s2 = fma(st.get(slot, k, 8), u2, s2);
  }
  // (explicit fmas: every kernel that calls this rounds the same way, whatever the
  // compiler would contract around it)
  const T v1 = un[13][1], v2 = un[13][2];
  r0 = fma(-st.get(slot, 13, 1), v1, fma(-st.get(slot, 13, 2), v2, -s0));
  r1 = fma(-st.get(slot, 13, 5), v2, -s1);
  r2 = -s2;
  if constexpr (ST::kScaled) {
    const T m = st.scale(slot);
    r0 *= m; r1 *= m; r2 *= m;
  }
}

// stored-stencil counterpart of k_fine_resid_z (nodes of colour ZC, later-colour blocks only)
template <class T, int ZC, class ST>
__global__ __launch_bounds__(256) void k_st_resid_z(LevelView L, const int4* __restrict__ cl, int nch, int64_t slot0,
                                                    ST st, const T* __restrict__ u, T* __restrict__ r) {
  constexpr unsigned after = nb_mask_after(ZC);
  NodeRef n;
  if (!chunk_node(cl, nch, 2, n)) return;
  const int64_t p = n.p;
  T un[27][3];
  load_nbhd<T, -1, after | (1u << 13)>(L, u, p, un);
  T r0, r1, r2;
  st_resid_after<T, after>(st, slot0 + n.slot, un, r0, r1, r2);
  r[p] = r0;
  r[L.NP + p] = r1;
  r[2 * L.NP + p] = r2;
}

// k_st_resid_z for the two colours ZA, ZB = ZA + 1 of a class in one pass, on the forward
// segments of k_st_gs_pairw (lane l: the ZA node x1 and the ZB node x1 + 1 from one load
// of x1's neighbourhood plus lane l+1's neighbour column; stencil rows by the slot map)
template <class T, int ZA, int ZB, class ST>
__global__ __launch_bounds__(256) void k_st_resid_pairw(LevelView L, const int4* __restrict__ seg, int nseg,
                                                        const uint8_t* __restrict__ fre,
                                                        const int* __restrict__ slotmap, ST st,
                                                        const T* __restrict__ u, T* __restrict__ r) {
  constexpr unsigned MA = nb_mask_after(ZA), MB = nb_mask_after(ZB);
  constexpr unsigned LA = MA | (1u << 13), LB = MB | (1u << 13);  // (own values too)
  const int si = (int)xcd_block() * 4 + (threadIdx.x >> 6);
  if (si >= nseg) return;  // whole wave
  const int lane = threadIdx.x & 63;
  const int4 S = seg[si];
  const int xs = (S.y & 0x7fff) - 1, xe = ((S.y >> 15) & 0x7fff) - 1;
  const int x1 = xs + 2 * lane;
  const bool live = x1 <= xe + 1;
  const int xa = live ? x1 : xe + 1;  // clamped as in k_st_gs_pairw
  const int64_t pa = (int64_t)S.x + (xa - xs);
  const bool f1 = live && fre[pa] != 0;
  const bool f2 = live && lane < 63 && x1 + 1 <= xe && fre[pa + 1] != 0;
  T un[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k % 3 - 1;
    const bool need = dx < 0 ? ((LA >> k) & 1u) != 0
                              : (((LA >> k) & 1u) || ((LB >> (k - 1)) & 1u) ||
                                 (dx == 0 && ((LB >> (k + 1)) & 1u)));
    const int64_t q = (dx < 0 ? (xa >= 0 ? pa - 1 : pa) : pa + dx) +
                      (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
#pragma unroll
    for (int c = 0; c < 3; ++c) un[k][c] = need ? u[c * L.NP + q] : T(0);
  }
  const int sl1 = slotmap[pa], sl2 = slotmap[pa + 1];
  if (f1) {
    T r0, r1, r2;
    st_resid_after<T, MA>(st, sl1, un, r0, r1, r2);
    r[pa] = r0; r[L.NP + pa] = r1; r[2 * L.NP + pa] = r2;
  }
  T vn[27][3];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k % 3 - 1;
    if (dx <= 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) vn[k][c] = un[k + 1][c];
    } else {
      const bool need = ((LB >> k) & 1u) != 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) vn[k][c] = need ? __shfl_down(un[k - 1][c], 1, 64) : T(0);
    }
  }
  if (f2) {
    T r0, r1, r2;
    st_resid_after<T, MB>(st, sl2, vn, r0, r1, r2);
    r[pa + 1] = r0; r[L.NP + pa + 1] = r1; r[2 * L.NP + pa + 1] = r2;
  }
}

// ---------------------------------------------------------------------------------------
// Transfers: trilinear P (weights 1, 1/2, 1/4, 1/8), restriction = P^T, both limited to
// free nodes (restrict_residual_kernel grid/Grid.cu:2456-2507, prolongate_correction_kernel
// :2594-2632)
// ---------------------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void restrict_node(const LevelView& F, const LevelView& C, const NodeRef& n,
                                              const T* __restrict__ rf, T* __restrict__ fc) {
  T a0 = 0, a1 = 0, a2 = 0;
  const int64_t c0 = F.pnode(2 * n.x, 2 * n.y, 2 * n.z);
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
    const T wgt = T((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0));
    const int64_t a = c0 + dx + (int64_t)F.NX * (dy + (int64_t)F.NY * dz);
    a0 += wgt * rf[a]; a1 += wgt * rf[F.NP + a]; a2 += wgt * rf[2 * F.NP + a];
  }
  fc[n.p] = a0; fc[C.NP + n.p] = a1; fc[2 * C.NP + n.p] = a2;
}

template <class T>
__global__ __launch_bounds__(256) void k_restrict(LevelView F, LevelView C, const int4* __restrict__ clc, int nch,
                                                  const T* __restrict__ rf, T* __restrict__ fc) {
  NodeRef n;
  if (!chunk_node(clc, nch, 1, n)) return;
  restrict_node(F, C, n, rf, fc);
}

// u_f += P u_c: a fine node takes the trilinear interpolation of the (up to 8) coarse
// nodes around it.  All 8 corner nodes (X..X+1, Y..Y+1, Z..Z+1, within the padding) are
// read with weight 0 where a fine coordinate is even, so the loads are issued together
// instead of as a divergent loop.
template <class T>
__device__ __forceinline__ void prolong_node(const LevelView& F, const LevelView& C, const NodeRef& n,
                                             const T* __restrict__ uc, T* __restrict__ uf) {
  const int x = n.x, y = n.y, z = n.z;
  const T hx = (x & 1) ? T(0.5) : T(0), hy = (y & 1) ? T(0.5) : T(0), hz = (z & 1) ? T(0.5) : T(0);
  const T wx[2] = {T(1) - hx, hx}, wy[2] = {T(1) - hy, hy}, wz[2] = {T(1) - hz, hz};
  const int64_t I0 = C.pnode(x >> 1, y >> 1, z >> 1);
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = k & 1, j = (k >> 1) & 1, l = k >> 2;
    const T wgt = wx[i] * wy[j] * wz[l];
    const int64_t I = I0 + i + (int64_t)C.NX * (j + (int64_t)C.NY * l);
    a0 = fma(wgt, uc[I], a0);
    a1 = fma(wgt, uc[C.NP + I], a1);
    a2 = fma(wgt, uc[2 * C.NP + I], a2);
  }
  const int64_t p = n.p;
  const T f0 = uf[p], f1 = uf[F.NP + p], f2 = uf[2 * F.NP + p];
  uf[p] = f0 + a0;
  uf[F.NP + p] = f1 + a1;
  uf[2 * F.NP + p] = f2 + a2;
}

template <class T>
__global__ __launch_bounds__(256) void k_prolong_add(LevelView F, LevelView C, const int4* __restrict__ clf, int nch,
                                                     const T* __restrict__ uc, T* __restrict__ uf) {
  NodeRef n;
  if (!chunk_node(clf, nch, 1, n)) return;
  prolong_node(F, C, n, uc, uf);
}

// ---------------------------------------------------------------------------------------
// Skip-layer transfers between level 0 and a 4x coarser level 1: the factor-4 trilinear hat,
// weight prod (4 - |d|) / 4 over the offsets d of a fine node from a coarse node
// (restrict_residual_nondyadic_kernel grid/Grid.cu:2509-2570, prolongate_correction_
// nondyadic_kernel :2635-2671), coarse node X at fine node 4X.
// ---------------------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_restrict4(LevelView F, LevelView C, const int4* __restrict__ clc, int nch,
                                                   const T* __restrict__ rf, T* __restrict__ fc) {
  NodeRef n;
  if (!chunk_node(clc, nch, 1, n)) return;
  T a0 = 0, a1 = 0, a2 = 0;
  const int x0 = 4 * n.x, y0 = 4 * n.y, z0 = 4 * n.z;
  for (int dz = -3; dz <= 3; ++dz) {
    const int z = z0 + dz;
    if (z < 0 || z >= F.nz) continue;
    const T wz = T(4 - (dz < 0 ? -dz : dz)) * T(0.25);
    for (int dy = -3; dy <= 3; ++dy) {
      const int y = y0 + dy;
      if (y < 0 || y >= F.ny) continue;
      const T wzy = wz * (T(4 - (dy < 0 ? -dy : dy)) * T(0.25));
      const int64_t row = F.pnode(0, y, z);
#pragma unroll
      for (int dx = -3; dx <= 3; ++dx) {
        const int x = x0 + dx;
        if (x < 0 || x >= F.nx) continue;
        const T w = wzy * (T(4 - (dx < 0 ? -dx : dx)) * T(0.25));
        const int64_t q = row + x;
        a0 = fma(w, rf[q], a0);
        a1 = fma(w, rf[F.NP + q], a1);
        a2 = fma(w, rf[2 * F.NP + q], a2);
      }
    }
  }
  fc[n.p] = a0;
  fc[C.NP + n.p] = a1;
  fc[2 * C.NP + n.p] = a2;
}

template <class T>
__global__ __launch_bounds__(256) void k_prolong4(LevelView F, LevelView C, const int4* __restrict__ clf, int nch,
                                                  const T* __restrict__ uc, T* __restrict__ uf) {
  NodeRef n;
  if (!chunk_node(clf, nch, 1, n)) return;
  const int tx = n.x & 3, ty = n.y & 3, tz = n.z & 3;
  const T wx[2] = {T(4 - tx) * T(0.25), T(tx) * T(0.25)};
  const T wy[2] = {T(4 - ty) * T(0.25), T(ty) * T(0.25)};
  const T wz[2] = {T(4 - tz) * T(0.25), T(tz) * T(0.25)};
  // corners X0..X0+1 (X0 + 1 lies in the padding when the weight is 0)
  const int64_t I0 = C.pnode(n.x >> 2, n.y >> 2, n.z >> 2);
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = k & 1, j = (k >> 1) & 1, l = k >> 2;
    const T wgt = wx[i] * wy[j] * wz[l];
    const int64_t I = I0 + i + (int64_t)C.NX * (j + (int64_t)C.NY * l);
    a0 = fma(wgt, uc[I], a0);
    a1 = fma(wgt, uc[C.NP + I], a1);
    a2 = fma(wgt, uc[2 * C.NP + I], a2);
  }
  const int64_t p = n.p;
  uf[p] += a0;
  uf[F.NP + p] += a1;
  uf[2 * F.NP + p] += a2;
}

// weight of node i (bits x, y, z) of fine child c of a coarse element from its corner m
__device__ __forceinline__ double p4_weight(int c, int i, int m) {
  double w = 1.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int t = ((c >> (2 * a)) & 3) + ((i >> a) & 1);
    w *= ((m >> a) & 1) ? t * 0.25 : (4 - t) * 0.25;
  }
  return w;
}

// Galerkin operator of the skip-layer level 1 (restrict_stencil_nondyadic_OTFA_WS_kernel
// grid/Grid.cu:1112-1225): for coarse node I and block row r (blockIdx.y), the 27 blocks
// are summed over the coarse elements E around I (I at corner a of E) and their 64 fine
// children c: rho_c^p Mc4[c] rows 3a + r, columns (corner b, component s) into the block of
// the neighbour I + b - a.  A child with a supported node takes its element matrix with
// those rows and columns dropped and 1e6 on their diagonal entries (not scaled by rho^p,
// grid/Grid.cu:14, 1199-1207), explicitly.  Mc4's rows of corner a for all children are
// staged in LDS per corner.
__global__ __launch_bounds__(256) void k_galerkin_skip(LevelView F, LevelView C, const int4* __restrict__ cl,
                                                       int nch, const double* __restrict__ w64,
                                                       const uint8_t* __restrict__ efix,
                                                       const double* __restrict__ Ke,
                                                       const double* __restrict__ Mc4, double* __restrict__ st64,
                                                       float* __restrict__ st32, int64_t NS) {
  __shared__ double sM[64 * 24];
  const int r = blockIdx.y;
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 2, n);  // colour list: every other x
  double acc[81];
#pragma unroll
  for (int t = 0; t < 81; ++t) acc[t] = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    __syncthreads();
    for (int t = threadIdx.x; t < 64 * 24; t += 256) sM[t] = Mc4[(t / 24) * 576 + (3 * a + r) * 24 + t % 24];
    __syncthreads();
    if (!ok) continue;
    const int EX = n.x - (a & 1), EY = n.y - ((a >> 1) & 1), EZ = n.z - (a >> 2);
    if (EX < 0 || EY < 0 || EZ < 0 || EX >= C.ex || EY >= C.ey || EZ >= C.ez) continue;
    for (int c = 0; c < 64; ++c) {
      const int fx = 4 * EX + (c & 3), fy = 4 * EY + ((c >> 2) & 3), fz = 4 * EZ + (c >> 4);
      if (fx >= F.ex || fy >= F.ey || fz >= F.ez) continue;
      const int64_t pe = F.pelem(fx, fy, fz);
      const double w = w64[pe];  // rho^p, 0 for inactive elements
      if (w == 0.0) continue;
      const unsigned fm = efix[pe];
      if (fm == 0) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const int k = ((b & 1) - (a & 1) + 1) + 3 * (((b >> 1) & 1) - ((a >> 1) & 1) + 1) +
                        9 * ((b >> 2) - (a >> 2) + 1);
#pragma unroll
          for (int s2 = 0; s2 < 3; ++s2) acc[k * 3 + s2] = fma(w, sM[c * 24 + 3 * b + s2], acc[k * 3 + s2]);
        }
      } else {
        for (int i = 0; i < 8; ++i) {
          const double wa = p4_weight(c, i, a);
          if (wa == 0.0) continue;
          const bool fi = (fm >> i) & 1u;
          for (int j = 0; j < 8; ++j) {
            const bool fj = (fm >> j) & 1u;
            double v[3];
#pragma unroll
            for (int s2 = 0; s2 < 3; ++s2)
              v[s2] = (fi || fj) ? ((i == j && r == s2) ? 1e6 : 0.0) : w * Ke[(3 * i + r) * 24 + 3 * j + s2];
#pragma unroll
            for (int b = 0; b < 8; ++b) {
              const double wab = wa * p4_weight(c, j, b);
              const int k = ((b & 1) - (a & 1) + 1) + 3 * (((b >> 1) & 1) - ((a >> 1) & 1) + 1) +
                            9 * ((b >> 2) - (a >> 2) + 1);
#pragma unroll
              for (int s2 = 0; s2 < 3; ++s2) acc[k * 3 + s2] = fma(wab, v[s2], acc[k * 3 + s2]);
            }
          }
        }
      }
    }
  }
  if (!ok) return;
#pragma unroll
  for (int k = 0; k < 27; ++k)
#pragma unroll
    for (int s2 = 0; s2 < 3; ++s2) {
      const int64_t e = (int64_t)(k * 9 + r * 3 + s2) * NS + n.slot;
      if (st64) st64[e] = acc[k * 3 + s2];
      if (st32) st32[e] = (float)acc[k * 3 + s2];
    }
}

// u_f += P u_c over row segments (even-x first, all rows): lane l takes the fine nodes
// x1 = xs + 2l (coarse column X = x1/2, weight 1 along x) and x1 + 1 (columns X and X+1,
// 1/2 each), so one wave reads each coarse column once for two fine nodes.  Lane 0 of a
// continuation segment repeats the previous segment's lane 63 and does not write.
template <class T>
__device__ __forceinline__ void prolong_pairw_lane(const LevelView& F, const LevelView& C, const int4 S, int lane,
                                                   const uint8_t* __restrict__ fre, const T* __restrict__ uc,
                                                   T* __restrict__ uf) {
  const int xs = (S.y & 0x7fff) - 1, xe = ((S.y >> 15) & 0x7fff) - 1;
  const bool cont = (S.y >> 30) & 1;
  const int x1 = xs + 2 * lane, y = S.z, z = S.w;
  const bool live = x1 <= xe + 1;
  const int xa = live ? x1 : xe + 1;  // (even: the segments start on even x)
  const int64_t pa = (int64_t)S.x + (xa - xs);
  const bool f1 = live && !(cont && lane == 0) && fre[pa] != 0;
  const bool f2 = live && lane < 63 && x1 + 1 <= xe && fre[pa + 1] != 0;
  const T hy = (y & 1) ? T(0.5) : T(0), hz = (z & 1) ? T(0.5) : T(0);
  const T wy[2] = {T(1) - hy, hy}, wz[2] = {T(1) - hz, hz};
  const int64_t I0 = C.pnode(xa >> 1, y >> 1, z >> 1);
  T a[2][3] = {{0, 0, 0}, {0, 0, 0}};  // columns X and X+1
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = k & 1, j = (k >> 1) & 1, l = k >> 2;
    const T wgt = wy[j] * wz[l];
    const int64_t I = I0 + i + (int64_t)C.NX * (j + (int64_t)C.NY * l);
#pragma unroll
    for (int c = 0; c < 3; ++c) a[i][c] = fma(wgt, uc[c * C.NP + I], a[i][c]);
  }
  T g1[3], g2[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    g1[c] = uf[c * F.NP + pa];
    g2[c] = uf[c * F.NP + pa + 1];
  }
  if (f1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) uf[c * F.NP + pa] = g1[c] + a[0][c];
  }
  if (f2) {
#pragma unroll
    for (int c = 0; c < 3; ++c) uf[c * F.NP + pa + 1] = g2[c] + T(0.5) * (a[0][c] + a[1][c]);
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_prolong_add_pairw(LevelView F, LevelView C, const int4* __restrict__ seg,
                                                           int nseg, const uint8_t* __restrict__ fre,
                                                           const T* __restrict__ uc, T* __restrict__ uf) {
  const int si = (int)xcd_block() * 4 + (threadIdx.x >> 6);
  if (si >= nseg) return;
  prolong_pairw_lane(F, C, seg[si], threadIdx.x & 63, fre, uc, uf);
}

template <class T>
__global__ __launch_bounds__(256) void k_zero_nodes(const int4* __restrict__ cl, int nch, int64_t NP, T* __restrict__ u) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
  u[n.p] = T(0); u[NP + n.p] = T(0); u[2 * NP + n.p] = T(0);
}

template <class A, class B>
__global__ __launch_bounds__(256) void k_cast_nodes(const int4* __restrict__ cl, int nch, int64_t NP,
                                                    const A* __restrict__ a, B* __restrict__ b) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
  b[n.p] = (B)a[n.p]; b[NP + n.p] = (B)a[NP + n.p]; b[2 * NP + n.p] = (B)a[2 * NP + n.p];
}

// ---------------------------------------------------------------------------------------
// Galerkin coarse operators, the coarse nodes taken from the colour-ordered chunk list,
// output in the compact layout.
// First coarse level: the coarse element matrices are sum over active children of
// rho^p P_c^T Ke P_c (8 precomputed 24x24 products Mc), evaluated with the eliminated
// rows/columns when a child touches a non-free fine node (allfree[e] = 0)
// (restrict_stencil_nondyadic_OTFA_WS_kernel grid/Grid.cu:1112-1225 computes the same
// triple product per coarse node directly from the fine densities).
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

// the two blocks (jx = 0, 1) of one child whose element has an eliminated fine node:
// sum over its free fine nodes i, j of we P[i, mI] P[j, n] Ke[i][j]
struct GalPair {
  double v[2][9];
};
__device__ __noinline__ GalPair gal_row_partial(const LevelView F, const uint8_t* __restrict__ ffre,
                                                const double* __restrict__ Ke, int fx, int fy, int fz, int ch,
                                                int mI, int jy, int jz, double we) {
  GalPair t;
  for (int jx = 0; jx < 2; ++jx)
    for (int e = 0; e < 9; ++e) t.v[jx][e] = 0;
  const int64_t fb = F.pnode(fx, fy, fz);
  bool fr[8];
  for (int i = 0; i < 8; ++i)
    fr[i] = ffre[fb + (i & 1) + (int64_t)F.NX * (((i >> 1) & 1) + (int64_t)F.NY * (i >> 2))];
#pragma unroll
  for (int jx = 0; jx < 2; ++jx) {
    const int n = jx + 2 * jy + 4 * jz;
    for (int i = 0; i < 8; ++i) {
      if (!fr[i]) continue;
      const double pi = pw(ch, i, mI);
      if (pi == 0) continue;
      for (int j = 0; j < 8; ++j) {
        if (!fr[j]) continue;
        const double pj = pw(ch, j, n);
        if (pj == 0) continue;
        const double sc = we * pi * pj;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) t.v[jx][r * 3 + c] += sc * Ke[(3 * i + r) * 24 + 3 * j + c];
      }
    }
  }
  return t;
}

template <int AX>
__device__ __forceinline__ void gal_elem_row(const LevelView& F, const LevelView& C, int X, int Y, int Z, int ay,
                                             int az, int jy, int jz, const uint8_t* __restrict__ ffre,
                                             const uint8_t* __restrict__ fallfree, const double* __restrict__ w,
                                             const double* __restrict__ Ke, const double* __restrict__ Mc,
                                             double (&acc)[3][9]) {
  const int Ex = X - 1 + AX, Ey = Y - 1 + ay, Ez = Z - 1 + az;
  if (Ex < 0 || Ey < 0 || Ez < 0 || Ex >= C.ex || Ey >= C.ey || Ez >= C.ez) return;
  const int mI = 7 - (AX + 2 * ay + 4 * az);
  // the eight children's weights and flags are loaded together before any is used
  double wc[8];
  uint8_t af[8];
#pragma unroll
  for (int ch = 0; ch < 8; ++ch) {
    const int fx = 2 * Ex + (ch & 1), fy = 2 * Ey + ((ch >> 1) & 1), fz = 2 * Ez + (ch >> 2);
    const bool in = fx < F.ex && fy < F.ey && fz < F.ez;
    const int64_t fe = F.pelem(in ? fx : 0, in ? fy : 0, in ? fz : 0);
    wc[ch] = w[fe];
    af[ch] = fallfree[fe];
    if (!in) wc[ch] = 0;
  }
#pragma unroll
  for (int ch = 0; ch < 8; ++ch) {
    const double we = wc[ch];
    if (we == 0) continue;
    if (af[ch]) {
      const double* m0 = Mc + ch * 576 + (3 * mI) * 24 + 3 * (2 * jy + 4 * jz);
#pragma unroll
      for (int jx = 0; jx < 2; ++jx) {
        const double* m = m0 + 3 * jx;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[jx + AX][r * 3 + c] += we * m[r * 24 + c];
      }
    } else {  // a fine node of the child is eliminated: rare, out of line
      const int fx = 2 * Ex + (ch & 1), fy = 2 * Ey + ((ch >> 1) & 1), fz = 2 * Ez + (ch >> 2);
      const GalPair t = gal_row_partial(F, ffre, Ke, fx, fy, fz, ch, mI, jy, jz, we);
#pragma unroll
      for (int jx = 0; jx < 2; ++jx)
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[jx + AX][e] += t.v[jx][e];
    }
  }
}

__global__ __launch_bounds__(256) void k_galerkin_first_row(LevelView F, LevelView C, const int4* __restrict__ clc,
                                                            int nch, const uint8_t* __restrict__ ffre,
                                                            const uint8_t* __restrict__ fallfree,
                                                            const uint8_t* __restrict__ cfre,
                                                            const double* __restrict__ w,
                                                            const double* __restrict__ Ke,
                                                            const double* __restrict__ Mc, double* __restrict__ st,
                                                            float* __restrict__ st32, int64_t NS) {
  NodeRef nr;
  if (!chunk_node(clc, nch, 2, nr)) return;
  const int dy = (int)blockIdx.y % 3 - 1, dz = (int)blockIdx.y / 3 - 1;
  const int X = nr.x, Y = nr.y, Z = nr.z;
  double acc[3][9];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[b][e] = 0;
  for (int az = 0; az < 2; ++az) {  // elements in a order: ax inner, then ay, then az
    const int jz = 1 + dz - az;
    if (jz < 0 || jz > 1) continue;
    for (int ay = 0; ay < 2; ++ay) {
      const int jy = 1 + dy - ay;
      if (jy < 0 || jy > 1) continue;
      gal_elem_row<0>(F, C, X, Y, Z, ay, az, jy, jz, ffre, fallfree, w, Ke, Mc, acc);
      gal_elem_row<1>(F, C, X, Y, Z, ay, az, jy, jz, ffre, fallfree, w, Ke, Mc, acc);
    }
  }
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const int dx = b - 1;
    const int k = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
    const bool keep = cfre[nr.p + dx + (int64_t)C.NX * (dy + (int64_t)C.NY * dz)] != 0;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const double v = keep ? acc[b][e] : 0.0;
      if (st) st[(int64_t)(k * 9 + e) * NS + nr.slot] = v;
      if (st32) st32[(int64_t)(k * 9 + e) * NS + nr.slot] = (float)v;
    }
  }
}

// ---------------------------------------------------------------------------------------
// Galerkin coarse operator by axes.  Trilinear P is the tensor product of 1-D node-aligned
// interpolations, so P^T A P = (Pz Py Px)^T A (Px Py Pz) is formed in three passes, each
// halving one axis of a 27-point 3x3-block stencil:
//   A'[I][dX] = sum_{a,b in {-1,0,1}, kin = 2 dX + b - a in {-1,0,1}} w(a) w(b) A[2I + a][kin]
// (w(0) = 1, w(+-1) = 1/2, offsets along the other axes unchanged).  The fine stencil is
// zero on eliminated rows and towards eliminated columns, so the staged product equals
// the masked one; the last pass drops coarse columns that are not free.  Every pass reads
// each input stencil a few times instead of the ~13^3 / 27 times of the direct sum.
// Input: the fine level's compact stencil (IN_CMP, slot map) or an intermediate box;
// output: an intermediate box (only the nodes `omask` marks as read by the next pass:
// the box spans the whole grid, the domain often a third of it) or the coarse level's
// compact stencil (OUT_CMP; with an FP64 pipeline optionally also its FP32 copy).
// TS = storage type of input, boxes and output (double, or float for the FP32-stored
// operator of the mixed-precision cycles); sums are FP64 either way.
// Thread = (output node, output neighbour k = blockIdx.y).
// ---------------------------------------------------------------------------------------
struct GBox {
  int X, Y, Z;
};

template <int AX, bool IN_CMP, bool OUT_CMP, class TS>
__global__ __launch_bounds__(256) void k_gal_axis(LevelView F, const int* __restrict__ fslot,
                                                  const TS* __restrict__ sin, int64_t NSI, GBox bi, GBox bo,
                                                  LevelView C, const int4* __restrict__ clc, int nch,
                                                  const uint8_t* __restrict__ cfre, TS* __restrict__ sout,
                                                  float* __restrict__ sout32, int64_t NSO,
                                                  const uint8_t* __restrict__ omask) {
  int o[3];
  int64_t oidx;
  const int k = blockIdx.y;
  const int d[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
  bool keep = true;
  if (OUT_CMP) {
    NodeRef nr;
    if (!chunk_node(clc, nch, 2, nr)) return;
    o[0] = nr.x; o[1] = nr.y; o[2] = nr.z;
    oidx = nr.slot;
    keep = cfre[nr.p + d[0] + (int64_t)C.NX * (d[1] + (int64_t)C.NY * d[2])] != 0;
  } else {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (int64_t)bo.X * bo.Y * bo.Z || !omask[v]) return;  // (entries no later pass reads)
    o[0] = (int)(v % bo.X);
    o[1] = (int)((v / bo.X) % bo.Y);
    o[2] = (int)(v / ((int64_t)bo.X * bo.Y));
    oidx = v;
  }
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int nin = AX == 0 ? bi.X : (AX == 1 ? bi.Y : bi.Z);
  if (keep) {
    for (int a = -1; a <= 1; ++a) {
      int q[3] = {o[0], o[1], o[2]};
      q[AX] = 2 * o[AX] + a;
      if (q[AX] < 0 || q[AX] >= nin) continue;
      int64_t src;
      if (IN_CMP) {
        const int sl = fslot[F.pnode(q[0], q[1], q[2])];
        if (sl < 0) continue;  // eliminated fine node: zero row
        src = sl;
      } else {
        src = q[0] + (int64_t)bi.X * (q[1] + (int64_t)bi.Y * q[2]);
      }
      const double wa = a ? 0.5 : 1.0;
      for (int b = -1; b <= 1; ++b) {
        const int kin = 2 * d[AX] + b - a;
        if (kin < -1 || kin > 1) continue;
        const int bq = 2 * (o[AX] + d[AX]) + b;
        if (bq < 0 || bq >= nin) continue;
        int e3[3] = {d[0], d[1], d[2]};
        e3[AX] = kin;
        const int ein = (e3[0] + 1) + 3 * (e3[1] + 1) + 9 * (e3[2] + 1);
        const double wgt = wa * (b ? 0.5 : 1.0);
        const TS* sp = sin + (int64_t)ein * 9 * NSI + src;
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] += wgt * (double)sp[e * NSI];
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    sout[(int64_t)(k * 9 + e) * NSO + oidx] = (TS)acc[e];
    if (OUT_CMP && sout32) sout32[(int64_t)(k * 9 + e) * NSO + oidx] = (float)acc[e];
  }
}


// ---------------------------------------------------------------------------------------
// Coarsest level: dense matrix over the free nodes, then an in-place blocked symmetric
// sweep (Gauss-Jordan) that leaves -A^{-1}; the V-cycle applies it as a dense matvec.
// ---------------------------------------------------------------------------------------
template <class TS>
__global__ void k_coarse_assemble(LevelView L, const int64_t* __restrict__ cidx, int nnode,
                                  const int* __restrict__ cmap_pad, const int* __restrict__ slot,
                                  const TS* __restrict__ st, int64_t NS, double* __restrict__ A, int n) {
  const int64_t t = thread_id();
  if (t >= (int64_t)nnode * 27) return;
  const int pnode = t / 27, k = t % 27;
  const int64_t i = cidx[pnode];
  const int64_t j = i + (k % 3 - 1) + (int64_t)L.NX * ((k / 3) % 3 - 1 + (int64_t)L.NY * (k / 9 - 1));
  const int q = cmap_pad[j];
  if (q < 0) return;
  const int si = slot[i];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[(int64_t)(3 * pnode + r) * n + 3 * q + c] = (double)st[(int64_t)(k * 9 + r * 3 + c) * NS + si];
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

// Invert the diagonal block K (size b <= 32) by a scalar sweep, one wave: lane l keeps
// the elements (i, j) = ((l >> 5) + 2 m, l & 31), m < 16, in registers and publishes them
// in LDS after every step, where the next step reads its pivot row and column (one wave
// issues its LDS operations in order, so no barrier is needed).
// A pivot (Schur complement) below 1e-12 of the unswept diagonal marks the system as
// singular -- a floating part or a structure without supports.
__global__ __launch_bounds__(64) void k_gj_pivot(double* __restrict__ A, int n, int K0, int b,
                                                 double* __restrict__ Dinv, const double* __restrict__ diag0,
                                                 int* __restrict__ flag) {
  constexpr int M = kGJB / 2;
  __shared__ double S[kGJB][kGJB + 1];
  const int l = threadIdx.x, j = l & 31, i0 = l >> 5;
  double v[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = i0 + 2 * m;
    v[m] = (i < b && j < b) ? A[(int64_t)(K0 + i) * n + K0 + j] : 0.0;
    S[i][j] = v[m];
  }
  __builtin_amdgcn_wave_barrier();
  for (int k = 0; k < b; ++k) {
    const double piv = S[k][k];  // same LDS word for every lane: the branch is uniform
    if (!(piv > 1e-12 * diag0[K0 + k])) {
      if (l == 0) *flag = 1;
      return;
    }
    const double dd = 1.0 / piv;
    const double skj = S[k][j];
    double sik[M];
#pragma unroll
    for (int m = 0; m < M; ++m) sik[m] = S[i0 + 2 * m][k];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = i0 + 2 * m;
      double w = v[m];
      if (i != k && j != k) w = w - sik[m] * skj * dd;
      else if (i == k && j == k) w = -dd;
      else w = w * dd;
      v[m] = w;
      S[i][j] = w;
    }
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = i0 + 2 * m;
    if (i < b && j < b) Dinv[i * b + j] = -v[m];  // +A_KK^{-1}
  }
}

// C = A[:, K] * Dinv  (n x b), and R = A[K, :] (b x n) before the update overwrites it
__global__ void k_gj_panel(const double* __restrict__ A, int n, int K0, int b, const double* __restrict__ Dinv,
                           double* __restrict__ Cp, double* __restrict__ R) {
  const int64_t t = thread_id();
  if (t >= (int64_t)n * b) return;
  const int i = t / b, j = t % b;
  double s = 0;
  for (int q = 0; q < b; ++q) s += A[(int64_t)i * n + K0 + q] * Dinv[q * b + j];
  Cp[t] = s;
  R[(int64_t)j * n + i] = A[(int64_t)(K0 + j) * n + i];
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
// PCG vector kernels over the fine level's chunk list (FP64, fused with their reductions)
// ---------------------------------------------------------------------------------------
// u += alpha p ; r -= alpha q ; partial ||r||^2 ; optional float copy of r
// (p in FP64, or FP32 in the mixed-precision solves)
template <class TP>
__global__ __launch_bounds__(256) void k_pcg_ur(const int4* __restrict__ cl, int nch, int64_t NP,
                                                double* __restrict__ u, double* __restrict__ r,
                                                const TP* __restrict__ p, const double* __restrict__ q,
                                                const double* __restrict__ scal, float* __restrict__ r32,
                                                double* __restrict__ partial) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double s = 0;
  if (ok) {
    // every load before the first store: the stores to u and r could alias the later
    // components' loads as far as the compiler knows, which would serialise them
    const double al = scal[S_ALPHA];
    double uv[3], pv[3], rv[3], qv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = c * NP + n.p;
      uv[c] = u[i]; pv[c] = (double)p[i]; rv[c] = r[i]; qv[c] = q[i];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = c * NP + n.p;
      const double rn = rv[c] - al * qv[c];
      u[i] = uv[c] + al * pv[c];
      r[i] = rn;
      if (r32) r32[i] = (float)rn;
      s += rn * rn;
    }
  }
  block_partial(s, partial, 0);
}

// r -= alpha q ; float copy of r ; partial ||r||^2 (the u update is deferred to k_pcg_up)
__global__ __launch_bounds__(256) void k_pcg_r(const int4* __restrict__ cl, int nch, int64_t NP,
                                               double* __restrict__ r, const double* __restrict__ q,
                                               const double* __restrict__ scal, float* __restrict__ r32,
                                               double* __restrict__ partial) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double s = 0;
  if (ok) {
    const double al = scal[S_ALPHA];
    double rv[3], qv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rv[c] = r[c * NP + n.p];
      qv[c] = q[c * NP + n.p];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double rn = rv[c] - al * qv[c];
      r[c * NP + n.p] = rn;
      r32[c * NP + n.p] = (float)rn;
      s += rn * rn;
    }
  }
  block_partial(s, partial, 0);
}

// FCG with FP32 residual and K p (solve precision 3): r -= alpha q (FP64 arithmetic), partial ||r||^2
__global__ __launch_bounds__(256) void k_pcg_r32(const int4* __restrict__ cl, int nch, int64_t NP,
                                                 float* __restrict__ r, const float* __restrict__ q,
                                                 const double* __restrict__ scal, double* __restrict__ partial) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double s = 0;
  if (ok) {
    const double al = scal[S_ALPHA];
    double rv[3], qv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rv[c] = (double)r[c * NP + n.p];
      qv[c] = (double)q[c * NP + n.p];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float rn = (float)(rv[c] - al * qv[c]);
      r[c * NP + n.p] = rn;
      s += (double)rn * (double)rn;
    }
  }
  block_partial(s, partial, 0);
}

// partials of z.r and z.q on FP32 vectors (q may be null)
__global__ __launch_bounds__(256) void k_pcg_zdots32(const int4* __restrict__ cl, int nch, int64_t NP,
                                                     const float* __restrict__ z, const float* __restrict__ r,
                                                     const float* __restrict__ q, double* __restrict__ partial) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double zr = 0, zq = 0;
  if (ok) {
    double zv[3], rv[3], qv[3] = {0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = c * NP + n.p;
      zv[c] = (double)z[i];
      rv[c] = (double)r[i];
      if (q) qv[c] = (double)q[i];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      zr += zv[c] * rv[c];
      zq += zv[c] * qv[c];
    }
  }
  block_partial(zr, partial, 0);
  block_partial(zq, partial, 1);
}

// u += alpha p (the previous iteration's step, unless first) ; p = z + beta p (first: p = z)
template <class TZ, class TP>
__global__ __launch_bounds__(256) void k_pcg_up(const int4* __restrict__ cl, int nch, int64_t NP,
                                                double* __restrict__ u, TP* __restrict__ p,
                                                const TZ* __restrict__ z, const double* __restrict__ scal,
                                                int first) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
  const double al = first ? 0.0 : scal[S_ALPHA], be = first ? 0.0 : scal[S_BETA];
  double zv[3], pv[3] = {0, 0, 0}, uv[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t i = c * NP + n.p;
    zv[c] = (double)z[i];
    if (!first) {
      pv[c] = (double)p[i];
      uv[c] = u[i];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t i = c * NP + n.p;
    if (!first) u[i] = uv[c] + al * pv[c];
    p[i] = (TP)(zv[c] + be * pv[c]);
  }
}

// Precision-3 FCG with four direction buffers: p_j goes to buffer j % 4, and u takes the
// steps alpha_k p_k of four directions in one pass (k_pcg_flush) instead of one read and
// write of the FP64 u per iteration.
// p_new = z + beta p_old
__global__ __launch_bounds__(256) void k_pcg_pnew(const int4* __restrict__ cl, int nch, int64_t NP,
                                                  float* __restrict__ pn, const float* __restrict__ z,
                                                  const float* __restrict__ po, const double* __restrict__ scal) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
  const double be = scal[S_BETA];
  double zv[3], pv[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t i = c * NP + n.p;
    zv[c] = (double)z[i];
    pv[c] = (double)po[i];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) pn[c * NP + n.p] = (float)(zv[c] + be * pv[c]);
}

// u += alpha_k p_k for the m directions of D in order (the same rounding per step as
// one step per iteration)
struct PcgDirs {
  const float* p[4];
  int aslot[4];  // scal slot of each direction's alpha
};
__global__ __launch_bounds__(256) void k_pcg_flush(const int4* __restrict__ cl, int nch, int64_t NP,
                                                   double* __restrict__ u, PcgDirs D, int m,
                                                   const double* __restrict__ scal) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
  double al[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) al[k] = k < m ? scal[D.aslot[k]] : 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t i = c * NP + n.p;
    double uv = u[i];
    float pv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) pv[k] = k < m ? D.p[k][i] : 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < m) uv = uv + al[k] * (double)pv[k];
    u[i] = uv;
  }
}

// u += alpha p (the last deferred step of a solve)
template <class TP>
__global__ __launch_bounds__(256) void k_pcg_ustep(const int4* __restrict__ cl, int nch, int64_t NP,
                                                   double* __restrict__ u, const TP* __restrict__ p,
                                                   const double* __restrict__ scal) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
  const double al = scal[S_ALPHA];
  double uv[3], pv[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uv[c] = u[c * NP + n.p];
    pv[c] = (double)p[c * NP + n.p];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) u[c * NP + n.p] = uv[c] + al * pv[c];
}

// partials of z.r and z.q (z in FP32 or FP64)
template <class TZ>
__global__ __launch_bounds__(256) void k_pcg_zdots(const int4* __restrict__ cl, int nch, int64_t NP,
                                                   const TZ* __restrict__ z, const double* __restrict__ r,
                                                   const double* __restrict__ q, double* __restrict__ partial) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double zr = 0, zq = 0;
  if (ok) {
    double zv[3], rv[3], qv[3] = {0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = c * NP + n.p;
      zv[c] = (double)z[i];
      rv[c] = r[i];
      if (q) qv[c] = q[i];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      zr += zv[c] * rv[c];
      zq += zv[c] * qv[c];
    }
  }
  block_partial(zr, partial, 0);
  block_partial(zq, partial, 1);
}

// p = z + beta p  (first != 0: p = z), formed in FP64 and stored as TP
template <class TZ, class TP>
__global__ __launch_bounds__(256) void k_pcg_p(const int4* __restrict__ cl, int nch, int64_t NP,
                                               TP* __restrict__ p, const TZ* __restrict__ z,
                                               const double* __restrict__ scal, int first) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
  const double be = first ? 0.0 : scal[S_BETA];
  double zv[3], pv[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t i = c * NP + n.p;
    zv[c] = (double)z[i];
    if (!first) pv[c] = (double)p[i];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) p[c * NP + n.p] = (TP)(zv[c] + be * pv[c]);
}

// ---- recycled start: Galerkin projection of the warm start onto earlier solutions ----
constexpr int kRcMax = 8;  // basis vectors: the current displacement and up to 7 earlier solutions
struct RcVecs {
  const double* v[kRcMax];   // basis
  const double* kv[kRcMax];  // K times the basis
};
struct RcCoef {
  double c[kRcMax];
};

// partial sums over the free nodes (grid-stride over the chunk list, gridDim.x partials per
// column): columns 0 .. M(M+1)/2-1 = v_a . K v_b for a <= b (row by row), then v_a . f
template <int M>
__global__ __launch_bounds__(256) void k_rc_gram(const int4* __restrict__ cl, int nch, int64_t NP, RcVecs V,
                                                 const double* __restrict__ f, double* __restrict__ partial) {
  constexpr int NG = M * (M + 1) / 2 + M;
  double g[NG];
#pragma unroll
  for (int i = 0; i < NG; ++i) g[i] = 0;
  const int lane = threadIdx.x & 63;
  for (int ch = blockIdx.x * 4 + (threadIdx.x >> 6); ch < nch; ch += gridDim.x * 4) {
    NodeRef n;
    if (!chunk_node_at(cl, nch, ch, lane, 1, n)) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = c * NP + n.p;
      double v[M], kv[M];
#pragma unroll
      for (int a = 0; a < M; ++a) {
        v[a] = V.v[a][i];
        kv[a] = V.kv[a][i];
      }
      const double fv = f[i];
      int q = 0;
#pragma unroll
      for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = a; b < M; ++b) g[q++] += v[a] * kv[b];
#pragma unroll
      for (int a = 0; a < M; ++a) g[q++] += v[a] * fv;
    }
  }
#pragma unroll
  for (int q = 0; q < NG; ++q) block_partial(g[q], partial, q);
}

// u = sum_a c_a v_a on the free nodes (v_0 may be u itself), and the solve's first
// residual r = f - sum_a c_a (K v_a) from the products the Gram pass read (K u by
// linearity, without another operator application), with its FP32 copy when r32 is set
template <int M>
__global__ __launch_bounds__(256) void k_rc_combine(const int4* __restrict__ cl, int nch, int64_t NP, RcVecs V,
                                                    RcCoef C, const double* __restrict__ f, double* __restrict__ u,
                                                    double* __restrict__ r, float* __restrict__ r32) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t i = c * NP + n.p;
    double s = 0, t = 0;
#pragma unroll
    for (int a = 0; a < M; ++a) {
      s = fma(C.c[a], V.v[a][i], s);
      t = fma(C.c[a], V.kv[a][i], t);
    }
    u[i] = s;
    const double rv = f[i] - t;
    r[i] = rv;
    if (r32) r32[i] = (float)rv;
  }
}

// dst = src on the free nodes
__global__ __launch_bounds__(256) void k_copy_nodes(const int4* __restrict__ cl, int nch, int64_t NP,
                                                    const double* __restrict__ src, double* __restrict__ dst) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) dst[c * NP + n.p] = src[c * NP + n.p];
}

// generic partials of a.b (and a2.b2 when a2 != null)
__global__ __launch_bounds__(256) void k_dot_nodes(const int4* __restrict__ cl, int nch, int64_t NP,
                                                   const double* __restrict__ a, const double* __restrict__ b,
                                                   const double* __restrict__ a2, const double* __restrict__ b2,
                                                   double* __restrict__ partial) {
  NodeRef n;
  const bool ok = chunk_node(cl, nch, 1, n);
  double s0 = 0, s1 = 0;
  if (ok) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = c * NP + n.p;
      s0 += a[i] * b[i];
      if (a2) s1 += a2[i] * b2[i];
    }
  }
  block_partial(s0, partial, 0);
  block_partial(s1, partial, 1);
}

// u += z (plain V-cycle iteration); z FP64, or the FP32 cycle's output read directly
template <class TZ>
__global__ __launch_bounds__(256) void k_add_nodes(const int4* __restrict__ cl, int nch, int64_t NP,
                                                   double* __restrict__ u, const TZ* __restrict__ z) {
  NodeRef n;
  if (!chunk_node(cl, nch, 1, n)) return;
  const double z0 = (double)z[n.p], z1 = (double)z[NP + n.p], z2 = (double)z[2 * NP + n.p];
  u[n.p] += z0; u[NP + n.p] += z1; u[2 * NP + n.p] += z2;
}

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
