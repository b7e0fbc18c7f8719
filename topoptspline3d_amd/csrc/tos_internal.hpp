// tos_internal.hpp -- device layout and context of the gfx950 hot path.
//
// Layout in HBM (per multigrid level):
//   nodes    padded box (nx+3) x (ny+3) x (nz+3), node (x,y,z) stored at (x+1,y+1,z+1),
//            x fastest; one ghost layer below and two above, so every 27-point
//            neighbourhood and every fine node 2I+-1 of a coarse node I addresses memory
//            inside the array.  The component stride NP is a multiple of 64.
//   vectors  3 components, structure of arrays, comp c at v[c*NP + p].  Ghost and
//            eliminated (supported / inactive / unreachable) nodes hold 0 in every
//            vector, so no kernel reads a mask for a neighbour.
//   elements padded box (ex+2) x (ey+2) x (ez+2), element (ex,ey,ez) at (ex+1,...);
//            ghost elements have weight 0.
//   stencil  Galerkin 27-point 3x3 blocks st[(k*9 + r*3 + c)*NP + p], k = (dx+1) +
//            3(dy+1) + 9(dz+1) (the reference's rxStencil order, grid/Grid.cu:847).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tos_api.h"

namespace tos {

struct LearnState;  // learned coarse-level correction (network, pair logging)
struct BatchState;  // K load cases solved together (interleaved vectors, graph, tuning)

// Index of entry (r, s) of block Ke[0][d] among the 14 distinct values of the first block
// row of the isotropic hex stiffness: diagonal entries depend on bit r of d and on the
// number of other set bits, off-diagonal entries on bits r, s and the third axis of d.
__host__ __device__ constexpr int ke_key(int d, int r, int s) {
  return r == s ? ((d >> r) & 1) * 3 + ((d >> ((r + 1) % 3)) & 1) + ((d >> ((r + 2) % 3)) & 1)
                : 6 + ((d >> r) & 1) * 4 + ((d >> s) & 1) * 2 + ((d >> (3 - r - s)) & 1);
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b -> XCD b % 8) and each XCD
// has its own L2.  Renumbering blocks so that XCD x gets a contiguous 1/8 of the logical
// block range keeps neighbouring grid rows (and their shared stencil data) in one L2.
// A bijection on [0, G): the first 8*floor(G/8) blocks are transposed, the tail kept.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t b = blockIdx.x, G = gridDim.x, G8 = G & ~(int64_t)7;
  return b < G8 ? (b & 7) * (G8 >> 3) + (b >> 3) : b;
}

// up to 8 chunk ranges of one launch (pieces of a colour-slab wavefront step)
struct SegTab {
  int n;
  int beg[8];      // first chunk of the range in the colour-ordered list
  int cnt[8];      // chunks in the range
  int blk[9];      // first block of each range (4 chunks per block), blk[n] = total
};

struct LevelView {  // passed by value to kernels
  int nx, ny, nz;     // node dims
  int NX, NY;         // padded node row / plane dims
  int64_t NP;         // padded node count (vector component stride)
  int ex, ey, ez;     // element dims
  int EX, EY;         // padded element row / plane dims
  int64_t NE;         // padded element count
  __host__ __device__ int64_t pnode(int x, int y, int z) const {
    return (int64_t)(x + 1) + (int64_t)NX * ((y + 1) + (int64_t)NY * (z + 1));
  }
  __host__ __device__ int64_t pelem(int x, int y, int z) const {
    return (int64_t)(x + 1) + (int64_t)EX * ((y + 1) + (int64_t)EY * (z + 1));
  }
};

struct Level {
  LevelView v;
  int64_t n_active = 0, n_free = 0;
  std::vector<uint8_t> h_eact;   // host copy (unpadded) of element activity
  std::vector<uint8_t> h_free;   // host copy (unpadded) of node freedom
  uint8_t* fre = nullptr;        // device, padded nodes
  double* st64 = nullptr;        // device stencil (levels >= 1)
  float* st32 = nullptr;
  uint32_t* st16 = nullptr;     // FP16 copy of st32 (122 words per slot) and its per-slot scale
  float* st_m = nullptr;
  double *u64 = nullptr, *f64 = nullptr, *r64 = nullptr;
  float *u32 = nullptr, *f32 = nullptr, *r32 = nullptr;
  // chunk lists, int4 {padded index of first node, x0 | count << 16, y, z}: `lin` holds
  // x-contiguous runs of free nodes, `col` the runs of each colour (x stride 2),
  // concatenated in colour order with col_off[c] the first chunk of colour c
  int4* cl_lin = nullptr;
  int n_lin = 0;
  int4* cl_col = nullptr;
  int n_col = 0;
  int col_off[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // colour-slab wavefront of the GS sweep: z-slabs of wf_T node planes, slab_off[c][s]
  // = first chunk (in the colour-ordered list) of colour c with z >= s * wf_T
  int wf_T = 0, wf_S = 0;
  std::vector<int> slab_off;    // 8 x (wf_S + 1)
  // grid rows of the (y%2, z%2) classes for the two-colour GS launches (fine level):
  // fwd {first chunk of the even-x colour, count, first chunk of the odd-x colour, count},
  // bwd the same with the two colours swapped
  int4* pair_fwd[4] = {nullptr, nullptr, nullptr, nullptr};
  int4* pair_bwd[4] = {nullptr, nullptr, nullptr, nullptr};
  int n_pair[4] = {0, 0, 0, 0};
  // row segments of the in-wave two-colour GS (k_fine_gs_pairw), per class and first-colour
  // x parity (0: forward, 1: backward), and the buffer of its deferred nodes
  int4* segw[4][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  int n_segw[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  float* ub32 = nullptr;
  bool pair_ok = false;          // stencil level: row segments built (k_st_gs_pairw)
  bool pair_single = false;      // ... and every row is one segment (no deferred nodes)
  void* ubuf = nullptr;          // deferred nodes of the stencil-level paired GS (3 NP doubles)
  int4* segp = nullptr;          // even-x segments of all rows (paired prolongation into this level)
  int n_segp = 0;
  int64_t NS = 0;                // stencil slots per entry = 64 * n_col
  int* slot = nullptr;           // padded node -> stencil slot, -1 if not free
  std::vector<int> h_slot;
  double* gal_b1 = nullptr;      // Galerkin intermediates of this (coarse) level, box layout
  double* gal_b2 = nullptr;
  uint8_t* gal_m1 = nullptr;     // which intermediate nodes a later pass reads (the rest are skipped)
  uint8_t* gal_m2 = nullptr;
};

// Projector onto the rigid motions of a node set: columns t_x, t_y, t_z, (0, z, -y),
// (-z, 0, x), (y, -x, 0) at node coordinates relative to o (the columns of [I | p^],
// optimization/projection.cpp:70-77, 212-219), optionally each node's 3-vector taken
// along its normal first.  With R those columns over the set and b = R^T v, the
// projection of v is R (M b); M is built on the host from the Gram matrix R^T R by
// Gram-Schmidt over the columns in order, dropping a column whose residual norm is below
// 1e-6 of its own norm or whose norm is below 1e-6 (projection.cpp:234-255).
struct RigidProj {
  double M[36];
  double o[3];
  int m;  // columns kept
};

struct Ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority): the FP16 stencil copies of a new operator, packed
  // while the coarsest level is factored on `stream` (ev_gal: stencils built, ev_pack: packed)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_gal = nullptr, ev_pack = nullptr;
  bool want16 = false;  // the last solve read FP16 stencils (pack them with the next operator)
  std::string err;
  int nx, ny, nz;
  double h;
  double E = 1e6, nu = 0.4, penal = 3.0, rho_min = 1e-3;
  std::vector<Level> L;
  std::vector<uint8_t> h_solid, h_support, h_shell;
  int64_t n_active = 0;
  // fine-level element data (padded element layout)
  uint8_t* eact = nullptr;
  uint8_t* allfree = nullptr;   // active element whose 8 nodes are all free
  uint8_t* shell = nullptr;
  float* rho = nullptr;
  float* rho_pre = nullptr;
  float* sens = nullptr;
  float* vsens = nullptr;
  float* tmpf = nullptr;        // element-sized scratch
  double* w64 = nullptr;        // rho^p
  float* w32 = nullptr;
  // template matrices
  double* Ke64 = nullptr;       // 576
  float* Ke32 = nullptr;
  double* Kc64 = nullptr;       // 16: the distinct entries of Ke's first block row (ke_key)
  float* Kc32 = nullptr;
  double* Mc64 = nullptr;       // 8 x 576 Galerkin child products
  double h_Ke[576];
  double h_Mc[8 * 576];
  // loads and solutions: K cases x 3 x NP (FP64)
  int K = 0;
  double* F = nullptr;
  double* U = nullptr;
  // PCG work vectors (FP64, 3*NP)
  double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  float* p32 = nullptr;         // FCG search direction of the FP32-cycle solves (0 off the free nodes)
  float* q32 = nullptr;         // K p stored in FP32 (solve precision 3; allocated on first use)
  float* p32s[3] = {nullptr, nullptr, nullptr};  // direction buffers 1..3 of the precision-3 FCG
  // z.r and z.q from the last fine post-smoothing sweep (precision-3 FCG): dots_q = q while
  // the FCG asks for them, dots_last while that sweep runs, dots_done once it has written
  // the per-wave sums (gsdots: zr of every backward segment, then zq)
  const float* dots_q = nullptr;
  bool dots_last = false, dots_done = false;
  double* gsdots = nullptr;
  int64_t gsdots_n = 0;
  // recycled start (tos_set_start_recycling): per case a ring of the last rc_depth
  // solutions (FP64, 3*NP each), K times the basis, the Gram partials and their sums
  int rc_depth = 0;
  int rc_K = 0;                 // cases the ring is sized for
  double* rc_ring = nullptr;    // [case][slot] x 3*NP
  double* rc_kv = nullptr;      // rc_depth x 3*NP
  bool r_ready = false;         // c.r (and its FP32 copy) = f - K u of the next solve, from the recycled start
  double* rc_part = nullptr;    // Gram columns x kRedBlocks
  double* rc_out = nullptr;     // Gram column sums
  std::vector<int> rc_head, rc_cnt;  // per case: newest slot, stored solutions
  // reductions
  double* partial = nullptr;    // 2 * max(kRedBlocks, blocks of the fine chunk list) + 2 * 256 stage
  int nb0 = 0;                  // blocks of a launch over the fine chunk list
  double* scal = nullptr;       // device scalars
  double* h_scal = nullptr;     // pinned host mirror
  // coarsest level
  int cn = 0;                   // dense size (3 * free nodes)
  int64_t* cidx = nullptr;      // padded node index of each coarse unknown node
  double* cinv = nullptr;       // -A^{-1} after the sweep (cn x cn), from the FP64 stencils
  double* cinv32 = nullptr;     // the same from the FP32-built stencils (Galerkin precision 1)
  double* cwork = nullptr;      // panel buffers
  int* cflag = nullptr;         // singular pivot flag
  int* cmap = nullptr;          // padded coarsest node -> compact node index or -1
  void* stage = nullptr;        // host<->device conversion buffer
  size_t stage_bytes = 0;
  bool op_ready = false;
  // worst-case load: load nodes (padded index), unit normals and the last force
  // (node-major, 3 per node), start force; grown on demand
  int64_t wc_cap = 0;
  int64_t* wc_idx = nullptr;
  double *wc_nrm = nullptr, *wc_fs = nullptr, *wc_src = nullptr;
  double* wc_pos = nullptr;     // load-node coordinates (3 per node), floating contexts
  // skip-layer hierarchy: level 1 is 4x coarser than level 0 (factor-4 transfers, Galerkin
  // with the Dirichlet penalty)
  bool skip = false;
  uint8_t* efix = nullptr;      // per padded fine element: bit k = its node k is supported
  double* Mc4 = nullptr;        // 64 x 576: T_c^T Ke T_c per child position c of a coarse element
  // floating body (no supported fine node): the operator's null space is the rigid motions
  bool floating = false;
  double* crig = nullptr;       // coarsest level: orthonormal rigid motions, column k at crig[k * cn]
  int crig_m = 0;
  RigidProj rig_all{};          // projector over the fine active nodes (displacementProject)
  double* rig_part = nullptr;   // 6 x kRedBlocks partial sums
  // spline
  int sp_mode = -1, sp_order = 0, sp_param = 0;
  int nc[3] = {0, 0, 0};
  int* sp_idx[3] = {nullptr, nullptr, nullptr};    // n_elem_axis * order
  float* sp_wt[3] = {nullptr, nullptr, nullptr};
  uint8_t* sp_ok[3] = {nullptr, nullptr, nullptr};
  int sp_nq[3] = {0, 0, 0};                          // list lengths (sp_cptr[a][nc[a]])
  int* sp_cptr[3] = {nullptr, nullptr, nullptr};   // CSR transpose: per coefficient
  int* sp_celm[3] = {nullptr, nullptr, nullptr};
  float* sp_cw[3] = {nullptr, nullptr, nullptr};
  float* coef = nullptr;
  float* coef_old = nullptr;
  float* csens = nullptr;
  float* cvsens = nullptr;
  double* sp_t1 = nullptr;  // separable adjoint stages
  double* sp_t2 = nullptr;
  double* sp_g = nullptr;   // last adjoint stage (per coefficient, FP64)
  // surface constraints (tos_surface_constraints): background lattice of sf_cells cells per
  // axis, its node values and point offsets, the points (capacity 3 per lattice node) with
  // gradient, Hessian, normal and gradient weights, and the coefficient gradients
  int sf_cells = 0;
  int64_t sf_n = 0;
  double* sf_val = nullptr;
  int *sf_cnt = nullptr, *sf_off = nullptr;
  int *sf_bsum = nullptr, *sf_boff = nullptr;  // scan block sums / offsets
  int* sf_tab_i = nullptr;   // per axis and lattice index: coefficient indices, basis values, flags
  double* sf_tab_v = nullptr;
  uint8_t* sf_tab_ok = nullptr;
  int* sf_bidx = nullptr;    // per point and axis: coefficient indices and basis values (grown on demand)
  double* sf_bN = nullptr;
  int* sf_key = nullptr;     // per point: knot cell (grown with sf_bidx)
  int* sf_list = nullptr;    // points grouped by knot cell
  int64_t sf_bcap = 0;
  int64_t sf_ncoef = 0;      // coefficient count the per-coefficient buffers below are sized for
  int *sf_ccnt = nullptr, *sf_cstart = nullptr;  // points per knot cell (then fill cursor), cell offsets
  int *sf_cbsum = nullptr, *sf_cboff = nullptr;  // scan block sums / offsets over the cells
  double *sf_pts = nullptr, *sf_grad = nullptr, *sf_hess = nullptr, *sf_nrm = nullptr;
  double *sf_wss = nullptr, *sf_wdr = nullptr;
  double* sf_knots[3] = {nullptr, nullptr, nullptr};  // clamped knots (spline mode 0)
  float* sscsens = nullptr;
  float* dripcsens = nullptr;
  double ss_value = 1.0, drip_value = 0.0;
  bool sf_valid = false;
  std::vector<void*> allocs;
  // kernel timing (HIP events on `stream`), see tos_profile
  bool prof_on = false;
  struct ProfRec {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<ProfRec> prof;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  LearnState* learn = nullptr;
  BatchState* batch = nullptr;
  bool half_st = false;         // the running FP32 cycle reads the FP16 stencils
  bool st16_valid = false;      // FP16 stencils match the current operator
  int gal_prec = 0;             // coarse operators: 0 built and stored in FP64 (FP32 copies
                                // rounded from them); 1 built into FP32 storage, FP64 built
                                // on first demand (FP64 cycles, tos_get_stencil)
  bool st64_valid = false;      // FP64 stencils and cinv match the current operator
  int sweep_mode = 3;           // fine GS: 0 colour passes, 1 colour-slab wavefront, 2 row pairs,
                                // 3 in-wave pairs (FP32 cycles)
  bool graph_replay = true;     // replay the captured FCG iteration body
  struct SolveGraph {           // captured FCG iteration body per solve configuration
    int precision, pre, post, sweep, gal;
    int slot;                   // precision 3: the direction buffer the body writes
    const double* u;
    uint64_t epoch;             // alloc_epoch at capture
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<SolveGraph> graphs;
  uint64_t alloc_epoch = 0;     // bumped by every context allocation or release after creation:
                                // a captured graph from an older epoch is destroyed, not replayed
};

constexpr int kRedBlocks = 1024;  // fixed partial count -> deterministic reductions

}  // namespace tos

// scalar slots in Ctx::scal
enum {
  S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RN2, S_ZR, S_ZQ, S_NF2,
  S_OC_LO, S_OC_HI, S_OC_G, S_OC_V, S_OC_DONE, S_OC_STEPS, S_OC_GMAX, S_VOL,
  S_TMP0, S_TMP1, S_TMP2, S_TMP3,
  S_NN0, S_NN1, S_NN2,               // learned correction: |r~|^2, e.r, e.(r - A e)
  S_NN_TRIED, S_NN_KEPT, S_NN_STEP,  // its statistics since the last attach
  S_RIG0,                            // 6 slots: R^T v of a rigid-motion projection
  S_AH0 = S_RIG0 + 6,                // 4 slots: alpha of the directions in the four direction
                                     // buffers of the precision-3 FCG
  S_COUNT = 40
};
static_assert(S_AH0 + 4 <= S_COUNT, "scalar slots");
