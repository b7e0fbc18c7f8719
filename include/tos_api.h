/*
 * tos_api.h -- C ABI of the MI355X (gfx950) topology-optimisation hot path.
 *
 * One context (tos_ctx) owns every device buffer of one voxel design domain: the
 * multigrid hierarchy, the load cases, densities, sensitivities, spline coefficients
 * and the MMA state.  All calls on one context run on that context's HIP stream and
 * are not thread-safe; different contexts may be used from different threads.
 * Every call returns 0 on success or a negative TOS_E* code; tos_last_error() gives
 * the message of the last failure on that context.  Host arrays are always
 * lexicographic with x fastest:
 *   element e = ex + nx*(ey + ny*ez)                      (nx*ny*nz entries)
 *   node    n = ix + (nx+1)*(iy + (ny+1)*iz)              ((nx+1)(ny+1)(nz+1) entries)
 *   vector  v[c*n_nodes + n] for component c in {x,y,z}   (structure of arrays)
 * FP64 at the boundary for displacements, forces and compliance; FP32 for densities,
 * sensitivities and spline coefficients (the reference's storage types).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   grid::HierarchyGrid / grid::Grid     grid/Grid.h:214-1037
 *   optimization free functions          optimization/optimization.h:41-126
 *   MMA::mma_t                           optimization/mma_t.h:159-266
 * Each entry point below names the reference function it stands in for.
 */
#ifndef TOS_API_H
#define TOS_API_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOS_OK 0
#define TOS_E_ARG -1        /* bad argument (null pointer, size out of range)          */
#define TOS_E_HIP -2        /* HIP runtime error (allocation, launch)                  */
#define TOS_E_NOCONV -3     /* solver did not reach the tolerance within maxit         */
#define TOS_E_NAN -4        /* NaN/Inf detected in an input or a reduction             */
#define TOS_E_SINGULAR -5   /* coarsest-level Cholesky found a non-positive pivot      */
#define TOS_E_STATE -6      /* call order violated (e.g. solve before update_operator) */
#define TOS_E_TOO_LARGE -7  /* input beyond the capacity of a fixed-size entry point    */

typedef struct tos_ctx tos_ctx;

/* Design domain.  Replaces HierarchyGrid::genFromMesh(bits, reso) (grid/Grid.cpp:977-1236)
 * plus the support/load flagging of Grid::build (grid/Grid.cpp:2939-3161). */
typedef struct tos_grid_desc {
  int nx, ny, nz;           /* elements per axis (any positive sizes)                  */
  double h;                 /* element edge length                                     */
  const uint8_t* solid;     /* nx*ny*nz, nonzero = active element                      */
  const uint8_t* support;   /* nodes, nonzero = node fixed in x, y and z (may be NULL) */
  const uint8_t* shell;     /* nx*ny*nz, nonzero = density pinned to 1 (may be NULL)   */
  int n_levels;             /* 0 = coarsen while a level has > coarsen_threshold active
                               elements (grid/Grid.cpp:747); >0 = exactly this many    */
  int coarsen_threshold;    /* 0 = 400                                                 */
  int skip_layer;           /* 1: the reference's skip-layer hierarchy (set_skip_layer,
                               optimization/optimization.cpp:20): level 1 of the dyadic
                               hierarchy is dropped, so the first coarse level is 4x
                               coarser, with factor-4 trilinear transfers (grid/Grid.cu:
                               2509-2570, 2635-2671), all nodes of its active elements as
                               unknowns, and its Galerkin operator formed from the fine
                               element products with the reference's Dirichlet penalty
                               (supported fine nodes: rows and columns dropped, 1e6 on the
                               diagonal per element, grid/Grid.cu:14, 1112-1225); needs 3
                               or more dyadic levels (ignored otherwise); 0: dyadic      */
} tos_grid_desc;

typedef struct tos_info {
  int n_levels;
  int level_dims[16][3];        /* element dims per level                              */
  int64_t n_active_elements[16];
  int64_t n_free_nodes[16];
  int64_t n_elements, n_nodes;  /* fine level, dense counts                            */
  int coarse_dofs;              /* size of the dense coarsest system                   */
  int n_cases;                  /* load cases                                          */
  int64_t n_coeffs;             /* spline coefficients (0 until tos_spline_init)       */
  int has_support;              /* 0: no fine node is supported (a floating body)      */
} tos_info;

typedef struct tos_solve_opts {
  double rtol;              /* stop when ||f - K u||_2 / ||f||_2 <= rtol (per case)     */
  int maxit;                /* outer (flexible PCG) iterations                          */
  int precision;            /* 0: FP64 V-cycle; 1: FP32 V-cycle inside FP64 PCG;
                               2: as 1 with the stored stencils of the coarse levels read
                               as FP16 (each node's 243 entries scaled by their largest
                               magnitude); 3 (tos_solve only, use_pcg = 1): as 2 with
                               the FCG residual and K p stored in FP32 (FP64 operator
                               arithmetic, FP64 u), restarted from the FP64 residual
                               if that misses rtol (no learned correction); no
                               counterpart in the reference                           */
  int pre, post;            /* Gauss-Seidel sweeps per level (reference: 1, 1)          */
  int warm_start;           /* 1: start from the current displacement                   */
  int use_pcg;              /* 1: flexible PCG (default); 0: plain V-cycle iteration as
                               grid/Grid.cpp:2399-2444 with utils/test_utils.cpp:659-663 */
} tos_solve_opts;

typedef struct tos_solve_stats {
  int iterations[64];       /* per case                                                */
  double rel_res[64];       /* final true relative residual (recomputed in FP64)       */
  double seconds;           /* wall time of the call                                   */
} tos_solve_stats;

/* ---- context ------------------------------------------------------------------ */
int tos_ctx_create(const tos_grid_desc* desc, tos_ctx** out);
int tos_ctx_destroy(tos_ctx* ctx);
const char* tos_last_error(const tos_ctx* ctx);
int tos_get_info(tos_ctx* ctx, tos_info* out);
/* Node flags of the fine level (bit0 active, bit1 support, bit2 free, bits 3-5 GS
 * colour x%2+2(y%2)+4(z%2)), compare Grid::Bitmask grid/Grid.h:376-416. */
int tos_get_node_flags(tos_ctx* ctx, int level, uint8_t* out);

/* ---- material / template matrix -------------------------------------------------
 * Replaces setParameters (optimization/optimization.cpp:35-65) and
 * uploadTemplateMatrix/initTemplateMatrix (optimization/optimization.cu:804-816,
 * element/templateMatrix.cpp:31-99). */
int tos_set_material(tos_ctx* ctx, double youngs, double poisson, double penal, double rho_min);
int tos_get_template_matrix(tos_ctx* ctx, double* ke576);

/* ---- loads ----------------------------------------------------------------------
 * Replaces setForceSupport/getPreloadForce (optimization/projection.cu:344-368).
 * f: n_cases x 3 x n_nodes.  Entries on non-free nodes are ignored (set to 0). */
int tos_set_loads(tos_ctx* ctx, int n_cases, const double* f);
int tos_get_loads(tos_ctx* ctx, int k, double* f);

/* ---- densities -------------------------------------------------------------------
 * Replaces initDensities (optimization.h:111), Grid::init_rho (Grid.h:610), getRho. */
int tos_set_density(tos_ctx* ctx, const float* rho);
int tos_get_density(tos_ctx* ctx, float* rho);

/* ---- operator --------------------------------------------------------------------
 * Replaces update_stencil (grid/Grid.cpp:4167-4178): element weights rho^p, Galerkin
 * coarse stencils on every level, and the factorisation of the coarsest system. */
int tos_update_operator(tos_ctx* ctx);
/* Galerkin stencil of level >= 1: 27*9*n_nodes(level), entry (k*9+r*3+c)*nn + node,
 * k = (dx+1)+3(dy+1)+9(dz+1) -- the reference's rxStencil layout (grid/Grid.cu:847). */
int tos_get_stencil(tos_ctx* ctx, int level, double* out);
/* y = K(rho) u on the fine level (applyK_OTFA_kernel, grid/Grid.cu:4278-4356). */
int tos_apply_operator(tos_ctx* ctx, const double* u, double* y);
/* z = B r, one symmetric V(pre,post)-cycle (HierarchyGrid::v_cycle, grid/Grid.cpp:2399-2444);
 * precision as in tos_solve_opts. */
int tos_vcycle(tos_ctx* ctx, const double* r, double* z, int precision, int pre, int post);

/* ---- solve -----------------------------------------------------------------------
 * Replaces the V-cycle loop of the drivers (utils/test_utils.cpp:659-663). */
int tos_solve(tos_ctx* ctx, const tos_solve_opts* opts, tos_solve_stats* stats);
int tos_get_displacement(tos_ctx* ctx, int k, double* u);
int tos_set_displacement(tos_ctx* ctx, int k, const double* u);
/* c_k = f_k . u_k for every case (Grid::compliance, grid/Grid.cpp:4035-4038). */
int tos_compliance(tos_ctx* ctx, double* c);
/* sum(rho)/N_active over active elements (Grid::volumeRatio, grid/Grid.cu:5462-5470). */
int tos_volume(tos_ctx* ctx, double* v);

/* ---- floating bodies ----------------------------------------------------------------
 * A domain whose fine level has no supported node is a floating body: K(rho) has the six
 * rigid motions as its null space (Grid::hasSupport false, grid/Grid.h:730).  Its
 * coarsest level is factored with alpha R R^T added (R the orthonormalised rigid motions
 * of the coarsest nodes, alpha the mean diagonal), so that for a right-hand side
 * orthogonal to them the coarsest solve returns the minimum-norm solution, as the
 * reference's BDCSVD solve does (grid/Grid.cpp:3654-3661, 4072); the loads of a solve must
 * then be balanced (tos_balance_loads), or the FCG stalls at their unbalanced part.
 *
 * forceProject (optimization/projection.cpp:289-350, optimization/projection.cu:193-250)
 * on load case k: the force is kept on the n given load nodes only (zero elsewhere); with
 * normals, only its component along each node's normal is kept; on a floating context its
 * rigid-motion components over those nodes (of the normal components, when normals are
 * given) are removed as well, so that its resultant force and moment vanish.  Rigid columns
 * whose norm after orthogonalisation falls below 1e-6 of their own are dropped, as in the
 * reference's Gram-Schmidt (projection.cpp:234-255).  Nodes: lexicographic ids of free
 * fine nodes, no repeats. */
int tos_balance_loads(tos_ctx* ctx, int k, int n, const int64_t* nodes, const double* normals);
/* displacementProject (optimization/projection.cpp:374-395): removes the rigid-motion
 * components of load case k's displacement over the active nodes.  Floating contexts only
 * (TOS_E_STATE otherwise). */
int tos_remove_rigid(tos_ctx* ctx, int k);

/* ---- worst-case load (modified power method) --------------------------------------
 * Replaces modifiedPM (optimization/optimization.cpp:213-351) with forceProject
 * (optimization/projection.cpp:289-350, optimization/projection.cu:193-250).  On a
 * floating context the force update also removes the force's rigid-motion components on
 * the load nodes (as tos_balance_loads) and every step removes the rigid motion of u over
 * the active nodes after the V-cycle (displacementProject, optimization.cpp:282-285).
 * The force lives on the load nodes only.  Start: a random force on the
 * load nodes, projected and scaled to unit 2-norm, u = 0.  Each step runs one
 * V(pre,post)-cycle on the displacement of load case 0 (u += B(f - K u)), takes the new
 * force from u on the load nodes (its component along each node's normal when normals
 * are given, all of it otherwise), scales it to unit 2-norm, and stops once
 * ||f_new - f_old||_2 / ||f_old||_2 <= fch_tol and the cycle's relative residual
 * ||f - K u|| / ||f|| <= res_tol, or after max_iter cycles.  On return load case 0
 * holds the last force and displacement and stats->compliance = f . u (the worst-case
 * compliance).  A cycle residual above 1e4 returns TOS_E_NOCONV (the reference exits). */
typedef struct tos_worst_opts {
  int n_load;               /* load nodes: free fine nodes, no repeats                  */
  const int64_t* nodes;     /* their lexicographic node ids                             */
  const double* normals;    /* 3*n_load (x, y, z per node; any nonzero length) or NULL:
                               the force direction is free                              */
  uint64_t seed;            /* start force: component c of load node i is the
                               (3i+c+1)-th splitmix64 value of seed mapped to [-1, 1)
                               (the reference draws it from a time-seeded cuRAND)       */
  int precision;            /* V-cycle precision as in tos_solve_opts                   */
  int pre, post;            /* reference: 1, 1                                          */
  int max_iter;             /* reference: 500                                           */
  double fch_tol, res_tol;  /* reference: 1e-4, 1e-2                                    */
} tos_worst_opts;

typedef struct tos_worst_stats {
  int iterations;           /* V-cycles run                                             */
  double compliance;        /* f . u on return                                          */
  double fch, rel_res;      /* last force change and last cycle residual                */
  double seconds;
} tos_worst_stats;

int tos_worst_compliance(tos_ctx* ctx, const tos_worst_opts* opts, tos_worst_stats* stats);

/* ---- sensitivities ---------------------------------------------------------------
 * Sensitivity of case k: s_e = -p rho_e^(p-1) u_e^T Ke u_e (computeSensitivity_kernel,
 * optimization/optimization.cu:96-159), into the context's sensitivity buffer. */
int tos_sensitivity(tos_ctx* ctx, int k);
/* which: 0 = compliance sensitivity, 1 = volume sensitivity. */
int tos_get_sensitivity(tos_ctx* ctx, int which, float* out);
int tos_set_sensitivity(tos_ctx* ctx, int which, const float* in);
/* volume sensitivity := 1/N_active on active elements (computeSensitivity2, :263-267). */
int tos_init_vol_sensitivity(tos_ctx* ctx);
/* Radius-R filter (Grid::filterSensitivity / filterVolSensitivity, grid/Grid.cu:4131-4276). */
int tos_filter(tos_ctx* ctx, int which, double radius);
/* Element compliance rho^p u_e^T Ke u_e of case k (elementCompliance, grid/Grid.cu:5389-5430). */
int tos_element_compliance(tos_ctx* ctx, int k, double* out);

/* ---- density update --------------------------------------------------------------
 * OC bisection (updateDensities, optimization/optimization.cu:383-460).  Returns the
 * threshold g and the number of bisection steps. */
int tos_oc_update(tos_ctx* ctx, double vgoal, double move, double damp,
                  double* g_out, int* steps_out);
/* OC update of the design variable behind the Heaviside projection (updateDensities2,
 * optimization/optimization.cu:462-557): the design variable is the projection input kept
 * by tos_project_density; each bisection trial updates it as tos_oc_update does and
 * projects it (beta) before the volume sum; the search stops at |V - vgoal| <= 1e-4, at a
 * relative bracket (hi - lo)/(lo + hi) <= 1e-3 or after 51 trials; the last trial becomes
 * the design variable and the densities its projection.  change (may be NULL) receives the
 * max and the mean |x_new - x_old| over the active elements (the reference divides its
 * sum by the padded element count). */
int tos_oc_update_projected(tos_ctx* ctx, double vgoal, double move, double damp, double beta,
                            double* g_out, int* steps_out, double* change);
/* OC update without any allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same update as tos_oc_update (threshold bisection
 * over [0, max|s|], at most 31 trials, stop at |V - vgoal| <= 1e-4, V = mean density of
 * the active elements) over caller-owned DEVICE arrays of n elements in any order:
 * active, shell (u8, may be NULL: all active / no shell), rho, sens (float) in,
 * rho_out (float, may equal rho) out; inactive elements are copied through.  Runs on
 * `stream` (a hipStream_t, NULL = the null stream); g_out / steps_out (host, may be NULL)
 * make the call wait for the stream.  The bisection's partials and scalars live in one
 * fixed device array per GPU (2*8192 + 32 doubles), so n > TOS_NOALLOC_MAX_ELEMENTS
 * returns TOS_E_TOO_LARGE; calls that overlap on one GPU (other threads or streams)
 * share that array and race. */
#define TOS_NOALLOC_MAX_ELEMENTS 67108864
int tos_oc_update_noalloc(int64_t n, const uint8_t* active, const uint8_t* shell, const float* rho,
                          const float* sens, float* rho_out, double vgoal, double move, double damp,
                          double rho_min, void* stream, double* g_out, int* steps_out);
/* Heaviside projection: keeps the current densities as the projection input and
 * replaces them by the projected field (projectDensities, optimization.cu:339-353). */
int tos_project_density(tos_ctx* ctx, double beta);
/* Multiplies sensitivity `which` by d(projected)/d(input) (computeProjectionSensitivity,
 * optimization.cu:161-183). */
int tos_projection_sensitivity(tos_ctx* ctx, int which, double beta);

/* ---- spline density --------------------------------------------------------------
 * mode 0: clamped knots of order `order` with `param` interior partitions per axis on
 *         the grid box enlarged by 2*box/(partition+1) per side (reference layout,
 *         HierarchyGrid::set_spline_partition grid/Grid.cu:3134-3203,
 *         Grid::set_spline_knot_series grid/Grid.cpp:3253-3377).
 * mode 1: uniform lattice of order `order`, coefficient (i,j,k) at node
 *         (param*i, param*j, param*k), boundary coefficients replicated. */
int tos_spline_init(tos_ctx* ctx, int mode, int order, int param);
int tos_spline_set_coeffs(tos_ctx* ctx, const float* c);
int tos_spline_get_coeffs(tos_ctx* ctx, float* c);
/* rho = clamp(B c, rho_min, 1), shell -> 1 (Grid::coeff2density, grid/Grid.cu:3361-3456). */
int tos_spline_eval(tos_ctx* ctx);
/* coefficient gradient of sensitivity `which` (Grid::ddensity2dcoeff_update /
 * dvol2dcoeff, grid/Grid.cu:3643-3859), deterministic gather. */
int tos_spline_adjoint(tos_ctx* ctx, int which);
/* which: 0 compliance, 1 volume, 2 self-supporting, 3 drip (tos_surface_constraints) */
int tos_spline_get_coeff_sens(tos_ctx* ctx, int which, float* out);
/* OC bisection on the spline coefficients with the compliance coefficient sensitivity
 * (updateCoeff / tryCSensMultiplier_kernel, optimization/optimization.cu:559-695): each
 * trial c' = clamp(clamp(c (max(0,-s)/g)^damp, c - move, c + move), cmin, cmax) is
 * evaluated to densities and the bisection stops at |V - vgoal| <= 1e-4 or 31 trials.
 * The coefficients end at the last trial (the reference applies the multiplier once more
 * on top of it). */
int tos_oc_update_coeffs(tos_ctx* ctx, double vgoal, double move, double damp, double cmin,
                         double cmax, double* g_out, int* steps_out);

/* ---- self-supporting and drip constraints on spline surface points --------------------
 * Replaces deal_surface_points (optimization/optimization.cpp:1103-1128) with
 * Grid::generate_spline_surface_nodes (grid/Grid.cpp:3984-4030), the surface-point normals
 * and their orientation test (grid/Grid.cu:5529-5763), the self-supporting constraint
 * (:6015-6320) and the drip constraint (:6485-6560, 6744-7156) with their coefficient
 * gradients.  Surface points: the marching-cubes vertices of the level set
 * {spline = 1/2} sampled on a background lattice of `cells` cells per axis over the grid
 * box plus one cell on each side, node values spline - 1/2 inside the box and -1/2
 * outside (compute_background_mcPoints_value, grid/Grid.cu:3986-4080): one point per
 * lattice edge whose end values change sign, linearly interpolated, kept when it lies in
 * an active element, ordered by lattice node then axis.  The normal is direction * grad
 * spline, direction = +1 when most points see void (Heaviside(spline, beta) < 1/2) a step
 * 1e-3 * grad spline ahead, else -1.  With inner = n_z / (|n| cos print_angle):
 *   ss_mode 0/1: (mean inner^p_norm)^(1/p_norm) - 1; 2/3: mean h(inner), h(x) =
 *   hfunction_c (x-1)^2 for x >= 1; 4/5: mean oh(inner), oh(x) = 1 + sigmoid((x - 1.5) /
 *   sigmoid_c); even modes over the points with n_z < 0 (value 0 if none), odd over all.
 *   drip_mode 6/7 (exp_drip / exp2_drip): mean over the n_z < 0 / all points of
 *   u(hinner) q1(H_xx) q2(H_xx H_yy - H_xy H_yx), hinner = n_z / (|n| cos drip_angle),
 *   u(x) = sigmoid((x - 1.2) drip_mu) for drip_angle > pi/2 else sigmoid(-(x + 1.2) drip_mu)
 *   (drip_version 1: exp(-drip_alpha ((n_x/n_z)^2 + (n_y/n_z)^2)) instead of u),
 *   q1(x) = drip_beta_scale x^2 / (1 + exp(drip_beta x)), q2(x) = drip_beta_scale x^2 /
 *   (1 + exp(-drip_beta x)), H the spline's Hessian; modes 0-5 give per-point value 0
 *   (the reference's compute_spline_drip_constraint_test evaluates only 6 and 7), so the
 *   global value is -1 for modes 0/1 and 0 otherwise.
 * The coefficient gradients (points held fixed) are exact derivatives of these values and
 * are read with tos_spline_get_coeff_sens(which = 2: self-support, 3: drip). */
typedef struct tos_surface_opts {
  int cells;                 /* background lattice cells per axis (reference: 64)          */
  double beta;               /* Heaviside sharpness of the orientation test                */
  double print_angle;        /* radians (reference default 3*pi/4, Grid.cpp:2576)          */
  int ss_mode;               /* GlobalSSMode 0..5 (reference flag default "oh" = 4)        */
  int drip_mode;             /* GlobalDripMode 0..7 (reference flag default "oh" = 4)      */
  double p_norm, hfunction_c, sigmoid_c;          /* 8, 10, 0.09 (grid/Grid.h:351-355)    */
  double drip_beta, drip_mu, drip_angle, drip_alpha, drip_beta_scale;
                             /* 128, 32, 160/180*pi, 4, 10 (grid/Grid.h:359-367)           */
  int drip_version;          /* 0 (grid/Grid.h:366)                                        */
} tos_surface_opts;

typedef struct tos_surface_stats {
  int64_t n_points;          /* surface points kept                                        */
  int64_t n_down;            /* of them with n_z < 0                                       */
  double direction;          /* normal orientation, +1 or -1                               */
  double ss_value;           /* global self-supporting value                               */
  double drip_value;         /* global drip value                                          */
} tos_surface_stats;

int tos_surface_constraints(tos_ctx* ctx, const tos_surface_opts* opts, tos_surface_stats* stats);
/* the last call's points and normals (3 per point, x y z), at most `cap` of them */
int tos_surface_points(tos_ctx* ctx, int64_t cap, double* xyz, double* normals);

/* All K load cases in one solve (no counterpart in the reference, which solves load
 * cases one at a time): case-interleaved vectors so every stencil / density read serves
 * all K cases, one flexible CG per case (FP64 outer vectors, per-case scalars and
 * convergence flags on the device) around one FP32 V-cycle for all cases.  The body of
 * an FCG iteration is captured once into a hipGraph and replayed; kernel block sizes
 * are chosen on the first call by timing the candidates (64/128/256 threads) on this
 * context's levels.  Requires precision = 1 and use_pcg = 1; same results contract as
 * tos_solve per case. */
int tos_solve_batched(tos_ctx* ctx, const tos_solve_opts* opts, tos_solve_stats* stats);
/* text description of the block sizes chosen for the batched solve ("" before it ran) */
int tos_batch_tuning(tos_ctx* ctx, char* buf, int len);

/* ---- MMA --------------------------------------------------------------------------
 * Method of moving asymptotes over n design variables with m constraints
 * (MMA::mma_t, optimization/mma_t.h:159-266; update mma_t.cpp:18-51).
 * The context-bound variant optimises the spline coefficients (n = n_coeffs) when
 * `on_coeffs` is 1, else the element densities. */
typedef struct tos_mma tos_mma;
int tos_mma_create(int n, int m, const double* xmin, const double* xmax, double move,
                   tos_mma** out);
int tos_mma_destroy(tos_mma* mma);
int tos_mma_set_x(tos_mma* mma, const double* x);
int tos_mma_get_x(tos_mma* mma, double* x);
/* df: n; dg: m*n (constraint-major); g: m.  converged != 0 after 10 consecutive
 * iterations with ||dx|| < 1e-3 (mma_t.cpp:40-48). */
int tos_mma_update(tos_mma* mma, const double* df, const double* dg, const double* g,
                   int* converged);
/* One MMA step of the spline driver (utils/test_utils.cpp:1135-1188): df = sens_scale *
 * coefficient sensitivity, dg = vol_scale * coefficient volume sensitivity,
 * g = vol_scale*(volume - vfrac); the new x is written to the coefficients. */
int tos_ctx_mma_step(tos_ctx* ctx, tos_mma* mma, double sens_scale, double vol_scale,
                     double vfrac, int* converged);
/* One MMA step of the self-supporting spline driver (utils/test_utils.cpp:1135-1188 with
 * ENABLE_SELFSUPPORT): m = 3, df = sens_scale * compliance coefficient sensitivity,
 * g = [vol_scale (volume - vfrac), ss_scale (ss_value - 1), drip_scale drip_value] with
 * their coefficient gradients (volume, self-support, drip; the last
 * tos_surface_constraints call's values); the new x is written to the coefficients. */
int tos_ctx_mma_step_surface(tos_ctx* ctx, tos_mma* mma, double sens_scale, double vol_scale, double vfrac,
                             double ss_scale, double drip_scale, int* converged);

/* Matrix-core element products against the matrix-free operator: a measurement, no
 * counterpart in the reference.  On the fine level and a seeded u (free nodes) it times
 * `reps` launches each of y = K u in FP32 by the matrix-free VALU kernel, by element
 * products on v_mfma_f32_16x16x4f32 (16 elements per wave as the columns of a 24x24 by
 * 24x16 product; gathers and products only), and by the same with a float-atomic scatter
 * into the nodes.  out[4] = {VALU ms, MFMA products ms, MFMA + scatter ms, relative
 * difference of the scattered result from the VALU one over the free nodes}.  Allocates
 * and frees its own buffers. */
int tos_probe_element_mfma(tos_ctx* ctx, int reps, double* out);

/* Fine-level Gauss-Seidel schedule (no counterpart in the reference, which launches one
 * pass per colour): 0 = eight colour passes; 1 = colour-slab wavefront, colour c of z-slab
 * s relaxed in step s + 2c (about 32 slabs); 2 = four launches of two colours each, one
 * workgroup per grid row relaxing the row's even-x colour then its odd-x colour; 3 = four
 * launches of two colours each, paired inside a wave (each lane relaxes a node of the
 * first colour, then the next node along x, using its neighbours' new values by lane
 * shuffles), also on the stored-stencil levels, FP32 cycles only (FP64 cycles use mode 0);
 * 3 is the default.  All perform exactly the same updates, so results are bit-identical
 * across modes. */
int tos_set_sweep_mode(tos_ctx* ctx, int mode);

/* Storage precision of the Galerkin coarse operators built by tos_update_operator (no
 * counterpart in the reference, whose rxStencil is FP64, grid/Grid.h:306):
 * 0 (default) = built and stored in FP64, the FP32 copies of the FP32 cycles rounded from
 * them; 1 = built with FP64 sums into FP32 storage (input stencils, intermediate boxes and
 * result), for contexts that solve with precision 1 or 2: half the Galerkin traffic.  With
 * 1 the FP64 stencils and their coarsest factor are built on first demand after each
 * update (a precision-0 solve or V-cycle, tos_get_stencil), from the same FP64 pipeline as
 * with 0, so FP64 results do not depend on the setting; the FP32 cycles keep the
 * FP32-built stencils.  A change takes effect at the next tos_update_operator, which is
 * then required. */
int tos_set_galerkin_precision(tos_ctx* ctx, int prec);

/* The body of an FCG iteration is captured once into a hipGraph and replayed (default
 * enable = 1); enable = 0 launches it kernel by kernel, e.g. for tools that trace
 * launches.  Kernel timing (tos_profile) always launches kernel by kernel. */
int tos_set_graph_replay(tos_ctx* ctx, int enable);

/* Recycled start for sequences of solves (no counterpart in the reference, whose solves
 * start from the previous displacement).  With depth m (1..8; 0 = off, the default) every
 * tos_solve with warm_start = 1 first replaces the current displacement u of each case by
 * the Galerkin projection of the solution onto span{u, s_2, ..., s_m}, where s_i is the
 * i-th newest solution a converged tos_solve of this case left (the newest, s_1, is u
 * itself unless the displacement was set since): u0 = V c with (V^T K V) c = V^T f, the
 * point of the span with the smallest energy-norm error, so never worse than u in that
 * norm.  Costs m FP64 operator applications, one pass over 2m + 1 vectors and one over
 * m + 1, and m stored solutions per case (FP64, the fine level's padded size).  The Gram
 * system is scaled to unit diagonal and solved through its eigenvalues; those below 1e-10
 * of the largest are dropped (nearly dependent solutions).  A changed depth starts an
 * empty history.  tos_solve_batched ignores it. */
int tos_set_start_recycling(tos_ctx* ctx, int depth);

/* ---- measurement ---------------------------------------------------------------------
 * Kernel timing with HIP events recorded on the context's stream around every launch of
 * the solver kernels.  enable = 1 clears and starts recording; enable = 0 stops, waits for
 * the stream and fills `out` (may be NULL).  No counterpart in the reference (its
 * tictoc timers, utils/tictoc.h:35-83, time host phases only). */
#define TOS_PROF_CLASSES 20
typedef struct tos_kernel_stats {
  int n_classes;
  char names[TOS_PROF_CLASSES][32];
  double total_ms[TOS_PROF_CLASSES];
  int64_t launches[TOS_PROF_CLASSES];
} tos_kernel_stats;
int tos_profile(tos_ctx* ctx, int enable, tos_kernel_stats* out);

/* ---- learned coarse-level correction --------------------------------------------------
 * No counterpart in the reference: a 3-level 3-D convolutional U-Net (16/32/64 channels,
 * 3x3x3 kernels, BF16 weights on MFMA, FP32 accumulation) that maps the residual of a
 * multigrid level (3 channels, scaled by the inverse square root of the Galerkin
 * diagonal and by its RMS) plus the nodal log-density (1 channel) to an error correction
 * (3 channels) on that level's node box.  Trained with these entry points on
 * residual / error pairs logged from the solver's own V-cycles.
 * Inputs of tos_unet_forward / tos_unet_loss_grad: feat[B][Z][Y][X][4] (x fastest),
 * target[B][Z][Y][X][3], mask[B][Z][Y][X]; loss = 1/B sum_b |m (y - t)|^2 / |m t|^2.
 * Parameters: per layer W[27][cin][cout] (tap = (dx+1) + 3(dy+1) + 9(dz+1)) then
 * b[cout], layers in the order enc0a, enc0b, enc1a, enc1b, enc2a, enc2b, dec1a, dec1b,
 * dec0a, dec0b, out. */
typedef struct tos_unet tos_unet;
typedef struct tos_dataset tos_dataset;
int tos_unet_create(int c0, int c1, int c2, uint64_t seed, tos_unet** out);  /* 16, 32, 64 */
int tos_unet_destroy(tos_unet* net);
const char* tos_unet_last_error(const tos_unet* net);
int tos_unet_num_params(const tos_unet* net, int64_t* n);
int tos_unet_get_params(tos_unet* net, float* out);
int tos_unet_set_params(tos_unet* net, const float* in);  /* also resets the Adam state */
int tos_unet_forward(tos_unet* net, int B, int X, int Y, int Z, const float* feat, float* out);
int tos_unet_loss_grad(tos_unet* net, int B, int X, int Y, int Z, const float* feat,
                       const float* target, const uint8_t* mask, double* loss, float* grad);
/* device-resident pair dataset on a fixed node box */
int tos_dataset_create(int X, int Y, int Z, int capacity, tos_dataset** out);
int tos_dataset_destroy(tos_dataset* ds);
int tos_dataset_size(const tos_dataset* ds, int* n);
int tos_dataset_get(tos_dataset* ds, int i, float* feat4, float* target, uint8_t* mask);
/* Adam (beta 0.9 / 0.999) on random batches of samples [first, first + count) */
int tos_unet_train(tos_unet* net, tos_dataset* ds, int first, int count, int steps, int batch,
                   double lr, uint64_t seed, double* loss_first, double* loss_last);
int tos_unet_eval(tos_unet* net, tos_dataset* ds, int first, int count, double* mean_loss);
/* mode 1: inside FP32 V-cycles the network proposes a correction e of level `level` after
 * its coarse-grid correction.  It is applied with the energy-optimal step
 * alpha = e.r / e.Ae and only when 0.2 <= alpha <= 5 (it lowers the A-norm error and is
 * roughly of the right size); otherwise plain multigrid continues unchanged.  After 3
 * rejections in a row the network pauses until the next solve; after a solve in which it
 * was tried and never kept it sits out the next 1, 3, 7, 15 solves (reset when a
 * correction is kept).  mode 0 detaches.  The context keeps the pointer: the network
 * must outlive the attachment. */
int tos_ctx_learn_attach(tos_ctx* ctx, tos_unet* net, int level, int mode);
/* log (features, target) pairs at `level` of every `stride`-th FP32 V-cycle into ds (grid =
 * the level's node box); target = the level's error after `n_inner` nested V-cycles.
 * ds = NULL stops logging. */
int tos_ctx_learn_log(tos_ctx* ctx, tos_dataset* ds, int level, int n_inner, int stride);
/* corrections tried / kept since the last attach and the mean step alpha of those kept */
int tos_ctx_learn_stats(tos_ctx* ctx, int* tried, int* accepted, double* mean_step);

#ifdef __cplusplus
}
#endif
#endif /* TOS_API_H */
