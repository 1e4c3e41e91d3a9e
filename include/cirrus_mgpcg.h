/*
 * cirrus_mgpcg.h -- C ABI of the MI355X-native MGPCG pressure-projection
 * library (libcirrus_mgpcg.so, HIP for gfx950).
 *
 * The library solves the volume-weighted Poisson problem of an adaptive octree
 * of 8^3-cell tiles (the reference's HAGrid of PoissonTiles) with a
 * multigrid-preconditioned conjugate gradient, on one GPU.  Plain C: opaque
 * handles, pointers and sizes; no exceptions cross the boundary; every call
 * returns a cg_status and cg_last_error() describes the last failure of the
 * calling thread.
 *
 * Conventions shared with the reference:
 *   tile types   CG_LEAF=1, CG_NONLEAF=2, CG_GHOST=4           (src/HATileInfo.h:19)
 *   cell types   CG_DIRICHLET=1, CG_INTERIOR=2, CG_NEUMANN=4   (src/PoissonTile.h:6)
 *   cell order   i-major inside a tile, offset = i<<6 | j<<3 | k (src/HAAccessor.h:32-42)
 *   cell size    h(level) = h0 / 2^level                        (src/HAAccessor.h:72-74)
 *   channels     CG_CH_X=0 (pressure), CG_CH_B=CG_CH_R=1 (rhs / residual),
 *                CG_CH_P=2, CG_CH_AP=3, CG_CH_Z=4, CG_CH_TMP=5 (src/PoissonTile.h:12-24)
 * Per-tile host arrays (channels, cell types) are passed in the order in
 * which the tiles were given to cg_problem_create.
 */
#ifndef CIRRUS_MGPCG_H
#define CIRRUS_MGPCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CG_LEAF = 1, CG_NONLEAF = 2, CG_GHOST = 4 };
enum { CG_DIRICHLET = 1, CG_INTERIOR = 2, CG_NEUMANN = 4 };
enum { CG_CH_X = 0, CG_CH_B = 1, CG_CH_R = 1, CG_CH_P = 2, CG_CH_AP = 3, CG_CH_Z = 4, CG_CH_TMP = 5, CG_NUM_CHANNELS = 6 };
/* Face velocities at the cells' minus faces (src/PoissonTile.h:12-24, channels u, v, w = 6, 7, 8);
 * moved with cg_set_channel / cg_get_channel, allocated (zero) on first use. */
enum { CG_CH_U = 6, CG_CH_V = 7, CG_CH_W = 8 };
enum { CG_SMOOTH_RBGS = 0, CG_SMOOTH_JACOBI = 1 };
/* Storage of the V-cycle vectors (its x on every level, its rhs and scratch);
 * arithmetic is FP32 either way.  BF16 rounds each stored value to nearest
 * even and halves their HBM traffic.  F16 does the same in float16 (three more
 * significand bits, range +-65504: stores clamp there); a solve in F16 scales
 * the right-hand side by a power of two 2^k, k = -floor(e/2) - 4 where
 * |b|^2 / cells = 2^e f (f in [1, 2)), and scales x and r back, so the result
 * is the solve of 2^k b divided by 2^k exactly.  F16 is available to cg_solve,
 * cg_solve_batch and cg_solve_refined, not to the server.  No reference
 * counterpart. */
enum { CG_STORE_F32 = 0, CG_STORE_BF16 = 1, CG_STORE_F16 = 2 };

typedef enum {
  CG_OK = 0,
  CG_ERR_INVALID = 1,   /* bad argument or malformed tree */
  CG_ERR_HIP = 2,       /* a HIP runtime call failed */
  CG_ERR_NOMEM = 3,     /* device or host allocation failed */
  CG_ERR_STATE = 4,     /* call not valid in the handle's current state */
  CG_ERR_BUSY = 5       /* no capacity now (HBM budget or problem slots); retry later */
} cg_status;

/* Opaque device-resident hierarchy of one octree (tiles, tables, channels). */
typedef struct cg_problem cg_problem;

/* A complete HAGrid as the solver receives it: every LEAF, NONLEAF and GHOST
 * tile with its cell types (what GMGSolver::solve reads from HADeviceGrid,
 * src/GMGSolver.cu:531, after CalcCellTypesFromLeafs, src/PoissonSolver.cu:7).
 * Neighbour tables (src/PoissonGrid.cu:121-134) and parent links are derived
 * from the coordinates. */
typedef struct {
  double h0;                 /* level-0 cell size (the reference uses 1/8, src/FluidEuler.h:415) */
  int32_t num_tiles;
  const int32_t* level;      /* [num_tiles] */
  const int32_t* coord;      /* [3*num_tiles] tile coordinate at its level */
  const uint8_t* type;       /* [num_tiles] CG_LEAF / CG_NONLEAF / CG_GHOST */
  const uint8_t* cell_type;  /* [512*num_tiles] */
} cg_tree_desc;

/* Solver parameters; cg_params_default() gives the reference's project() call
 * solve(grid, false, 1000, 1e-6, 1, 10, 1, pure) with GMGSolver(1, 1)
 * (src/FluidEuler.h:591,598). */
typedef struct {
  int32_t max_iters;        /* 1000 */
  double rel_tol;           /* 1e-6: stop when |r|^2 < rel_tol^2 |b|^2 (recursive r) */
  int32_t level_iters;      /* smoothing sweeps per level, pre and post (1) */
  int32_t coarsest_iters;   /* red-black GS iterations on level 0 (10) */
  int32_t smoother;         /* CG_SMOOTH_RBGS (reference) or CG_SMOOTH_JACOBI */
  float omega;              /* damped-Jacobi weight (2/3) */
  int32_t sync_stride;      /* the residual is tested after every sync_stride-th iteration (1); <= 0:
                               never, run max_iters and report rel_res = -1 (src/GMGSolver.cu:611,655-665) */
  int32_t pure_neumann;     /* re-centre r before each V-cycle (src/GMGSolver.cu:556,621) */
  float one_over_alpha;     /* restriction scale (1) */
  float prolong_coeff;      /* prolongation scale (1) */
  int32_t use_graph;        /* 1: capture one CG iteration in a hipGraph and replay it; 0: launch each
                               kernel on the handle's stream; -1: replay for the AMG preconditioner and
                               handles of fewer than 4,096 tiles, stream launches otherwise (-1;
                               results identical in every mode) */
  int32_t storage;          /* CG_STORE_F32 (0, the reference), CG_STORE_BF16 or CG_STORE_F16 */
  int32_t verify_f64;       /* recompute |b - A x| / |b| in FP64 at the end (cg_result.rel_res_f64) (0) */
} cg_params;

typedef struct {
  int32_t iters;            /* CG iterations run */
  int32_t converged;        /* 1 if |r| < rel_tol |b| was reached */
  double rel_res;           /* sqrt(|r|^2/|b|^2) of the recursive residual */
  double bb;                /* |b|^2 over LEAF tiles */
  double rr;                /* final |r|^2 */
  double seconds;           /* wall time of the solve, host clock, device synchronised */
  double rel_res_f64;       /* sqrt(|b - A x|^2 / |b|^2) of the FP32 x recomputed in FP64 over LEAF tiles
                               (re-centred for pure Neumann) when verify_f64 is set, else -1; no reference
                               counterpart */
} cg_result;

void cg_params_default(cg_params* p);
const char* cg_last_error(void);
/* Number of visible HIP devices (0 when there is none); no reference
 * counterpart, used by callers to check for a GPU before creating problems. */
int32_t cg_device_count(void);

/* Upload a tree; allocates the device hierarchy.  Replaces building an
 * HADeviceGrid (src/HAGrid.h:387-561) plus CalculateNeighborTiles
 * (src/PoissonGrid.cu:121-134). */
cg_status cg_problem_create(const cg_tree_desc* desc, cg_problem** out);
/* Batched form, no reference counterpart: num_problems independent trees in
 * one handle, solved together by cg_solve_batch (every kernel of a CG
 * iteration covers the tiles of all problems; each problem keeps its own
 * scalars and convergence test, so its result is bit-identical to solving it
 * alone).  Per-tile host arrays of the handle concatenate the problems in
 * order, each in the tile order of its description; topology indices refer to
 * that concatenated order. */
cg_status cg_problem_create_batch(const cg_tree_desc* descs, int32_t num_problems, cg_problem** out);
cg_status cg_problem_num_problems(const cg_problem* p, int32_t* num_problems);
cg_status cg_problem_destroy(cg_problem* p);
/* Tile counts and the deepest level. */
cg_status cg_problem_info(const cg_problem* p, int32_t* num_tiles, int32_t* num_leaf_tiles, int32_t* max_level);
/* Derived topology in the caller's tile order: nbr[6*t+f] is the caller index
 * of the same-level face neighbour f (x-,y-,z-,x+,y+,z+) or -1, exactly the
 * reference's tile.mNeighbors (src/PoissonGrid.cu:121-134); parent[t] likewise. */
cg_status cg_problem_topology(const cg_problem* p, int32_t* nbr, int32_t* parent);

/* Channel transfer (host <-> device), caller's tile order, 512 floats per tile;
 * channels 0-5 and the face velocities 6-8. */
cg_status cg_set_channel(cg_problem* p, int32_t channel, const float* host);
cg_status cg_get_channel(cg_problem* p, int32_t channel, float* host);
/* The same moves between device buffers and the handle, replacing the
 * reference's in-place channel access on its device tiles (GMGSolver::solve
 * reads b and writes x on HADeviceGrid, src/GMGSolver.cu:531-666, which
 * FluidEuler::project fills and reads on the device, src/FluidEuler.h:581-612).
 * The buffer is device-accessible memory of 512 floats per tile in the
 * caller's tile order (checked: memory known to HIP, large enough).  `stream`
 * is the caller's hipStream_t (NULL: the null stream): the move waits for the
 * work already enqueued there, and work the caller enqueues there afterwards
 * sees its result; the call itself does not wait on the host. */
cg_status cg_set_channel_device(cg_problem* p, int32_t channel, const float* device_src, void* stream);
cg_status cg_get_channel_device(cg_problem* p, int32_t channel, float* device_dst, void* stream);

/* GMGSolver::solve (src/GMGSolver.cu:531-666): solves A x = b with b in
 * CG_CH_B, pressure returned in CG_CH_X on LEAF tiles.  Channels 1-5 are
 * clobbered, as in the reference. */
cg_status cg_solve(cg_problem* p, const cg_params* prm, cg_result* res);
/* All problems of a batched handle; results[num_problems]. */
cg_status cg_solve_batch(cg_problem* p, const cg_params* prm, cg_result* results);

/* ---- Mixed-precision form (no reference counterpart) ----------------------
 * Iterative refinement: x and the residual r = b - A x are FP64 (the operator
 * of cg_full_negative_laplacian evaluated in FP64); each refinement step solves
 * A d = float(r) with the MGPCG of cg_solve (FP32 arithmetic, FP64 dot
 * products, V-cycle storage as prm->storage) and adds d to x.  The stop test
 * is |r|^2 < rel_tol^2 |b|^2 in FP64 over LEAF tiles, with the true residual.
 * A step's inner solve stops when its recursive |r_in|^2 < max(inner_rel_tol^2
 * |float(r)|^2, rel_tol^2 |b|^2 / 4); prm->max_iters caps each inner solve
 * (prm->rel_tol is not used).  Every
 * problem of a batched handle runs its own outer loop; results[num_problems].
 * On return CG_CH_X holds float(x) and cg_get_solution_f64 gives x. */
typedef struct {
  double rel_tol;           /* 1e-8 */
  double inner_rel_tol;     /* 1e-4 */
  int32_t max_outer;        /* refinement steps at most (20) */
} cg_refine_params;
typedef struct {
  int32_t outer;            /* refinement steps (inner solves) run */
  int32_t iters;            /* CG iterations over all inner solves */
  int32_t converged;        /* 1 if |b - A x| < rel_tol |b| (FP64) was reached */
  double rel_res;           /* sqrt(|b - A x|^2 / |b|^2), FP64 */
  double bb;                /* |b|^2 over LEAF tiles */
  double rr;                /* final |b - A x|^2 */
  double seconds;           /* wall time, host clock, device synchronised */
} cg_refine_result;
void cg_refine_params_default(cg_refine_params* p);
cg_status cg_solve_refined(cg_problem* p, const cg_params* prm, const cg_refine_params* rp, cg_refine_result* results);
/* The FP64 x of the last cg_solve_refined (caller's tile order, 512 per tile;
 * LEAF tiles carry the solution). */
cg_status cg_get_solution_f64(cg_problem* p, double* host);

/* ---- Schedules and the autotuner (no reference counterpart) ---------------
 * A schedule changes how the kernels cover the tiles, never a result:
 * coarse_shift = how many of each problem's LDS-resident coarse levels (as many
 * as fit by default) go back to grid-wide launches; swizzle = XCD-aware tile
 * order (1, default) or the plain order (0); tile_block = how the grid-wide
 * colour sweeps take their tiles: 1 (default) one tile per 64-thread
 * workgroup, 2 two tiles per wavefront (the loads of both in flight together),
 * 4 four wavefronts of one tile each per 256-thread workgroup, 8 as many
 * wavefronts as the device holds at once, each walking its share of a level's
 * tiles with the next tile's loads issued before the current one is updated
 * (on levels with more tiles than that; others as 1). */
cg_status cg_set_schedule(cg_problem* p, int32_t coarse_shift, int32_t swizzle, int32_t tile_block);
/* Two more kernel forms chosen by level size, also without effect on results:
 * levels (all problems of the handle together) of at most eager_max_tiles
 * tiles run the prolongation-fused colour pass and the restriction with every
 * load issued before one wait (shorter latency chain; more registers); with
 * 16-bit storage, levels of at most fused16_max_tiles tiles run the first
 * down-sweep iteration as one launch (it reads the halo's diagonals).  0: never,
 * a negative value: the default 16384.  pair_sweeps: 1 runs every red-black
 * iteration after the first as one launch of both colours (the halo's first
 * colour formed again from the neighbours' data; bit-identical, but slower
 * than the two passes on every level measured), 0 or a negative value as two
 * colour passes (the default).  No reference counterpart. */
cg_status cg_set_level_forms(cg_problem* p, int32_t eager_max_tiles, int32_t fused16_max_tiles, int32_t pair_sweeps);
typedef struct {
  int32_t candidates;        /* parameter sets timed */
  double base_seconds;       /* one solve with the caller's parameters, default schedule */
  double best_seconds;       /* one solve with the chosen schedule and parameters */
  int32_t coarse_shift;      /* the chosen schedule (kept by the handle) */
  int32_t swizzle;
  int32_t tile_block;
} cg_tune_report;
/* Times solves of the handle's current right-hand sides (restored afterwards)
 * under candidate schedules (swizzle on/off x tile_block 1/2/4, then coarse_shift
 * 1-2 with the best of those) with the caller's
 * smoother, then under candidate smoother settings (RB-GS with 1-3 sweeps per
 * level, damped Jacobi with 2-3; with allow_16bit also bfloat16 and float16 V-cycle storage),
 * keeping the fastest whose every problem converges (a later candidate must be
 * more than 2% faster to replace an earlier one).  The handle keeps the
 * chosen schedule; *best receives the chosen parameters (base with smoother,
 * level_iters and storage replaced). */
cg_status cg_autotune(cg_problem* p, const cg_params* base, int32_t allow_16bit, cg_params* best, cg_tune_report* report);
/* The same search with every candidate timed as the FP64-refined solve
 * (cg_solve_refined with rp) that the caller will run, whose best settings can
 * differ from a single solve's (fewer sweeps per level win a single solve to
 * rel_tol, more win the refined one).  No counterpart in the reference. */
cg_status cg_autotune_refined(cg_problem* p, const cg_params* base, const cg_refine_params* rp, int32_t allow_16bit,
                              cg_params* best, cg_tune_report* report);

/* ---- The projection around the solve (src/FluidEuler.h:581-612) ---------- */
/* VelocityVolumeDivergenceOnLeafs (src/PoissonSolver.cu:66-142) into div_channel: with
 * propagate_ghosts, the ghost velocities are first taken from the parents (PropagateToChildren,
 * src/PoissonGrid.cu:635-661, as FluidEuler::project does); INTERIOR cells of LEAF|GHOST tiles
 * sum u h^2 over LEAF-LEAF and LEAF-GHOST faces, ghost sums are added to their LEAF parents. */
cg_status cg_velocity_divergence(cg_problem* p, int32_t div_channel, int32_t propagate_ghosts);
/* AddGradientToFaceCenters (src/PoissonSolver.cu:144-209): ghost p from the parents, u += grad p
 * on the minus faces of LEAF|GHOST tiles (1/h, 1/(1.5 h) across coarse/fine faces, none across
 * NEUMANN cells), ghost velocities added to their LEAF parents with weight 1/4. */
cg_status cg_add_gradient(cg_problem* p, int32_t p_channel);
/* FluidEuler::project without the velocity boundary conditions: divergence of the face
 * velocities into CG_CH_B, MGPCG solve (every problem of a batched handle) into CG_CH_X, the
 * pressure gradient added to the face velocities; results[num_problems]. */
cg_status cg_project(cg_problem* p, const cg_params* prm, cg_result* results);

/* ---- AMGSolver (src/AMGSolver.h:17-53, src/AMGSolver.cu) ------------------
 * The stored-coefficient preconditioner.  cg_amg_prepare is
 * AMGSolver::prepareTypesAndCoeffs (CoarsenTypesAndAMGCoeffs,
 * src/AMGSolver.cu:110-209): ghost cell types from the parents, the LEAF/GHOST
 * coefficients (CalculateAMGCoefficients, :40-107; a LEAF-GHOST face weighs
 * 1/(1.5h) here), then, finest level first, the Galerkin sums of the children
 * into NONLEAF cells (retyped INTERIOR if a child is, else DIRICHLET) and into
 * LEAF cells with GHOST children.  The AMG cell types are kept beside the
 * handle's own (the GMG path keeps using those).  cg_amg_solve is
 * AMGSolver::solve (:722-855): the MGPCG loop of cg_solve with VCycleAMG
 * (:610-715) as the preconditioner (prm: level_iters, coarsest_iters,
 * one_over_alpha, prolong_coeff, sync_stride, pure_neumann, max_iters,
 * rel_tol, use_graph; FP32 storage only; smoother ignored).  One result per
 * problem of the handle.  cg_amg_coefficients: which = 0..2 the cells'
 * minus-face entries along x, y, z, 3 the diagonal, 4 the AMG cell types (as
 * floats), 512 per tile in the caller's tile order. */
cg_status cg_amg_prepare(cg_problem* p, float one_over_alpha);
cg_status cg_amg_solve(cg_problem* p, const cg_params* prm, cg_result* results);
cg_status cg_amg_coefficients(cg_problem* p, int32_t which, float* host);

/* GMGSolver::dampedJacobiSolve (src/GMGSolver.cu:449-529). */
cg_status cg_damped_jacobi_solve(cg_problem* p, int32_t max_iters, double rel_tol, float omega, cg_result* res);

/* GMGSolver::VCycle (src/GMGSolver.cu:383-447): x_channel <- M^-1 f_channel
 * on LEAF tiles (one V-cycle with the given smoother). */
cg_status cg_vcycle(cg_problem* p, int32_t x_channel, int32_t f_channel, int32_t level_iters,
                    int32_t coarsest_iters, int32_t smoother, float omega);

/* FullNegativeLaplacian (src/GMGSolver.h:13, src/GMGSolver.cu:203-213):
 * ghost values from parents, the 7-point operator on LEAF|GHOST tiles, ghost
 * rows summed into their LEAF parents.  calc_diag computes the diagonal. */
cg_status cg_full_negative_laplacian(cg_problem* p, int32_t x_channel, int32_t ax_channel, int32_t calc_diag);

/* NegativeLaplacianSameLevel128 (src/GMGSolver.h:7, src/GMGSolver.cu:145-200)
 * on the tiles of `level` (all levels if level < 0) whose subtree type at
 * subtree_level is in tile_types. */
cg_status cg_negative_laplacian_same_level(cg_problem* p, int32_t level, int32_t subtree_level, int32_t tile_types,
                                           int32_t x_channel, int32_t ax_channel, int32_t calc_diag);

/* One red-black Gauss-Seidel colour sweep on the LEAF|NONLEAF tiles of level
 * (NegativeLaplacianAndGaussSeidelSameLevel128, src/GMGSolver.cu:216-274). */
cg_status cg_gauss_seidel_color(cg_problem* p, int32_t level, int32_t x_channel, int32_t rhs_channel, int32_t color);
/* One damped-Jacobi sweep on the LEAF|NONLEAF tiles of level, the update of
 * src/GMGSolver.cu:508-525 with the level's operator. */
cg_status cg_jacobi_sweep(cg_problem* p, int32_t level, int32_t x_channel, int32_t rhs_channel, float omega);
/* ResidualAndRestrict (src/GMGSolver.cu:284-343) from level to level-1. */
cg_status cg_residual_restrict(cg_problem* p, int32_t level, int32_t x_channel, int32_t rhs_channel,
                               int32_t coarse_channel, float one_over_alpha);
/* ProlongateAndUpdate128 (src/GMGSolver.h:15, src/GMGSolver.cu:345-380) from level-1 to level. */
cg_status cg_prolongate(cg_problem* p, int32_t level, int32_t coarse_channel, int32_t fine_channel, float prolong_coeff);

/* Dot over all cells of the tiles of type in tile_types, FP64 accumulation in
 * a fixed order (Dot, src/PoissonGrid.cu:217-228). */
cg_status cg_dot(cg_problem* p, int32_t a_channel, int32_t b_channel, int32_t tile_types, double* out);
/* cg_dot with no allocation at all (no hipMalloc, hipFree or other allocating
 * call): the same result bit for bit.  No counterpart in the reference; for
 * callers that cannot allocate (tight serving loops, harnesses that forbid
 * it).  Its tile list, partials and result live in a fixed device array of
 * 3 MiB per GPU (held until the process ends), so the problem may have at most
 * CG_NOALLOC_MAX_TILES tiles (CG_ERR_INVALID beyond).  Every call on one GPU
 * shares that array: calls that overlap on one GPU (from several threads, or
 * several handles at once) race. */
#define CG_NOALLOC_MAX_TILES 262144
cg_status cg_dot_noalloc(cg_problem* p, int32_t a_channel, int32_t b_channel, int32_t tile_types, double* out);
/* ReCenterLeafVoxels (src/PoissonSolver.cu:356-381). */
cg_status cg_recenter_leaf_voxels(cg_problem* p, int32_t channel);

/* Average device duration (microseconds, hipEvents on the problem's stream)
 * of `reps` (at most 100,000) back-to-back launches of one kernel of the
 * solve on its real worklist, for roofline measurement; no reference
 * counterpart.
 * kernel: CG_KERNEL_GS_SWEEP (one colour sweep of `level`, V-cycle mode),
 * CG_KERNEL_APPLY (CG operator on LEAF+GHOST tiles), CG_KERNEL_RESTRICT and
 * CG_KERNEL_PROLONG (from/to `level`), CG_KERNEL_JACOBI (one sweep of `level`),
 * CG_KERNEL_UPDATE_XR (x/r update and |r|^2), CG_KERNEL_COARSE (the LDS-resident
 * coarse levels of a V-cycle, RB-GS 2+2, `level` bottom iterations), CG_KERNEL_EMPTY (an empty kernel of
 * `level` workgroups: the launch floor), CG_KERNEL_COPY (copy of one channel
 * over all tiles: 4096 bytes moved per tile, the reference point that scales
 * hardware byte counters), CG_KERNEL_GS_PROLONG (prolongation to `level` fused
 * with the first up-sweep colour pass), CG_KERNEL_GS_SWEEP_BF16 and
 * CG_KERNEL_GS_PROLONG_BF16 (the same with bfloat16 V-cycle storage),
 * CG_KERNEL_GS_SWEEP_F16 (the colour sweep with float16 storage),
 * CG_KERNEL_AMG_GS (one colour of the AMG stored-coefficient sweep on `level`;
 * after cg_amg_prepare), CG_KERNEL_GS_PAIR and CG_KERNEL_GS_PAIR_BF16 (one
 * red-black iteration of `level` in one launch, FP32 / bfloat16 storage).
 * Channels are overwritten. */
enum { CG_KERNEL_GS_SWEEP = 0, CG_KERNEL_APPLY = 1, CG_KERNEL_RESTRICT = 2, CG_KERNEL_PROLONG = 3,
       CG_KERNEL_JACOBI = 4, CG_KERNEL_UPDATE_XR = 5, CG_KERNEL_COARSE = 6, CG_KERNEL_EMPTY = 7,
       CG_KERNEL_COPY = 8, CG_KERNEL_GS_PROLONG = 9, CG_KERNEL_GS_SWEEP_BF16 = 10, CG_KERNEL_GS_PROLONG_BF16 = 11,
       CG_KERNEL_GS_SWEEP_F16 = 12, CG_KERNEL_AMG_GS = 13, CG_KERNEL_GS_PAIR = 14, CG_KERNEL_GS_PAIR_BF16 = 15 };
cg_status cg_benchmark_kernel(cg_problem* p, int32_t kernel, int32_t level, int32_t reps, double* avg_us,
                              int64_t* tiles);

/* CalcCellTypesFromLeafs (src/PoissonSolver.cu:7-61) on a host tree: fills
 * the NONLEAF and GHOST cell types of desc from its LEAF cell types.
 * cell_type_inout has 512*num_tiles entries in desc's tile order. */
cg_status cg_calc_cell_types_from_leafs(const cg_tree_desc* desc, uint8_t* cell_type_inout);

/* ---- Serving (no reference counterpart) ------------------------------------
 * A request loop over resident hierarchies.  cg_server_admit places a tree in
 * device memory sized from the HBM budget (CG_ERR_BUSY when it does not fit:
 * release instances and retry); it stays resident until released, and its
 * pressure stays on the device as the next request's warm start.
 * cg_server_submit queues a right-hand side for an idle instance;
 * cg_server_step enqueues one CG iteration over every running request and, one
 * iteration behind, starts queued requests and retires finished ones, so
 * requests join and leave at iteration boundaries while the others keep
 * running (continuous batching).  A finished request is read with
 * cg_server_next_result; each result equals the same request solved alone.
 * Warm start: x0 = the instance's last pressure, r0 = b - A x0, tolerance
 * relative to |b| (a cold start uses x0 = 0, as GMGSolver::solve).
 * cg_server_release returns at once: the instance's memory and slot are
 * reused once the device has passed the point of release (an admission that
 * needs them waits for that point).  One handle must not be used from two
 * threads at once. */
typedef struct cg_server cg_server;
typedef struct {
  int64_t budget_bytes;      /* device memory for channels, cell types and worklists */
  int32_t max_problems;      /* resident instances at most */
  cg_params params;          /* solver settings for every request (verify_f64: each result's
                                |b - A x| / |b| recomputed in FP64 before it is reported) */
  int32_t refine;            /* 1: every request is an FP64 iterative refinement -- the instance's x
                                is kept in FP64 across requests (the warm start), each step solves
                                A d = float(b - A x) with the MGPCG below half the tolerance and adds
                                d to x, and the request is reported once |b - A x| < rel_tol |b| in
                                FP64 (at most 5 steps); the pressure returned is float(x) (0) */
} cg_server_config;
cg_status cg_server_create(const cg_server_config* cfg, cg_server** out);
cg_status cg_server_destroy(cg_server* s);
cg_status cg_server_admit(cg_server* s, const cg_tree_desc* desc, int32_t* instance);
cg_status cg_server_release(cg_server* s, int32_t instance);
/* rhs: 512 floats per tile in the tile order of the admitted description, or
 * null to use the right-hand side already in the instance's device memory.
 * flags: CG_SUBMIT_WARM (start from the instance's last pressure),
 * CG_SUBMIT_NO_DOWNLOAD (the pressure stays on the device only). */
enum { CG_SUBMIT_WARM = 1, CG_SUBMIT_NO_DOWNLOAD = 2 };
cg_status cg_server_submit(cg_server* s, int32_t instance, const float* rhs, int32_t flags, int64_t tag);
/* Synthetic client: b = rho b + sigma u on the INTERIOR cells of an idle
 * instance, u seeded unit-variance noise, on the device (for benchmarks). */
cg_status cg_server_perturb_rhs(cg_server* s, int32_t instance, float rho, float sigma, uint64_t seed);
cg_status cg_server_step(cg_server* s, int32_t* num_active);
/* *got = 1 and the oldest finished request (its tag, instance, result and, if x
 * is not null, pressure in the caller's tile order) if one is ready; with wait
 * set, waits for that request's pressure to arrive. */
cg_status cg_server_next_result(cg_server* s, int32_t wait, int32_t* got, int64_t* tag, int32_t* instance, cg_result* res,
                                float* x);
cg_status cg_server_info(const cg_server* s, int64_t* capacity_tiles, int64_t* free_tiles, int32_t* num_instances);

/* ---- Synthetic scenes (the benchmark workloads) --------------------------
 * A scene is an octree built the reference's way (IterativeRefine with 2:1
 * face balance and ghost spawning, src/PoissonGrid.h:35-162,317-326 and
 * src/PoissonGrid.cu:30-60), typed (walls/solid), and a seeded face-velocity
 * field whose divergence is the right-hand side (src/PoissonSolver.cu:66-142). */
enum { CG_SCENE_UNIFORM = 0, CG_SCENE_SPHERE_BAND = 1, CG_SCENE_PLANE = 2, CG_SCENE_VORTEX = 3 };
typedef struct {
  int32_t kind;
  int32_t nroot[3];          /* root tiles per axis at level 0 */
  double h0;
  int32_t uniform_level;
  int32_t has_solid;         /* cells whose centre is inside the sphere are NEUMANN */
  double solid_c[3];
  double solid_r;
  double band_c[3];
  double band_r;
  double band_w[2];
  int32_t band_level[3];
  double plane_z;
  int32_t plane_level[2];
  int32_t n_tubes;
  double tube_r_lo, tube_r_hi;
  double vortex_thresh;
  int32_t vortex_level[2];
  int32_t face_bc[6];        /* x-,x+,y-,y+,z-,z+ : CG_NEUMANN or CG_DIRICHLET */
  uint64_t seed;
  int32_t num_layers;        /* maximum number of levels (10) */
} cg_scene_params;

/* The octree and cell types of a scene built on the device, every step a
 * kernel: refinement rounds to the scene's target levels with 2:1 face
 * balance (RefineLeafsOneStep / IterativeRefine, src/PoissonGrid.h:35-162,
 * 317-326), ghost spawning (SpawnGhostTiles, src/PoissonGrid.cu:30-60), the
 * walls and solid of the scene on LEAF cells, then CalcCellTypesFromLeafs
 * (src/PoissonSolver.cu:7-61) -- the tree cg_scene_create builds on the host,
 * tile for tile.  Scene kinds CG_SCENE_UNIFORM, CG_SCENE_SPHERE_BAND and
 * CG_SCENE_PLANE.  max_tiles: the build's capacity (<= 0: 4,194,304;
 * CG_ERR_NOMEM when the tree needs more).  The result is a tile set (tree and
 * cell types, no channels; cg_tileset_tree, cg_tileset_destroy). */
typedef struct cg_tileset cg_tileset;
cg_status cg_scene_tree_device(const cg_scene_params* sp, int64_t max_tiles, cg_tileset** out);

/* Device-side adaptation of a tree to a scene's target levels, as a
 * simulation step adapts its grid (src/FluidEuler.h:820-835;
 * RefineWithPointValue then CoarsenWithPointValue, src/FluidEuler.h:68-111):
 * with CG_ADAPT_REFINE, refinement rounds with 2:1 face balance until none
 * refines (RefineLeafsOneStep / SpawnGhostTiles, src/PoissonGrid.h:35-162,
 * src/PoissonGrid.cu:30-60); with CG_ADAPT_COARSEN, then CoarsenStep
 * (src/PoissonGrid.h:165-314) until a step deletes nothing; then ghost
 * spawning and the scene's cell types (walls and solid on LEAF cells,
 * CalcCellTypesFromLeafs).  The tree's LEAF and NONLEAF tiles are the input
 * (its ghosts and cell types are not read); its h0 must equal the scene's and
 * its tiles must lie in the scene's root tiles and layers.  Scene kinds as
 * cg_scene_tree_device; max_tiles likewise.  The result is a tile set. */
enum { CG_ADAPT_REFINE = 1, CG_ADAPT_COARSEN = 2 };
cg_status cg_tree_adapt_device(const cg_tree_desc* tree, const cg_scene_params* sp, int32_t ops, int64_t max_tiles,
                               cg_tileset** out);

typedef struct cg_scene cg_scene;
/* Build the tree, cell types and RHS on the host. */
cg_status cg_scene_create(const cg_scene_params* sp, cg_scene** out);
cg_status cg_scene_destroy(cg_scene* s);
/* Borrowed views valid until cg_scene_destroy: the tree (tiles in level,
 * Morton order) and the divergence RHS (512 floats per tile). */
cg_status cg_scene_tree(const cg_scene* s, cg_tree_desc* desc, const float** rhs);
/* The face velocities (3 x 512 floats per tile, minus faces, after ghost
 * propagation) from which the RHS was formed. */
cg_status cg_scene_velocity(const cg_scene* s, const float** uvw);

/* ---- Tile files ------------------------------------------------------------
 * A portable on-disk form of a tile octree (no reference counterpart: the
 * reference's own dump, WriteHAHostTileHolderToFile in
 * src/PoissonIOFunc.cu:38-75, writes raw structs with pointers and is tied to
 * its compiler's layout).  Format, version 1, little-endian: a 64-byte header
 * (magic "CIRTILES", uint32 version, uint32 channel count m, int64 tile count
 * n, float64 h0, uint64 FNV-1a 64 checksum of the rest, zero padding), then
 * int32 level[n], int32 coord[3n], uint8 type[n], uint8 cell_type[512n] and per
 * channel an int32 channel number (0-10, the reference's channels,
 * src/PoissonTile.h:12-23), an int32 element type (1 = float32) and
 * float32 data[512n]; cells i-major.  Readers check the magic, the version,
 * the size, the checksum and every type value.  (cg_tileset is declared with
 * cg_scene_tree_device above.) */
cg_status cg_tiles_save(const char* path, const cg_tree_desc* desc, int32_t num_channels, const int32_t* channel_ids,
                        const float* const* channels);
cg_status cg_tiles_load(const char* path, cg_tileset** out);
/* ReadHAHostTileHolderFromFile (src/PoissonIOFunc.cu:79-124) for a dump of
 * HAHostTileHolder<PoissonTile<float>> written on x86-64: every tile a level
 * list points to, with its cell types and the requested channels (e.g. x = 0,
 * b = 1, u/v/w = 6/7/8). */
cg_status cg_tiles_import_holder(const char* path, int32_t num_channels, const int32_t* channel_ids, cg_tileset** out);
/* Borrowed views valid until cg_tileset_destroy: the tree (tiles in file
 * order), the channel numbers and the channel data ([num_channels][512 n]). */
cg_status cg_tileset_tree(const cg_tileset* s, cg_tree_desc* desc);
cg_status cg_tileset_channels(const cg_tileset* s, int32_t* num_channels, const int32_t** channel_ids,
                              const float** data);
cg_status cg_tileset_destroy(cg_tileset* s);

#ifdef __cplusplus
}
#endif
#endif /* CIRRUS_MGPCG_H */
