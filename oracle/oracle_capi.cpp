// TEST INFRASTRUCTURE ONLY -- C API of the oracle (cirrus_oracle.hpp) for the
// Python tests and the CPU-baseline leg of bench.py (loaded with ctypes).
// All per-tile arrays are in canonical order: level, then Morton code of the
// tile coordinate; cells in the reference's i-major order (HAAccessor.h:32).
#include <omp.h>

#include <chrono>
#include <cstdio>

#include "cirrus_oracle.hpp"

using namespace orc;

struct OrcHandle {
  int fp64 = 0;
  std::unique_ptr<Problem<float>> pf;
  std::unique_ptr<Problem<double>> pd;
  std::unique_ptr<Problem<double>> mirror;  // FP64 copy of pf for the refinement's outer loop
};

// Plain-C mirror of orc::SceneParams (same field order and types).
extern "C" {
typedef struct {
  int kind;
  int nroot[3];
  double h0;
  int uniform_level;
  int has_solid;
  double solid_c[3];
  double solid_r;
  double band_c[3];
  double band_r;
  double band_w[2];
  int band_level[3];
  double plane_z;
  int plane_level[2];
  int n_tubes;
  double tube_r_lo, tube_r_hi;
  double vortex_thresh;
  int vortex_level[2];
  int face_bc[6];
  unsigned long long seed;
  int num_layers;
} orc_scene_params;
}

static SceneParams to_params(const orc_scene_params* c) {
  SceneParams p;
  p.kind = c->kind;
  for (int i = 0; i < 3; i++) {
    p.nroot[i] = c->nroot[i];
    p.solid_c[i] = c->solid_c[i];
    p.band_c[i] = c->band_c[i];
    p.band_level[i] = c->band_level[i];
  }
  p.h0 = c->h0;
  p.uniform_level = c->uniform_level;
  p.has_solid = c->has_solid;
  p.solid_r = c->solid_r;
  p.band_r = c->band_r;
  p.band_w[0] = c->band_w[0]; p.band_w[1] = c->band_w[1];
  p.plane_z = c->plane_z;
  p.plane_level[0] = c->plane_level[0]; p.plane_level[1] = c->plane_level[1];
  p.n_tubes = c->n_tubes;
  p.tube_r_lo = c->tube_r_lo; p.tube_r_hi = c->tube_r_hi;
  p.vortex_thresh = c->vortex_thresh;
  p.vortex_level[0] = c->vortex_level[0]; p.vortex_level[1] = c->vortex_level[1];
  for (int i = 0; i < 6; i++) p.face_bc[i] = c->face_bc[i];
  p.seed = c->seed;
  p.num_layers = c->num_layers;
  return p;
}

template <class F>
static auto with(void* h, F f) {
  OrcHandle* o = static_cast<OrcHandle*>(h);
  if (o->fp64) return f(*o->pd);
  return f(*o->pf);
}

extern "C" {

void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }

// Build the scene grid (refinement, ghosts, cell types, neighbours) and its
// divergence RHS in channel b.  fp64 selects T=double.
void* orc_create(const orc_scene_params* cp, int fp64) {
  OrcHandle* o = new OrcHandle();
  o->fp64 = fp64;
  SceneParams p = to_params(cp);
  try {
    if (fp64) { o->pd = BuildScene<double>(p); SceneRHS(*o->pd); }
    else { o->pf = BuildScene<float>(p); SceneRHS(*o->pf); }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "oracle: %s\n", e.what());
    delete o;
    return nullptr;
  }
  return o;
}
// The tree of scene `from` adapted to scene `to`'s target levels (AdaptScene;
// ops bit 0 refine, bit 1 coarsen), typed and given its RHS by scene `to`.
void* orc_create_adapted(const orc_scene_params* cfrom, const orc_scene_params* cto, int ops, int fp64) {
  OrcHandle* o = new OrcHandle();
  o->fp64 = fp64;
  SceneParams from = to_params(cfrom), to = to_params(cto);
  try {
    if (fp64) { o->pd = AdaptScene<double>(from, to, ops); SceneRHS(*o->pd); }
    else { o->pf = AdaptScene<float>(from, to, ops); SceneRHS(*o->pf); }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "oracle: %s\n", e.what());
    delete o;
    return nullptr;
  }
  return o;
}
void orc_destroy(void* h) { delete static_cast<OrcHandle*>(h); }

int orc_num_tiles(void* h) { return with(h, [](auto& P) { return int(P.canon.size()); }); }
int orc_num_leaf_tiles(void* h) { return with(h, [](auto& P) { return int(P.leafCanon.size()); }); }
int orc_max_level(void* h) { return with(h, [](auto& P) { return P.g().maxLevel; }); }

void orc_export_topology(void* h, int32_t* level, int32_t* coord, uint8_t* type, int32_t* nbr) {
  with(h, [&](auto& P) {
    auto& G = P.g();
    for (size_t i = 0; i < P.canon.size(); i++) {
      const Info& in = P.canon[i];
      level[i] = in.level;
      for (int d = 0; d < 3; d++) coord[3 * i + d] = in.coord[d];
      type[i] = in.type;
      auto& t = G.tile(in);
      for (int f = 0; f < 6; f++) nbr[6 * i + f] = t.nbr[f].empty() ? -1 : P.canonOfStore[t.nbr[f].tile];
    }
    return 0;
  });
}
void orc_export_celltypes(void* h, uint8_t* ct) {
  with(h, [&](auto& P) {
    for (size_t i = 0; i < P.canon.size(); i++) std::memcpy(ct + 512 * i, P.g().tile(P.canon[i]).ctype, 512);
    return 0;
  });
}
void orc_get_channel(void* h, int ch, double* out) {
  with(h, [&](auto& P) {
    for (size_t i = 0; i < P.canon.size(); i++) {
      auto& t = P.g().tile(P.canon[i]);
      for (int c = 0; c < 512; c++) out[512 * i + c] = double(t.ch[ch][c]);
    }
    return 0;
  });
}
void orc_set_channel(void* h, int ch, const double* in) {
  with(h, [&](auto& P) {
    using T = std::remove_reference_t<decltype(P.g().tile(P.canon[0]).ch[0][0])>;
    for (size_t i = 0; i < P.canon.size(); i++) {
      auto& t = P.g().tile(P.canon[i]);
      for (int c = 0; c < 512; c++) t.ch[ch][c] = T(in[512 * i + c]);
    }
    return 0;
  });
}
// Set the cell types of every tile (canonical order), e.g. for hand-made cases.
void orc_set_celltypes(void* h, const uint8_t* ct) {
  with(h, [&](auto& P) {
    for (size_t i = 0; i < P.canon.size(); i++) std::memcpy(P.g().tile(P.canon[i]).ctype, ct + 512 * i, 512);
    return 0;
  });
}
// Re-derive NONLEAF/GHOST cell types from the LEAF cell types.
void orc_calc_celltypes_from_leafs(void* h) { with(h, [&](auto& P) { CalcCellTypesFromLeafs(P.g()); return 0; }); }

void orc_full_negative_laplacian(void* h, int xch, int axch, int diag) {
  with(h, [&](auto& P) { Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P); S.FullNegativeLaplacian(xch, axch, diag != 0); return 0; });
}
// one NegativeLaplacianSameLevel pass on the tiles of `level` (or all tiles
// when level < 0) with subtree view `subtree_level` and tile-type mask `types`
void orc_negative_laplacian_same_level(void* h, int level, int subtree_level, int types, int xch, int axch, int diag) {
  with(h, [&](auto& P) {
    Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P);
    const auto& tiles = level < 0 ? P.g().all : P.g().arrays[level];
    S.NegLapSameLevel(tiles, subtree_level, uint8_t(types), xch, axch, diag != 0);
    return 0;
  });
}
void orc_gs_color(void* h, int level, int xch, int rhsch, int color) {
  with(h, [&](auto& P) { Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P); S.GSColor(level, xch, rhsch, color); return 0; });
}
void orc_jacobi_sweep(void* h, int level, int xch, int rhsch, double omega) {
  with(h, [&](auto& P) {
    using T = std::remove_reference_t<decltype(P.g().h0)>;
    Solver<T> S(P); S.JacobiSweep(level, xch, rhsch, T(omega)); return 0;
  });
}
void orc_residual_restrict(void* h, int level, int xch, int rhsch, int cch) {
  with(h, [&](auto& P) { Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P); S.ResidualAndRestrict(level, xch, rhsch, cch, LEAF | NONLEAF | GHOST); return 0; });
}
void orc_prolongate(void* h, int level, int cxch, int fxch) {
  with(h, [&](auto& P) { Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P); S.ProlongateAndUpdate(level, cxch, fxch); return 0; });
}
void orc_vcycle(void* h, int xch, int fch, int rhsch, int level_iters, int coarsest_iters, int smoother, double omega,
                int storage) {
  with(h, [&](auto& P) {
    using T = std::remove_reference_t<decltype(P.g().h0)>;
    Solver<T> S(P);
    S.vround = storage;
    S.VCycle(xch, fch, rhsch, level_iters, coarsest_iters, smoother, T(omega));
    return 0;
  });
}
double orc_dot(void* h, int ach, int bch, int types) {
  return with(h, [&](auto& P) { Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P); return S.Dot(ach, bch, uint8_t(types)); });
}
void orc_recenter(void* h, int ch) {
  with(h, [&](auto& P) { Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P); S.ReCenterLeafVoxels(ch); return 0; });
}
// GMGSolver::solve.  out[0]=relative residual (recursive r), out[1]=|b|^2,
// out[2]=|r|^2, out[3]=wall seconds of the solve.  Returns the iteration count.
// storage 1: V-cycle stores rounded to bfloat16; 2: to float16, with the
// right-hand side scaled by a power of two (FP32 handles only).
int orc_solve(void* h, int max_iters, double tol, int level_iters, int coarsest_iters, int pure_neumann,
              int smoother, double omega, int sync_stride, int storage, double* out) {
  return with(h, [&](auto& P) {
    using T = std::remove_reference_t<decltype(P.g().h0)>;
    Solver<T> S(P);
    S.vround = storage;
    auto t0 = std::chrono::steady_clock::now();
    auto r = S.solve(max_iters, tol, level_iters, coarsest_iters, pure_neumann != 0, smoother, T(omega), sync_stride);
    auto t1 = std::chrono::steady_clock::now();
    if (out) { out[0] = r.rel_res; out[1] = r.bb; out[2] = r.rr; out[3] = std::chrono::duration<double>(t1 - t0).count(); }
    return r.iters;
  });
}
// AMGSolver (AMGSolver.cu): prepareTypesAndCoeffs, then solve / one V-cycle
// (z from r) / the stored-coefficient operator on one level; the coefficients
// out in canonical tile order (a = 0..2 minus-face off-diagonals, 3 diagonal).
void orc_amg_prepare(void* h) {
  with(h, [&](auto& P) { Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P); S.AMGPrepare(); return 0; });
}
int orc_amg_solve(void* h, int max_iters, double tol, int level_iters, int coarsest_iters, int pure_neumann,
                  int sync_stride, double* out) {
  return with(h, [&](auto& P) {
    using T = std::remove_reference_t<decltype(P.g().h0)>;
    if (P.amg[0].empty()) throw std::runtime_error("orc_amg_solve: orc_amg_prepare first");
    Solver<T> S(P);
    S.amg = true;
    auto r = S.solve(max_iters, tol, level_iters, coarsest_iters, pure_neumann != 0, SMOOTH_RBGS, T(2) / T(3), sync_stride);
    if (out) { out[0] = r.rel_res; out[1] = r.bb; out[2] = r.rr; }
    return r.iters;
  });
}
void orc_amg_vcycle(void* h, int level_iters, int coarsest_iters) {
  with(h, [&](auto& P) {
    Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P);
    S.VCycleAMG(CH_Z, CH_R, CH_AP, CH_TMP, level_iters, coarsest_iters);
    return 0;
  });
}
void orc_amg_apply(void* h, int level, int xch, int axch) {
  with(h, [&](auto& P) {
    Solver<std::remove_reference_t<decltype(P.g().h0)>> S(P);
    S.NegLapAMG(level, LEAF | NONLEAF | GHOST, xch, axch);
    return 0;
  });
}
void orc_amg_coeffs(void* h, int a, double* out) {
  with(h, [&](auto& P) {
    for (size_t i = 0; i < P.canon.size(); i++) {
      const auto& c = P.amg[a][size_t(P.canon[i].tile)];
      for (int k = 0; k < 512; k++) out[512 * i + k] = double(c[k]);
    }
    return 0;
  });
}
// The projection's neighbours of the solve (FluidEuler.h:581-612): ghost
// velocities from the parents, the volume divergence into div_ch, and the
// pressure gradient added to the face velocities (channels 6-8).
void orc_velocity_divergence(void* h, int div_ch, int propagate) {
  with(h, [&](auto& P) {
    if (propagate) PropagateVelocityToGhosts(P.g(), CH_U);
    VelocityVolumeDivergenceOnLeafs(P.g(), CH_U, div_ch);
    return 0;
  });
}
void orc_add_gradient(void* h, int p_ch) {
  with(h, [&](auto& P) { AddGradientToFaceCenters(P.g(), p_ch, CH_U); return 0; });
}

// Mixed-precision iterative refinement on an FP32 handle (SolveRefined).
// out[0]=relative FP64 residual, [1]=|b|^2, [2]=|r|^2, [3]=CG iterations,
// [4]=converged, [5]=wall seconds.  Returns the number of refinement steps;
// -1 on an FP64 handle.
int orc_solve_refined(void* h, double tol, double inner_tol, int max_outer, int max_iters, int level_iters,
                      int coarsest_iters, int pure_neumann, int smoother, double omega, int sync_stride, int storage,
                      double* out) {
  OrcHandle* o = static_cast<OrcHandle*>(h);
  if (o->fp64) return -1;
  auto t0 = std::chrono::steady_clock::now();
  o->mirror = ConvertProblem<double>(*o->pf);
  auto r = SolveRefined(*o->pf, *o->mirror, tol, inner_tol, max_outer, max_iters, level_iters, coarsest_iters,
                        pure_neumann != 0, smoother, float(omega), sync_stride, storage);
  auto t1 = std::chrono::steady_clock::now();
  if (out) {
    out[0] = r.rel_res; out[1] = r.bb; out[2] = r.rr; out[3] = r.iters; out[4] = r.converged;
    out[5] = std::chrono::duration<double>(t1 - t0).count();
  }
  return r.outer;
}
// The FP64 x of the last orc_solve_refined (canonical order); 0 if none ran.
int orc_get_refined_x(void* h, double* out) {
  OrcHandle* o = static_cast<OrcHandle*>(h);
  if (!o->mirror) return 0;
  auto& P = *o->mirror;
  for (size_t i = 0; i < P.canon.size(); i++) std::memcpy(out + 512 * i, P.g().tile(P.canon[i]).ch[CH_X], 512 * sizeof(double));
  return 1;
}
// Round FP32 values to bfloat16 as the V-cycle stores do (for pinning).
void orc_fp16_round(const float* in, float* out, long n) {
  for (long i = 0; i < n; i++) out[i] = fp16_round(in[i]);
}
int orc_fp16_scale_exp(double bb, double cells) { return fp16_scale_exp(bb, cells); }
void orc_bf16_round(const float* in, float* out, long n) {
  for (long i = 0; i < n; i++) out[i] = bf16_round(in[i]);
}

int orc_damped_jacobi_solve(void* h, int max_iters, double tol, double omega, double* out) {
  return with(h, [&](auto& P) {
    using T = std::remove_reference_t<decltype(P.g().h0)>;
    Solver<T> S(P);
    auto r = S.dampedJacobiSolve(max_iters, tol, T(omega));
    if (out) { out[0] = r.rel_res; out[1] = r.bb; out[2] = r.rr; }
    return r.iters;
  });
}

}  // extern "C"
