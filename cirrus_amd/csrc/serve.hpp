// Serving engine of the MGPCG library (host side; included by mgpcg.hip).
//
// Instances are resident hierarchies: each admitted tree gets a slab of tiles,
// leaf slots and ghost slots in device arrays sized from an HBM budget, its own
// tile records on the device and a pinned staging buffer.  A request is a right
// hand side for an instance (optionally warm-started from the instance's
// current pressure).  Every cg_server_step enqueues one CG iteration over the
// instances whose request is running -- their worklists assembled on the device
// whenever that set changes -- so requests that arrive while others run join at
// the next iteration boundary (continuous batching) and each result is the same
// as solving the request alone.
//
// verify_f64: each finished request's |b - A x| / |b| is recomputed in FP64 from
// the instance's kept b before the request is reported.
// refine: every request is an FP64 iterative refinement -- the instance's x is
// kept in FP64 across requests; a request computes r = b - A x in FP64, solves
// A d = float(r) with the MGPCG until its residual is below half the tolerance,
// adds d to x, and checks |b - A x| in FP64; a request whose check fails runs
// another such step (kMaxRefineSteps at most) before it is reported.
#pragma once

#include <deque>
#include <map>

#include "range_alloc.hpp"

namespace {

enum InstState { INST_IDLE = 0, INST_QUEUED = 1, INST_ACTIVE = 2, INST_DONE = 3, INST_VERIFY = 4 };
constexpr int kMaxRefineSteps = 5;

struct Instance {
  int slot = -1;
  int64_t tile_begin = 0, leaf_begin = 0, ghost_begin = 0;
  int ntiles = 0, nleaf = 0, nghost = 0;
  int max_level = 0, C = -1;
  bool rz_ok = false;    // every LEAF tile is on a grid-wide level >= 1
  std::vector<int> ext;  // internal tile i -> caller index
  TileRec* d_recs = nullptr;
  DList leaf, gather, applyl, ghosts, all;
  std::vector<DList> actg, lvlg;
  CoarseSlot* d_slots = nullptr;
  CoarsePlan plan;
  float* h_buf = nullptr;  // pinned: rhs on the way in, pressure on the way out
  hipEvent_t ev = nullptr;
  hipEvent_t ev_ver = nullptr;  // the FP64 check's result has arrived
  int state = INST_IDLE;
  int steps = 0;        // refinement steps run for the current request
  int ir_cont = 0;      // the next start continues a refinement (its rhs is in place)
  double bb_b = 0.0;    // |b|^2 of the current request (refinement)
  int iters_acc = 0;    // CG iterations of the request's earlier steps
  int warm = 0;
  int dev_rhs = 0;      // the right-hand side is already on the device
  int no_download = 0;  // the pressure stays on the device
  int64_t tag = 0;
  std::chrono::steady_clock::time_point t_submit;
  cg_result res{};
  const TileRec* L(const DList& d) const { return d_recs + d.off; }
};

}  // namespace

struct cg_server {
  cg_params prm{};
  int refine = 0;  // every request is an FP64 iterative refinement
  VParams v;
  int max_probs = 0;
  int64_t cap_tiles = 0, cap_leaf = 0, cap_ghost = 0;
  RangeAlloc tiles_alloc, leaf_alloc, ghost_alloc;
  std::vector<int> free_slots;
  std::vector<std::unique_ptr<Instance>> inst;  // by slot
  std::vector<int> active;                      // slots whose request runs, assembly order
  std::vector<int> snap_active[2];              // active set of the iteration each snapshot follows
  bool dirty = false;
  std::deque<int> done_q;                       // slots with a finished request, in completion order
  int64_t step_no = 0;
  bool snap_pending[2] = {false, false};

  // device
  cg_problem view;  // the kernels' view of the assembled active set (owner = false)
  uint8_t* d_ctype = nullptr;
  uint8_t* d_fmask = nullptr;   // face masks, then (same allocation) the colour-split masks d_fsplit
  uint32_t* d_fsplit = nullptr;
  float* d_diag = nullptr;
  float* d_ch[CG_NUM_CHANNELS] = {};
  float* d_zb = nullptr;
  float* d_b = nullptr;  // each instance's right-hand side, kept across its requests
  double* d_x64 = nullptr;  // FP64 check of finished requests (verify_f64): x, residual, per-slot |r|^2
  double* d_r64 = nullptr;
  double* d_ver = nullptr;
  double* h_ver = nullptr;  // pinned
  bf16* d_zh = nullptr;  // bfloat16 V-cycle vectors (storage CG_STORE_BF16)
  bf16* d_rhsvh = nullptr;
  bf16* d_zbh = nullptr;
  float* d_gsum = nullptr;
  ProbState* d_ps = nullptr;
  int32_t* d_dflag = nullptr;
  int32_t* d_dflag_init = nullptr;
  int32_t* d_pmap = nullptr;
  int32_t* d_newmap = nullptr;
  double* d_part = nullptr;
  double* d_part2 = nullptr;
  TileRec* d_asm = nullptr;
  int64_t asm_cap = 0;
  CoarsePlan* d_plans = nullptr;
  const TileRec** d_srcs = nullptr;
  int32_t* d_offs = nullptr;
  int32_t* d_counts = nullptr;
  int runs_cap = 0;
  hipStream_t stream = nullptr;
  // pinned host memory: ProbState snapshots, per-slot ProbState staging, per-step small tables
  ProbState* h_snap[2] = {nullptr, nullptr};
  hipEvent_t ev_snap[2] = {nullptr, nullptr};
  ProbState* h_ps_stage = nullptr;
  char* h_tab[2] = {nullptr, nullptr};
  size_t tab_cap = 0;
  ~cg_server();
};

cg_server::~cg_server() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& I : inst) {
    if (!I) continue;
    if (I->d_recs) (void)hipFree(I->d_recs);
    if (I->d_slots) (void)hipFree(I->d_slots);
    if (I->h_buf) (void)hipHostFree(I->h_buf);
    if (I->ev) (void)hipEventDestroy(I->ev);
    if (I->ev_ver) (void)hipEventDestroy(I->ev_ver);
  }
  if (h_ver) (void)hipHostFree(h_ver);
  void* ptrs[] = {d_ctype, d_fmask, d_diag, d_zb, d_b, d_x64, d_r64, d_ver, d_zh, d_rhsvh, d_zbh, d_gsum, d_ps, d_dflag, d_dflag_init, d_pmap, d_newmap, d_part, d_part2, d_asm,
                  d_plans, (void*)d_srcs, d_offs, d_counts};
  for (void* q : ptrs) if (q) (void)hipFree(q);
  for (auto q : d_ch) if (q) (void)hipFree(q);
  for (auto e : ev_snap) if (e) (void)hipEventDestroy(e);
  for (auto q : h_snap) if (q) (void)hipHostFree(q);
  for (auto q : h_tab) if (q) (void)hipHostFree(q);
  if (h_ps_stage) (void)hipHostFree(h_ps_stage);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

// Device bytes one tile costs: 6 channels, the V-cycle scratch, the kept rhs
// and the colour sweep's diagonal (FP32), cell types and face masks, ghost row sums, dot partials and room
// for its records in the assembled worklists; BF16 storage adds three bfloat16
// vectors, verify_f64 two FP64 vectors.
constexpr int kRecKindsPerTile = 5;  // leaf, gather, applyl, actg, lvlg: at most 5 records per tile
constexpr int64_t kBytesPerTile =
    (CG_NUM_CHANNELS + 3) * 512 * 4 + 2 * 512 + 64 * 4 + 2 * 8 + kRecKindsPerTile * int64_t(sizeof(TileRec));

void server_create(const cg_server_config* cfg, cg_server& S) {
  if (!cfg || cfg->budget_bytes <= 0 || cfg->max_problems <= 0) fail(CG_ERR_INVALID, "cg_server_create: bad configuration");
  const cg_params& prm = cfg->params;
  check_params(prm);
  if (prm.storage == CG_STORE_F16) fail(CG_ERR_INVALID, "cg_server_create: float16 storage is not available to the server");
  S.prm = prm;
  S.v.nu = prm.level_iters;
  S.v.coarsest = prm.coarsest_iters;
  S.v.smoother = prm.smoother;
  S.v.omega = prm.omega;
  S.v.ooa = prm.one_over_alpha;
  S.v.pcoef = prm.prolong_coeff;
  S.max_probs = cfg->max_problems;
  S.refine = cfg->refine ? 1 : 0;
  const bool f64 = prm.verify_f64 || S.refine;
  const int64_t per_tile = kBytesPerTile + (prm.storage == CG_STORE_BF16 ? 3 * 512 * int64_t(sizeof(bf16)) : 0) +
                           (f64 ? 2 * 512 * int64_t(sizeof(double)) : 0);
  S.cap_tiles = std::max<int64_t>(64, cfg->budget_bytes / per_tile);
  if (S.cap_tiles * 512 > int64_t(INT32_MAX)) S.cap_tiles = int64_t(INT32_MAX) / 512;  // cell offsets stay 32-bit safe
  S.cap_leaf = S.cap_tiles;
  S.cap_ghost = S.cap_tiles;
  S.tiles_alloc.init(S.cap_tiles);
  S.leaf_alloc.init(S.cap_leaf);
  S.ghost_alloc.init(S.cap_ghost);
  S.inst.resize(S.max_probs);
  for (int q = S.max_probs - 1; q >= 0; q--) S.free_slots.push_back(q);

  HIPC(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
  const size_t cells = size_t(S.cap_tiles) * 512;
  HIPC(hipMalloc(&S.d_ctype, cells));
  HIPC(hipMalloc(&S.d_fmask, 2 * cells));
  S.d_fsplit = reinterpret_cast<uint32_t*>(S.d_fmask + cells);
  HIPC(hipMalloc(&S.d_diag, cells * sizeof(float)));
  for (int c = 0; c < CG_NUM_CHANNELS; c++) HIPC(hipMalloc(&S.d_ch[c], cells * sizeof(float)));
  HIPC(hipMalloc(&S.d_zb, cells * sizeof(float)));
  HIPC(hipMalloc(&S.d_b, cells * sizeof(float)));
  HIPC(hipMemsetAsync(S.d_b, 0, cells * sizeof(float), S.stream));
  if (f64) {
    HIPC(hipMalloc(&S.d_x64, cells * sizeof(double)));
    HIPC(hipMalloc(&S.d_r64, cells * sizeof(double)));
    HIPC(hipMemsetAsync(S.d_x64, 0, cells * sizeof(double), S.stream));
    HIPC(hipMemsetAsync(S.d_r64, 0, cells * sizeof(double), S.stream));
    HIPC(hipMalloc(&S.d_ver, 2 * sizeof(double) * S.max_probs));
    HIPC(hipHostMalloc(&S.h_ver, 2 * sizeof(double) * S.max_probs, hipHostMallocDefault));
  }
  if (S.prm.storage == CG_STORE_BF16) {
    bf16** bufs[3] = {&S.d_zh, &S.d_rhsvh, &S.d_zbh};
    for (bf16** b : bufs) {
      HIPC(hipMalloc(b, cells * sizeof(bf16)));
      HIPC(hipMemsetAsync(*b, 0, cells * sizeof(bf16), S.stream));
    }
  }
  HIPC(hipMalloc(&S.d_gsum, size_t(S.cap_ghost) * 64 * sizeof(float)));
  HIPC(hipMalloc(&S.d_ps, sizeof(ProbState) * S.max_probs));
  HIPC(hipMemsetAsync(S.d_ps, 0, sizeof(ProbState) * S.max_probs, S.stream));
  HIPC(hipMalloc(&S.d_dflag, sizeof(int32_t) * S.max_probs));
  HIPC(hipMalloc(&S.d_dflag_init, sizeof(int32_t) * S.max_probs));
  HIPC(hipMalloc(&S.d_pmap, sizeof(int32_t) * S.max_probs));
  HIPC(hipMalloc(&S.d_newmap, sizeof(int32_t) * S.max_probs));
  HIPC(hipMalloc(&S.d_part, sizeof(double) * S.cap_leaf));
  HIPC(hipMalloc(&S.d_part2, sizeof(double) * S.cap_leaf));
  S.asm_cap = S.cap_tiles * kRecKindsPerTile;
  HIPC(hipMalloc(&S.d_asm, sizeof(TileRec) * S.asm_cap));
  HIPC(hipMalloc(&S.d_plans, sizeof(CoarsePlan) * S.max_probs));
  S.runs_cap = S.max_probs * (3 + 2 * 32);
  HIPC(hipMalloc(&S.d_srcs, sizeof(TileRec*) * S.runs_cap));
  HIPC(hipMalloc(&S.d_offs, sizeof(int32_t) * S.runs_cap));
  HIPC(hipMalloc(&S.d_counts, sizeof(int32_t) * S.runs_cap));
  for (int k = 0; k < 2; k++) {
    HIPC(hipHostMalloc(&S.h_snap[k], sizeof(ProbState) * S.max_probs, hipHostMallocDefault));
    HIPC(hipEventCreateWithFlags(&S.ev_snap[k], hipEventDisableTiming));
  }
  HIPC(hipHostMalloc(&S.h_ps_stage, sizeof(ProbState) * S.max_probs, hipHostMallocDefault));
  // per-step tables: runs (srcs, offs, counts), active map, new map, coarse plans
  S.tab_cap = size_t(S.runs_cap) * (sizeof(TileRec*) + 8) + size_t(S.max_probs) * (8 + sizeof(CoarsePlan)) + 4096;
  for (int k = 0; k < 2; k++) HIPC(hipHostMalloc(&S.h_tab[k], S.tab_cap, hipHostMallocDefault));

  // the kernels' view: shares the server's arrays
  cg_problem& P = S.view;
  P.owner = false;
  P.stream = S.stream;
  P.d_ctype = S.d_ctype;
  P.d_fmask = S.d_fmask;
  P.d_fsplit = S.d_fsplit;
  P.d_diag = S.d_diag;
  for (int c = 0; c < CG_NUM_CHANNELS; c++) P.d_ch[c] = S.d_ch[c];
  P.d_zb = S.d_zb;
  P.d_zh = S.d_zh;
  P.d_rhsvh = S.d_rhsvh;
  P.d_zbh = S.d_zbh;
  P.d_gsum = S.d_gsum;
  P.d_ps = S.d_ps;
  P.d_dflag = S.d_dflag;
  P.d_pmap = S.d_pmap;
  P.d_part = S.d_part;
  P.d_part2 = S.d_part2;
  P.d_lists = S.d_asm;
  P.d_plans = S.d_plans;
  P.nprob = 0;
  HIPC(hipStreamSynchronize(S.stream));  // the zero fills above are done before any request
}

void server_admit(cg_server& S, const cg_tree_desc* d, int32_t* out) {
  if (!d || !out) fail(CG_ERR_INVALID, "cg_server_admit: null argument");
  if (d->num_tiles <= 0) fail(CG_ERR_INVALID, "cg_server_admit: empty tree");
  if (S.free_slots.empty()) fail(CG_ERR_BUSY, "cg_server_admit: no free problem slot");
  const int n = d->num_tiles;
  const int64_t tb = S.tiles_alloc.alloc(n);
  if (tb < 0) fail(CG_ERR_BUSY, "cg_server_admit: HBM budget exhausted");
  // leaf and ghost counts are known only after the build: reserve the tile count
  // first, trim after
  const int64_t lb = S.leaf_alloc.alloc(n), gb = S.ghost_alloc.alloc(n);
  if (lb < 0 || gb < 0) {
    S.tiles_alloc.release(tb, n);
    if (lb >= 0) S.leaf_alloc.release(lb, n);
    if (gb >= 0) S.ghost_alloc.release(gb, n);
    fail(CG_ERR_BUSY, "cg_server_admit: HBM budget exhausted");
  }
  const int slot = S.free_slots.back();
  auto I = std::make_unique<Instance>();
  TreeBuild B;
  try {
    build_tree(d, slot, int(tb), int(lb), int(gb), B);
  } catch (...) {
    S.tiles_alloc.release(tb, n);
    S.leaf_alloc.release(lb, n);
    S.ghost_alloc.release(gb, n);
    throw;
  }
  S.leaf_alloc.release(lb + B.nleaf, n - B.nleaf);
  S.ghost_alloc.release(gb + B.nghost, n - B.nghost);
  S.free_slots.pop_back();
  I->slot = slot;
  I->tile_begin = tb;
  I->leaf_begin = lb;
  I->ghost_begin = gb;
  I->ntiles = n;
  I->nleaf = B.nleaf;
  I->nghost = B.nghost;
  I->ext = B.ext;
  I->C = coarse_depth(B.lev, B.ty);
  I->max_level = 0;
  for (int l : B.lev) I->max_level = std::max(I->max_level, l);
  // records of the instance's lists, back to back
  std::vector<TileRec> W;
  auto add_list = [&](auto pred) {
    DList dl;
    dl.off = int(W.size());
    for (int i = 0; i < n; i++) if (pred(i)) W.push_back(B.tiles[i]);
    dl.n = int(W.size()) - dl.off;
    return dl;
  };
  auto grid_level = [&](int i) { return B.lev[i] > I->C; };
  I->leaf = add_list([&](int i) { return B.ty[i] == LEAF; });
  I->rz_ok = true;
  for (int i = 0; i < n; i++)
    if (B.ty[i] == LEAF && !(B.lev[i] >= 1 && grid_level(i))) I->rz_ok = false;
  I->gather = add_list([&](int i) { return B.ty[i] == LEAF && tr_has_gchild(B.tiles[i]); });
  I->applyl = add_list([&](int i) { return B.ty[i] == LEAF || B.ty[i] == GHOST; });
  I->ghosts = add_list([&](int i) { return B.ty[i] == GHOST; });
  I->all = add_list([&](int) { return true; });
  I->actg.assign(I->max_level + 1, DList());
  I->lvlg.assign(I->max_level + 1, DList());
  for (int l = 0; l <= I->max_level; l++) {
    I->actg[l] = add_list([&](int i) { return B.lev[i] == l && B.ty[i] != GHOST && grid_level(i); });
    I->lvlg[l] = add_list([&](int i) { return B.lev[i] == l && grid_level(i); });
  }
  HIPC(hipMalloc(&I->d_recs, sizeof(TileRec) * W.size()));
  HIPC(hipMemcpy(I->d_recs, W.data(), sizeof(TileRec) * W.size(), hipMemcpyHostToDevice));
  if (I->C >= 0) {
    std::vector<CoarseSlot> slots;
    coarse_plan(B.tiles, int(tb), I->C, slot, I->plan, slots);
    HIPC(hipMalloc(&I->d_slots, sizeof(CoarseSlot) * slots.size()));
    HIPC(hipMemcpy(I->d_slots, slots.data(), sizeof(CoarseSlot) * slots.size(), hipMemcpyHostToDevice));
    I->plan.slots = I->d_slots;
  }
  // cell types in internal order; x = 0 so that a first warm start is a cold one
  std::vector<uint8_t> ct(size_t(n) * 512);
  for (int i = 0; i < n; i++) std::memcpy(&ct[size_t(i) * 512], d->cell_type + size_t(B.ext[i]) * 512, 512);
  HIPC(hipMemcpy(S.d_ctype + size_t(tb) * 512, ct.data(), ct.size(), hipMemcpyHostToDevice));
  // face masks on the server's stream, ahead of the instance's first request
  k_face_mask<<<I->all.n, 64, 0, S.stream>>>(S.d_ctype, S.d_fmask, S.d_fsplit, I->L(I->all), I->all.n);
  k_diag<<<I->all.n, 64, 0, S.stream>>>(S.view.ctx(), I->L(I->all), I->all.n, S.d_diag);
  HIPC(hipGetLastError());
  HIPC(hipMemsetAsync(S.d_ch[CG_CH_X] + size_t(tb) * 512, 0, size_t(n) * 512 * sizeof(float), S.stream));
  HIPC(hipMemsetAsync(S.d_b + size_t(tb) * 512, 0, size_t(n) * 512 * sizeof(float), S.stream));
  if (S.d_x64) HIPC(hipMemsetAsync(S.d_x64 + size_t(tb) * 512, 0, size_t(n) * 512 * sizeof(double), S.stream));
  HIPC(hipHostMalloc(&I->h_buf, size_t(n) * 512 * sizeof(float), hipHostMallocDefault));
  HIPC(hipEventCreateWithFlags(&I->ev, hipEventDisableTiming));
  HIPC(hipEventCreateWithFlags(&I->ev_ver, hipEventDisableTiming));
  *out = slot;
  S.inst[slot] = std::move(I);
}

Instance& instance(cg_server& S, int32_t id) {
  if (id < 0 || id >= S.max_probs || !S.inst[id]) fail(CG_ERR_INVALID, "unknown instance");
  return *S.inst[id];
}

void server_release(cg_server& S, int32_t id) {
  Instance& I = instance(S, id);
  if (I.state == INST_QUEUED || I.state == INST_ACTIVE || I.state == INST_VERIFY)
    fail(CG_ERR_STATE, "cg_server_release: the instance has a request running");
  HIPC(hipStreamSynchronize(S.stream));  // no kernel may still read its slab
  if (I.state == INST_DONE)
    for (auto it = S.done_q.begin(); it != S.done_q.end(); ++it)
      if (*it == id) { S.done_q.erase(it); break; }
  if (!S.tiles_alloc.release(I.tile_begin, I.ntiles) || !S.leaf_alloc.release(I.leaf_begin, I.nleaf) ||
      !S.ghost_alloc.release(I.ghost_begin, I.nghost))
    fail(CG_ERR_STATE, "cg_server_release: the instance's ranges were already free");
  HIPC(hipFree(I.d_recs));
  I.d_recs = nullptr;
  if (I.d_slots) HIPC(hipFree(I.d_slots));
  I.d_slots = nullptr;
  HIPC(hipHostFree(I.h_buf));
  I.h_buf = nullptr;
  HIPC(hipEventDestroy(I.ev));
  I.ev = nullptr;
  HIPC(hipEventDestroy(I.ev_ver));
  I.ev_ver = nullptr;
  S.inst[id].reset();
  S.free_slots.push_back(id);
}

void server_submit(cg_server& S, int32_t id, const float* rhs, int32_t flags, int64_t tag) {
  Instance& I = instance(S, id);
  if (I.state != INST_IDLE) fail(CG_ERR_STATE, "cg_server_submit: the instance has a request running or an unread result");
  if (rhs)
    for (int i = 0; i < I.ntiles; i++) std::memcpy(I.h_buf + size_t(i) * 512, rhs + size_t(I.ext[i]) * 512, 512 * sizeof(float));
  I.dev_rhs = rhs ? 0 : 1;
  I.no_download = (flags & CG_SUBMIT_NO_DOWNLOAD) ? 1 : 0;
  I.warm = (flags & CG_SUBMIT_WARM) ? 1 : 0;
  I.tag = tag;
  I.steps = 0;
  I.ir_cont = 0;
  I.iters_acc = 0;
  I.t_submit = std::chrono::steady_clock::now();
  I.state = INST_QUEUED;
}

// A Ctx whose ProbState array starts at the instance's slot, for one-workgroup
// per-problem reductions.
Ctx one_problem(cg_server& S, const Instance& I) {
  Ctx cv = S.view.ctx();
  cv.ps = S.d_ps + I.slot;
  cv.pmap = nullptr;
  return cv;
}

// r64 = b - A x64 on the instance's LEAF tiles in FP64 (re-centred for pure
// Neumann) with |r64|^2 partials; write_r: float(r64) into the r channel.
void instance_residual64(cg_server& S, Instance& I, bool write_r) {
  const hipStream_t st = S.stream;
  const Ctx c = S.view.ctx();
  const unsigned gl = grid_for(I.leaf.n, kSwizzle);
  if (I.ghosts.n) k_propagate_ghosts64<<<I.ghosts.n, 64, 0, st>>>(c, I.L(I.ghosts), I.ghosts.n, S.d_x64);
  k_apply64<<<I.applyl.n, 64, 0, st>>>(c, I.L(I.applyl), I.applyl.n, S.d_x64, S.d_r64);
  if (I.ghosts.n) k_accumulate_ghosts64<<<I.ghosts.n, 64, 0, st>>>(c, I.L(I.ghosts), I.ghosts.n, S.d_r64);
  if (S.prm.pure_neumann) {
    k_ir_residual<<<gl, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, S.d_b, S.d_r64, 1);
    k_finalize<<<1, FIN_THREADS, 0, st>>>(one_problem(S, I), FIN_MEAN, 0);
    k_ir_recenter<<<gl, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, S.d_r64, write_r ? 1 : 0);
  } else {
    k_ir_residual<<<gl, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, S.d_b, S.d_r64, write_r ? 0 : 2);
  }
}

// Start queued requests: ProbState, rhs upload, x0 (zero for a cold start),
// r0 = b - A x0 on the instance's own lists, |b|^2.
void server_start_queued(cg_server& S, int k) {
  std::vector<int> fresh;
  for (int q = 0; q < S.max_probs; q++)
    if (S.inst[q] && S.inst[q]->state == INST_QUEUED) fresh.push_back(q);
  if (fresh.empty()) return;
  const hipStream_t st = S.stream;
  int32_t* hmap = reinterpret_cast<int32_t*>(S.h_tab[k] + S.tab_cap - 2048);  // tail of this step's table
  const int nf = std::min<int>(int(fresh.size()), 2048 / 4);
  Ctx c = S.view.ctx();
  for (int f = 0; f < nf; f++) {
    Instance& I = *S.inst[fresh[f]];
    ProbState& ps = S.h_ps_stage[I.slot];
    std::memset(&ps, 0, sizeof(ps));
    const double tol2 = S.prm.rel_tol * S.prm.rel_tol;
    // refinement: the inner solve stops below half the tolerance of |b|
    ps.tol2 = S.refine ? 0.0 : tol2;
    ps.tol_abs2 = S.refine ? 0.25 * tol2 : 0.0;
    if (S.refine && I.ir_cont) {
      ps.bb_b = I.bb_b;
      ps.thr_abs = 0.25 * tol2 * I.bb_b;
    }
    ps.max_iters = S.prm.max_iters;
    ps.leaf_begin = int32_t(I.leaf_begin);
    ps.leaf_count = I.nleaf;
    ps.pure_neumann = S.prm.pure_neumann;
    ps.check_stride = S.prm.sync_stride;
    HIPC(hipMemcpyAsync(S.d_ps + I.slot, &ps, sizeof(ps), hipMemcpyHostToDevice, st));
    if (!I.dev_rhs && !I.ir_cont)
      HIPC(hipMemcpyAsync(S.d_b + size_t(I.tile_begin) * 512, I.h_buf, size_t(I.ntiles) * 512 * sizeof(float),
                          hipMemcpyHostToDevice, st));
    hmap[f] = I.slot;
    I.state = INST_ACTIVE;
    S.active.push_back(I.slot);
    S.dirty = true;
    if (S.refine) {
      if (!I.ir_cont) {  // r = b - A x64 in FP64 (x64 = 0 for a cold start); |b|^2 and the threshold floor
        if (!I.warm) HIPC(hipMemsetAsync(S.d_x64 + size_t(I.tile_begin) * 512, 0, size_t(I.ntiles) * 512 * sizeof(double), st));
        instance_residual64(S, I, true);
        k_ir_norm<<<grid_for(I.leaf.n, kSwizzle), 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, S.d_b);
        k_finalize<<<1, FIN_THREADS, 0, st>>>(one_problem(S, I), FIN_IRBB, 0);
      }
      // the step solves A d = float(r) from d = 0
      if (S.prm.storage == CG_STORE_BF16) k_cg_init<bf16><<<I.leaf.n, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, 0);
      else k_cg_init<float><<<I.leaf.n, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, 0);
      continue;
    }
    float* X = S.d_ch[CG_CH_X];
    if (!I.warm) k_fill_tiles<float><<<I.all.n, 64, 0, st>>>(I.L(I.all), I.all.n, X, 0.f);
    if (I.ghosts.n) k_propagate_ghosts<<<I.ghosts.n, 64, 0, st>>>(c, I.L(I.ghosts), I.ghosts.n, X, X);
    k_apply<<<I.applyl.n, 64, 0, st>>>(c, I.L(I.applyl), I.applyl.n, X, c.ap, -1, 0, 0, 0);
    if (I.ghosts.n) k_accumulate_ghosts<<<I.ghosts.n, 64, 0, st>>>(c, I.L(I.ghosts), I.ghosts.n, c.ap);
    if (S.prm.storage == CG_STORE_BF16) k_cg_init_warm<bf16><<<I.leaf.n, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, S.d_b, 0);
    else k_cg_init_warm<float><<<I.leaf.n, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, S.d_b, 0);
  }
  HIPC(hipMemcpyAsync(S.d_newmap, hmap, sizeof(int32_t) * nf, hipMemcpyHostToDevice, st));
  Ctx ci = c;
  ci.pmap = S.d_newmap;
  ci.dflag = S.d_dflag_init;
  k_finalize<<<unsigned(nf), FIN_THREADS, 0, st>>>(ci, FIN_BB, 0);
  HIPC(hipGetLastError());
}

// Assemble the worklists of the active set on the device (records copied from
// each instance's own lists), with its problem map and coarse plans.
void server_assemble(cg_server& S, int k) {
  cg_problem& P = S.view;
  const int na = int(S.active.size());
  P.nprob = na;
  P.plans.clear();
  if (na == 0) return;
  int L = 0;
  for (int q : S.active) L = std::max(L, S.inst[q]->max_level);
  P.max_level = L;
  char* tab = S.h_tab[k];
  auto** srcs = reinterpret_cast<const TileRec**>(tab);
  auto* offs = reinterpret_cast<int32_t*>(tab + sizeof(TileRec*) * S.runs_cap);
  auto* cnts = offs + S.runs_cap;
  auto* pmap = cnts + S.runs_cap;
  auto* plans = reinterpret_cast<CoarsePlan*>(reinterpret_cast<char*>(pmap + S.max_probs) + 64 -
                                              (reinterpret_cast<uintptr_t>(pmap + S.max_probs) & 63));
  int nr = 0;
  int64_t pos = 0;
  auto kind = [&](auto get) {
    DList out;
    out.off = int(pos);
    for (int q : S.active) {
      const Instance& I = *S.inst[q];
      const DList* dl = get(I);
      if (!dl || dl->n == 0) continue;
      if (nr >= S.runs_cap) fail(CG_ERR_BUSY, "cg_server_step: too many worklist runs");
      srcs[nr] = I.L(*dl);
      offs[nr] = int32_t(pos);
      cnts[nr] = dl->n;
      nr++;
      pos += dl->n;
    }
    out.n = int(pos) - out.off;
    return out;
  };
  P.leaf = kind([](const Instance& I) { return &I.leaf; });
  P.rz_in_vcycle = true;
  for (int q : S.active)
    if (!S.inst[q]->rz_ok) P.rz_in_vcycle = false;
  P.gather = kind([](const Instance& I) { return &I.gather; });
  P.applyl = kind([](const Instance& I) { return &I.applyl; });
  P.actg.assign(L + 1, DList());
  P.lvlg.assign(L + 1, DList());
  for (int l = 0; l <= L; l++) {
    P.actg[l] = kind([l](const Instance& I) { return l <= I.max_level ? &I.actg[l] : nullptr; });
    P.lvlg[l] = kind([l](const Instance& I) { return l <= I.max_level ? &I.lvlg[l] : nullptr; });
  }
  if (pos > S.asm_cap) fail(CG_ERR_BUSY, "cg_server_step: worklists exceed their reserved space");
  for (int a = 0; a < na; a++) pmap[a] = S.active[a];
  int np = 0;
  for (int q : S.active)
    if (S.inst[q]->C >= 0) plans[np++] = S.inst[q]->plan;
  for (int a = 0; a < np; a++) P.plans.push_back(plans[a]);
  const hipStream_t st = S.stream;
  HIPC(hipMemcpyAsync(S.d_srcs, srcs, sizeof(TileRec*) * nr, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(S.d_offs, offs, sizeof(int32_t) * nr, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(S.d_counts, cnts, sizeof(int32_t) * nr, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(S.d_pmap, pmap, sizeof(int32_t) * na, hipMemcpyHostToDevice, st));
  if (np) HIPC(hipMemcpyAsync(S.d_plans, plans, sizeof(CoarsePlan) * np, hipMemcpyHostToDevice, st));
  if (nr) k_gather_recs<<<dim3(64, unsigned(std::min(nr, 65535))), 256, 0, st>>>(S.d_srcs, S.d_offs, S.d_counts, nr, S.d_asm);
  HIPC(hipGetLastError());
  S.dirty = false;
}

// A request is reported: its pressure goes to the pinned buffer (unless it stays
// on the device) and it joins the completion queue.
void finish_request(cg_server& S, Instance& I) {
  if (!I.no_download)
    HIPC(hipMemcpyAsync(I.h_buf, S.d_ch[CG_CH_X] + size_t(I.tile_begin) * 512, size_t(I.ntiles) * 512 * sizeof(float),
                        hipMemcpyDeviceToHost, S.stream));
  HIPC(hipEventRecord(I.ev, S.stream));
  I.state = INST_DONE;
  S.done_q.push_back(I.slot);
}

// FP64 check of a finished request: x64 = x (verify_f64) or x64 += d (a
// refinement step, whose r = float(b - A x64) is left ready for another step),
// then |b - A x64|^2 and |b|^2 to the host.
void enqueue_verify(cg_server& S, Instance& I) {
  const hipStream_t st = S.stream;
  const Ctx c = S.view.ctx();
  const unsigned gl = grid_for(I.leaf.n, kSwizzle);
  if (S.refine) k_ir_accumulate<<<gl, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, S.d_x64, 0);
  else k_ir_load_x<<<gl, 64, 0, st>>>(c, I.L(I.leaf), I.leaf.n, S.d_ch[CG_CH_X], S.d_x64);
  instance_residual64(S, I, S.refine != 0);
  k_ir_verdict<<<1, FIN_THREADS, 0, st>>>(one_problem(S, I), S.d_ver + 2 * I.slot);
  HIPC(hipMemcpyAsync(S.h_ver + 2 * I.slot, S.d_ver + 2 * I.slot, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPC(hipEventRecord(I.ev_ver, st));
  I.state = INST_VERIFY;
}

// Finished requests of the iteration snapshot k: results, pressure download.
void server_collect(cg_server& S, int k) {
  if (!S.snap_pending[k]) return;
  HIPC(hipEventSynchronize(S.ev_snap[k]));
  S.snap_pending[k] = false;
  const ProbState* st = S.h_snap[k];
  const auto now = std::chrono::steady_clock::now();
  for (int q : S.snap_active[k]) {
    if (!S.inst[q]) continue;
    Instance& I = *S.inst[q];
    if (I.state != INST_ACTIVE || !st[q].done) continue;
    const ProbState& s = st[q];
    I.res.iters = I.iters_acc + s.iter;
    I.res.converged = s.converged;
    I.res.rel_res_f64 = -1.0;
    I.res.seconds = std::chrono::duration<double>(now - I.t_submit).count();
    S.active.erase(std::find(S.active.begin(), S.active.end(), q));
    S.dirty = true;
    // the last iteration's x update (k_cg_update_p applies each earlier one)
    k_cg_finish_x<<<grid_for(I.leaf.n, kSwizzle), 64, 0, S.stream>>>(S.view.ctx(), I.L(I.leaf), I.leaf.n, kSwizzle);
    if (S.refine) {
      // the step's recursive residual, relative to the request's b
      I.bb_b = s.bb_b;
      I.steps++;
      I.res.bb = s.bb_b;
      I.res.rr = s.rr;
      I.res.rel_res = s.bb_b > 0 ? (S.prm.sync_stride > 0 ? std::sqrt(s.rr / s.bb_b) : -1.0) : 0.0;
      if (s.bb_b > 0) {
        enqueue_verify(S, I);
      } else {  // b = 0: x = 0
        HIPC(hipMemsetAsync(S.d_x64 + size_t(I.tile_begin) * 512, 0, size_t(I.ntiles) * 512 * sizeof(double), S.stream));
        HIPC(hipMemsetAsync(S.d_ch[CG_CH_X] + size_t(I.tile_begin) * 512, 0, size_t(I.ntiles) * 512 * sizeof(float), S.stream));
        I.res.converged = 1;
        I.res.rr = 0.0;
        I.res.rel_res_f64 = 0.0;
        finish_request(S, I);
      }
      continue;
    }
    I.res.bb = s.bb;
    I.res.rr = s.rr;
    I.res.rel_res = s.bb > 0 ? (S.prm.sync_stride > 0 ? std::sqrt(s.rr / s.bb) : -1.0) : 0.0;
    if (S.prm.verify_f64 && s.bb > 0) {
      enqueue_verify(S, I);
    } else {
      if (S.prm.verify_f64) I.res.rel_res_f64 = 0.0;  // b = 0: x = 0 is exact
      finish_request(S, I);
    }
  }
}

// Requests whose FP64 check has arrived.  verify_f64: reported with the
// checked residual.  refine: reported once |b - A x|^2 < rel_tol^2 |b|^2 (x =
// float(x64) then), or after kMaxRefineSteps steps; otherwise another step.
void server_verified(cg_server& S) {
  for (int q = 0; q < S.max_probs; q++) {
    if (!S.inst[q] || S.inst[q]->state != INST_VERIFY) continue;
    Instance& I = *S.inst[q];
    const hipError_t e = hipEventQuery(I.ev_ver);
    if (e == hipErrorNotReady) continue;
    HIPC(e);
    const double rr = S.h_ver[2 * q];
    const double bb = S.refine ? S.h_ver[2 * q + 1] : I.res.bb;
    I.res.rel_res_f64 = std::sqrt(rr / bb);
    if (!S.refine) {
      finish_request(S, I);
      continue;
    }
    const double thr = std::max(S.prm.rel_tol * S.prm.rel_tol * bb, std::numeric_limits<double>::min());
    const bool pass = rr < thr;
    if (pass || I.steps >= kMaxRefineSteps || !I.res.converged) {
      I.res.converged = pass ? 1 : 0;
      I.res.rr = rr;
      k_ir_accumulate<<<grid_for(I.leaf.n, kSwizzle), 64, 0, S.stream>>>(S.view.ctx(), I.L(I.leaf), I.leaf.n, S.d_x64, 1);
      finish_request(S, I);
    } else {
      I.ir_cont = 1;
      I.iters_acc = I.res.iters;
      I.state = INST_QUEUED;
    }
  }
}

// One step: an iteration over the current active set, then (one iteration
// behind, so the device keeps working) finished requests, new requests and a
// reassembly of the worklists for the next step.
void server_step(cg_server& S, int32_t* n_active) {
  const int k = int(S.step_no & 1);
  cg_problem& P = S.view;
  if (!S.active.empty() && !S.dirty) {
    enqueue_cg_iteration(P, S.v, S.prm.pure_neumann, S.prm.storage);
    HIPC(hipGetLastError());
    int maxslot = 0;
    for (int q : S.active) maxslot = std::max(maxslot, q + 1);
    HIPC(hipMemcpyAsync(S.h_snap[k], S.d_ps, sizeof(ProbState) * maxslot, hipMemcpyDeviceToHost, S.stream));
    HIPC(hipEventRecord(S.ev_snap[k], S.stream));
    S.snap_active[k] = S.active;
    S.snap_pending[k] = true;
  }
  server_collect(S, 1 - k);
  if (S.prm.verify_f64 || S.refine) server_verified(S);
  server_start_queued(S, k);
  if (S.dirty) server_assemble(S, k);
  S.step_no++;
  if (n_active) *n_active = int32_t(S.active.size());
}

// Seeded on-device update of an idle instance's right-hand side (synthetic client).
void server_perturb(cg_server& S, int32_t id, float rho, float sigma, uint64_t seed) {
  Instance& I = instance(S, id);
  if (I.state != INST_IDLE) fail(CG_ERR_STATE, "cg_server_perturb_rhs: the instance has a request running or an unread result");
  k_perturb_rhs<<<I.leaf.n, 64, 0, S.stream>>>(S.view.ctx(), I.L(I.leaf), I.leaf.n, S.d_b, rho, sigma, seed);
  HIPC(hipGetLastError());
}

// Pop a finished request whose pressure has arrived (or wait for it if `wait`).
bool server_next(cg_server& S, int wait, int64_t* tag, int32_t* id, cg_result* res, float* x) {
  if (S.done_q.empty()) return false;
  Instance& I = *S.inst[S.done_q.front()];
  if (wait) HIPC(hipEventSynchronize(I.ev));
  else {
    const hipError_t e = hipEventQuery(I.ev);
    if (e == hipErrorNotReady) return false;
    HIPC(e);
  }
  S.done_q.pop_front();
  if (tag) *tag = I.tag;
  if (id) *id = I.slot;
  if (res) *res = I.res;
  if (x && !I.no_download)
    for (int i = 0; i < I.ntiles; i++) std::memcpy(x + size_t(I.ext[i]) * 512, I.h_buf + size_t(i) * 512, 512 * sizeof(float));
  I.state = INST_IDLE;
  return true;
}

}  // namespace
