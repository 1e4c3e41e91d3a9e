// Serving engine of the MGPCG library (host side; included by mgpcg.hip).
//
// Instances are resident hierarchies: each admitted tree gets a slab of tiles,
// leaf slots and ghost slots in device arrays sized from an HBM budget, its own
// tile records in a device arena and a pinned staging buffer.  A request is a right
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
//
// The per-request work around the iterations (a start's residual and state, a
// finish's last x update and FP64 check, a passed check's x) runs once per step
// for all the requests that start, finish or pass in that step, over worklists
// assembled like the active set's.  A release hands the instance's memory back
// once the stream has passed it (an event), so it neither waits for the device
// nor frees device memory: records and coarse slots live in server-wide arenas
// and pinned staging buffers are kept for later admissions.
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
  TileRec* d_recs = nullptr;  // the instance's records: rec_count records of the server's arena from rec_begin
  int64_t rec_begin = 0;
  int rec_count = 0;
  DList leaf, gather, applyl, ghosts, all;
  std::vector<DList> actg, lvlg;
  CoarseSlot* d_slots = nullptr;  // the slot's part of the server's coarse-slot arena
  CoarsePlan plan;
  float* h_buf = nullptr;  // pinned: rhs on the way in, pressure on the way out
  size_t h_bytes = 0;      // its capacity (a kept buffer may be larger than the instance needs)
  hipEvent_t ev = nullptr;      // the slot's events (the server's)
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

// A released instance's resources, handed back once the stream passes `ev`.
struct Released {
  hipEvent_t ev = nullptr;
  int slot = -1;
  int64_t tile_begin = 0, leaf_begin = 0, ghost_begin = 0, rec_begin = 0;
  int ntiles = 0, nleaf = 0, nghost = 0, rec_count = 0;
  float* h_buf = nullptr;
  size_t h_bytes = 0;
};

// Worklists of a group of instances (the requests that start, finish or pass
// their FP64 check in one step), assembled in the server's group arena: leaf
// over all of them, leaf_full / ghosts / applyl over the first nfull.
struct Group {
  int n = 0, nfull = 0;
  DList leaf, leaf_full, ghosts, applyl;
};
enum : int { GRP_START = 0, GRP_FINISH = 1, GRP_PASSED = 2, GRP_REGIONS = 3 };

}  // namespace

struct cg_server {
  cg_params prm{};
  int refine = 0;  // every request is an FP64 iterative refinement
  VParams v;
  int max_probs = 0;  // resident instances at most
  int nslots = 0;     // problem slots: max_probs and room for releases the stream has not yet passed
  int nres = 0;       // resident instances
  int64_t cap_tiles = 0, cap_leaf = 0, cap_ghost = 0;
  RangeAlloc tiles_alloc, leaf_alloc, ghost_alloc, rec_alloc;
  std::vector<int> free_slots;
  std::vector<Released> pending;  // releases the stream has not yet passed
  std::vector<hipEvent_t> ev_free;
  std::multimap<size_t, float*> hbuf_cache;  // pinned staging buffers of released instances, by size
  // pinned ring for admission uploads (records, coarse slots, cell types),
  // copied on the stream; a segment is reused once its event has passed
  char* h_ring = nullptr;
  size_t ring_cap = 0, ring_head = 0;
  struct RingSeg { size_t begin, end; hipEvent_t ev; };
  std::deque<RingSeg> ring_busy;
  std::vector<hipEvent_t> slot_ev, slot_ev_ver;
  std::vector<std::unique_ptr<Instance>> inst;  // by slot
  std::vector<int> active;                      // slots whose request runs, assembly order
  std::vector<int> snap_active[2];              // active set of the iteration each snapshot follows
  bool dirty = false;
  // the iteration over an assembled active set is captured in a graph on its
  // second step and replayed while the set stays; replaced graphs are destroyed
  // once the stream has passed their last launch
  int64_t asm_ver = 0, graph_ver = -1, last_iter_ver = -1;
  hipGraphExec_t graph = nullptr;
  std::vector<std::pair<hipGraphExec_t, hipEvent_t>> retired;
  std::deque<int> done_q;                       // slots with a finished request, in completion order
  int64_t step_no = 0;
  bool snap_pending[2] = {false, false};

  // device
  cg_problem view;  // the kernels' view of the assembled active set (owner = false)
  uint8_t* d_ctype = nullptr;
  uint8_t* d_inert = nullptr;   // per tile: the running request leaves it untouched (k_cg_init)
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
  double* d_part = nullptr;
  double* d_part2 = nullptr;
  TileRec* d_asm = nullptr;
  int64_t asm_cap = 0;
  TileRec* d_recs = nullptr;  // arena of the instances' own records
  CoarseSlot* d_cslots = nullptr;  // arena of the instances' coarse slots (CL_MAX_TILES per slot)
  TileRec* d_gasm = nullptr;  // group worklists
  int64_t gasm_cap = 0;
  // group tables (runs, map, staged problem states): pinned, per step parity and
  // region; on the device one region, reused in stream order
  char* h_gtab[2] = {nullptr, nullptr};
  char* d_gtab = nullptr;
  size_t gstride = 0, goff_map = 0, goff_ps = 0;
  hipEvent_t ev_gtab[2][GRP_REGIONS] = {};
  bool gtab_used[2][GRP_REGIONS] = {};
  CoarsePlan* d_plans = nullptr;
  const TileRec** d_srcs = nullptr;
  int32_t* d_offs = nullptr;
  int32_t* d_counts = nullptr;
  int runs_cap = 0;
  hipStream_t stream = nullptr;
  // pinned host memory: ProbState snapshots, per-step small tables
  ProbState* h_snap[2] = {nullptr, nullptr};
  hipEvent_t ev_snap[2] = {nullptr, nullptr};
  char* h_tab[2] = {nullptr, nullptr};
  size_t tab_cap = 0;
  ~cg_server();
};

cg_server::~cg_server() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (graph) (void)hipGraphExecDestroy(graph);
  for (auto& r : retired) {
    (void)hipGraphExecDestroy(r.first);
    (void)hipEventDestroy(r.second);
  }
  for (auto& I : inst)
    if (I && I->h_buf) (void)hipHostFree(I->h_buf);
  for (auto& R : pending) {
    if (R.h_buf) (void)hipHostFree(R.h_buf);
    if (R.ev) (void)hipEventDestroy(R.ev);
  }
  for (auto& kv : hbuf_cache) (void)hipHostFree(kv.second);
  for (auto& g : ring_busy) (void)hipEventDestroy(g.ev);
  if (h_ring) (void)hipHostFree(h_ring);
  for (auto e : ev_free) (void)hipEventDestroy(e);
  for (auto e : slot_ev) if (e) (void)hipEventDestroy(e);
  for (auto e : slot_ev_ver) if (e) (void)hipEventDestroy(e);
  for (auto& row : ev_gtab)
    for (auto e : row) if (e) (void)hipEventDestroy(e);
  if (h_ver) (void)hipHostFree(h_ver);
  void* ptrs[] = {d_ctype, d_inert, d_fmask, d_diag, d_zb, d_b, d_x64, d_r64, d_ver, d_zh, d_rhsvh, d_zbh, d_gsum, d_ps, d_dflag,
                  d_dflag_init, d_pmap, d_part, d_part2, d_asm, d_recs, d_cslots, d_gasm, d_gtab, d_plans,
                  (void*)d_srcs, d_offs, d_counts};
  for (void* q : ptrs) if (q) (void)hipFree(q);
  for (auto q : d_ch) if (q) (void)hipFree(q);
  for (auto e : ev_snap) if (e) (void)hipEventDestroy(e);
  for (auto q : h_snap) if (q) (void)hipHostFree(q);
  for (auto q : h_tab) if (q) (void)hipHostFree(q);
  for (auto q : h_gtab) if (q) (void)hipHostFree(q);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

// Device bytes one tile costs: 6 channels, the V-cycle scratch, the kept rhs
// and the colour sweep's diagonal (FP32), cell types, inert flags and face masks, ghost row
// sums, dot partials and room for its records: the instance's own lists, the
// assembled active worklists and the group worklists; BF16 storage adds three
// bfloat16 vectors, verify_f64 two FP64 vectors.
constexpr int kRecKindsPerTile = 5;  // leaf, gather, applyl, actg, lvlg: at most 5 records per tile
// An instance's own lists (leaf, gather, applyl, ghosts, all, actg, lvlg) hold
// 3 leaf + ghost + 3 n <= 6 n records; a group's (leaf, ghosts, applyl) <= 2 n.
constexpr int kOwnRecsPerTile = 6, kGroupRecsPerTile = 2;
constexpr int64_t kBytesPerTile = (CG_NUM_CHANNELS + 3) * 512 * 4 + 2 * 512 + 1 + 64 * 4 + 2 * 8 +
                                  (kRecKindsPerTile + kOwnRecsPerTile + kGroupRecsPerTile) * int64_t(sizeof(TileRec));

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
  S.nslots = S.max_probs + std::max(16, S.max_probs / 8);
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
  S.rec_alloc.init(S.cap_tiles * kOwnRecsPerTile);
  S.inst.resize(S.nslots);
  S.slot_ev.assign(S.nslots, nullptr);
  S.slot_ev_ver.assign(S.nslots, nullptr);
  for (int q = S.nslots - 1; q >= 0; q--) S.free_slots.push_back(q);

  HIPC(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
  const size_t cells = size_t(S.cap_tiles) * 512;
  HIPC(hipMalloc(&S.d_ctype, cells));
  HIPC(hipMalloc(&S.d_inert, size_t(S.cap_tiles)));
  HIPC(hipMemsetAsync(S.d_inert, 0, size_t(S.cap_tiles), S.stream));
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
    HIPC(hipMalloc(&S.d_ver, 2 * sizeof(double) * S.nslots));
    HIPC(hipHostMalloc(&S.h_ver, 2 * sizeof(double) * S.nslots, hipHostMallocDefault));
  }
  if (S.prm.storage == CG_STORE_BF16) {
    bf16** bufs[3] = {&S.d_zh, &S.d_rhsvh, &S.d_zbh};
    for (bf16** b : bufs) {
      HIPC(hipMalloc(b, cells * sizeof(bf16)));
      HIPC(hipMemsetAsync(*b, 0, cells * sizeof(bf16), S.stream));
    }
  }
  HIPC(hipMalloc(&S.d_gsum, size_t(S.cap_ghost) * 64 * sizeof(float)));
  HIPC(hipMalloc(&S.d_ps, sizeof(ProbState) * S.nslots));
  HIPC(hipMemsetAsync(S.d_ps, 0, sizeof(ProbState) * S.nslots, S.stream));
  HIPC(hipMalloc(&S.d_dflag, sizeof(int32_t) * S.nslots));
  HIPC(hipMalloc(&S.d_dflag_init, sizeof(int32_t) * S.nslots));
  HIPC(hipMalloc(&S.d_pmap, sizeof(int32_t) * S.max_probs));
  HIPC(hipMalloc(&S.d_part, sizeof(double) * S.cap_leaf));
  HIPC(hipMalloc(&S.d_part2, sizeof(double) * S.cap_leaf));
  S.asm_cap = S.cap_tiles * kRecKindsPerTile;
  HIPC(hipMalloc(&S.d_asm, sizeof(TileRec) * S.asm_cap));
  HIPC(hipMalloc(&S.d_recs, sizeof(TileRec) * S.cap_tiles * kOwnRecsPerTile));
  HIPC(hipMalloc(&S.d_cslots, sizeof(CoarseSlot) * CL_MAX_TILES * S.nslots));
  S.gasm_cap = S.cap_tiles * kGroupRecsPerTile;
  HIPC(hipMalloc(&S.d_gasm, sizeof(TileRec) * S.gasm_cap));
  {
    const size_t R = size_t(3) * S.nslots;  // runs: leaf, ghosts and applyl of each instance
    S.goff_map = R * (sizeof(TileRec*) + 8);
    S.goff_ps = (S.goff_map + sizeof(int32_t) * S.nslots + 63) & ~size_t(63);
    S.gstride = (S.goff_ps + sizeof(ProbState) * S.nslots + 255) & ~size_t(255);
    for (int k = 0; k < 2; k++) {
      HIPC(hipHostMalloc(&S.h_gtab[k], S.gstride * GRP_REGIONS, hipHostMallocDefault));
      for (int g = 0; g < GRP_REGIONS; g++) HIPC(hipEventCreateWithFlags(&S.ev_gtab[k][g], hipEventDisableTiming));
    }
    HIPC(hipMalloc(&S.d_gtab, S.gstride));
  }
  HIPC(hipMalloc(&S.d_plans, sizeof(CoarsePlan) * S.max_probs));
  S.runs_cap = S.max_probs * (3 + 2 * 32);
  HIPC(hipMalloc(&S.d_srcs, sizeof(TileRec*) * S.runs_cap));
  HIPC(hipMalloc(&S.d_offs, sizeof(int32_t) * S.runs_cap));
  HIPC(hipMalloc(&S.d_counts, sizeof(int32_t) * S.runs_cap));
  for (int k = 0; k < 2; k++) {
    HIPC(hipHostMalloc(&S.h_snap[k], sizeof(ProbState) * S.nslots, hipHostMallocDefault));
    HIPC(hipEventCreateWithFlags(&S.ev_snap[k], hipEventDisableTiming));
  }
  // per-step tables: runs (srcs, offs, counts), active map, coarse plans
  S.tab_cap = size_t(S.runs_cap) * (sizeof(TileRec*) + 8) + size_t(S.max_probs) * (8 + sizeof(CoarsePlan)) + 4096;
  for (int k = 0; k < 2; k++) HIPC(hipHostMalloc(&S.h_tab[k], S.tab_cap, hipHostMallocDefault));
  S.ring_cap = size_t(64) << 20;
  HIPC(hipHostMalloc(&S.h_ring, S.ring_cap, hipHostMallocDefault));

  // the kernels' view: shares the server's arrays
  cg_problem& P = S.view;
  P.owner = false;
  P.stream = S.stream;
  P.d_ctype = S.d_ctype;
  P.d_inert = S.d_inert;
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

// Releases whose point the stream has passed (all of them, waiting, if `wait`)
// give their slot, ranges and staging buffer back.
void reap_releases(cg_server& S, bool wait) {
  for (size_t i = 0; i < S.pending.size();) {
    Released& R = S.pending[i];
    if (wait) {
      HIPC(hipEventSynchronize(R.ev));
    } else {
      const hipError_t e = hipEventQuery(R.ev);
      if (e == hipErrorNotReady) { i++; continue; }
      HIPC(e);
    }
    if (!S.tiles_alloc.release(R.tile_begin, R.ntiles) || !S.leaf_alloc.release(R.leaf_begin, R.nleaf) ||
        !S.ghost_alloc.release(R.ghost_begin, R.nghost) || !S.rec_alloc.release(R.rec_begin, R.rec_count))
      fail(CG_ERR_STATE, "cg_server_release: the instance's ranges were already free");
    S.free_slots.push_back(R.slot);
    if (R.h_buf) {
      if (S.hbuf_cache.size() < size_t(S.max_probs)) S.hbuf_cache.emplace(R.h_bytes, R.h_buf);
      else HIPC(hipHostFree(R.h_buf));
    }
    S.ev_free.push_back(R.ev);
    S.pending.erase(S.pending.begin() + i);
  }
}

// A range of `n` from `A`; when none is free and releases are pending, after
// waiting for them.
int64_t alloc_range(cg_server& S, RangeAlloc& A, int64_t n) {
  int64_t at = A.alloc(n);
  if (at < 0 && !S.pending.empty()) {
    reap_releases(S, true);
    at = A.alloc(n);
  }
  return at;
}

void admit_upload(cg_server& S, const cg_tree_desc* d, const TreeBuild& B, const std::vector<TileRec>& W,
                  const std::vector<CoarseSlot>& cslots, Instance& inst, int slot, int64_t tb, int64_t rb);

void server_admit(cg_server& S, const cg_tree_desc* d, int32_t* out) {
  if (!d || !out) fail(CG_ERR_INVALID, "cg_server_admit: null argument");
  if (d->num_tiles <= 0) fail(CG_ERR_INVALID, "cg_server_admit: empty tree");
  if (!S.pending.empty()) reap_releases(S, false);
  if (S.nres >= S.max_probs) fail(CG_ERR_BUSY, "cg_server_admit: no free problem slot");
  if (S.free_slots.empty()) reap_releases(S, true);
  if (S.free_slots.empty()) fail(CG_ERR_BUSY, "cg_server_admit: no free problem slot");
  const int n = d->num_tiles;
  const int64_t tb = alloc_range(S, S.tiles_alloc, n);
  if (tb < 0) fail(CG_ERR_BUSY, "cg_server_admit: HBM budget exhausted");
  // leaf and ghost counts are known only after the build: reserve the tile count
  // first, trim after
  const int64_t lb = alloc_range(S, S.leaf_alloc, n), gb = alloc_range(S, S.ghost_alloc, n);
  if (lb < 0 || gb < 0) {
    S.tiles_alloc.release(tb, n);
    if (lb >= 0) S.leaf_alloc.release(lb, n);
    if (gb >= 0) S.ghost_alloc.release(gb, n);
    fail(CG_ERR_BUSY, "cg_server_admit: HBM budget exhausted");
  }
  const int slot = S.free_slots.back();
  S.free_slots.pop_back();
  auto I = std::make_unique<Instance>();
  TreeBuild B;
  try {
    build_tree(d, slot, int(tb), int(lb), int(gb), B);
  } catch (...) {
    S.tiles_alloc.release(tb, n);
    S.leaf_alloc.release(lb, n);
    S.ghost_alloc.release(gb, n);
    S.free_slots.push_back(slot);
    throw;
  }
  S.leaf_alloc.release(lb + B.nleaf, n - B.nleaf);
  S.ghost_alloc.release(gb + B.nghost, n - B.nghost);
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
  std::vector<CoarseSlot> cslots;
  if (I->C >= 0) coarse_plan(B.tiles, int(tb), I->C, slot, I->plan, cslots);
  const bool fits = W.size() <= size_t(n) * kOwnRecsPerTile && cslots.size() <= size_t(CL_MAX_TILES);
  const int64_t rb = fits ? alloc_range(S, S.rec_alloc, int64_t(W.size())) : -2;
  auto undo = [&] {
    S.tiles_alloc.release(tb, n);
    S.leaf_alloc.release(lb, B.nleaf);
    S.ghost_alloc.release(gb, B.nghost);
    if (rb >= 0) S.rec_alloc.release(rb, int64_t(W.size()));
    S.free_slots.push_back(slot);
  };
  if (rb < 0) {
    undo();
    if (rb == -2) fail(CG_ERR_INVALID, "cg_server_admit: more records or coarse slots than the tree allows");
    fail(CG_ERR_BUSY, "cg_server_admit: HBM budget exhausted");
  }
  try {
    admit_upload(S, d, B, W, cslots, *I, slot, tb, rb);
  } catch (...) {
    undo();
    if (I->h_buf) (void)hipHostFree(I->h_buf);
    throw;
  }
  *out = slot;
  S.inst[slot] = std::move(I);
  S.nres++;
}

// `bytes` of the admission ring (nullptr if it never fits): the oldest
// segments in the way are waited for (they precede everything enqueued since).
char* ring_take(cg_server& S, size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (bytes > S.ring_cap) return nullptr;
  size_t at = S.ring_head;
  if (at + bytes > S.ring_cap) at = 0;
  auto overlaps = [&](const cg_server::RingSeg& g) { return g.begin < at + bytes && at < g.end; };
  while (!S.ring_busy.empty() && std::any_of(S.ring_busy.begin(), S.ring_busy.end(), overlaps)) {
    HIPC(hipEventSynchronize(S.ring_busy.front().ev));
    S.ev_free.push_back(S.ring_busy.front().ev);
    S.ring_busy.pop_front();
  }
  S.ring_head = at + bytes;
  return S.h_ring + at;
}
// The segment [p, p + bytes) of the ring is read by copies enqueued so far.
void ring_release_after_copies(cg_server& S, const char* p, size_t bytes) {
  cg_server::RingSeg g;
  g.begin = size_t(p - S.h_ring);
  g.end = g.begin + ((bytes + 255) & ~size_t(255));
  if (S.ev_free.empty()) {
    HIPC(hipEventCreateWithFlags(&g.ev, hipEventDisableTiming));
  } else {
    g.ev = S.ev_free.back();
    S.ev_free.pop_back();
  }
  HIPC(hipEventRecord(g.ev, S.stream));
  S.ring_busy.push_back(g);
}

// The admitted instance's records, coarse slots and cell types to the device,
// its face masks, diagonal and zeroed vectors (on the server's stream), its
// staging buffer and events.
void admit_upload(cg_server& S, const cg_tree_desc* d, const TreeBuild& B, const std::vector<TileRec>& W,
                  const std::vector<CoarseSlot>& cslots, Instance& inst, int slot, int64_t tb, int64_t rb) {
  Instance* I = &inst;
  const int n = d->num_tiles;
  // the instance's pinned staging buffer (right-hand sides in, pressures out)
  const size_t need = size_t(n) * 512 * sizeof(float);
  auto kept = S.hbuf_cache.lower_bound(need);
  if (kept != S.hbuf_cache.end() && kept->first <= 2 * need) {
    I->h_bytes = kept->first;
    I->h_buf = kept->second;
    S.hbuf_cache.erase(kept);
  } else {
    HIPC(hipHostMalloc(&I->h_buf, need, hipHostMallocDefault));
    I->h_bytes = need;
  }
  I->rec_begin = rb;
  I->rec_count = int(W.size());
  I->d_recs = S.d_recs + rb;
  // records, coarse slots and cell types through the admission ring, copied on
  // the stream (the host does not wait for a device that is busy with other
  // requests); a tree too large for the ring goes through its own staging
  // buffer with synchronous copies
  const size_t rec_bytes = sizeof(TileRec) * W.size(), ct_bytes = size_t(n) * 512;
  const size_t slot_bytes = sizeof(CoarseSlot) * cslots.size();
  const size_t o_ct = (rec_bytes + 255) & ~size_t(255), o_sl = o_ct + ((ct_bytes + 255) & ~size_t(255));
  const size_t total = o_sl + slot_bytes;
  char* ring = ring_take(S, total);
  char* stage = ring ? ring : reinterpret_cast<char*>(I->h_buf);
  if (!ring && total > I->h_bytes) fail(CG_ERR_INVALID, "cg_server_admit: staging buffer too small");
  std::memcpy(stage, W.data(), rec_bytes);
  for (int i = 0; i < n; i++) std::memcpy(stage + o_ct + size_t(i) * 512, d->cell_type + size_t(B.ext[i]) * 512, 512);
  if (slot_bytes) std::memcpy(stage + o_sl, cslots.data(), slot_bytes);
  auto up = [&](void* dst, size_t off, size_t bytes) {
    if (ring) HIPC(hipMemcpyAsync(dst, stage + off, bytes, hipMemcpyHostToDevice, S.stream));
    else HIPC(hipMemcpy(dst, stage + off, bytes, hipMemcpyHostToDevice));
  };
  up(I->d_recs, 0, rec_bytes);
  if (I->C >= 0) {
    I->d_slots = S.d_cslots + size_t(slot) * CL_MAX_TILES;
    up(I->d_slots, o_sl, slot_bytes);
    I->plan.slots = I->d_slots;
  }
  // cell types in internal order; x = 0 so that a first warm start is a cold one
  up(S.d_ctype + size_t(tb) * 512, o_ct, ct_bytes);
  if (ring) ring_release_after_copies(S, ring, total);
  // face masks on the server's stream, ahead of the instance's first request
  k_face_mask<<<I->all.n, 64, 0, S.stream>>>(S.d_ctype, S.d_fmask, S.d_fsplit, I->L(I->all), I->all.n);
  k_diag<<<I->all.n, 64, 0, S.stream>>>(S.view.ctx(), I->L(I->all), I->all.n, S.d_diag);
  HIPC(hipGetLastError());
  HIPC(hipMemsetAsync(S.d_ch[CG_CH_X] + size_t(tb) * 512, 0, size_t(n) * 512 * sizeof(float), S.stream));
  HIPC(hipMemsetAsync(S.d_b + size_t(tb) * 512, 0, size_t(n) * 512 * sizeof(float), S.stream));
  if (S.d_x64) HIPC(hipMemsetAsync(S.d_x64 + size_t(tb) * 512, 0, size_t(n) * 512 * sizeof(double), S.stream));
  if (!S.slot_ev[slot]) HIPC(hipEventCreateWithFlags(&S.slot_ev[slot], hipEventDisableTiming));
  if (!S.slot_ev_ver[slot]) HIPC(hipEventCreateWithFlags(&S.slot_ev_ver[slot], hipEventDisableTiming));
  I->ev = S.slot_ev[slot];
  I->ev_ver = S.slot_ev_ver[slot];
}

Instance& instance(cg_server& S, int32_t id) {
  if (id < 0 || id >= S.nslots || !S.inst[id]) fail(CG_ERR_INVALID, "unknown instance");
  return *S.inst[id];
}

// The instance leaves at once; its slot, ranges and staging buffer come back
// once the stream has passed this point (no kernel may still read its slab).
void server_release(cg_server& S, int32_t id) {
  Instance& I = instance(S, id);
  if (I.state == INST_QUEUED || I.state == INST_ACTIVE || I.state == INST_VERIFY)
    fail(CG_ERR_STATE, "cg_server_release: the instance has a request running");
  Released R;
  if (S.ev_free.empty()) {
    HIPC(hipEventCreateWithFlags(&R.ev, hipEventDisableTiming));
  } else {
    R.ev = S.ev_free.back();
    S.ev_free.pop_back();
  }
  const hipError_t e = hipEventRecord(R.ev, S.stream);
  if (e != hipSuccess) {
    S.ev_free.push_back(R.ev);
    HIPC(e);
  }
  if (I.state == INST_DONE)
    for (auto it = S.done_q.begin(); it != S.done_q.end(); ++it)
      if (*it == id) { S.done_q.erase(it); break; }
  R.slot = id;
  R.tile_begin = I.tile_begin;
  R.leaf_begin = I.leaf_begin;
  R.ghost_begin = I.ghost_begin;
  R.rec_begin = I.rec_begin;
  R.ntiles = I.ntiles;
  R.nleaf = I.nleaf;
  R.nghost = I.nghost;
  R.rec_count = I.rec_count;
  R.h_buf = I.h_buf;
  R.h_bytes = I.h_bytes;
  S.pending.push_back(R);
  S.inst[id].reset();
  S.nres--;
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

const TileRec* GL(const cg_server& S, const DList& d) { return S.d_gasm + d.off; }
const int32_t* gmap(const cg_server& S) { return reinterpret_cast<const int32_t*>(S.d_gtab + S.goff_map); }
// The kernels' Ctx for one workgroup per group member (workgroup a -> slot map[a]).
Ctx group_ctx(cg_server& S) {
  Ctx c = S.view.ctx();
  c.pmap = gmap(S);
  return c;
}

// A group's worklists: each listed instance's runs copied from its own lists
// into the group arena (the leaf runs of every instance, the ghost and operator
// runs of the first nfull), with the slots as the group map and, for a start,
// the instances' new problem states scattered to their slots.  The host table
// of (step parity k, region) is reused only after its last upload has run.
Group group_lists(cg_server& S, int k, int region, const std::vector<int>& slots, int nfull, const ProbState* ps) {
  Group G;
  G.n = int(slots.size());
  G.nfull = nfull;
  const hipStream_t st = S.stream;
  if (S.gtab_used[k][region]) HIPC(hipEventSynchronize(S.ev_gtab[k][region]));
  char* tab = S.h_gtab[k] + size_t(region) * S.gstride;
  const int R = 3 * S.nslots;
  auto** srcs = reinterpret_cast<const TileRec**>(tab);
  auto* offs = reinterpret_cast<int32_t*>(tab + sizeof(TileRec*) * R);
  auto* cnts = offs + R;
  auto* map = reinterpret_cast<int32_t*>(tab + S.goff_map);
  int nr = 0;
  int64_t pos = 0;
  auto kind = [&](int a0, int a1, auto get) {
    DList out;
    out.off = int(pos);
    for (int a = a0; a < a1; a++) {
      const Instance& I = *S.inst[slots[a]];
      const DList& dl = get(I);
      if (dl.n == 0) continue;
      srcs[nr] = I.L(dl);
      offs[nr] = int32_t(pos);
      cnts[nr] = dl.n;
      nr++;
      pos += dl.n;
    }
    out.n = int(pos) - out.off;
    return out;
  };
  G.leaf_full = kind(0, nfull, [](const Instance& I) -> const DList& { return I.leaf; });
  const DList more = kind(nfull, G.n, [](const Instance& I) -> const DList& { return I.leaf; });
  G.leaf = DList{G.leaf_full.off, G.leaf_full.n + more.n};
  G.ghosts = kind(0, nfull, [](const Instance& I) -> const DList& { return I.ghosts; });
  G.applyl = kind(0, nfull, [](const Instance& I) -> const DList& { return I.applyl; });
  if (pos > S.gasm_cap) fail(CG_ERR_BUSY, "cg_server_step: group worklists exceed their reserved space");
  for (int a = 0; a < G.n; a++) map[a] = slots[a];
  size_t bytes = S.goff_map + sizeof(int32_t) * G.n;
  if (ps) {
    std::memcpy(tab + S.goff_ps, ps, sizeof(ProbState) * G.n);
    bytes = S.goff_ps + sizeof(ProbState) * G.n;
  }
  HIPC(hipMemcpyAsync(S.d_gtab, tab, bytes, hipMemcpyHostToDevice, st));
  HIPC(hipEventRecord(S.ev_gtab[k][region], st));
  S.gtab_used[k][region] = true;
  if (ps && G.n)
    k_ps_scatter<<<unsigned(G.n), 64, 0, st>>>(S.d_ps, reinterpret_cast<const ProbState*>(S.d_gtab + S.goff_ps), gmap(S), G.n);
  if (nr)
    k_gather_recs<<<dim3(64, unsigned(std::min(nr, 65535))), 256, 0, st>>>(
        reinterpret_cast<const TileRec* const*>(S.d_gtab), reinterpret_cast<const int32_t*>(S.d_gtab + sizeof(TileRec*) * R),
        reinterpret_cast<const int32_t*>(S.d_gtab + sizeof(TileRec*) * R + sizeof(int32_t) * R), nr, S.d_gasm);
  HIPC(hipGetLastError());
  return G;
}

// r64 = b - A x64 on the LEAF tiles of the group's first nfull instances in
// FP64 (re-centred for pure Neumann) with |r64|^2 partials; write_r:
// float(r64) into the r channel.
void group_residual64(cg_server& S, const Group& G, bool write_r) {
  const hipStream_t st = S.stream;
  const Ctx c = S.view.ctx();
  const unsigned gl = grid_for(G.leaf_full.n, kSwizzle);
  if (G.ghosts.n) k_propagate_ghosts64<<<G.ghosts.n, 64, 0, st>>>(c, GL(S, G.ghosts), G.ghosts.n, S.d_x64);
  k_apply64<<<G.applyl.n, 64, 0, st>>>(c, GL(S, G.applyl), G.applyl.n, S.d_x64, S.d_r64);
  if (G.ghosts.n) k_accumulate_ghosts64<<<G.ghosts.n, 64, 0, st>>>(c, GL(S, G.ghosts), G.ghosts.n, S.d_r64);
  if (S.prm.pure_neumann) {
    k_ir_residual<<<gl, 64, 0, st>>>(c, GL(S, G.leaf_full), G.leaf_full.n, S.d_b, S.d_r64, 1);
    k_finalize<<<unsigned(G.nfull), FIN_THREADS, 0, st>>>(group_ctx(S), FIN_MEAN, 0);
    k_ir_recenter<<<gl, 64, 0, st>>>(c, GL(S, G.leaf_full), G.leaf_full.n, S.d_r64, write_r ? 1 : 0);
  } else {
    k_ir_residual<<<gl, 64, 0, st>>>(c, GL(S, G.leaf_full), G.leaf_full.n, S.d_b, S.d_r64, write_r ? 0 : 2);
  }
}

// Start queued requests: ProbState, rhs upload, x0 (zero for a cold start),
// r0 = b - A x0, |b|^2 -- one group for all of them.
void server_start_queued(cg_server& S, int k) {
  std::vector<int> fresh, cont;  // refinement: requests that continue one are put last (no residual)
  for (int q = 0; q < S.nslots; q++)
    if (S.inst[q] && S.inst[q]->state == INST_QUEUED) (S.refine && S.inst[q]->ir_cont ? cont : fresh).push_back(q);
  const int nfull = int(fresh.size());
  fresh.insert(fresh.end(), cont.begin(), cont.end());
  if (fresh.empty()) return;
  const hipStream_t st = S.stream;
  std::vector<ProbState> pss(fresh.size());
  for (size_t f = 0; f < fresh.size(); f++) {
    Instance& I = *S.inst[fresh[f]];
    ProbState& ps = pss[f];
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
    if (!I.dev_rhs && !I.ir_cont)
      HIPC(hipMemcpyAsync(S.d_b + size_t(I.tile_begin) * 512, I.h_buf, size_t(I.ntiles) * 512 * sizeof(float),
                          hipMemcpyHostToDevice, st));
    I.state = INST_ACTIVE;
    S.active.push_back(I.slot);
    S.dirty = true;
    if (S.refine) {  // r = b - A x64 in FP64 (x64 = 0 for a cold start)
      if (!I.ir_cont && !I.warm)
        HIPC(hipMemsetAsync(S.d_x64 + size_t(I.tile_begin) * 512, 0, size_t(I.ntiles) * 512 * sizeof(double), st));
    } else if (!I.warm) {
      k_fill_tiles<float><<<I.all.n, 64, 0, st>>>(I.L(I.all), I.all.n, S.d_ch[CG_CH_X], 0.f);
    }
  }
  const Group G = group_lists(S, k, GRP_START, fresh, nfull, pss.data());
  const Ctx c = S.view.ctx();
  if (S.refine) {
    if (G.nfull) {  // |b|^2 and the threshold floor
      group_residual64(S, G, true);
      k_ir_norm<<<grid_for(G.leaf_full.n, kSwizzle), 64, 0, st>>>(c, GL(S, G.leaf_full), G.leaf_full.n, S.d_b);
      k_finalize<<<unsigned(G.nfull), FIN_THREADS, 0, st>>>(group_ctx(S), FIN_IRBB, 0);
    }
    // the step solves A d = float(r) from d = 0
    if (S.prm.storage == CG_STORE_BF16) k_cg_init<bf16><<<G.leaf.n, 64, 0, st>>>(c, GL(S, G.leaf), G.leaf.n, 0);
    else k_cg_init<float><<<G.leaf.n, 64, 0, st>>>(c, GL(S, G.leaf), G.leaf.n, 0);
  } else {
    float* X = S.d_ch[CG_CH_X];
    if (G.ghosts.n) k_propagate_ghosts<<<G.ghosts.n, 64, 0, st>>>(c, GL(S, G.ghosts), G.ghosts.n, X, X);
    k_apply<<<G.applyl.n, 64, 0, st>>>(c, GL(S, G.applyl), G.applyl.n, X, c.ap, -1, 0, 0, 0);
    if (G.ghosts.n) k_accumulate_ghosts<<<G.ghosts.n, 64, 0, st>>>(c, GL(S, G.ghosts), G.ghosts.n, c.ap);
    if (S.prm.storage == CG_STORE_BF16) k_cg_init_warm<bf16><<<G.leaf.n, 64, 0, st>>>(c, GL(S, G.leaf), G.leaf.n, S.d_b, 0);
    else k_cg_init_warm<float><<<G.leaf.n, 64, 0, st>>>(c, GL(S, G.leaf), G.leaf.n, S.d_b, 0);
  }
  Ctx ci = group_ctx(S);
  ci.dflag = S.d_dflag_init;
  k_finalize<<<unsigned(G.n), FIN_THREADS, 0, st>>>(ci, FIN_BB, 0);
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
  S.asm_ver++;
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

// FP64 check of the finished requests of a group: x64 = x (verify_f64) or
// x64 += d (a refinement step, whose r = float(b - A x64) is left ready for
// another step), then |b - A x64|^2 and |b|^2 to the host.
void enqueue_verify(cg_server& S, const Group& G, const std::vector<int>& slots) {
  const hipStream_t st = S.stream;
  const Ctx c = S.view.ctx();
  const unsigned gl = grid_for(G.leaf.n, kSwizzle);
  if (S.refine) k_ir_accumulate<<<gl, 64, 0, st>>>(c, GL(S, G.leaf), G.leaf.n, S.d_x64, 0);
  else k_ir_load_x<<<gl, 64, 0, st>>>(c, GL(S, G.leaf), G.leaf.n, S.d_ch[CG_CH_X], S.d_x64);
  group_residual64(S, G, S.refine != 0);
  k_ir_verdict<<<unsigned(G.n), FIN_THREADS, 0, st>>>(group_ctx(S), S.d_ver);
  // one copy over the group's slot range (the other slots' entries in it are
  // rewritten with the values their own checks left)
  const int lo = *std::min_element(slots.begin(), slots.end()), hi = *std::max_element(slots.begin(), slots.end());
  HIPC(hipMemcpyAsync(S.h_ver + 2 * lo, S.d_ver + 2 * lo, 2 * sizeof(double) * (hi - lo + 1), hipMemcpyDeviceToHost, st));
  for (int q : slots) {
    Instance& I = *S.inst[q];
    HIPC(hipEventRecord(I.ev_ver, st));
    I.state = INST_VERIFY;
  }
}

// Finished requests of the iteration snapshot k: results, the last x update,
// the FP64 checks (one group), pressure download.
void server_collect(cg_server& S, int k) {
  if (!S.snap_pending[k]) return;
  HIPC(hipEventSynchronize(S.ev_snap[k]));
  S.snap_pending[k] = false;
  const ProbState* st = S.h_snap[k];
  const auto now = std::chrono::steady_clock::now();
  std::vector<int> check, rest;  // to the FP64 check; reported now (b = 0, or no check)
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
    if (S.refine) {
      // the step's recursive residual, relative to the request's b
      I.bb_b = s.bb_b;
      I.steps++;
      I.res.bb = s.bb_b;
      I.res.rr = s.rr;
      I.res.rel_res = s.bb_b > 0 ? (S.prm.sync_stride > 0 ? std::sqrt(s.rr / s.bb_b) : -1.0) : 0.0;
      (s.bb_b > 0 ? check : rest).push_back(q);
    } else {
      I.res.bb = s.bb;
      I.res.rr = s.rr;
      I.res.rel_res = s.bb > 0 ? (S.prm.sync_stride > 0 ? std::sqrt(s.rr / s.bb) : -1.0) : 0.0;
      (S.prm.verify_f64 && s.bb > 0 ? check : rest).push_back(q);
    }
  }
  const hipStream_t stm = S.stream;
  const Ctx c = S.view.ctx();
  // the last iteration's x update (k_cg_update_p applies each earlier one)
  for (int q : rest) {
    const Instance& I = *S.inst[q];
    k_cg_finish_x<<<grid_for(I.leaf.n, kSwizzle), 64, 0, stm>>>(c, I.L(I.leaf), I.leaf.n, kSwizzle | F_INERT);
  }
  if (!check.empty()) {
    const Group G = group_lists(S, k, GRP_FINISH, check, int(check.size()), nullptr);
    k_cg_finish_x<<<grid_for(G.leaf.n, kSwizzle), 64, 0, stm>>>(c, GL(S, G.leaf), G.leaf.n, kSwizzle | F_INERT);
    enqueue_verify(S, G, check);
  }
  for (int q : rest) {
    Instance& I = *S.inst[q];
    if (S.refine) {  // b = 0: x = 0
      HIPC(hipMemsetAsync(S.d_x64 + size_t(I.tile_begin) * 512, 0, size_t(I.ntiles) * 512 * sizeof(double), stm));
      HIPC(hipMemsetAsync(S.d_ch[CG_CH_X] + size_t(I.tile_begin) * 512, 0, size_t(I.ntiles) * 512 * sizeof(float), stm));
      I.res.converged = 1;
      I.res.rr = 0.0;
      I.res.rel_res_f64 = 0.0;
    } else if (S.prm.verify_f64) {
      I.res.rel_res_f64 = 0.0;  // b = 0: x = 0 is exact
    }
    finish_request(S, I);
  }
  HIPC(hipGetLastError());
}

// Requests whose FP64 check has arrived.  verify_f64: reported with the
// checked residual.  refine: reported once |b - A x|^2 < rel_tol^2 |b|^2 (x =
// float(x64) then, one group), or after kMaxRefineSteps steps; otherwise
// another step.
void server_verified(cg_server& S, int k) {
  std::vector<int> passed;
  for (int q = 0; q < S.nslots; q++) {
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
      passed.push_back(q);
    } else {
      I.ir_cont = 1;
      I.iters_acc = I.res.iters;
      I.state = INST_QUEUED;
    }
  }
  if (passed.empty()) return;
  const Group G = group_lists(S, k, GRP_PASSED, passed, 0, nullptr);
  k_ir_accumulate<<<grid_for(G.leaf.n, kSwizzle), 64, 0, S.stream>>>(S.view.ctx(), GL(S, G.leaf), G.leaf.n, S.d_x64, 1);
  HIPC(hipGetLastError());
  for (int q : passed) finish_request(S, *S.inst[q]);
}

// One step: an iteration over the current active set, then (one iteration
// behind, so the device keeps working) finished requests, new requests and a
// reassembly of the worklists for the next step.
// One CG iteration over the assembled active set: launched kernel by kernel on
// a set's first step, captured on its second, replayed after that.
void server_iteration(cg_server& S) {
  cg_problem& P = S.view;
  for (size_t i = 0; i < S.retired.size();) {
    const hipError_t e = hipEventQuery(S.retired[i].second);
    if (e == hipErrorNotReady) { i++; continue; }
    HIPC(e);
    HIPC(hipGraphExecDestroy(S.retired[i].first));
    S.ev_free.push_back(S.retired[i].second);
    S.retired.erase(S.retired.begin() + i);
  }
  if (S.graph && S.graph_ver == S.asm_ver) {
    HIPC(hipGraphLaunch(S.graph, S.stream));
  } else if (S.last_iter_ver == S.asm_ver) {
    if (S.graph) {  // every launch of the old graph is ahead of this event
      hipEvent_t ev;
      if (S.ev_free.empty()) {
        HIPC(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      } else {
        ev = S.ev_free.back();
        S.ev_free.pop_back();
      }
      S.retired.emplace_back(S.graph, ev);
      S.graph = nullptr;
      HIPC(hipEventRecord(ev, S.stream));
    }
    hipGraph_t g;
    HIPC(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
    enqueue_cg_iteration(P, S.v, S.prm.pure_neumann, S.prm.storage);
    HIPC(hipStreamEndCapture(S.stream, &g));
    const hipError_t e = hipGraphInstantiate(&S.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPC(e);
    S.graph_ver = S.asm_ver;
    HIPC(hipGraphLaunch(S.graph, S.stream));
  } else {
    enqueue_cg_iteration(P, S.v, S.prm.pure_neumann, S.prm.storage);
  }
  S.last_iter_ver = S.asm_ver;
}

void server_step(cg_server& S, int32_t* n_active) {
  const int k = int(S.step_no & 1);
  if (!S.pending.empty()) reap_releases(S, false);
  if (!S.active.empty() && !S.dirty) {
    server_iteration(S);
    HIPC(hipGetLastError());
    int maxslot = 0;
    for (int q : S.active) maxslot = std::max(maxslot, q + 1);
    HIPC(hipMemcpyAsync(S.h_snap[k], S.d_ps, sizeof(ProbState) * maxslot, hipMemcpyDeviceToHost, S.stream));
    HIPC(hipEventRecord(S.ev_snap[k], S.stream));
    S.snap_active[k] = S.active;
    S.snap_pending[k] = true;
  }
  server_collect(S, 1 - k);
  if (S.prm.verify_f64 || S.refine) server_verified(S, k);
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
