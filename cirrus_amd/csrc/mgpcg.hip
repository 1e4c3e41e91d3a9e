// Host side of the MI355X MGPCG library: problem upload, worklists, the
// V-cycle and CG launch sequences, hipGraph replay of one CG iteration with
// device-side convergence flags, and the C ABI of include/cirrus_mgpcg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <type_traits>
#include <vector>

#include "cirrus_mgpcg.h"
#include "kernels.hpp"
#include "tilefile.hpp"
#include "tree.hpp"

namespace {

thread_local std::string g_err;

struct Fail {
  cg_status st;
};
[[noreturn]] void fail(cg_status st, const std::string& msg) {
  g_err = msg;
  throw Fail{st};
}
#define HIPC(expr)                                                                                     \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      fail(e_ == hipErrorOutOfMemory ? CG_ERR_NOMEM : CG_ERR_HIP,                                     \
           std::string(#expr) + ": " + hipGetErrorString(e_));                                         \
  } while (0)

template <class F>
cg_status guard(F f) {
  try {
    f();
    return CG_OK;
  } catch (const Fail& e) {
    return e.st;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return CG_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return CG_ERR_INVALID;
  }
}

struct DList {
  int off = 0, n = 0;  // in TileRec units
};

}  // namespace

using namespace cg;

// A pool of one or more independent problems (octrees) solved together: tiles
// of all problems share the channel arrays, each worklist spans all problems,
// and every problem has its own ProbState, leaf-slot range and coarse plan.
struct cg_problem {
  bool owner = true;  // false: a view whose device arrays belong to someone else
  int nprob = 0;
  int n = 0;  // tiles of all problems
  int max_level = 0;
  int n_leaf = 0, n_ghost = 0;
  std::vector<int> ext_of;       // internal -> caller index (problems concatenated)
  std::vector<int> leaf_begin;   // per problem: first leaf slot
  std::vector<int> leaf_count;   // per problem: LEAF tiles
  std::vector<int> coarse_C;     // per problem: deepest LDS-resident level (-1: none)
  std::vector<TileRec> tiles;    // one record per tile (internal order)
  DList leaf, gather, applyl, ghosts, all;
  std::vector<DList> act, lvl;   // per level, all problems (operator API)
  std::vector<DList> actg, lvlg; // per level, tiles the V-cycle smooths with grid-wide launches
  std::vector<CoarsePlan> plans; // LDS-resident coarse levels, one per problem with C >= 0
  std::vector<CoarseSlot> coarse_slots;
  // schedule (no effect on results): XCD-aware tile order, LDS-resident coarse
  // levels given up to grid-wide launches, tiles per wavefront of the colour sweeps
  int swz = F_SWIZZLE;
  int coarse_shift = 0;
  bool rz_in_vcycle = false;     // every LEAF tile is on a grid-wide level >= 1 (its up-sweep can sum r.z)
  int gs_block = 1;
  int device = 0;           // the HIP device current when the handle was created
  int eager_max = 16384;    // levels of at most this many tiles: every load before one wait (prolongation pass, restriction)
  int fused16_max = 16384;  // 16-bit levels of at most this many tiles: first down-sweep iteration in one launch
  int gs_pair = 0;          // RB-GS iterations after the first as one launch (k_gs_pair; needs d_tiles; slower, off)
  std::vector<int> lev, tbase, coarse_auto;
  std::vector<uint8_t> ty;

  // device
  uint8_t* d_ctype = nullptr;
  uint8_t* d_inert = nullptr;  // per tile: the current solve leaves it untouched (k_cg_init)
  uint8_t* d_fmask = nullptr;   // face masks, then (same allocation) the colour-split masks d_fsplit
  uint32_t* d_fsplit = nullptr;
  float* d_diag = nullptr;  // the colour sweep's diagonal per cell (k_diag), for k_gs_first
  float* d_amgc = nullptr;    // AMG stored coefficients: off x, off y, off z, diag (4 x cells; cg_amg_prepare)
  uint8_t* d_amgct = nullptr; // AMG cell types
  bool amg_ready = false;
  float* d_ch[CG_NUM_CHANNELS] = {};
  float* d_zb = nullptr;
  bf16* d_zh = nullptr;     // 16-bit z, V-cycle rhs and scratch (bfloat16 or float16 values; allocated at the first such solve)
  bf16* d_rhsvh = nullptr;
  bf16* d_zbh = nullptr;
  double* d_x64 = nullptr;  // iterative refinement: FP64 x and residual, the caller's b, per-problem sums
  double* d_r64 = nullptr;
  float* d_bsave = nullptr;
  double* d_psum = nullptr;
  float* d_api[3] = {nullptr, nullptr, nullptr};  // colour-split copies of channels for the operator API
  float* d_vel[3] = {nullptr, nullptr, nullptr};  // face velocities u, v, w (channels 6-8, allocated on first use)
  float* d_gsum = nullptr;
  ProbState* d_ps = nullptr;
  int32_t* d_dflag = nullptr;
  const int32_t* d_pmap = nullptr;  // finalize workgroup -> problem (pools whose problems are a subset)
  CoarsePlan* d_plans = nullptr;
  double* d_part = nullptr;
  double* d_part2 = nullptr;
  double* d_scalar = nullptr;
  TileRec* d_lists = nullptr;
  TileRec* d_tiles = nullptr;    // every tile's record in canonical order (the worklists are gathered from it)
  TileRec* d_tmp_list = nullptr;
  int tmp_list_cap = 0;
  int32_t* d_ext = nullptr;      // internal tile -> caller index (channel moves from and to device buffers)
  CoarseSlot* d_coarse = nullptr;
  hipStream_t stream = nullptr;
  int* h_pin = nullptr;
  ProbState* h_ps = nullptr;     // pinned staging of the ProbState uploads
  hipEvent_t ev_ps = nullptr;    // recorded after the last upload from h_ps
  bool ps_pending = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipEvent_t ev_io[2] = {nullptr, nullptr};  // orders device channel moves against the caller's stream
  hipGraphExec_t graph = nullptr;
  std::string graph_key;

  const TileRec* L(const DList& d) const { return d_lists + d.off; }
  Ctx ctx(float* z, float* r) const {
    Ctx c;
    c.ctype = d_ctype;
    c.fmask = d_fmask;
    c.fsplit = d_fsplit;
    c.diag = d_diag;
    c.x = d_ch[CG_CH_X];
    c.r = r;
    c.p = d_ch[CG_CH_P];
    c.ap = d_ch[CG_CH_AP];
    c.z = z;
    c.rhsv = d_ch[CG_CH_TMP];
    c.zb = d_zb;
    c.zh = d_zh;
    c.rhsvh = d_rhsvh;
    c.zbh = d_zbh;
    c.zf = reinterpret_cast<f16*>(d_zh);
    c.rhsvf = reinterpret_cast<f16*>(d_rhsvh);
    c.zbf = reinterpret_cast<f16*>(d_zbh);
    c.gsum = d_gsum;
    c.ps = d_ps;
    c.dflag = d_dflag;
    c.pmap = d_pmap;
    c.part = d_part;
    c.part2 = d_part2;
    c.tiles = gs_pair ? d_tiles : nullptr;
    c.inert = d_inert;
    return c;
  }
  Ctx ctx() const { return ctx(d_ch[CG_CH_Z], d_ch[CG_CH_R]); }
  ~cg_problem();
};

cg_problem::~cg_problem() {
  if (!owner) return;
  if (graph) (void)hipGraphExecDestroy(graph);
  for (auto e : ev) if (e) (void)hipEventDestroy(e);
  for (auto e : ev_io) if (e) (void)hipEventDestroy(e);
  if (h_pin) (void)hipHostFree(h_pin);
  if (h_ps) (void)hipHostFree(h_ps);
  if (ev_ps) (void)hipEventDestroy(ev_ps);
  void* ptrs[] = {d_ctype, d_inert, d_fmask, d_diag, d_zb, d_zh, d_rhsvh, d_zbh, d_x64, d_r64, d_bsave, d_psum, d_gsum, d_ps,
                  d_dflag, d_plans, d_part, d_part2, d_scalar, d_lists, d_tiles, d_tmp_list, d_coarse, d_amgc, d_amgct, d_ext};
  for (void* q : ptrs) if (q) (void)hipFree(q);
  for (auto q : d_ch) if (q) (void)hipFree(q);
  for (auto q : d_api) if (q) (void)hipFree(q);
  for (auto q : d_vel) if (q) (void)hipFree(q);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

constexpr int kSwizzle = F_SWIZZLE;

// A device buffer freed when it goes out of scope (also when a call throws).
struct DeviceBuffer {
  void* p = nullptr;
  explicit DeviceBuffer(size_t bytes) { HIPC(hipMalloc(&p, bytes)); }
  ~DeviceBuffer() { if (p) (void)hipFree(p); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

inline unsigned grid_for(int n, int flags) { return (flags & F_SWIZZLE) ? unsigned((n + 7) & ~7) : unsigned(n); }

// ---------------------------------------------------------------------------
// Problem construction
// ---------------------------------------------------------------------------
// One tree: validation, canonical order (level, then Morton code of the tile
// coordinate), neighbour / parent / ghost-child links and tile records.  Output
// arrays are local to the tree (tiles[i], ext[i] = caller index of tile i,
// lev[i], ty[i]); indices stored in the records are global (offset by `base`,
// leaf and ghost slots by leaf_base / ghost_base).
struct TreeBuild {
  std::vector<TileRec> tiles;
  std::vector<int> ext, lev;
  std::vector<uint8_t> ty;
  int nleaf = 0, nghost = 0;
};
void build_tree(const cg_tree_desc* d, int prob, int base, int leaf_base, int ghost_base, TreeBuild& B) {
  if (!d || d->num_tiles <= 0 || !d->level || !d->coord || !d->type || !d->cell_type)
    fail(CG_ERR_INVALID, "cg_problem_create: empty or incomplete tree description");
  if (!(d->h0 > 0)) fail(CG_ERR_INVALID, "cg_problem_create: h0 must be positive");
  const int n = d->num_tiles;
  B.tiles.assign(n, TileRec());
  B.ext.assign(n, 0);
  B.lev.assign(n, 0);
  B.ty.assign(n, 0);
  std::vector<int>& lev = B.lev;
  std::vector<uint8_t>& ty = B.ty;
  std::vector<uint64_t> mk(n);
  for (int t = 0; t < n; t++) {
    int l = d->level[t];
    const int32_t* c = d->coord + 3 * t;
    if (l < 0 || l > 30) fail(CG_ERR_INVALID, "tile " + std::to_string(t) + ": level out of range");
    if (!key_ok(c[0], c[1], c[2])) fail(CG_ERR_INVALID, "tile " + std::to_string(t) + ": coordinate out of range");
    uint8_t tt = d->type[t];
    if (tt != LEAF && tt != NONLEAF && tt != GHOST) fail(CG_ERR_INVALID, "tile " + std::to_string(t) + ": bad tile type");
    mk[t] = morton3(uint32_t(c[0]), uint32_t(c[1]), uint32_t(c[2]));
  }
  std::vector<int> order(n);
  for (int t = 0; t < n; t++) order[t] = t;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    if (d->level[a] != d->level[b]) return d->level[a] < d->level[b];
    return mk[a] < mk[b];
  });
  std::unordered_map<uint64_t, int> idx;
  idx.reserve(size_t(n) * 2);
  std::vector<std::array<int32_t, 3>> cd(n);
  for (int i = 0; i < n; i++) {
    const int e = order[i];
    B.ext[i] = e;
    lev[i] = d->level[e];
    cd[i] = {d->coord[3 * e], d->coord[3 * e + 1], d->coord[3 * e + 2]};
    ty[i] = d->type[e];
    if (!idx.emplace(tile_key(lev[i], cd[i][0], cd[i][1], cd[i][2]), i).second)
      fail(CG_ERR_INVALID, "duplicate tile at level " + std::to_string(lev[i]));
  }
  auto find = [&](int l, int x, int y, int z) -> int {
    if (l < 0 || !key_ok(x, y, z)) return -1;
    auto it = idx.find(tile_key(l, x, y, z));
    return it == idx.end() ? -1 : it->second;
  };
  std::vector<int> leaf_slot(n, -1), gslot(n, -1);
  int nleaf = 0, nghost = 0;
  for (int i = 0; i < n; i++) {
    if (ty[i] == LEAF) leaf_slot[i] = leaf_base + nleaf++;
    if (ty[i] == GHOST) gslot[i] = ghost_base + nghost++;
  }
  B.nleaf = nleaf;
  B.nghost = nghost;
  if (nleaf == 0) fail(CG_ERR_INVALID, "tree has no LEAF tiles");
  for (int i = 0; i < n; i++) {
    TileRec& T = B.tiles[i];
    std::memset(&T, 0, sizeof(T));
    const int l = lev[i];
    const uint8_t my = ty[i];
    int nbl[6];  // face neighbours, local indices
    for (int f = 0; f < 6; f++) {
      int q[3] = {cd[i][0], cd[i][1], cd[i][2]};
      q[f % 3] += (f < 3) ? -1 : 1;
      const int nb = find(l, q[0], q[1], q[2]);
      nbl[f] = nb;
      T.nbr[f] = nb < 0 ? -1 : base + nb;
      T.ntypes |= uint32_t(nb >= 0 ? ty[nb] : 0) << (4 * f);
    }
    T.parent = -1;
    if (l > 0) {
      const int p = find(l - 1, cd[i][0] >> 1, cd[i][1] >> 1, cd[i][2] >> 1);
      if (p < 0) fail(CG_ERR_INVALID, "tile at level " + std::to_string(l) + " has no parent tile");
      if (my == GHOST && ty[p] != LEAF) fail(CG_ERR_INVALID, "a GHOST tile's parent must be a LEAF");
      if (my != GHOST && ty[p] != NONLEAF) fail(CG_ERR_INVALID, "a LEAF/NONLEAF tile's parent must be a NONLEAF");
      T.parent = base + p;
    }
    const uint32_t octant = uint32_t((cd[i][0] & 1) | ((cd[i][1] & 1) << 1) | ((cd[i][2] & 1) << 2));
    T.prob = prob;
    T.leaf_slot = leaf_slot[i];
    T.gslot = gslot[i];
    T.self = base + i;
    T.h = float(d->h0 / double(1u << l));
    T.meta = uint32_t(l) | (uint32_t(my) << 8) | (octant << 16);
    for (int o = 0; o < 8; o++) { T.gch[o] = -1; T.gcs[o] = -1; }
    if (my != GHOST) {
      for (int o = 0; o < 8; o++) {
        const int ch = find(l + 1, 2 * cd[i][0] + (o & 1), 2 * cd[i][1] + ((o >> 1) & 1), 2 * cd[i][2] + ((o >> 2) & 1));
        if (my == NONLEAF && (ch < 0 || ty[ch] == GHOST))
          fail(CG_ERR_INVALID, "a NONLEAF tile must have 8 LEAF/NONLEAF children");
        if (my == LEAF && ch >= 0) {
          if (ty[ch] != GHOST) fail(CG_ERR_INVALID, "a LEAF tile can only have GHOST children");
          T.gch[o] = base + ch;
          T.gcs[o] = gslot[ch];
          T.meta |= 1u << 24;
        }
      }
    }
    // smoothing-view coefficients (subtree at the tile's own level)
    view_coeffs(T, l, T.bc);
    // the cell types as 64 words of 8 cells (cell index i*64 + j*8 + k)
    constexpr uint64_t kOnes = 0x0101010101010101ull;
    auto words = [&](int t) {
      std::array<uint64_t, 64> w;
      std::memcpy(w.data(), d->cell_type + size_t(B.ext[t]) * 512, 512);
      return w;
    };
    const auto ct = words(i);
    bool all_int = true, no_int = true;
    for (int q = 0; q < 64; q++) {
      all_int = all_int && ct[q] == kOnes * C_INT;
      no_int = no_int && (ct[q] & (kOnes * C_INT)) == 0;
    }
    if (all_int) T.meta |= 1u << 25;
    if (no_int && my == LEAF && !tr_has_gchild(T)) T.meta |= 1u << 26;
    // every face mask M_OPEN_INT (as k_face_mask forms them): all cells
    // INTERIOR and no NEUMANN cell across a tile face (no neighbour: open).
    // The neighbour's face layer: x faces i = 7 / 0 (words 56..63 / 0..7),
    // y faces j = 7 / 0 (word i*8 + 7 / i*8), z faces k = 7 / 0 (byte 7 / 0 of
    // every word).
    bool open = all_int;
    for (int f = 0; f < 6 && open; f++) {
      if (nbl[f] < 0) continue;
      const auto nw = words(nbl[f]);
      const bool minus = f < 3;
      uint64_t neu = 0;
      for (int u = 0; u < 8; u++) {
        if (f % 3 == 0) neu |= nw[(minus ? 56 : 0) + u];
        else if (f % 3 == 1) neu |= nw[u * 8 + (minus ? 7 : 0)];
      }
      if (f % 3 == 2)
        for (int q = 0; q < 64; q++) neu |= (nw[q] >> (minus ? 56 : 0)) & 0xffu;
      if (neu & (kOnes * C_NEU)) open = false;
    }
    if (open) T.meta |= 1u << 27;
  }
  for (int i = 0; i < n; i++) {
    TileRec& T = B.tiles[i];
    for (int f = 0; f < 6; f++) T.nbr_parent[f] = T.nbr[f] < 0 ? -1 : B.tiles[T.nbr[f] - base].parent;
  }
}

// Deepest level C of a tree kept in LDS by the coarse kernel: levels 0..C hold
// at most CL_MAX_TILES tiles in total, at most CL_MAX_PER_LEVEL on any level,
// and no GHOST tile (-1 if level 0 alone is larger).  (A GHOST tile needs a LEAF
// tile one level up beside a NONLEAF one, i.e. at least 14 tiles on two levels,
// so the size bound alone already excludes them; the check keeps the kernel's
// assumption explicit.)  CIRRUS_COARSE_LEVELS=k caps it at k levels, for
// measurement.
int coarse_depth(const std::vector<int>& lev, const std::vector<uint8_t>& ty) {
  int L = 0;
  for (int l : lev) L = std::max(L, l);
  std::vector<int> cnt(L + 1, 0), cnt_act(L + 1, 0);
  for (size_t i = 0; i < lev.size(); i++) {
    cnt[lev[i]]++;
    if (ty[i] != GHOST) cnt_act[lev[i]]++;
  }
  int C = -1, total = 0;
  while (C + 1 <= L && C + 1 < CL_MAX_LEVELS && total + cnt[C + 1] <= CL_MAX_TILES && cnt[C + 1] <= CL_MAX_PER_LEVEL &&
         cnt[C + 1] == cnt_act[C + 1]) {
    C++;
    total += cnt[C];
  }
  if (const char* e = std::getenv("CIRRUS_COARSE_LEVELS")) C = std::max(-1, std::min(C, std::atoi(e) - 1));
  return C;
}

// Coarse plan of tiles [first, first + count) of `tiles` (global indices start at
// base): slots of levels 0..C in level order; `slots` is filled in by the caller.
void coarse_plan(const std::vector<TileRec>& tiles, int base, int C, int prob, CoarsePlan& cp, std::vector<CoarseSlot>& out) {
  std::memset(&cp, 0, sizeof(cp));
  cp.prob = prob;
  cp.C = C;
  const int n = int(tiles.size());
  std::unordered_map<int, int> slot_of;
  const int first = int(out.size());
  for (int l = 0; l <= C; l++) {
    cp.lvl_begin[l] = int(out.size()) - first;
    for (int i = 0; i < n; i++) {
      if (tr_level(tiles[i]) != l) continue;
      slot_of[base + i] = int(out.size()) - first;
      CoarseSlot cs;
      std::memset(&cs, 0, sizeof(cs));
      cs.tile = base + i;
      out.push_back(cs);
    }
  }
  cp.lvl_begin[C + 1] = int(out.size()) - first;
  cp.nslots = cp.lvl_begin[C + 1];
  for (int s = first; s < int(out.size()); s++) {
    CoarseSlot& cs = out[s];
    const TileRec& T = tiles[cs.tile - base];
    for (int f = 0; f < 6; f++) cs.nbr[f] = T.nbr[f] < 0 ? -1 : slot_of.at(T.nbr[f]);
    cs.parent = T.parent < 0 ? -1 : slot_of.at(T.parent);
    cs.level = tr_level(T);
    cs.h = T.h;
    cs.type = tr_type(T);
    cs.octant = uint8_t(tr_octant(T));
    for (int f = 0; f < 7; f++) cs.bc[f] = T.bc[f];
  }
}

// Worklists and coarse plans of a pool for its schedule (host side only):
// each problem keeps its deepest LDS-resident coarse level coarse_auto -
// coarse_shift (-1: none), the levels above it go to grid-wide launches.  The
// lists' tile indices come back in `ids` (every list's DList indexes it); the
// coarse plans hold slot offsets until upload_schedule patches them.
void plan_schedule(cg_problem& P, std::vector<int32_t>& ids) {
  const int n = P.n, nprob = P.nprob;
  const std::vector<int>& lev = P.lev;
  const std::vector<uint8_t>& ty = P.ty;
  const std::vector<int>& tbase = P.tbase;
  for (int q = 0; q < nprob; q++) P.coarse_C[q] = std::max(-1, P.coarse_auto[q] - P.coarse_shift);
  const int L = P.max_level;
  // worklists (canonical order, problems one after another): the tile indices
  // of each list first, then one sized array of records
  std::vector<std::vector<int>> bucket(L + 1);  // tiles of each level, in order
  for (int i = 0; i < n; i++) bucket[lev[i]].push_back(i);
  std::vector<std::pair<DList*, std::vector<int>>> sel;
  auto add_list = [&](DList* dl, const std::vector<int>& from, auto pred) {
    std::vector<int> ids;
    for (int i : from) if (pred(i)) ids.push_back(i);
    sel.emplace_back(dl, std::move(ids));
  };
  std::vector<int> every(n);
  for (int i = 0; i < n; i++) every[i] = i;
  auto grid_level = [&](int i) { return lev[i] > P.coarse_C[P.tiles[i].prob]; };
  add_list(&P.leaf, every, [&](int i) { return ty[i] == LEAF; });
  P.rz_in_vcycle = true;
  for (int i = 0; i < n; i++)
    if (ty[i] == LEAF && !(lev[i] >= 1 && grid_level(i))) P.rz_in_vcycle = false;
  add_list(&P.gather, every, [&](int i) { return ty[i] == LEAF && tr_has_gchild(P.tiles[i]); });
  add_list(&P.applyl, every, [&](int i) { return ty[i] == LEAF || ty[i] == GHOST; });
  add_list(&P.ghosts, every, [&](int i) { return ty[i] == GHOST; });
  add_list(&P.all, every, [&](int) { return true; });
  P.act.assign(L + 1, DList());
  P.lvl.assign(L + 1, DList());
  P.actg.assign(L + 1, DList());
  P.lvlg.assign(L + 1, DList());
  for (int l = 0; l <= L; l++) {
    add_list(&P.act[l], bucket[l], [&](int i) { return ty[i] != GHOST; });
    add_list(&P.lvl[l], bucket[l], [&](int) { return true; });
    add_list(&P.actg[l], bucket[l], [&](int i) { return ty[i] != GHOST && grid_level(i); });
    add_list(&P.lvlg[l], bucket[l], [&](int i) { return grid_level(i); });
  }
  // the index lists go to the device, which gathers the records from d_tiles
  ids.clear();
  size_t total = 0;
  for (auto& s : sel) total += s.second.size();
  ids.reserve(total);
  for (auto& s : sel) {
    s.first->off = int(ids.size());
    s.first->n = int(s.second.size());
    ids.insert(ids.end(), s.second.begin(), s.second.end());
  }

  // coarse plans and slots
  P.plans.clear();
  P.coarse_slots.clear();
  for (int q = 0; q < nprob; q++) {
    if (P.coarse_C[q] < 0) continue;
    CoarsePlan cp;
    const int first = int(P.coarse_slots.size());
    std::vector<TileRec> mine(P.tiles.begin() + tbase[q], P.tiles.begin() + tbase[q + 1]);
    coarse_plan(mine, tbase[q], P.coarse_C[q], q, cp, P.coarse_slots);
    cp.slots = reinterpret_cast<const CoarseSlot*>(size_t(first));  // offset, patched after upload
    P.plans.push_back(cp);
  }
  for (const CoarsePlan& cp : P.plans) {  // every slot link stays inside its own plan (the kernel's LDS)
    const size_t first = size_t(cp.slots);
    if (cp.nslots > CL_MAX_TILES || first + size_t(cp.nslots) > P.coarse_slots.size())
      fail(CG_ERR_INVALID, "coarse plan: too many slots");
    for (int k = 0; k < cp.nslots; k++) {
      const CoarseSlot& cs = P.coarse_slots[first + k];
      for (int f = 0; f < 6; f++)
        if (cs.nbr[f] >= cp.nslots) fail(CG_ERR_INVALID, "coarse plan: neighbour slot out of range");
      if (cs.parent >= cp.nslots || cs.tile < 0 || cs.tile >= P.n) fail(CG_ERR_INVALID, "coarse plan: slot out of range");
    }
  }
  for (int32_t t : ids)
    if (t < 0 || t >= P.n) fail(CG_ERR_INVALID, "worklist: tile index out of range");
}

// The planned lists and coarse plans to the device (records gathered there
// from d_tiles); the captured iteration is dropped.
void upload_schedule(cg_problem& P, const std::vector<int32_t>& ids) {
  const size_t total = ids.size();
  HIPC(hipStreamSynchronize(P.stream));  // no kernel may still read the old lists
  if (P.d_lists) HIPC(hipFree(P.d_lists));
  if (P.d_coarse) HIPC(hipFree(P.d_coarse));
  if (P.d_plans) HIPC(hipFree(P.d_plans));
  P.d_lists = nullptr;
  P.d_coarse = nullptr;
  P.d_plans = nullptr;
  HIPC(hipMalloc(&P.d_lists, sizeof(TileRec) * std::max<size_t>(1, total)));
  if (total) {
    DeviceBuffer d_ids(sizeof(int32_t) * total);
    HIPC(hipMemcpyAsync(d_ids.p, ids.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice, P.stream));
    const int64_t words = int64_t(total) * int64_t(sizeof(TileRec) / 16);
    const unsigned blocks = unsigned(std::min<int64_t>((words + 255) / 256, 65536));
    k_gather_tiles<<<blocks, 256, 0, P.stream>>>(P.d_tiles, static_cast<const int32_t*>(d_ids.p), int64_t(total), P.d_lists);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(P.stream));  // d_ids is freed at the end of this scope
  }
  if (!P.plans.empty()) {
    HIPC(hipMalloc(&P.d_coarse, sizeof(CoarseSlot) * P.coarse_slots.size()));
    HIPC(hipMemcpyAsync(P.d_coarse, P.coarse_slots.data(), sizeof(CoarseSlot) * P.coarse_slots.size(),
                        hipMemcpyHostToDevice, P.stream));
    for (auto& cp : P.plans) cp.slots = P.d_coarse + size_t(cp.slots);
    HIPC(hipMalloc(&P.d_plans, sizeof(CoarsePlan) * P.plans.size()));
    HIPC(hipMemcpyAsync(P.d_plans, P.plans.data(), sizeof(CoarsePlan) * P.plans.size(), hipMemcpyHostToDevice, P.stream));
  }
  HIPC(hipStreamSynchronize(P.stream));
  if (P.graph) HIPC(hipGraphExecDestroy(P.graph));
  P.graph = nullptr;
  P.graph_key.clear();
}

void build_schedule(cg_problem& P) {
  std::vector<int32_t> ids;
  plan_schedule(P, ids);
  upload_schedule(P, ids);
}

void build_pool(const cg_tree_desc* descs, int nprob, cg_problem& P) {
  if (!descs || nprob <= 0) fail(CG_ERR_INVALID, "cg_problem_create: no tree description");
  P.nprob = nprob;
  std::vector<int> tbase(nprob + 1, 0);
  for (int q = 0; q < nprob; q++) {
    if (descs[q].num_tiles <= 0) fail(CG_ERR_INVALID, "cg_problem_create: empty or incomplete tree description");
    tbase[q + 1] = tbase[q] + descs[q].num_tiles;
  }
  const int n = tbase[nprob];
  P.n = n;
  P.ext_of.assign(n, 0);
  P.tiles.assign(n, TileRec());
  std::vector<int> lev(n);
  std::vector<uint8_t> ty(n);
  P.leaf_begin.assign(nprob, 0);
  P.leaf_count.assign(nprob, 0);
  P.coarse_C.assign(nprob, -1);
  // leaf and ghost slot bases from the type counts, then the trees are built
  // independently on up to 16 host threads (each writes its own tile range)
  std::vector<int> leaf_base(nprob + 1, 0), ghost_base(nprob + 1, 0);
  for (int q = 0; q < nprob; q++) {
    int nl = 0, ng = 0;
    if (descs[q].type)
      for (int t = 0; t < descs[q].num_tiles; t++) { nl += descs[q].type[t] == LEAF; ng += descs[q].type[t] == GHOST; }
    leaf_base[q + 1] = leaf_base[q] + nl;
    ghost_base[q + 1] = ghost_base[q] + ng;
  }
  auto build_one = [&](int q) {
    TreeBuild B;
    build_tree(&descs[q], q, tbase[q], leaf_base[q], ghost_base[q], B);
    for (int i = 0; i < descs[q].num_tiles; i++) {
      P.tiles[tbase[q] + i] = B.tiles[i];
      P.ext_of[tbase[q] + i] = tbase[q] + B.ext[i];
      lev[tbase[q] + i] = B.lev[i];
      ty[tbase[q] + i] = B.ty[i];
    }
    P.leaf_begin[q] = leaf_base[q];
    P.leaf_count[q] = B.nleaf;
    P.coarse_C[q] = coarse_depth(B.lev, B.ty);
  };
  const int nthreads = std::min(nprob, std::max(1, std::min(16, int(std::thread::hardware_concurrency()))));
  if (nthreads <= 1) {
    for (int q = 0; q < nprob; q++) build_one(q);
  } else {
    std::atomic<int> next{0};
    std::mutex mu;
    cg_status first_st = CG_OK;
    std::string first_msg;
    int first_q = nprob;
    auto worker = [&]() {
      for (int q = next++; q < nprob; q = next++) {
        try {
          build_one(q);
        } catch (const Fail& e) {
          std::lock_guard<std::mutex> lk(mu);
          if (q < first_q) { first_q = q; first_st = e.st; first_msg = g_err; }
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> lk(mu);
          if (q < first_q) { first_q = q; first_st = CG_ERR_INVALID; first_msg = e.what(); }
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    if (first_st != CG_OK) fail(first_st, first_msg);  // the failure of the first bad tree, as a serial build reports it
  }
  const int leaf_total = leaf_base[nprob], ghost_total = ghost_base[nprob];
  P.n_leaf = leaf_total;
  P.n_ghost = ghost_total;
  P.max_level = 0;
  for (int i = 0; i < n; i++) P.max_level = std::max(P.max_level, lev[i]);

  P.lev = lev;
  P.ty = ty;
  P.tbase = tbase;
  P.coarse_auto = P.coarse_C;
  // every host-side plan is made before the first device call
  std::vector<int32_t> sched_ids;
  plan_schedule(P, sched_ids);

  // device buffers
  HIPC(hipGetDevice(&P.device));
  HIPC(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
  const size_t cells = size_t(n) * 512;
  HIPC(hipMalloc(&P.d_ctype, cells));
  HIPC(hipMalloc(&P.d_fmask, 2 * cells));
  P.d_fsplit = reinterpret_cast<uint32_t*>(P.d_fmask + cells);
  HIPC(hipMalloc(&P.d_diag, cells * sizeof(float)));
  for (int c = 0; c < CG_NUM_CHANNELS; c++) {
    HIPC(hipMalloc(&P.d_ch[c], cells * sizeof(float)));
    HIPC(hipMemsetAsync(P.d_ch[c], 0, cells * sizeof(float), P.stream));
  }
  HIPC(hipMalloc(&P.d_zb, cells * sizeof(float)));
  HIPC(hipMemsetAsync(P.d_zb, 0, cells * sizeof(float), P.stream));
  HIPC(hipMalloc(&P.d_inert, size_t(n)));
  HIPC(hipMemsetAsync(P.d_inert, 0, size_t(n), P.stream));
  size_t g64 = size_t(std::max(1, ghost_total)) * 64 * sizeof(float);
  HIPC(hipMalloc(&P.d_gsum, g64));
  HIPC(hipMemsetAsync(P.d_gsum, 0, g64, P.stream));
  HIPC(hipMalloc(&P.d_ps, sizeof(ProbState) * nprob));
  HIPC(hipMemsetAsync(P.d_ps, 0, sizeof(ProbState) * nprob, P.stream));
  HIPC(hipMalloc(&P.d_dflag, sizeof(int32_t) * nprob));
  HIPC(hipMemsetAsync(P.d_dflag, 0, sizeof(int32_t) * nprob, P.stream));
  size_t np = size_t(std::max(n, 64));
  HIPC(hipMalloc(&P.d_part, np * sizeof(double)));
  HIPC(hipMalloc(&P.d_part2, np * sizeof(double)));
  HIPC(hipMalloc(&P.d_scalar, 8 * sizeof(double)));
  HIPC(hipMalloc(&P.d_tiles, sizeof(TileRec) * size_t(n)));
  HIPC(hipMemcpyAsync(P.d_tiles, P.tiles.data(), sizeof(TileRec) * size_t(n), hipMemcpyHostToDevice, P.stream));
  upload_schedule(P, sched_ids);
  // cell types: each tree's array goes up as given (into d_fmask, free until the
  // face masks are formed) and is permuted to the canonical order on the device
  for (int q = 0; q < nprob; q++)
    HIPC(hipMemcpyAsync(P.d_fmask + size_t(tbase[q]) * 512, descs[q].cell_type, size_t(descs[q].num_tiles) * 512,
                        hipMemcpyHostToDevice, P.stream));
  HIPC(hipMalloc(&P.d_ext, sizeof(int32_t) * size_t(n)));
  HIPC(hipMemcpyAsync(P.d_ext, P.ext_of.data(), sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, P.stream));
  k_permute_ctype<<<unsigned(n), 64, 0, P.stream>>>(P.d_fmask, P.d_ext, n, P.d_ctype);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(P.stream));
  k_face_mask<<<P.all.n, 64, 0, P.stream>>>(P.d_ctype, P.d_fmask, P.d_fsplit, P.L(P.all), P.all.n);
  k_diag<<<P.all.n, 64, 0, P.stream>>>(P.ctx(), P.L(P.all), P.all.n, P.d_diag);
  HIPC(hipGetLastError());
  HIPC(hipHostMalloc(&P.h_pin, 2 * sizeof(int) * nprob, hipHostMallocDefault));
  HIPC(hipHostMalloc(&P.h_ps, sizeof(ProbState) * nprob, hipHostMallocDefault));
  HIPC(hipEventCreateWithFlags(&P.ev_ps, hipEventDisableTiming));
  HIPC(hipEventCreateWithFlags(&P.ev[0], hipEventDisableTiming));
  HIPC(hipEventCreateWithFlags(&P.ev[1], hipEventDisableTiming));
  HIPC(hipEventCreateWithFlags(&P.ev_io[0], hipEventDisableTiming));
  HIPC(hipEventCreateWithFlags(&P.ev_io[1], hipEventDisableTiming));
  HIPC(hipStreamSynchronize(P.stream));
}

void check_channel(int ch) {
  if (ch < 0 || ch >= CG_NUM_CHANNELS) fail(CG_ERR_INVALID, "channel index out of range");
}
// Any channel the caller may move: 0-5, or the face velocities 6-8 (allocated
// and zeroed on first use).
float* any_channel(cg_problem& P, int ch) {
  if (ch >= 0 && ch < CG_NUM_CHANNELS) return P.d_ch[ch];
  if (ch < CG_CH_U || ch > CG_CH_W) fail(CG_ERR_INVALID, "channel index out of range");
  if (!P.owner) fail(CG_ERR_INVALID, "velocity channels need an owning problem handle");
  float*& v = P.d_vel[ch - CG_CH_U];
  if (!v) {
    HIPC(hipMalloc(&v, size_t(P.n) * 512 * sizeof(float)));
    HIPC(hipMemsetAsync(v, 0, size_t(P.n) * 512 * sizeof(float), P.stream));
  }
  return v;
}

// The caller's device buffer of one channel (512 floats per tile, caller's
// tile order) must be device-accessible memory holding the whole channel.
void check_device_buffer(const void* ptr, size_t bytes, const char* what) {
  if (!ptr) fail(CG_ERR_INVALID, std::string(what) + ": null device pointer");
  hipPointerAttribute_t a;
  const hipError_t e = hipPointerGetAttributes(&a, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // the query's error is the answer, not a pending failure
    fail(CG_ERR_INVALID, std::string(what) + ": the pointer is not memory known to HIP");
  }
  if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged && a.type != hipMemoryTypeHost)
    fail(CG_ERR_INVALID, std::string(what) + ": the pointer is not device-accessible memory");
  void* base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, const_cast<void*>(ptr)) == hipSuccess && base) {
    const size_t off = size_t(static_cast<const char*>(ptr) - static_cast<const char*>(base));
    if (off > size || size - off < bytes)
      fail(CG_ERR_INVALID, std::string(what) + ": the buffer is smaller than 512 floats per tile");
  } else {
    (void)hipGetLastError();
  }
}

// A channel moved between the caller's device buffer and the handle, on the
// handle's stream, ordered after the work already enqueued on the caller's
// stream and before the work the caller enqueues there next; no host wait.
void channel_device(cg_problem& P, int ch, const float* src, float* dst, hipStream_t cs) {
  const bool set = src != nullptr;
  float* d = any_channel(P, ch);
  check_device_buffer(set ? static_cast<const void*>(src) : static_cast<const void*>(dst),
                      size_t(P.n) * 512 * sizeof(float), set ? "cg_set_channel_device" : "cg_get_channel_device");
  HIPC(hipEventRecord(P.ev_io[0], cs));
  HIPC(hipStreamWaitEvent(P.stream, P.ev_io[0], 0));
  k_permute_channel<<<unsigned(P.n), 64, 0, P.stream>>>(set ? src : d, P.d_ext, P.n, set ? d : dst, set ? 1 : 0);
  HIPC(hipGetLastError());
  HIPC(hipEventRecord(P.ev_io[1], P.stream));
  HIPC(hipStreamWaitEvent(cs, P.ev_io[1], 0));
}

// upload every problem's ProbState (synchronous)
// tol_q / maxit_q: per-problem tolerance and iteration cap (null: tol / max_iters for all)
void set_prob_state(cg_problem& P, double tol, int max_iters, int pure_neumann, int check_stride = 1,
                    const double* tol_q = nullptr, const int* maxit_q = nullptr, const double* thr_abs_q = nullptr) {
  // staged in pinned memory, so the upload is ordered on the stream with no
  // host wait (only for the previous upload from the same buffer)
  std::vector<ProbState> local;
  ProbState* st = P.h_ps;
  if (st) {
    if (P.ps_pending) HIPC(hipEventSynchronize(P.ev_ps));
  } else {
    local.resize(P.nprob);
    st = local.data();
  }
  for (int q = 0; q < P.nprob; q++) {
    ProbState& s = st[q];
    std::memset(&s, 0, sizeof(s));
    const double t = tol_q ? tol_q[q] : tol;
    s.tol2 = t * t;
    s.thr_abs = thr_abs_q ? thr_abs_q[q] : 0.0;
    s.max_iters = maxit_q ? maxit_q[q] : max_iters;
    s.leaf_begin = P.leaf_begin[q];
    s.leaf_count = P.leaf_count[q];
    s.pure_neumann = pure_neumann;
    s.check_stride = check_stride;
  }
  HIPC(hipMemcpyAsync(P.d_ps, st, sizeof(ProbState) * P.nprob, hipMemcpyHostToDevice, P.stream));
  HIPC(hipMemsetAsync(P.d_dflag, 0, sizeof(int32_t) * P.nprob, P.stream));
  if (P.h_ps) {
    HIPC(hipEventRecord(P.ev_ps, P.stream));
    P.ps_pending = true;
  } else {
    HIPC(hipStreamSynchronize(P.stream));  // a local buffer
  }
}

// ---------------------------------------------------------------------------
// Launch sequences
// ---------------------------------------------------------------------------
struct VParams {
  int nu = 1, coarsest = 10, smoother = CG_SMOOTH_RBGS;
  float omega = 2.f / 3.f, ooa = 1.f, pcoef = 1.f;
  int check_done = 1;
  int amg = 0;    // the preconditioner is VCycleAMG (cg_amg_solve)
  int inert = 0;  // a CG iteration's V-cycle: skip the tiles the solve's init found inert (F_INERT)
};

// Workgroups of a looped colour sweep (tile block 8): as many as the device
// holds at once (the kernel's occupancy per CU times the CUs), rounded down
// to a multiple of the 8 XCDs; found once per kernel instance.
template <class V, bool DOT>
int loop_waves() {
  static const int w = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_gs_color_loop<V, DOT>, 64, 0) != hipSuccess || per <= 0)
      return 0;
    return std::max(8, (cus * per) & ~7);
  }();
  return w;
}
template <class V, bool DOT>
void launch_gs_as(cg_problem& P, const DList& a, const Ctx& c, int l, V* Z, const V* R, int color, int flags) {
  const int lw = P.gs_block == 8 ? loop_waves<V, DOT>() : 0;
  if (lw > 0 && a.n > lw) {
    k_gs_color_loop<V, DOT><<<unsigned(lw), 64, 0, P.stream>>>(c, P.L(a), a.n, l, Z, R, color, flags);
  } else if (P.gs_block == 2 && a.n > 1)
    k_gs_color2<V, DOT><<<grid_for((a.n + 1) / 2, flags), 64, 0, P.stream>>>(c, P.L(a), a.n, l, Z, R, color, flags);
  else if (P.gs_block == 4 && a.n > 1)
    k_gs_color4<V, DOT><<<grid_for((a.n + 3) / 4, flags), 256, 0, P.stream>>>(c, P.L(a), a.n, l, Z, R, color, flags);
  else
    k_gs_color<V, DOT><<<grid_for(a.n, flags), 64, 0, P.stream>>>(c, P.L(a), a.n, l, Z, R, color, flags);
}
template <class V>
void launch_gs(cg_problem& P, const DList& a, const Ctx& c, int l, V* Z, const V* R, int color, int flags) {
  if (a.n == 0) return;
  if (flags & F_DOT_RZ) {
    launch_gs_as<V, true>(P, a, c, l, Z, R, color, flags);
    return;
  }
  launch_gs_as<V, false>(P, a, c, l, Z, R, color, flags);
}
// One red-black iteration of level list a, colour c1 then 1 - c1, in one
// k_gs_pair launch: the listed tiles' values are read from XI, ghosts' from XG,
// the results written to XO (neither of the others).
template <class V>
void launch_gs_pair(cg_problem& P, const DList& a, const Ctx& c, const V* XI, const V* XG, V* XO, int c1, int flags) {
  if (a.n == 0) return;
  if (flags & F_DOT_RZ) k_gs_pair<V, true><<<grid_for(a.n, flags), 64, 0, P.stream>>>(c, P.L(a), a.n, XI, XG, XO, c1, flags);
  else k_gs_pair<V, false><<<grid_for(a.n, flags), 64, 0, P.stream>>>(c, P.L(a), a.n, XI, XG, XO, c1, flags);
}
// The pool's V-cycle vectors of storage type V.
template <class V> V* pool_z(cg_problem& P);
template <class V> V* pool_zb(cg_problem& P);
template <> float* pool_z<float>(cg_problem& P) { return P.d_ch[CG_CH_Z]; }
template <> bf16* pool_z<bf16>(cg_problem& P) { return P.d_zh; }
template <> float* pool_zb<float>(cg_problem& P) { return P.d_zb; }
template <> bf16* pool_zb<bf16>(cg_problem& P) { return P.d_zbh; }
template <> f16* pool_z<f16>(cg_problem& P) { return reinterpret_cast<f16*>(P.d_zh); }
template <> f16* pool_zb<f16>(cg_problem& P) { return reinterpret_cast<f16*>(P.d_zbh); }
// Scratch k (0..2) of the operator API: a channel in the colour-split order of
// the V-cycle kernels; src (natural, may be null) is converted into it.
float* api_split(cg_problem& P, int k, const float* src) {
  if (!P.d_api[k]) HIPC(hipMalloc(&P.d_api[k], size_t(P.n) * 512 * sizeof(float)));
  if (src) k_to_split<<<P.n, 64, 0, P.stream>>>(P.L(P.all), P.n, src, P.d_api[k]);
  return P.d_api[k];
}
void api_unsplit(cg_problem& P, const DList& l, const float* split, float* dst) {
  if (l.n) k_from_split<<<l.n, 64, 0, P.stream>>>(P.L(l), l.n, split, dst);
}
// bfloat16 V-cycle vectors, allocated once per pool when first needed
void ensure_bf16(cg_problem& P) {
  if (P.d_zh) return;
  if (!P.owner) fail(CG_ERR_INVALID, "bfloat16 storage was not configured for this pool");
  const size_t cells = size_t(P.n) * 512;
  bf16** bufs[3] = {&P.d_zh, &P.d_rhsvh, &P.d_zbh};
  for (bf16** b : bufs) {
    HIPC(hipMalloc(b, cells * sizeof(bf16)));
    HIPC(hipMemsetAsync(*b, 0, cells * sizeof(bf16), P.stream));
  }
  HIPC(hipStreamSynchronize(P.stream));
  if (P.graph) HIPC(hipGraphExecDestroy(P.graph));
  P.graph = nullptr;
  P.graph_key.clear();
}
// One V-cycle Z <- M^-1 R (GMGSolver::VCycle, GMGSolver.cu:383-447).  The
// V-cycle rhs of LEAF tiles (rhsv) must hold R on entry.  The zero initial
// guess and the zero ghost values of the reference's init pass are implicit in
// the kernels' flags.
// Z is the down-sweep buffer, ZO receives z (they are the same buffer for the
// damped-Jacobi smoother, which ping-pongs with the scratch buffer instead).
// V is the storage type of Z, ZO, the scratch buffer and the V-cycle rhs (R is
// only a placeholder: in V-cycle mode the rhs comes from the V-cycle rhs array).
// The fused first iteration (k_gs_first) saves a launch and the first pass's
// traffic, but reads the halo's diagonals, which no other kernel brings into
// the cache; with 16-bit storage on a level of many tiles that costs more than
// the second launch saves (measured on the 256-problem pool), so large 16-bit
// levels keep the two passes.  Results are the same either way.
constexpr int kFirstFused16Max = 16384;
// Levels of at most this many tiles use k_gs_prolong_eager and the eager
// k_restrict (all loads before one wait: faster where a launch waits on
// latency, slower on big levels).
constexpr int kProlongEagerMax = 16384;
// Both limits are per handle (cg_set_level_forms), defaulting to these.

// last_flags is added to the last colour pass of each grid-wide level's
// up-sweep (RB-GS with nu >= 1), the pass after which that level's z is final.
template <class V>
void enqueue_vcycle(cg_problem& P, V* Z, V* ZO, const V* R, const VParams& v, int last_flags = 0) {
  const Ctx c = P.ctx();
  const int base = F_VCYCLE | P.swz | (v.check_done ? F_CHECK_DONE : 0) | (v.inert ? F_INERT : 0);
  const int L = P.max_level;  // a problem's levels above its coarse_C use the grid-wide lists actg / lvlg
  V* ZB = pool_zb<V>(P);
  for (int l = L; l >= 1; l--) {
    const DList& a = P.actg[l];
    if (v.smoother == CG_SMOOTH_RBGS) {
      // With one-launch iterations (k_gs_pair, which must not write the array
      // it reads) the level's values alternate between ZO's level region,
      // unused until the up-sweep, and Z, so that they end in Z.
      const bool pairs = c.tiles != nullptr && v.nu >= 2;
      V* cur = (pairs && v.nu % 2 == 0) ? ZO : Z;
      for (int s = 0; s < v.nu; s++) {
        if (s == 0 && P.d_diag && (std::is_same<V, float>::value || a.n <= P.fused16_max)) {
          // both colours of the first iteration from the zero guess in one launch
          if (a.n) k_gs_first<V><<<grid_for(a.n, base), 64, 0, P.stream>>>(c, P.L(a), a.n, cur, base | F_GHOST_ZERO);
          continue;
        }
        if (s == 0 || !pairs) {  // two colour passes in place (s == 0: from the zero guess)
          launch_gs(P, a, c, l, cur, R, 0, base | F_GHOST_ZERO | (s == 0 ? F_ZERO_X : 0));
          launch_gs(P, a, c, l, cur, R, 1, base | F_GHOST_ZERO);
          continue;
        }
        V* nxt = cur == Z ? ZO : Z;
        launch_gs_pair<V>(P, a, c, cur, cur, nxt, 0, base | F_GHOST_ZERO);  // ghosts read as 0
        cur = nxt;
      }
    } else {
      const V* src = Z;
      for (int s = 0; s < v.nu; s++) {
        V* dst = ((v.nu - 1 - s) % 2 == 0) ? Z : ZB;
        int fl = base | F_GHOST_ZERO | (s == 0 ? F_ZERO_X : 0);
        if (a.n) k_jacobi<V><<<grid_for(a.n, fl), 64, 0, P.stream>>>(c, P.L(a), a.n, l, src, Z, dst, R, v.omega, fl);
        src = dst;
      }
    }
    const DList& al = P.lvlg[l];
    int fl = base | F_GHOST_ZERO | (v.nu == 0 ? F_ZERO_X : 0);
    if (al.n > P.eager_max)
      k_restrict<V><<<grid_for(al.n, fl), 64, 0, P.stream>>>(c, P.L(al), al.n, l, Z, R, nullptr, v.ooa, fl);
    else if (al.n)
      k_restrict<V, true><<<grid_for(al.n, fl), 64, 0, P.stream>>>(c, P.L(al), al.n, l, Z, R, nullptr, v.ooa, fl);
  }
  if (!P.plans.empty()) {
    VCycleArgs va{v.nu, v.coarsest, v.smoother, v.omega, v.ooa, v.pcoef};
    k_coarse_lds<V><<<unsigned(P.plans.size()), CL_THREADS, 0, P.stream>>>(c, P.d_plans, va, ZO, base & ~F_SWIZZLE);
  }
  {
    // problems whose level 0 is too large for LDS: bottom sweeps as grid-wide launches
    const DList& a0 = P.actg[0];
    const int it0 = v.coarsest / 2;
    if (it0 == 0 && a0.n) k_fill_tiles<V><<<a0.n, 64, 0, P.stream>>>(P.L(a0), a0.n, ZO, 0.f);
    for (int i = 0; i < it0; i++) {
      launch_gs(P, a0, c, 0, ZO, R, 0, base | ((i == 0) ? F_ZERO_X : 0));
      launch_gs(P, a0, c, 0, ZO, R, 1, base);
    }
    for (int i = 0; i < it0; i++) {
      launch_gs(P, a0, c, 0, ZO, R, 1, base);
      launch_gs(P, a0, c, 0, ZO, R, 0, base);
    }
  }
  for (int l = 1; l <= L; l++) {
    const DList& al = P.lvlg[l];
    const DList& a = P.actg[l];
    const bool jac = v.smoother != CG_SMOOTH_RBGS;
    if (!jac && v.nu > 0) {
      // prolongation fused into the first (black) pass; the level's result goes to ZO
      if (al.n > P.eager_max)
        k_gs_prolong<V><<<grid_for(al.n, base), 64, 0, P.stream>>>(c, P.L(al), al.n, Z, ZO, v.pcoef, 1, base);
      else if (al.n)
        k_gs_prolong_eager<V><<<grid_for(al.n, base), 64, 0, P.stream>>>(c, P.L(al), al.n, Z, ZO, v.pcoef, 1, base);
      if (c.tiles == nullptr || v.nu < 2) {
        launch_gs(P, a, c, l, ZO, R, 0, base | (v.nu == 1 ? last_flags : 0));
        for (int s = 1; s < v.nu; s++) {
          launch_gs(P, a, c, l, ZO, R, 1, base);
          launch_gs(P, a, c, l, ZO, R, 0, base | (s == v.nu - 1 ? last_flags : 0));
        }
        continue;
      }
      // one-launch iterations: the red pass and the iterations alternate the
      // level's values between Z's level region (its down-sweep values are
      // consumed by the prolongation pass) and ZO, ending in ZO; ghosts keep
      // their prolongated values in ZO throughout
      V* cur = ZO;
      if ((v.nu - 1) % 2 == 1) {
        if (a.n) k_gs_color_to<V><<<grid_for(a.n, base), 64, 0, P.stream>>>(c, P.L(a), a.n, ZO, Z, 0, base);
        cur = Z;
      } else {
        launch_gs(P, a, c, l, ZO, R, 0, base);
      }
      for (int s = 1; s < v.nu; s++) {
        V* nxt = cur == ZO ? Z : ZO;
        launch_gs_pair<V>(P, a, c, cur, ZO, nxt, 1, base | (s == v.nu - 1 ? last_flags : 0));
        cur = nxt;
      }
      continue;
    }
    V* XO = (jac && (v.nu % 2 == 1)) ? ZB : ZO;
    int fl = base | (v.nu == 0 ? F_ZERO_X : 0);
    if (al.n) k_prolong<V><<<grid_for(al.n, fl), 64, 0, P.stream>>>(c, P.L(al), al.n, l, ZO, Z, XO, ZO, v.pcoef, fl);
    if (!jac) continue;
    {
      const V* src = XO;
      for (int s = 0; s < v.nu; s++) {
        V* dst = ((v.nu - 1 - s) % 2 == 0) ? Z : ZB;
        if (a.n) k_jacobi<V><<<grid_for(a.n, base), 64, 0, P.stream>>>(c, P.L(a), a.n, l, src, Z, dst, R, v.omega, base);
        src = dst;
      }
    }
  }
}

// One CG iteration of the rotated loop (GMGSolver.cu:573-653 shifted so that it
// ---------------------------------------------------------------------------
// AMGSolver (AMGSolver.cu): stored coefficients and VCycleAMG
// ---------------------------------------------------------------------------
AmgC amg_arrays(const cg_problem& P) {
  const size_t cells = size_t(P.n) * 512;
  AmgC A;
  for (int a = 0; a < 3; a++) A.off[a] = P.d_amgc + size_t(a) * cells;
  A.diag = P.d_amgc + 3 * cells;
  A.ct = P.d_amgct;
  return A;
}
// AMGSolver::prepareTypesAndCoeffs (CoarsenTypesAndAMGCoeffs, AMGSolver.cu:110-209).
// The AMG cell types are kept beside the handle's own, which the GMG path keeps using.
void amg_prepare(cg_problem& P, float one_over_alpha) {
  if (!P.owner) fail(CG_ERR_INVALID, "cg_amg_prepare: not an owning handle");
  const size_t cells = size_t(P.n) * 512;
  if (!P.d_amgc) {
    HIPC(hipMalloc(&P.d_amgc, 4 * cells * sizeof(float)));
    HIPC(hipMalloc(&P.d_amgct, cells));
  }
  HIPC(hipMemsetAsync(P.d_amgc, 0, 4 * cells * sizeof(float), P.stream));
  HIPC(hipMemcpyAsync(P.d_amgct, P.d_ctype, cells, hipMemcpyDeviceToDevice, P.stream));
  const AmgC A = amg_arrays(P);
  if (P.ghosts.n) k_amg_ghost_types<<<P.ghosts.n, 64, 0, P.stream>>>(P.L(P.ghosts), P.ghosts.n, P.d_ctype, P.d_amgct);
  if (P.applyl.n) k_amg_coeffs<<<P.applyl.n, 64, 0, P.stream>>>(P.L(P.applyl), P.applyl.n, A);
  for (int l = P.max_level; l >= 1; l--)  // finest level first: children fill their parents
    if (P.lvl[l].n) k_amg_coarsen<<<P.lvl[l].n, 64, 0, P.stream>>>(P.L(P.lvl[l]), P.lvl[l].n, A, one_over_alpha);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(P.stream));
  P.amg_ready = true;
}
// VCycleAMG (AMGSolver.cu:610-715): x in the FP32 scratch array, rhs in the
// TMP channel, the residual in the AP channel (as the reference uses its
// channels), all in natural cell order; z for the CG kernels is then put in
// the split order on LEAF tiles.
void enqueue_amg_vcycle(cg_problem& P, const VParams& v) {
  const Ctx c = P.ctx();
  const AmgC A = amg_arrays(P);
  const int fl = P.swz | (v.check_done ? F_CHECK_DONE : 0);
  float* X = P.d_zb;
  float* RHS = P.d_ch[CG_CH_TMP];
  float* TMP = P.d_ch[CG_CH_AP];
  const float* F = c.r;
  k_amg_init<<<grid_for(P.all.n, fl), 64, 0, P.stream>>>(P.L(P.all), P.all.n, X, F, RHS, fl);
  auto gs = [&](int l, int color) {
    if (P.act[l].n) k_amg_gs<<<grid_for(P.act[l].n, fl), 64, 0, P.stream>>>(c, P.L(P.act[l]), P.act[l].n, X, RHS, A, color, fl);
  };
  const int L = P.max_level;
  for (int l = L; l >= 1; l--) {
    for (int it = 0; it < v.nu; it++) { gs(l, 0); gs(l, 1); }
    if (P.lvl[l].n)
      k_amg_residual<<<grid_for(P.lvl[l].n, fl), 64, 0, P.stream>>>(c, P.L(P.lvl[l]), P.lvl[l].n, X, RHS, TMP, A, fl);
    if (P.act[l].n)
      k_amg_restrict<<<grid_for(P.act[l].n, fl), 64, 0, P.stream>>>(c, P.L(P.act[l]), P.act[l].n, TMP, RHS, A, v.ooa, fl);
    if (P.lvl[l - 1].n)
      k_amg_leaf_rhs<<<grid_for(P.lvl[l - 1].n, fl), 64, 0, P.stream>>>(c, P.L(P.lvl[l - 1]), P.lvl[l - 1].n, TMP, F, RHS, fl);
  }
  for (int it = 0; it < v.coarsest / 2; it++) { gs(0, 0); gs(0, 1); }
  for (int it = 0; it < v.coarsest / 2; it++) { gs(0, 1); gs(0, 0); }
  for (int l = 1; l <= L; l++) {
    if (P.lvl[l].n) k_amg_prolong<<<grid_for(P.lvl[l].n, fl), 64, 0, P.stream>>>(c, P.L(P.lvl[l]), P.lvl[l].n, X, A, v.pcoef, fl);
    for (int it = 0; it < v.nu; it++) { gs(l, 1); gs(l, 0); }
  }
  k_to_split<<<unsigned(P.leaf.n), 64, 0, P.stream>>>(P.L(P.leaf), P.leaf.n, X, c.z);
}

// starts with the V-cycle): [re-centre r], z = M^-1 r, gamma/beta, p, A p and
// alpha, x/r and the convergence test.
template <class V>
void enqueue_cg_iteration(cg_problem& P, const VParams& vp, int pure_neumann) {
  const Ctx c = P.ctx();
  // tiles the solve's (or the request's) init found inert are skipped (not with
  // AMG, whose V-cycle kernels do not take part)
  VParams v = vp;
  v.inert = !v.amg && c.inert != nullptr;
  const int fl = P.swz | F_CHECK_DONE | (v.inert ? F_INERT : 0);
  const unsigned np = unsigned(P.nprob);  // one finalize workgroup per problem
  // one problem and no problem map: its partials are slots 0 .. leaf tiles - 1
  const bool one = P.nprob == 1 && !c.pmap && P.leaf_count.size() == 1 && P.leaf_begin.size() == 1;
  const int rb = one ? P.leaf_begin[0] : 0, rc = one ? P.leaf_count[0] : -1;
  if (pure_neumann) {
    k_mean<<<grid_for(P.leaf.n, fl), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, c.r, fl & ~F_INERT);
    k_finalize<<<np, FIN_THREADS, 0, P.stream>>>(c, FIN_MEAN, fl);
    k_subtract_mean<V><<<grid_for(P.leaf.n, fl), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, c.r, 1, fl & ~F_INERT);
  }
  // RB-GS: down-sweep in the scratch buffer, z assembled in z by the up-sweep
  V* Z = pool_z<V>(P);
  // r.z: summed by the last up-sweep pass of each level where every LEAF tile
  // gets its final z (same partials, same order), else by its own pass
  const bool rz_fused = !v.amg && v.smoother == CG_SMOOTH_RBGS && v.nu > 0 && P.rz_in_vcycle;
  if (v.amg) enqueue_amg_vcycle(P, v);
  else enqueue_vcycle<V>(P, v.smoother == CG_SMOOTH_RBGS ? pool_zb<V>(P) : Z, Z, nullptr, v, rz_fused ? F_DOT_RZ : 0);
  if (!rz_fused) k_cg_dot_rz<V><<<grid_for(P.leaf.n, fl), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, fl);
  k_finalize<<<np, FIN_THREADS, 0, P.stream>>>(c, FIN_RZ, fl, rb, rc);
  k_cg_update_p<V><<<grid_for(P.leaf.n, fl), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, fl);
  k_apply<true><<<grid_for(P.applyl.n, fl), 64, 0, P.stream>>>(c, P.L(P.applyl), P.applyl.n, c.p, c.ap, -1, 0, 1, fl);
  if (P.gather.n) k_ghost_gather<<<grid_for(P.gather.n, fl), 64, 0, P.stream>>>(c, P.L(P.gather), P.gather.n, fl);
  k_finalize<<<np, FIN_THREADS, 0, P.stream>>>(c, FIN_PAP, fl, rb, rc);
  k_cg_update_xr<V><<<grid_for(P.leaf.n, fl), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, fl);
  k_finalize<<<np, FIN_THREADS, 0, P.stream>>>(c, FIN_RR, fl, rb, rc);
}
void enqueue_cg_iteration(cg_problem& P, const VParams& v, int pure_neumann, int storage) {
  if (storage == CG_STORE_BF16) enqueue_cg_iteration<bf16>(P, v, pure_neumann);
  else if (storage == CG_STORE_F16) enqueue_cg_iteration<f16>(P, v, pure_neumann);
  else enqueue_cg_iteration<float>(P, v, pure_neumann);
}

// A temporary worklist of the records of the given tiles.
const TileRec* tmp_list(cg_problem& P, const std::vector<int>& tiles) {
  if (int(tiles.size()) > P.tmp_list_cap) {
    if (P.d_tmp_list) HIPC(hipFree(P.d_tmp_list));
    P.d_tmp_list = nullptr;
    P.tmp_list_cap = std::max<int>(int(tiles.size()), 64);
    HIPC(hipMalloc(&P.d_tmp_list, sizeof(TileRec) * P.tmp_list_cap));
  }
  std::vector<TileRec> recs(tiles.size());
  for (size_t q = 0; q < tiles.size(); q++) recs[q] = P.tiles[tiles[q]];
  if (!recs.empty()) {
    HIPC(hipStreamSynchronize(P.stream));
    HIPC(hipMemcpy(P.d_tmp_list, recs.data(), sizeof(TileRec) * recs.size(), hipMemcpyHostToDevice));
  }
  return P.d_tmp_list;
}

double dot_types(cg_problem& P, const float* A, const float* B, int types) {
  std::vector<int> l;
  for (int i = 0; i < P.n; i++) if (tr_type(P.tiles[i]) & types) l.push_back(i);
  if (l.empty()) return 0.0;
  const TileRec* dl = tmp_list(P, l);
  const Ctx c = P.ctx();
  k_dot_tiles<<<unsigned(l.size()), 64, 0, P.stream>>>(c, dl, int(l.size()), A, B, P.d_part);
  k_sum_partials<<<1, FIN_THREADS, 0, P.stream>>>(P.d_part, int(l.size()), P.d_scalar);
  double out = 0;
  HIPC(hipMemcpyAsync(&out, P.d_scalar, sizeof(double), hipMemcpyDeviceToHost, P.stream));
  HIPC(hipStreamSynchronize(P.stream));
  return out;
}

void full_negative_laplacian(cg_problem& P, float* X, float* AX, int diag) {
  const Ctx c = P.ctx();
  if (P.ghosts.n) k_propagate_ghosts<<<P.ghosts.n, 64, 0, P.stream>>>(c, P.L(P.ghosts), P.ghosts.n, X, X);
  k_apply<<<P.applyl.n, 64, 0, P.stream>>>(c, P.L(P.applyl), P.applyl.n, X, AX, -1, diag, 0, 0);
  if (P.ghosts.n) k_accumulate_ghosts<<<P.ghosts.n, 64, 0, P.stream>>>(c, P.L(P.ghosts), P.ghosts.n, AX);
}

std::vector<ProbState> read_states(cg_problem& P) {
  std::vector<ProbState> s(P.nprob);
  HIPC(hipMemcpyAsync(s.data(), P.d_ps, sizeof(ProbState) * P.nprob, hipMemcpyDeviceToHost, P.stream));
  HIPC(hipStreamSynchronize(P.stream));
  return s;
}

// All problems of the pool, one graph replay per CG iteration for all of them;
// a converged problem's tiles drop out of every kernel at their first load.
void check_params(const cg_params& prm) {
  if (prm.max_iters < 0 || !(prm.rel_tol >= 0) || prm.level_iters < 0 || prm.coarsest_iters < 0)
    fail(CG_ERR_INVALID, "cg_solve: invalid parameters");
  if (prm.smoother != CG_SMOOTH_RBGS && prm.smoother != CG_SMOOTH_JACOBI) fail(CG_ERR_INVALID, "cg_solve: unknown smoother");
  if (prm.storage != CG_STORE_F32 && prm.storage != CG_STORE_BF16 && prm.storage != CG_STORE_F16)
    fail(CG_ERR_INVALID, "cg_solve: unknown storage type");
}

// ---------------------------------------------------------------------------
// FP64 residuals (iterative refinement, checks of finished solves)
// ---------------------------------------------------------------------------
void ensure_f64(cg_problem& P) {
  if (P.d_x64) return;
  if (!P.owner) fail(CG_ERR_INVALID, "FP64 residuals need an owning problem handle");
  const size_t cells = size_t(P.n) * 512;
  HIPC(hipMalloc(&P.d_x64, cells * sizeof(double)));
  HIPC(hipMalloc(&P.d_r64, cells * sizeof(double)));
  HIPC(hipMalloc(&P.d_bsave, cells * sizeof(float)));
  HIPC(hipMalloc(&P.d_psum, size_t(P.nprob) * sizeof(double)));
  HIPC(hipMemsetAsync(P.d_x64, 0, cells * sizeof(double), P.stream));
  HIPC(hipMemsetAsync(P.d_r64, 0, cells * sizeof(double), P.stream));
  HIPC(hipStreamSynchronize(P.stream));
}

// Per-problem totals of the leaf-slot partials in c.part, to the host.
std::vector<double> prob_sums(cg_problem& P) {
  const Ctx c = P.ctx();
  k_prob_sums<<<unsigned(P.nprob), FIN_THREADS, 0, P.stream>>>(c, P.d_psum);
  std::vector<double> out(P.nprob);
  HIPC(hipMemcpyAsync(out.data(), P.d_psum, sizeof(double) * P.nprob, hipMemcpyDeviceToHost, P.stream));
  HIPC(hipStreamSynchronize(P.stream));
  return out;
}

// r64 = B - A x64 on LEAF tiles in FP64 (re-centred for pure Neumann) and
// |r64|^2 per problem; write_r: also the FP32 copy in the r channel.
std::vector<double> residual64(cg_problem& P, const float* B, int pure_neumann, bool write_r) {
  const Ctx c = P.ctx();
  const unsigned gl = grid_for(P.leaf.n, kSwizzle);
  if (P.ghosts.n) k_propagate_ghosts64<<<P.ghosts.n, 64, 0, P.stream>>>(c, P.L(P.ghosts), P.ghosts.n, P.d_x64);
  k_apply64<<<P.applyl.n, 64, 0, P.stream>>>(c, P.L(P.applyl), P.applyl.n, P.d_x64, P.d_r64);
  if (P.ghosts.n) k_accumulate_ghosts64<<<P.ghosts.n, 64, 0, P.stream>>>(c, P.L(P.ghosts), P.ghosts.n, P.d_r64);
  if (pure_neumann) {
    k_ir_residual<<<gl, 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, B, P.d_r64, 1);
    k_finalize<<<unsigned(P.nprob), FIN_THREADS, 0, P.stream>>>(c, FIN_MEAN, 0);
    k_ir_recenter<<<gl, 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, P.d_r64, write_r ? 1 : 0);
  } else {
    k_ir_residual<<<gl, 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, B, P.d_r64, write_r ? 0 : 2);
  }
  return prob_sums(P);
}

// tol_q / maxit_q / thr_abs_q: per-problem tolerances, iteration caps and
// threshold floors (null: prm's for all, no floor).
// tail: work enqueued after the last iteration, before the host waits for the
// solve (cg_project's gradient), so the device does not idle in between.
void solve(cg_problem& P, const cg_params& prm, cg_result* res, const double* tol_q = nullptr, const int* maxit_q = nullptr,
           const double* thr_abs_q = nullptr, bool amg = false, const std::function<void()>* tail = nullptr) {
  check_params(prm);
  if (amg && !P.amg_ready) fail(CG_ERR_STATE, "cg_amg_solve: call cg_amg_prepare first");
  if (amg && prm.storage != CG_STORE_F32) fail(CG_ERR_INVALID, "cg_amg_solve: FP32 storage only");
  if (prm.storage != CG_STORE_F32) ensure_bf16(P);
  if (prm.verify_f64) {  // keep b for the FP64 check at the end
    ensure_f64(P);
    HIPC(hipMemcpyAsync(P.d_bsave, P.d_ch[CG_CH_R], size_t(P.n) * 512 * sizeof(float), hipMemcpyDeviceToDevice, P.stream));
  }
  VParams v;
  v.nu = prm.level_iters;
  v.coarsest = prm.coarsest_iters;
  v.smoother = prm.smoother;
  v.omega = prm.omega;
  v.ooa = prm.one_over_alpha;
  v.pcoef = prm.prolong_coeff;
  const int stride = std::max(1, prm.sync_stride);
  auto t0 = std::chrono::steady_clock::now();
  set_prob_state(P, prm.rel_tol, prm.max_iters, prm.pure_neumann, prm.sync_stride, tol_q, maxit_q, thr_abs_q);
  int max_iters = prm.max_iters;
  if (maxit_q) for (int q = 0; q < P.nprob; q++) max_iters = std::max(max_iters, maxit_q[q]);
  Ctx c = P.ctx();
  if (amg) c.inert = nullptr;  // the AMG V-cycle skips no tile
  const bool f16s = prm.storage == CG_STORE_F16;
  if (f16s) {  // float16 storage: solve for 2^k b, k from |b|^2 per problem (FIN_SCALE)
    k_ir_norm<<<grid_for(P.leaf.n, kSwizzle), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, c.r);
    k_finalize<<<unsigned(P.nprob), FIN_THREADS, 0, P.stream>>>(c, FIN_SCALE, 0);
    k_scale_leaf<<<grid_for(P.leaf.n, kSwizzle), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, c.r, 1);
  }
  if (prm.storage == CG_STORE_BF16)
    k_cg_init<bf16><<<grid_for(P.leaf.n, P.swz), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, P.swz);
  else if (f16s)
    k_cg_init<f16><<<grid_for(P.leaf.n, P.swz), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, P.swz);
  else
    k_cg_init<float><<<grid_for(P.leaf.n, P.swz), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, P.swz);
  k_finalize<<<unsigned(P.nprob), FIN_THREADS, 0, P.stream>>>(c, FIN_BB, 0);
  HIPC(hipGetLastError());

  char keybuf[160];
  v.amg = amg ? 1 : 0;
  std::snprintf(keybuf, sizeof(keybuf), "%d/%d/%d/%a/%a/%a/%d/%d/%d/%d", v.nu, v.coarsest, v.smoother, double(v.omega),
                double(v.ooa), double(v.pcoef), prm.pure_neumann, prm.storage, P.swz, v.amg);
  // -1: replay where launches dominate an iteration (the AMG V-cycle; handles
  // of fewer than kStreamMinTiles tiles), stream launches otherwise, which
  // measured faster on larger single problems (a graph node costs less when
  // empty, more once its kernel moves data)
  constexpr int kStreamMinTiles = 4096;
  const bool use_graph = prm.use_graph > 0 || (prm.use_graph < 0 && (amg || P.n < kStreamMinTiles));
  if (use_graph && P.graph_key != keybuf) {
    if (P.graph) HIPC(hipGraphExecDestroy(P.graph));
    P.graph = nullptr;
    hipGraph_t g;
    HIPC(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal));
    enqueue_cg_iteration(P, v, prm.pure_neumann, prm.storage);
    HIPC(hipStreamEndCapture(P.stream, &g));
    HIPC(hipGraphInstantiate(&P.graph, g, nullptr, nullptr, 0));
    HIPC(hipGraphDestroy(g));
    P.graph_key = keybuf;
  }
  auto one_iter = [&]() {
    if (use_graph) HIPC(hipGraphLaunch(P.graph, P.stream));
    else enqueue_cg_iteration(P, v, prm.pure_neumann, prm.storage);
  };
  // Poll the done flags one chunk behind the enqueue front, so the device never
  // waits for the host; iterations after convergence exit at their first load.
  const int npb = P.nprob;
  auto all_done = [&](int sl) {
    for (int q = 0; q < npb; q++) if (!P.h_pin[sl * npb + q]) return false;
    return true;
  };
  int enq = 0, slot = 0;
  bool have_prev = false;
  while (true) {
    if (max_iters == 0) break;
    for (int s = 0; s < stride; s++) { one_iter(); enq++; }
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(&P.h_pin[slot * npb], P.d_dflag, sizeof(int32_t) * npb, hipMemcpyDeviceToHost, P.stream));
    HIPC(hipEventRecord(P.ev[slot], P.stream));
    if (have_prev) {
      HIPC(hipEventSynchronize(P.ev[1 - slot]));
      if (all_done(1 - slot)) break;
    }
    if (enq >= max_iters) {
      HIPC(hipEventSynchronize(P.ev[slot]));
      break;
    }
    have_prev = true;
    slot = 1 - slot;
  }
  // the last iteration's x update (k_cg_update_p applies each earlier one)
  k_cg_finish_x<<<grid_for(P.leaf.n, P.swz), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n,
                                                                 P.swz | (c.inert ? F_INERT : 0));
  if (f16s) {  // back to the caller's scale: x and the residual r
    k_scale_leaf<<<grid_for(P.leaf.n, kSwizzle), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, c.x, -1);
    k_scale_leaf<<<grid_for(P.leaf.n, kSwizzle), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, c.r, -1);
  }
  if (tail) (*tail)();
  HIPC(hipStreamSynchronize(P.stream));
  auto t1 = std::chrono::steady_clock::now();
  std::vector<ProbState> st = read_states(P);
  for (ProbState& s : st) {  // |b|^2 and |r|^2 at the caller's scale (exact: powers of two)
    s.bb = std::ldexp(s.bb, -2 * s.scale_exp);
    s.rr = std::ldexp(s.rr, -2 * s.scale_exp);
  }
  std::vector<double> rr64;
  if (prm.verify_f64) {
    k_ir_load_x<<<grid_for(P.leaf.n, kSwizzle), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, c.x, P.d_x64);
    rr64 = residual64(P, P.d_bsave, prm.pure_neumann, false);
  }
  for (int q = 0; q < P.nprob; q++) {
    const ProbState& s = st[q];
    cg_result& r = res[q];
    r.rel_res_f64 = prm.verify_f64 ? (s.bb > 0 ? std::sqrt(rr64[q] / s.bb) : 0.0) : -1.0;
    r.iters = s.iter;
    r.converged = s.converged;
    r.bb = s.bb;
    r.rr = s.rr;
    // GMGSolver.cu:553-556, 655-665: (0, 0) for b = 0; relative residual -1
    // when convergence is never tested (sync_stride <= 0)
    r.rel_res = s.bb > 0 ? (prm.sync_stride > 0 ? std::sqrt(s.rr / s.bb) : -1.0) : 0.0;
    r.seconds = std::chrono::duration<double>(t1 - t0).count();
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision form: iterative refinement with an FP64 outer loop
// ---------------------------------------------------------------------------
// Iterative refinement to |b - A x|_2 < rel_tol |b|_2 measured in FP64, every
// problem of the pool with its own outer loop: the residual and x are FP64; each
// correction solves A d = float(r) with the FP32 MGPCG (V-cycle storage as
// prm.storage) until its recursive |r_in|^2 < max(inner_rel_tol^2 |float(r)|^2,
// thr / 4), thr = rel_tol^2 |b|^2, so a step does not solve further than the
// outer tolerance needs.
void solve_refined(cg_problem& P, const cg_params& prm, const cg_refine_params& rp, cg_refine_result* res) {
  check_params(prm);
  if (!(rp.rel_tol >= 0) || !(rp.inner_rel_tol >= 0) || rp.max_outer < 0)
    fail(CG_ERR_INVALID, "cg_solve_refined: invalid refinement parameters");
  ensure_f64(P);
  if (prm.storage != CG_STORE_F32) ensure_bf16(P);
  auto t0 = std::chrono::steady_clock::now();
  const int np = P.nprob;
  const size_t cells = size_t(P.n) * 512;
  const Ctx c = P.ctx();
  set_prob_state(P, 0.0, 0, prm.pure_neumann, 1);  // leaf ranges for the per-problem sums
  HIPC(hipMemcpyAsync(P.d_bsave, c.r, cells * sizeof(float), hipMemcpyDeviceToDevice, P.stream));
  HIPC(hipMemsetAsync(P.d_x64, 0, cells * sizeof(double), P.stream));
  k_ir_norm<<<grid_for(P.leaf.n, kSwizzle), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, P.d_bsave);
  const std::vector<double> bb = prob_sums(P);
  std::vector<double> thr(np), rr(bb);
  std::vector<int> done(np, 0), conv(np, 0), outer(np, 0), iters(np, 0);
  for (int q = 0; q < np; q++) {
    thr[q] = std::max(rp.rel_tol * rp.rel_tol * bb[q], std::numeric_limits<double>::min());
    if (bb[q] == 0.0) done[q] = conv[q] = 1;  // b = 0: x = 0
  }
  std::vector<double> tol_q(np, rp.inner_rel_tol), floor_q(np);
  std::vector<int> maxit_q(np);
  for (int q = 0; q < np; q++) floor_q[q] = 0.25 * thr[q];
  std::vector<cg_result> inner(np);
  cg_params inner_prm = prm;
  inner_prm.verify_f64 = 0;
  bool first = true;
  while (true) {
    // from x64 = 0 the first residual is b itself (A 0 = 0, r64 = b exactly, its
    // float is the b already in the r channel and |r64|^2 the |b|^2 above, summed
    // the same way), so that step's FP64 operator is skipped; the re-centred
    // pure-Neumann residual is formed as usual
    const bool from_zero = first && !prm.pure_neumann && rp.max_outer > 0;
    first = false;
    const std::vector<double> r2 = from_zero ? bb : residual64(P, P.d_bsave, prm.pure_neumann, true);
    bool any = false;
    for (int q = 0; q < np; q++) {
      if (done[q]) continue;
      rr[q] = r2[q];
      if (rr[q] < thr[q]) done[q] = conv[q] = 1;
      else if (outer[q] >= rp.max_outer) done[q] = 1;
      any = any || !done[q];
    }
    if (!any) break;
    for (int q = 0; q < np; q++) maxit_q[q] = done[q] ? 0 : prm.max_iters;
    solve(P, inner_prm, inner.data(), tol_q.data(), maxit_q.data(), floor_q.data());
    for (int q = 0; q < np; q++)
      if (!done[q]) { outer[q]++; iters[q] += inner[q].iters; }
    k_ir_accumulate<<<grid_for(P.leaf.n, kSwizzle), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, P.d_x64, 0);
  }
  k_ir_accumulate<<<grid_for(P.leaf.n, kSwizzle), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, P.d_x64, 1);
  HIPC(hipStreamSynchronize(P.stream));
  auto t1 = std::chrono::steady_clock::now();
  for (int q = 0; q < np; q++) {
    cg_refine_result& r = res[q];
    r.outer = outer[q];
    r.iters = iters[q];
    r.converged = conv[q];
    r.bb = bb[q];
    r.rr = bb[q] > 0 ? rr[q] : 0.0;
    r.rel_res = bb[q] > 0 ? std::sqrt(rr[q] / bb[q]) : 0.0;
    r.seconds = std::chrono::duration<double>(t1 - t0).count();
  }
}

// ---------------------------------------------------------------------------
// Autotuner: schedules timed on the pool's own right-hand sides
// ---------------------------------------------------------------------------
void set_schedule(cg_problem& P, int coarse_shift, int swizzle, int tile_block) {
  if (coarse_shift < 0 || coarse_shift > CL_MAX_LEVELS) fail(CG_ERR_INVALID, "cg_set_schedule: coarse_shift out of range");
  if (tile_block != 1 && tile_block != 2 && tile_block != 4 && tile_block != 8)
    fail(CG_ERR_INVALID, "cg_set_schedule: tile_block must be 1, 2, 4 or 8");
  if (!P.owner) fail(CG_ERR_INVALID, "cg_set_schedule: not an owning handle");
  P.swz = swizzle ? F_SWIZZLE : 0;
  P.gs_block = tile_block;
  if (coarse_shift != P.coarse_shift) {
    P.coarse_shift = coarse_shift;
    build_schedule(P);
  }
  if (P.graph) HIPC(hipGraphExecDestroy(P.graph));
  P.graph = nullptr;
  P.graph_key.clear();
}

// Time candidate schedules on the pool's current right-hand sides (restored
// before every solve): first the ones that leave every result unchanged (tile
// order and tile block, then the LDS coarse depth with the best of those) with
// the caller's smoother, then the smoother,
// sweeps per level and (if allowed) 16-bit storage with the best of those.
// A candidate counts only if every problem converges; each is timed as the
// best of two solves after a warm-up (graph capture).  The pool keeps the
// chosen schedule; *best receives the chosen solver parameters.
void solve_refined(cg_problem& P, const cg_params& prm, const cg_refine_params& rp, cg_refine_result* res);
// rp non-null: every candidate is timed as the FP64-refined solve it would run
// (cg_solve_refined), not as a single solve to base.rel_tol.
void autotune(cg_problem& P, const cg_params& base, int allow_16bit, cg_params* best, cg_tune_report* rep,
              const cg_refine_params* rp = nullptr) {
  check_params(base);
  const size_t bytes = size_t(P.n) * 512 * sizeof(float);
  DeviceBuffer keep(bytes);
  float* bkeep = static_cast<float*>(keep.p);
  HIPC(hipMemcpyAsync(bkeep, P.d_ch[CG_CH_R], bytes, hipMemcpyDeviceToDevice, P.stream));
  std::vector<cg_result> res(P.nprob);
  std::vector<cg_refine_result> rres(P.nprob);
  int candidates = 0;
  auto run = [&](const cg_params& prm) {
    double t = std::numeric_limits<double>::infinity();
    for (int rep_i = 0; rep_i < 3; rep_i++) {
      HIPC(hipMemcpyAsync(P.d_ch[CG_CH_R], bkeep, bytes, hipMemcpyDeviceToDevice, P.stream));
      double secs = 0.0;
      if (rp) {
        solve_refined(P, prm, *rp, rres.data());
        for (const cg_refine_result& r : rres)
          if (!r.converged) return std::numeric_limits<double>::infinity();
        secs = rres[0].seconds;
      } else {
        solve(P, prm, res.data());
        for (const cg_result& r : res)
          if (!r.converged) return std::numeric_limits<double>::infinity();
        secs = res[0].seconds;
      }
      if (rep_i > 0) t = std::min(t, secs);
    }
    candidates++;
    return t;
  };
  // a candidate replaces the current best only when it is faster by more than
  // the timing noise of one solve, so that near ties keep the earlier (simpler)
  // choice: FP32 before bfloat16 before float16, which also rescales each solve
  constexpr double kMargin = 0.98;
  set_schedule(P, 0, 1, 1);
  const double t_base = run(base);
  double t_best = t_base;
  int best_shift = 0, best_swz = 1, best_block = 1;
  int max_auto = -1;
  for (int q = 0; q < P.nprob; q++) max_auto = std::max(max_auto, P.coarse_auto[q]);
  for (int swz = 1; swz >= 0; swz--)
    for (int block : {1, 2, 4, 8}) {
      if (swz == 1 && block == 1) continue;
      set_schedule(P, 0, swz, block);
      const double t = run(base);
      if (t < t_best * kMargin) { t_best = t; best_swz = swz; best_block = block; }
    }
  for (int shift = 1; shift <= std::min(2, max_auto + 1); shift++) {
    set_schedule(P, shift, best_swz, best_block);
    const double t = run(base);
    if (t < t_best * kMargin) { t_best = t; best_shift = shift; }
  }
  set_schedule(P, best_shift, best_swz, best_block);
  cg_params chosen = base;
  struct Cand { int smoother, nu; };
  const Cand cands[] = {{CG_SMOOTH_RBGS, 1}, {CG_SMOOTH_RBGS, 2}, {CG_SMOOTH_RBGS, 3}, {CG_SMOOTH_JACOBI, 2}, {CG_SMOOTH_JACOBI, 3}};
  for (int storage = CG_STORE_F32; storage <= (allow_16bit ? CG_STORE_F16 : CG_STORE_F32); storage++)
    for (const Cand& cd : cands) {
      cg_params prm = base;
      prm.smoother = cd.smoother;
      prm.level_iters = cd.nu;
      prm.storage = storage;
      if (prm.smoother == base.smoother && prm.level_iters == base.level_iters && prm.storage == base.storage) continue;
      const double t = run(prm);
      if (t < t_best * kMargin) { t_best = t; chosen = prm; }
    }
  HIPC(hipMemcpyAsync(P.d_ch[CG_CH_R], bkeep, bytes, hipMemcpyDeviceToDevice, P.stream));
  HIPC(hipStreamSynchronize(P.stream));
  *best = chosen;
  if (rep) {
    rep->candidates = candidates;
    rep->base_seconds = t_base;
    rep->best_seconds = t_best;
    rep->coarse_shift = best_shift;
    rep->swizzle = best_swz;
    rep->tile_block = best_block;
  }
}

// ---------------------------------------------------------------------------
// The projection around the solve (FluidEuler::project, FluidEuler.h:581-612)
// ---------------------------------------------------------------------------
Vel velocities(cg_problem& P) {
  Vel U;
  for (int a = 0; a < 3; a++) U.u[a] = any_channel(P, CG_CH_U + a);
  return U;
}
// [ghost velocities from the parents,] divergence into DIV, ghost sums into
// their LEAF parents (VelocityVolumeDivergenceOnLeafs, PoissonSolver.cu:66-142)
void velocity_divergence(cg_problem& P, float* DIV, bool propagate) {
  const Ctx c = P.ctx();
  const Vel U = velocities(P);
  if (propagate && P.ghosts.n) k_propagate_vel<<<P.ghosts.n, 64, 0, P.stream>>>(c, P.L(P.ghosts), P.ghosts.n, U);
  k_divergence<<<P.applyl.n, 64, 0, P.stream>>>(c, P.L(P.applyl), P.applyl.n, U, DIV);
  k_accumulate_children<<<P.leaf.n, 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, DIV, C_INT, 1.0f);
}
// ghost p from the parents, u += grad p, ghost velocities into their LEAF
// parents with weight 1/4 (AddGradientToFaceCenters, PoissonSolver.cu:144-209)
void add_gradient(cg_problem& P, float* Pch) {
  const Ctx c = P.ctx();
  const Vel U = velocities(P);
  if (P.ghosts.n) k_propagate_ghosts<<<P.ghosts.n, 64, 0, P.stream>>>(c, P.L(P.ghosts), P.ghosts.n, Pch, Pch);
  k_gradient<<<P.applyl.n, 64, 0, P.stream>>>(c, P.L(P.applyl), P.applyl.n, Pch, U);
  k_accumulate_children_vel<<<P.leaf.n, 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, U, C_INT | C_DIR, 1.0f / 4.0f);
}

#include "devbuild.hpp"

// The device buffers of one DevTree (devbuild.hpp) of `cap` tiles; its hash
// table has at least twice as many slots.
struct DevTreeStore {
  DeviceBuffer count, flags, level, coord, type, mark, hkeys, hvals;
  DevTree B;
  static uint32_t slots(int cap) {
    uint32_t h = 1;
    while (h < 2u * uint32_t(cap)) h <<= 1;
    return h;
  }
  explicit DevTreeStore(int cap)
      : count{sizeof(int)}, flags{2 * sizeof(int)}, level{sizeof(int32_t) * size_t(cap)},
        coord{3 * sizeof(int32_t) * size_t(cap)}, type{size_t(cap)}, mark{size_t(cap)},
        hkeys{sizeof(unsigned long long) * slots(cap)}, hvals{sizeof(int32_t) * slots(cap)} {
    B.cap = cap;
    B.hmask = slots(cap) - 1;
    B.count = static_cast<int*>(count.p);
    B.flags = static_cast<int*>(flags.p);
    B.level = static_cast<int32_t*>(level.p);
    B.coord = static_cast<int32_t*>(coord.p);
    B.type = static_cast<uint8_t*>(type.p);
    B.mark = static_cast<uint8_t*>(mark.p);
    B.hkeys = static_cast<unsigned long long*>(hkeys.p);
    B.hvals = static_cast<int32_t*>(hvals.p);
  }
  void clear(hipStream_t s) {
    HIPC(hipMemsetAsync(B.hkeys, 0xff, sizeof(unsigned long long) * (size_t(B.hmask) + 1), s));
    HIPC(hipMemsetAsync(B.hvals, 0xff, sizeof(int32_t) * (size_t(B.hmask) + 1), s));
    HIPC(hipMemsetAsync(B.count, 0, sizeof(int), s));
    HIPC(hipMemsetAsync(B.flags, 0, 2 * sizeof(int), s));
  }
};

// The host driver of the device-side tree build and adaptation (devbuild.hpp):
// the tree lives in one of two DevTreeStores (the second receives the
// survivors of a coarsening step); every step is a kernel, the host reads the
// tile count and the flags between steps.
class DevBuilder {
 public:
  DevBuilder(const cg_scene_params& p, int64_t max_tiles, const char* who) : p_(p), who_(who) {
    if (p.kind != CG_SCENE_UNIFORM && p.kind != CG_SCENE_SPHERE_BAND && p.kind != CG_SCENE_PLANE)
      fail(CG_ERR_INVALID, std::string(who) + ": uniform, sphere band and plane scenes only");
    if (p.nroot[0] <= 0 || p.nroot[1] <= 0 || p.nroot[2] <= 0 || !(p.h0 > 0) || p.num_layers < 1 || p.num_layers > 19)
      fail(CG_ERR_INVALID, std::string(who) + ": bad scene parameters");
    if (int64_t(p.nroot[0]) * p.nroot[1] * p.nroot[2] > (1 << 19))
      fail(CG_ERR_INVALID, std::string(who) + ": too many root tiles");
    cap_ = int(max_tiles > 0 ? std::min<int64_t>(max_tiles, int64_t(1) << 26) : int64_t(1) << 22);
    HIPC(hipStreamCreateWithFlags(&st_.s, hipStreamNonBlocking));
    cur_ = std::make_unique<DevTreeStore>(cap_);
    cur_->clear(st_.s);
  }
  int tiles() const { return h_count_; }

  // Level 0 of the scene (FluidEuler.h:415-447).
  void roots() {
    const int nroots = p_.nroot[0] * p_.nroot[1] * p_.nroot[2];
    k_db_roots<<<unsigned((nroots + 255) / 256), 256, 0, st_.s>>>(cur_->B, p_);
    HIPC(hipGetLastError());
    read_state();
  }

  // The active (LEAF and NONLEAF) tiles of a host tree; its ghosts are left
  // out and spawned again at the end.
  void load(const cg_tree_desc& d) {
    if (d.num_tiles < 0 || (d.num_tiles && (!d.level || !d.coord || !d.type)))
      fail(CG_ERR_INVALID, who_ + ": bad tree");
    if (d.h0 != p_.h0) fail(CG_ERR_INVALID, who_ + ": the tree's h0 differs from the scene's");
    if (d.num_tiles > cap_) fail(CG_ERR_NOMEM, who_ + ": more tiles than max_tiles");
    const int n = d.num_tiles;
    for (int i = 0; i < n; i++) {
      const int l = d.level[i], x = d.coord[3 * i], y = d.coord[3 * i + 1], z = d.coord[3 * i + 2];
      const uint8_t t = d.type[i];
      if (t != T_LEAF && t != T_NONLEAF && t != T_GHOST) fail(CG_ERR_INVALID, who_ + ": bad tile type");
      if (l < 0 || l >= p_.num_layers || !db_key_ok(x, y, z))
        fail(CG_ERR_INVALID, who_ + ": tile level or coordinate outside the scene's layers");
      if (t != T_GHOST && ((x >> l) >= p_.nroot[0] || (y >> l) >= p_.nroot[1] || (z >> l) >= p_.nroot[2]))
        fail(CG_ERR_INVALID, who_ + ": active tile outside the scene's root tiles");
    }
    if (n == 0) return;
    DeviceBuffer lv{sizeof(int32_t) * size_t(n)}, cd{3 * sizeof(int32_t) * size_t(n)}, ty{size_t(n)};
    HIPC(hipMemcpyAsync(lv.p, d.level, sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, st_.s));
    HIPC(hipMemcpyAsync(cd.p, d.coord, 3 * sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, st_.s));
    HIPC(hipMemcpyAsync(ty.p, d.type, size_t(n), hipMemcpyHostToDevice, st_.s));
    k_db_load<<<unsigned((n + 255) / 256), 256, 0, st_.s>>>(cur_->B, static_cast<const int32_t*>(lv.p),
                                                           static_cast<const int32_t*>(cd.p),
                                                           static_cast<const uint8_t*>(ty.p), n);
    HIPC(hipGetLastError());
    read_state();
  }

  // Refinement rounds until no tile is marked (each round adds at least one
  // level somewhere, so the rounds are bounded by the layers times the ripple).
  void refine() {
    DevTree& B = cur_->B;
    for (int round = 0; round < 64 * p_.num_layers; round++) {
      const int n = h_count_;
      HIPC(hipMemsetAsync(B.flags, 0, sizeof(int), st_.s));
      k_db_mark<<<unsigned((n + 255) / 256), 256, 0, st_.s>>>(B, p_, n);
      HIPC(hipGetLastError());
      read_state();
      if (!h_flags_[0]) return;
      k_db_refine<<<unsigned((n + 255) / 256), 256, 0, st_.s>>>(B, n);
      HIPC(hipGetLastError());
      read_state();
    }
    fail(CG_ERR_INVALID, who_ + ": refinement did not settle");
  }

  // CoarsenStep until a step deletes nothing (CoarsenWithPointValue,
  // FluidEuler.h:90-111); returns the number of steps that deleted tiles.
  int coarsen() {
    int steps = 0;
    while (true) {
      DevTree& B = cur_->B;
      const int n = h_count_;
      if (n == 0) return steps;
      const unsigned g = unsigned((n + 255) / 256);
      for (int l = p_.num_layers - 1; l >= 0; l--) {
        k_dc_nonleaf<<<g, 256, 0, st_.s>>>(B, n, l);
        k_dc_leaf<<<g, 256, 0, st_.s>>>(B, p_, n, l);
      }
      k_dc_down<<<g, 256, 0, st_.s>>>(B, n);
      if (!next_) next_ = std::make_unique<DevTreeStore>(cap_);
      next_->clear(st_.s);
      k_dc_compact<<<g, 256, 0, st_.s>>>(B, next_->B, n);
      HIPC(hipGetLastError());
      std::swap(cur_, next_);
      read_state();
      const int deleted = h_flags_[0];
      if (h_count_ + deleted != n) fail(CG_ERR_INVALID, who_ + ": coarsening lost tiles");
      if (deleted == 0) return steps;
      steps++;
    }
  }

  // Ghosts, cell types, and the tree to the host.
  void finish(TileSet& out) {
    DevTree& B = cur_->B;
    const int n0 = h_count_;
    if (n0) k_db_ghosts<<<unsigned((6 * int64_t(n0) + 255) / 256), 256, 0, st_.s>>>(B, n0);
    HIPC(hipGetLastError());
    read_state();
    const int n = h_count_;
    out.h0 = p_.h0;
    out.level.resize(n);
    out.coord.resize(3 * size_t(n));
    out.type.resize(n);
    out.ctype.resize(size_t(n) * 512);
    out.channel_id.clear();
    out.channel.clear();
    if (n == 0) return;
    HIPC(hipMemcpyAsync(out.level.data(), B.level, sizeof(int32_t) * size_t(n), hipMemcpyDeviceToHost, st_.s));
    HIPC(hipStreamSynchronize(st_.s));
    int maxl = 0;
    for (int v : out.level) maxl = std::max(maxl, v);
    DeviceBuffer b_ctype{size_t(n) * 512};
    B.ctype = static_cast<uint8_t*>(b_ctype.p);
    HIPC(hipMemsetAsync(B.ctype, 0, size_t(n) * 512, st_.s));
    k_db_leaf_types<<<unsigned(n), 512, 0, st_.s>>>(B, p_, n);
    for (int l = maxl; l >= 0; l--) k_db_nonleaf_types<<<unsigned(n), 512, 0, st_.s>>>(B, n, l);
    k_db_ghost_types<<<unsigned(n), 512, 0, st_.s>>>(B, n);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out.coord.data(), B.coord, 3 * sizeof(int32_t) * size_t(n), hipMemcpyDeviceToHost, st_.s));
    HIPC(hipMemcpyAsync(out.type.data(), B.type, size_t(n), hipMemcpyDeviceToHost, st_.s));
    HIPC(hipMemcpyAsync(out.ctype.data(), B.ctype, size_t(n) * 512, hipMemcpyDeviceToHost, st_.s));
    HIPC(hipStreamSynchronize(st_.s));
    B.ctype = nullptr;
  }

 private:
  void read_state() {
    HIPC(hipMemcpyAsync(&h_count_, cur_->B.count, sizeof(int), hipMemcpyDeviceToHost, st_.s));
    HIPC(hipMemcpyAsync(h_flags_, cur_->B.flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st_.s));
    HIPC(hipStreamSynchronize(st_.s));
    if (h_flags_[1]) fail(CG_ERR_NOMEM, who_ + ": more tiles than max_tiles");
  }
  struct Stream {
    hipStream_t s = nullptr;
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
  };
  const cg_scene_params p_;
  const std::string who_;
  int cap_ = 0;
  Stream st_;  // destroyed after the stores (declared before them)
  std::unique_ptr<DevTreeStore> cur_, next_;
  int h_count_ = 0, h_flags_[2] = {0, 0};
};

// Device-side build of a synthetic scene's octree and cell types; the result
// goes to the host as a tile set (tree, cell types, no channels).
void scene_tree_device(const cg_scene_params& p, int64_t max_tiles, TileSet& out) {
  DevBuilder D(p, max_tiles, "cg_scene_tree_device");
  D.roots();
  D.refine();
  D.finish(out);
}

// Device-side adaptation of a tree to a scene's target levels: refinement
// rounds, then coarsening steps, then ghosts and the scene's cell types.
void tree_adapt_device(const cg_tree_desc& d, const cg_scene_params& p, int ops, int64_t max_tiles, TileSet& out) {
  if (ops & ~(CG_ADAPT_REFINE | CG_ADAPT_COARSEN)) fail(CG_ERR_INVALID, "cg_tree_adapt_device: bad ops");
  DevBuilder D(p, max_tiles, "cg_tree_adapt_device");
  D.load(d);
  if (ops & CG_ADAPT_REFINE) D.refine();
  if (ops & CG_ADAPT_COARSEN) D.coarsen();
  D.finish(out);
}

}  // namespace

// ---------------------------------------------------------------------------
// Device scratch of cg_dot_noalloc: the selected tile indices, their FP64
// partials and the result live in one file-scope device array, so that entry
// point calls no allocator.  Every device holds its own copy until the process
// ends; host code reaches a device's copy only through the address
// hipGetSymbolAddress returns for it (the array's name in host code is a
// host-side stand-in, not a device address).
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxTiles = CG_NOALLOC_MAX_TILES;
constexpr size_t kScrPart = 0;                                                  // double[kNoAllocMaxTiles]
constexpr size_t kScrSel = kScrPart + sizeof(double) * kNoAllocMaxTiles;        // int32[kNoAllocMaxTiles]
constexpr size_t kScrOut = kScrSel + sizeof(int32_t) * kNoAllocMaxTiles;        // double
constexpr size_t kScrBytes = kScrOut + 256;
static __device__ __align__(256) unsigned char scratch[kScrBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on device `dev`, looked up once per device with
// that device current (the caller's current device is restored).
unsigned char* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) fail(CG_ERR_INVALID, "cg_dot_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPC(hipGetDevice(&cur));
    if (cur != dev) HIPC(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIPC(hipSetDevice(cur));
    if (e != hipSuccess || !ptr)
      fail(CG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

#include "serve.hpp"

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct cg_scene {
  Scene s;
  std::vector<int32_t> coord_flat;
};
struct cg_tileset {
  TileSet s;
};

extern "C" {

const char* cg_last_error(void) { return g_err.c_str(); }

int32_t cg_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void cg_params_default(cg_params* p) {
  if (!p) return;
  p->max_iters = 1000;
  p->rel_tol = 1e-6;
  p->level_iters = 1;
  p->coarsest_iters = 10;
  p->smoother = CG_SMOOTH_RBGS;
  p->omega = 2.0f / 3.0f;
  p->sync_stride = 1;
  p->pure_neumann = 0;
  p->one_over_alpha = 1.0f;
  p->prolong_coeff = 1.0f;
  p->use_graph = -1;  // by the handle: replay for AMG and small handles, stream launches for large ones
  p->storage = CG_STORE_F32;
  p->verify_f64 = 0;
}

void cg_refine_params_default(cg_refine_params* p) {
  if (!p) return;
  p->rel_tol = 1e-8;
  p->inner_rel_tol = 1e-4;
  p->max_outer = 20;
}

cg_status cg_solve_refined(cg_problem* p, const cg_params* prm, const cg_refine_params* rp, cg_refine_result* results) {
  return guard([&] {
    if (!p || !prm || !rp || !results) fail(CG_ERR_INVALID, "cg_solve_refined: null argument");
    solve_refined(*p, *prm, *rp, results);
  });
}

cg_status cg_get_solution_f64(cg_problem* p, double* host) {
  return guard([&] {
    if (!p || !host) fail(CG_ERR_INVALID, "cg_get_solution_f64: null argument");
    if (!p->d_x64) fail(CG_ERR_STATE, "cg_get_solution_f64: no refined solve has run on this handle");
    std::vector<double> buf(size_t(p->n) * 512);
    HIPC(hipMemcpyAsync(buf.data(), p->d_x64, buf.size() * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    for (int i = 0; i < p->n; i++) std::memcpy(host + size_t(p->ext_of[i]) * 512, &buf[size_t(i) * 512], 512 * sizeof(double));
  });
}

cg_status cg_problem_create_batch(const cg_tree_desc* descs, int32_t num_problems, cg_problem** out) {
  if (!out) { g_err = "cg_problem_create: null output"; return CG_ERR_INVALID; }
  *out = nullptr;
  auto P = std::make_unique<cg_problem>();
  cg_status st = guard([&] { build_pool(descs, num_problems, *P); });
  if (st == CG_OK) *out = P.release();
  return st;
}

cg_status cg_problem_create(const cg_tree_desc* desc, cg_problem** out) { return cg_problem_create_batch(desc, 1, out); }

cg_status cg_problem_num_problems(const cg_problem* p, int32_t* num_problems) {
  if (!p || !num_problems) { g_err = "null argument"; return CG_ERR_INVALID; }
  *num_problems = p->nprob;
  return CG_OK;
}

cg_status cg_problem_destroy(cg_problem* p) {
  delete p;
  return CG_OK;
}

cg_status cg_problem_info(const cg_problem* p, int32_t* num_tiles, int32_t* num_leaf_tiles, int32_t* max_level) {
  if (!p) { g_err = "null problem"; return CG_ERR_INVALID; }
  if (num_tiles) *num_tiles = p->n;
  if (num_leaf_tiles) *num_leaf_tiles = p->n_leaf;
  if (max_level) *max_level = p->max_level;
  return CG_OK;
}

cg_status cg_problem_topology(const cg_problem* p, int32_t* nbr, int32_t* parent) {
  if (!p) { g_err = "null problem"; return CG_ERR_INVALID; }
  for (int i = 0; i < p->n; i++) {
    int e = p->ext_of[i];
    if (nbr)
      for (int f = 0; f < 6; f++) nbr[6 * e + f] = p->tiles[i].nbr[f] < 0 ? -1 : p->ext_of[p->tiles[i].nbr[f]];
    if (parent) parent[e] = p->tiles[i].parent < 0 ? -1 : p->ext_of[p->tiles[i].parent];
  }
  return CG_OK;
}

cg_status cg_set_channel(cg_problem* p, int32_t channel, const float* host) {
  return guard([&] {
    if (!p || !host) fail(CG_ERR_INVALID, "cg_set_channel: null argument");
    float* dst = any_channel(*p, channel);
    std::vector<float> buf(size_t(p->n) * 512);
    for (int i = 0; i < p->n; i++) std::memcpy(&buf[size_t(i) * 512], host + size_t(p->ext_of[i]) * 512, 512 * sizeof(float));
    HIPC(hipMemcpyAsync(dst, buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_get_channel(cg_problem* p, int32_t channel, float* host) {
  return guard([&] {
    if (!p || !host) fail(CG_ERR_INVALID, "cg_get_channel: null argument");
    const float* src = any_channel(*p, channel);
    std::vector<float> buf(size_t(p->n) * 512);
    HIPC(hipMemcpyAsync(buf.data(), src, buf.size() * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    HIPC(hipStreamSynchronize(p->stream));
    for (int i = 0; i < p->n; i++) std::memcpy(host + size_t(p->ext_of[i]) * 512, &buf[size_t(i) * 512], 512 * sizeof(float));
  });
}

cg_status cg_set_channel_device(cg_problem* p, int32_t channel, const float* device_src, void* stream) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "cg_set_channel_device: null problem");
    if (!device_src) fail(CG_ERR_INVALID, "cg_set_channel_device: null device pointer");
    channel_device(*p, channel, device_src, nullptr, static_cast<hipStream_t>(stream));
  });
}

cg_status cg_get_channel_device(cg_problem* p, int32_t channel, float* device_dst, void* stream) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "cg_get_channel_device: null problem");
    if (!device_dst) fail(CG_ERR_INVALID, "cg_get_channel_device: null device pointer");
    channel_device(*p, channel, nullptr, device_dst, static_cast<hipStream_t>(stream));
  });
}

cg_status cg_amg_prepare(cg_problem* p, float one_over_alpha) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    amg_prepare(*p, one_over_alpha);
  });
}

cg_status cg_amg_solve(cg_problem* p, const cg_params* prm, cg_result* results) {
  return guard([&] {
    if (!p || !prm || !results) fail(CG_ERR_INVALID, "cg_amg_solve: null argument");
    solve(*p, *prm, results, nullptr, nullptr, nullptr, true);
  });
}

cg_status cg_amg_coefficients(cg_problem* p, int32_t which, float* host) {
  return guard([&] {
    if (!p || !host) fail(CG_ERR_INVALID, "cg_amg_coefficients: null argument");
    if (!p->amg_ready) fail(CG_ERR_STATE, "cg_amg_coefficients: call cg_amg_prepare first");
    if (which < 0 || which > 4) fail(CG_ERR_INVALID, "cg_amg_coefficients: which must be 0..4");
    const size_t cells = size_t(p->n) * 512;
    std::vector<float> buf(cells);
    if (which < 4) {
      HIPC(hipMemcpyAsync(buf.data(), p->d_amgc + size_t(which) * cells, cells * sizeof(float), hipMemcpyDeviceToHost, p->stream));
      HIPC(hipStreamSynchronize(p->stream));
    } else {
      std::vector<uint8_t> ct(cells);
      HIPC(hipMemcpyAsync(ct.data(), p->d_amgct, cells, hipMemcpyDeviceToHost, p->stream));
      HIPC(hipStreamSynchronize(p->stream));
      for (size_t k = 0; k < cells; k++) buf[k] = float(ct[k]);
    }
    for (int i = 0; i < p->n; i++) std::memcpy(host + size_t(p->ext_of[i]) * 512, &buf[size_t(i) * 512], 512 * sizeof(float));
  });
}

cg_status cg_set_schedule(cg_problem* p, int32_t coarse_shift, int32_t swizzle, int32_t tile_block) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    set_schedule(*p, coarse_shift, swizzle, tile_block);
  });
}

cg_status cg_set_level_forms(cg_problem* p, int32_t eager_max_tiles, int32_t fused16_max_tiles, int32_t pair_sweeps) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    if (!p->owner) fail(CG_ERR_INVALID, "cg_set_level_forms: not an owning handle");
    p->eager_max = eager_max_tiles < 0 ? kProlongEagerMax : eager_max_tiles;
    p->fused16_max = fused16_max_tiles < 0 ? kFirstFused16Max : fused16_max_tiles;
    p->gs_pair = pair_sweeps > 0 ? 1 : 0;
    HIPC(hipStreamSynchronize(p->stream));
    if (p->graph) HIPC(hipGraphExecDestroy(p->graph));
    p->graph = nullptr;
    p->graph_key.clear();
  });
}

cg_status cg_autotune(cg_problem* p, const cg_params* base, int32_t allow_16bit, cg_params* best, cg_tune_report* report) {
  return guard([&] {
    if (!p || !base || !best) fail(CG_ERR_INVALID, "cg_autotune: null argument");
    autotune(*p, *base, allow_16bit, best, report);
  });
}

cg_status cg_autotune_refined(cg_problem* p, const cg_params* base, const cg_refine_params* rp, int32_t allow_16bit,
                              cg_params* best, cg_tune_report* report) {
  return guard([&] {
    if (!p || !base || !rp || !best) fail(CG_ERR_INVALID, "cg_autotune_refined: null argument");
    if (!(rp->rel_tol >= 0) || !(rp->inner_rel_tol >= 0) || rp->max_outer < 0)
      fail(CG_ERR_INVALID, "cg_autotune_refined: invalid refinement parameters");
    autotune(*p, *base, allow_16bit, best, report, rp);
  });
}

cg_status cg_velocity_divergence(cg_problem* p, int32_t div_channel, int32_t propagate_ghosts) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(div_channel);
    velocity_divergence(*p, p->d_ch[div_channel], propagate_ghosts != 0);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_add_gradient(cg_problem* p, int32_t p_channel) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(p_channel);
    add_gradient(*p, p->d_ch[p_channel]);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_project(cg_problem* p, const cg_params* prm, cg_result* results) {
  return guard([&] {
    if (!p || !prm || !results) fail(CG_ERR_INVALID, "cg_project: null argument");
    velocity_divergence(*p, p->d_ch[CG_CH_B], true);
    const std::function<void()> gradient = [p] { add_gradient(*p, p->d_ch[CG_CH_X]); };
    solve(*p, *prm, results, nullptr, nullptr, nullptr, false, &gradient);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_solve(cg_problem* p, const cg_params* prm, cg_result* res) {
  return guard([&] {
    if (!p || !prm || !res) fail(CG_ERR_INVALID, "cg_solve: null argument");
    if (p->nprob != 1) fail(CG_ERR_INVALID, "cg_solve: the handle holds several problems; use cg_solve_batch");
    solve(*p, *prm, res);
  });
}

cg_status cg_solve_batch(cg_problem* p, const cg_params* prm, cg_result* results) {
  return guard([&] {
    if (!p || !prm || !results) fail(CG_ERR_INVALID, "cg_solve_batch: null argument");
    solve(*p, *prm, results);
  });
}

cg_status cg_damped_jacobi_solve(cg_problem* p, int32_t max_iters, double rel_tol, float omega, cg_result* res) {
  return guard([&] {
    if (!p || !res) fail(CG_ERR_INVALID, "cg_damped_jacobi_solve: null argument");
    if (p->nprob != 1) fail(CG_ERR_INVALID, "cg_damped_jacobi_solve: one problem per handle");
    auto t0 = std::chrono::steady_clock::now();
    cg_problem& P = *p;
    const Ctx c = P.ctx();
    float* X = P.d_ch[CG_CH_X];
    float* B = P.d_ch[CG_CH_B];
    float* AX = P.d_ch[CG_CH_P];
    float* D = P.d_ch[CG_CH_AP];
    k_fill_tiles<float><<<P.n, 64, 0, P.stream>>>(P.L(P.all), P.n, X, 0.f);
    double bn = std::sqrt(dot_types(P, B, B, LEAF | NONLEAF | GHOST));
    double thr = std::max(rel_tol * bn, 2.2250738585072014e-308);
    int i = 0;
    double rn = bn;
    for (i = 0; i < max_iters; i++) {
      full_negative_laplacian(P, X, AX, 0);
      k_leaf_residual<<<P.leaf.n, 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, B, AX);
      rn = std::sqrt(dot_types(P, AX, AX, LEAF));
      if (rn < thr) break;
      full_negative_laplacian(P, X, D, 1);
      k_leaf_jacobi_update<<<P.leaf.n, 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, X, AX, D, omega);
    }
    HIPC(hipStreamSynchronize(P.stream));
    res->iters = i;
    res->converged = rn < thr;
    res->rel_res = bn > 0 ? rn / bn : 0.0;
    res->bb = bn * bn;
    res->rr = rn * rn;
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->rel_res_f64 = -1.0;
  });
}

cg_status cg_vcycle(cg_problem* p, int32_t x_channel, int32_t f_channel, int32_t level_iters, int32_t coarsest_iters,
                    int32_t smoother, float omega) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(x_channel);
    check_channel(f_channel);
    if (x_channel == f_channel || x_channel == CG_CH_TMP || f_channel == CG_CH_TMP)
      fail(CG_ERR_INVALID, "cg_vcycle: x and f must differ and may not be the scratch channel CG_CH_TMP");
    if (smoother != CG_SMOOTH_RBGS && smoother != CG_SMOOTH_JACOBI) fail(CG_ERR_INVALID, "cg_vcycle: unknown smoother");
    VParams v;
    v.nu = level_iters;
    v.coarsest = coarsest_iters;
    v.smoother = smoother;
    v.omega = omega;
    v.check_done = 0;
    // the V-cycle rhs of LEAF tiles starts as f (GMGSolver.cu:393-400)
    k_to_split<<<p->leaf.n, 64, 0, p->stream>>>(p->L(p->leaf), p->leaf.n, p->d_ch[f_channel], p->d_ch[CG_CH_TMP]);
    float* Z = api_split(*p, 0, nullptr);
    enqueue_vcycle<float>(*p, smoother == CG_SMOOTH_RBGS ? p->d_zb : Z, Z, nullptr, v);
    api_unsplit(*p, p->all, Z, p->d_ch[x_channel]);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_full_negative_laplacian(cg_problem* p, int32_t x_channel, int32_t ax_channel, int32_t calc_diag) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(x_channel);
    check_channel(ax_channel);
    full_negative_laplacian(*p, p->d_ch[x_channel], p->d_ch[ax_channel], calc_diag);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_negative_laplacian_same_level(cg_problem* p, int32_t level, int32_t subtree_level, int32_t tile_types,
                                           int32_t x_channel, int32_t ax_channel, int32_t calc_diag) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(x_channel);
    check_channel(ax_channel);
    std::vector<int> l;
    for (int i = 0; i < p->n; i++) {
      const TileRec& T = p->tiles[i];
      if (level >= 0 && tr_level(T) != level) continue;
      if (subtree_type(tr_type(T), tr_level(T), subtree_level) & tile_types) l.push_back(i);
    }
    if (l.empty()) return;
    const TileRec* dl = tmp_list(*p, l);
    const Ctx c = p->ctx();
    k_apply<<<unsigned(l.size()), 64, 0, p->stream>>>(c, dl, int(l.size()), p->d_ch[x_channel], p->d_ch[ax_channel],
                                                       subtree_level, calc_diag, 0, 0);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_gauss_seidel_color(cg_problem* p, int32_t level, int32_t x_channel, int32_t rhs_channel, int32_t color) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(x_channel);
    check_channel(rhs_channel);
    if (level < 0 || level > p->max_level) fail(CG_ERR_INVALID, "level out of range");
    float* X = api_split(*p, 0, p->d_ch[x_channel]);
    const float* F = api_split(*p, 1, p->d_ch[rhs_channel]);
    launch_gs(*p, p->act[level], p->ctx(), level, X, F, color & 1, 0);
    api_unsplit(*p, p->act[level], X, p->d_ch[x_channel]);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_jacobi_sweep(cg_problem* p, int32_t level, int32_t x_channel, int32_t rhs_channel, float omega) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(x_channel);
    check_channel(rhs_channel);
    if (level < 0 || level > p->max_level) fail(CG_ERR_INVALID, "level out of range");
    const DList& a = p->act[level];
    if (!a.n) return;
    const float* X = api_split(*p, 0, p->d_ch[x_channel]);
    const float* F = api_split(*p, 1, p->d_ch[rhs_channel]);
    float* Y = api_split(*p, 2, nullptr);
    k_jacobi<float><<<a.n, 64, 0, p->stream>>>(p->ctx(), p->L(a), a.n, level, X, X, Y, F, omega, 0);
    api_unsplit(*p, a, Y, p->d_ch[x_channel]);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_residual_restrict(cg_problem* p, int32_t level, int32_t x_channel, int32_t rhs_channel, int32_t coarse_channel,
                               float one_over_alpha) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(x_channel);
    check_channel(rhs_channel);
    check_channel(coarse_channel);
    if (level < 1 || level > p->max_level) fail(CG_ERR_INVALID, "level out of range");
    const DList& al = p->lvl[level];
    if (!al.n) return;
    const float* X = api_split(*p, 0, p->d_ch[x_channel]);
    const float* F = api_split(*p, 1, p->d_ch[rhs_channel]);
    float* CO = api_split(*p, 2, p->d_ch[coarse_channel]);
    k_restrict<float><<<al.n, 64, 0, p->stream>>>(p->ctx(), p->L(al), al.n, level, X, F, CO, one_over_alpha, 0);
    api_unsplit(*p, p->all, CO, p->d_ch[coarse_channel]);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_prolongate(cg_problem* p, int32_t level, int32_t coarse_channel, int32_t fine_channel, float prolong_coeff) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(coarse_channel);
    check_channel(fine_channel);
    if (level < 1 || level > p->max_level) fail(CG_ERR_INVALID, "level out of range");
    const DList& al = p->lvl[level];
    if (!al.n) return;
    const float* XC = api_split(*p, 0, p->d_ch[coarse_channel]);
    float* F = api_split(*p, 1, p->d_ch[fine_channel]);
    k_prolong<float><<<al.n, 64, 0, p->stream>>>(p->ctx(), p->L(al), al.n, level, XC, F, F, F, prolong_coeff, 0);
    api_unsplit(*p, al, F, p->d_ch[fine_channel]);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_dot(cg_problem* p, int32_t a_channel, int32_t b_channel, int32_t tile_types, double* out) {
  return guard([&] {
    if (!p || !out) fail(CG_ERR_INVALID, "null argument");
    check_channel(a_channel);
    check_channel(b_channel);
    *out = dot_types(*p, p->d_ch[a_channel], p->d_ch[b_channel], tile_types);
  });
}

// cg_dot without any allocation: the tile list, partials and result in the
// device's `scratch` (every slot read is written first by the same call, so
// nothing needs clearing); the same kernels' arithmetic and summation order.
cg_status cg_dot_noalloc(cg_problem* p, int32_t a_channel, int32_t b_channel, int32_t tile_types, double* out) {
  return guard([&] {
    if (!p || !out) fail(CG_ERR_INVALID, "null argument");
    check_channel(a_channel);
    check_channel(b_channel);
    cg_problem& P = *p;
    if (!P.owner || !P.d_tiles) fail(CG_ERR_INVALID, "cg_dot_noalloc: not an owning handle");
    if (P.n > kNoAllocMaxTiles)
      fail(CG_ERR_INVALID, "cg_dot_noalloc: the problem has more tiles than the scratch holds (CG_NOALLOC_MAX_TILES)");
    int count = 0;
    for (int i = 0; i < P.n; i++) count += (tr_type(P.tiles[i]) & tile_types) ? 1 : 0;
    if (count == 0) {
      *out = 0.0;
      return;
    }
    unsigned char* s = scratch_on(P.device);
    double* part = reinterpret_cast<double*>(s + kScrPart);
    int32_t* sel = reinterpret_cast<int32_t*>(s + kScrSel);
    double* res = reinterpret_cast<double*>(s + kScrOut);
    k_select_tiles<<<1, SEL_THREADS, 0, P.stream>>>(P.d_tiles, P.n, tile_types, sel);
    k_dot_selected<<<unsigned(count), 64, 0, P.stream>>>(sel, count, P.d_ch[a_channel], P.d_ch[b_channel], part);
    k_sum_partials<<<1, FIN_THREADS, 0, P.stream>>>(part, count, res);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, P.stream));
    HIPC(hipStreamSynchronize(P.stream));
  });
}

cg_status cg_recenter_leaf_voxels(cg_problem* p, int32_t channel) {
  return guard([&] {
    if (!p) fail(CG_ERR_INVALID, "null problem");
    check_channel(channel);
    set_prob_state(*p, 0.0, 0, 1);
    const Ctx c = p->ctx();
    k_mean<<<p->leaf.n, 64, 0, p->stream>>>(c, p->L(p->leaf), p->leaf.n, p->d_ch[channel], 0);
    k_finalize<<<1, FIN_THREADS, 0, p->stream>>>(c, FIN_MEAN, 0);
    k_subtract_mean<float><<<p->leaf.n, 64, 0, p->stream>>>(c, p->L(p->leaf), p->leaf.n, p->d_ch[channel], 0, 0);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(p->stream));
  });
}

cg_status cg_benchmark_kernel(cg_problem* p, int32_t kernel, int32_t level, int32_t reps, double* avg_us, int64_t* tiles) {
  return guard([&] {
    if (!p || !avg_us || reps <= 0 || reps > 100000) fail(CG_ERR_INVALID, "cg_benchmark_kernel: bad argument");
    cg_problem& P = *p;
    if (kernel == CG_KERNEL_AMG_GS && !P.amg_ready) fail(CG_ERR_STATE, "cg_benchmark_kernel: call cg_amg_prepare first");
    if ((kernel == CG_KERNEL_GS_PAIR || kernel == CG_KERNEL_GS_PAIR_BF16) && !P.d_tiles)
      fail(CG_ERR_STATE, "cg_benchmark_kernel: the handle has no tile-indexed records");
    if ((kernel == CG_KERNEL_GS_SWEEP || kernel == CG_KERNEL_JACOBI || kernel == CG_KERNEL_GS_SWEEP_BF16 ||
         kernel == CG_KERNEL_GS_SWEEP_F16 || kernel == CG_KERNEL_AMG_GS || kernel == CG_KERNEL_GS_PAIR ||
         kernel == CG_KERNEL_GS_PAIR_BF16) &&
        (level < 0 || level > P.max_level))
      fail(CG_ERR_INVALID, "level out of range");
    if ((kernel == CG_KERNEL_RESTRICT || kernel == CG_KERNEL_PROLONG || kernel == CG_KERNEL_GS_PROLONG ||
         kernel == CG_KERNEL_GS_PROLONG_BF16) &&
        (level < 1 || level > P.max_level))
      fail(CG_ERR_INVALID, "level out of range");
    if (kernel == CG_KERNEL_GS_SWEEP_BF16 || kernel == CG_KERNEL_GS_PROLONG_BF16 || kernel == CG_KERNEL_GS_SWEEP_F16 ||
        kernel == CG_KERNEL_GS_PAIR_BF16)
      ensure_bf16(P);
    if ((kernel == CG_KERNEL_COARSE || kernel == CG_KERNEL_EMPTY) && level < 0) fail(CG_ERR_INVALID, "negative count");
    set_prob_state(P, 0.0, 1 << 30, 0);
    const Ctx c = P.ctx();
    const int base = F_VCYCLE | kSwizzle | F_CHECK_DONE;
    const bf16* nob = nullptr;
    int64_t nt = 0;
    auto launch = [&]() {
      switch (kernel) {
        case CG_KERNEL_GS_SWEEP:
          nt = P.act[level].n;
          launch_gs(P, P.act[level], c, level, c.z, c.r, 0, base | F_GHOST_ZERO);
          break;
        case CG_KERNEL_GS_SWEEP_BF16: {
          const DList& a = P.act[level];
          nt = a.n;
          launch_gs(P, a, c, level, c.zh, nob, 0, base | F_GHOST_ZERO);
          break;
        }
        case CG_KERNEL_GS_PAIR: {  // one red-black iteration in one launch (z read, the scratch written)
          const DList& a = P.act[level];
          nt = a.n;
          launch_gs_pair<float>(P, a, c, c.z, c.z, c.zb, 0, base | F_GHOST_ZERO);
          break;
        }
        case CG_KERNEL_GS_PAIR_BF16: {
          const DList& a = P.act[level];
          nt = a.n;
          launch_gs_pair<bf16>(P, a, c, c.zh, c.zh, c.zbh, 0, base | F_GHOST_ZERO);
          break;
        }
        case CG_KERNEL_AMG_GS: {
          const DList& a = P.act[level];
          nt = a.n;
          if (a.n)
            k_amg_gs<<<grid_for(a.n, base), 64, 0, P.stream>>>(c, P.L(a), a.n, P.d_zb, P.d_ch[CG_CH_TMP], amg_arrays(P), 0,
                                                               base);
          break;
        }
        case CG_KERNEL_GS_SWEEP_F16: {
          const DList& a = P.act[level];
          nt = a.n;
          launch_gs(P, a, c, level, c.zf, static_cast<const f16*>(nullptr), 0, base | F_GHOST_ZERO);
          break;
        }
        case CG_KERNEL_GS_PROLONG_BF16: {
          const DList& al = P.lvl[level];
          nt = al.n;
          k_gs_prolong<bf16><<<grid_for(al.n, base), 64, 0, P.stream>>>(c, P.L(al), al.n, c.zbh, c.zh, 1.f, 1, base);
          break;
        }
        case CG_KERNEL_JACOBI: {
          const DList& a = P.act[level];
          nt = a.n;
          if (a.n) k_jacobi<float><<<grid_for(a.n, base), 64, 0, P.stream>>>(c, P.L(a), a.n, level, c.z, c.z, c.zb, c.r, 0.5f, base);
          break;
        }
        case CG_KERNEL_APPLY:
          nt = P.applyl.n;
          k_apply<true><<<grid_for(P.applyl.n, base), 64, 0, P.stream>>>(c, P.L(P.applyl), P.applyl.n, c.p, c.ap, -1, 0, 1,
                                                                        base);
          break;
        case CG_KERNEL_RESTRICT: {
          const DList& al = P.lvl[level];
          nt = al.n;
          int fl = base | F_GHOST_ZERO;
          k_restrict<float><<<grid_for(al.n, fl), 64, 0, P.stream>>>(c, P.L(al), al.n, level, c.z, c.r, nullptr, 1.f, fl);
          break;
        }
        case CG_KERNEL_PROLONG: {
          const DList& al = P.lvl[level];
          nt = al.n;
          k_prolong<float><<<grid_for(al.n, base), 64, 0, P.stream>>>(c, P.L(al), al.n, level, c.z, c.z, c.zb, c.zb, 1.f, base);
          break;
        }
        case CG_KERNEL_GS_PROLONG: {
          const DList& al = P.lvl[level];
          nt = al.n;
          k_gs_prolong<float><<<grid_for(al.n, base), 64, 0, P.stream>>>(c, P.L(al), al.n, c.zb, c.z, 1.f, 1, base);
          break;
        }
        case CG_KERNEL_UPDATE_XR:
          nt = P.leaf.n;
          k_cg_update_xr<float><<<grid_for(P.leaf.n, base), 64, 0, P.stream>>>(c, P.L(P.leaf), P.leaf.n, base);
          break;
        case CG_KERNEL_COARSE: {
          nt = int64_t(P.coarse_slots.size());
          VCycleArgs va{2, level, 0, 2.f / 3.f, 1.f, 1.f};  // `level` = bottom iterations
          if (!P.plans.empty())
            k_coarse_lds<float><<<unsigned(P.plans.size()), CL_THREADS, 0, P.stream>>>(c, P.d_plans, va, c.z, base & ~F_SWIZZLE);
          break;
        }
        case CG_KERNEL_COPY:
          nt = P.n;
          k_copy_tiles<<<P.n, 64, 0, P.stream>>>(P.L(P.all), P.n, c.x, c.ap);
          break;
        case CG_KERNEL_EMPTY:
          nt = level;
          k_empty<<<unsigned(std::max(1, level)), 64, 0, P.stream>>>(nullptr);
          break;
        default:
          fail(CG_ERR_INVALID, "unknown kernel id");
      }
    };
    launch();  // warm-up
    // back-to-back launches between one pair of events (an event pair around
    // each launch would add its own few microseconds to every launch)
    struct Events {
      hipEvent_t e[2] = {nullptr, nullptr};
      ~Events() { for (hipEvent_t x : e) if (x) (void)hipEventDestroy(x); }
    } ev;
    HIPC(hipEventCreate(&ev.e[0]));
    HIPC(hipEventCreate(&ev.e[1]));
    HIPC(hipEventRecord(ev.e[0], P.stream));
    for (int i = 0; i < reps; i++) launch();
    HIPC(hipEventRecord(ev.e[1], P.stream));
    HIPC(hipEventSynchronize(ev.e[1]));
    HIPC(hipGetLastError());
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    *avg_us = double(ms) * 1e3 / reps;
    if (tiles) *tiles = nt;
  });
}

cg_status cg_calc_cell_types_from_leafs(const cg_tree_desc* desc, uint8_t* cell_type_inout) {
  return guard([&] {
    if (!desc || !cell_type_inout || desc->num_tiles <= 0) fail(CG_ERR_INVALID, "null argument");
    Tree T;
    for (int t = 0; t < desc->num_tiles; t++) {
      const int32_t* c = desc->coord + 3 * t;
      if (!key_ok(c[0], c[1], c[2])) fail(CG_ERR_INVALID, "coordinate out of range");
      T.add(desc->level[t], c[0], c[1], c[2], desc->type[t]);
    }
    std::memcpy(T.ctype.data(), cell_type_inout, size_t(desc->num_tiles) * 512);
    T.cell_types_from_leafs();
    std::memcpy(cell_type_inout, T.ctype.data(), size_t(desc->num_tiles) * 512);
  });
}

cg_status cg_server_create(const cg_server_config* cfg, cg_server** out) {
  if (!out) { g_err = "cg_server_create: null output"; return CG_ERR_INVALID; }
  *out = nullptr;
  auto S = std::make_unique<cg_server>();
  cg_status st = guard([&] { server_create(cfg, *S); });
  if (st == CG_OK) *out = S.release();
  return st;
}

cg_status cg_server_destroy(cg_server* s) {
  delete s;
  return CG_OK;
}

cg_status cg_server_admit(cg_server* s, const cg_tree_desc* desc, int32_t* instance) {
  return guard([&] {
    if (!s) fail(CG_ERR_INVALID, "null server");
    server_admit(*s, desc, instance);
  });
}

cg_status cg_server_release(cg_server* s, int32_t instance) {
  return guard([&] {
    if (!s) fail(CG_ERR_INVALID, "null server");
    server_release(*s, instance);
  });
}

cg_status cg_server_submit(cg_server* s, int32_t instance, const float* rhs, int32_t flags, int64_t tag) {
  return guard([&] {
    if (!s) fail(CG_ERR_INVALID, "null server");
    server_submit(*s, instance, rhs, flags, tag);
  });
}

cg_status cg_server_perturb_rhs(cg_server* s, int32_t instance, float rho, float sigma, uint64_t seed) {
  return guard([&] {
    if (!s) fail(CG_ERR_INVALID, "null server");
    server_perturb(*s, instance, rho, sigma, seed);
  });
}

cg_status cg_server_step(cg_server* s, int32_t* num_active) {
  return guard([&] {
    if (!s) fail(CG_ERR_INVALID, "null server");
    server_step(*s, num_active);
  });
}

cg_status cg_server_next_result(cg_server* s, int32_t wait, int32_t* got, int64_t* tag, int32_t* instance, cg_result* res,
                                float* x) {
  return guard([&] {
    if (!s || !got) fail(CG_ERR_INVALID, "null argument");
    *got = server_next(*s, wait, tag, instance, res, x) ? 1 : 0;
  });
}

cg_status cg_server_info(const cg_server* s, int64_t* capacity_tiles, int64_t* free_tiles, int32_t* num_instances) {
  if (!s) { g_err = "null server"; return CG_ERR_INVALID; }
  if (capacity_tiles) *capacity_tiles = s->cap_tiles;
  if (free_tiles) {
    int64_t f = 0;
    for (auto& kv : s->tiles_alloc.free_) f += kv.second;
    *free_tiles = f;
  }
  if (num_instances) *num_instances = s->nres;
  return CG_OK;
}

cg_status cg_scene_create(const cg_scene_params* sp, cg_scene** out) {
  if (!out || !sp) { g_err = "null argument"; return CG_ERR_INVALID; }
  *out = nullptr;
  auto S = std::make_unique<cg_scene>();
  cg_status st = guard([&] {
    if (sp->nroot[0] <= 0 || sp->nroot[1] <= 0 || sp->nroot[2] <= 0 || sp->num_layers <= 0 || sp->num_layers > 20 || !(sp->h0 > 0))
      fail(CG_ERR_INVALID, "cg_scene_create: invalid scene parameters");
    scene_build(*sp, S->s);
    const Tree& T = S->s.tree;
    S->coord_flat.resize(size_t(T.size()) * 3);
    for (int t = 0; t < T.size(); t++) for (int d = 0; d < 3; d++) S->coord_flat[3 * t + d] = T.coord[t][d];
  });
  if (st == CG_OK) *out = S.release();
  return st;
}

cg_status cg_scene_destroy(cg_scene* s) {
  delete s;
  return CG_OK;
}

cg_status cg_scene_tree(const cg_scene* s, cg_tree_desc* desc, const float** rhs) {
  if (!s || !desc) { g_err = "null argument"; return CG_ERR_INVALID; }
  const Tree& T = s->s.tree;
  desc->h0 = T.h0;
  desc->num_tiles = T.size();
  desc->level = T.level.data();
  desc->coord = s->coord_flat.data();
  desc->type = T.type.data();
  desc->cell_type = T.ctype.data();
  if (rhs) *rhs = s->s.rhs.data();
  return CG_OK;
}

cg_status cg_scene_velocity(const cg_scene* s, const float** uvw) {
  if (!s || !uvw) { g_err = "null argument"; return CG_ERR_INVALID; }
  *uvw = s->s.uvw.data();
  return CG_OK;
}

cg_status cg_tiles_save(const char* path, const cg_tree_desc* desc, int32_t num_channels, const int32_t* channel_ids,
                        const float* const* channels) {
  return guard([&] {
    if (!path || !desc || desc->num_tiles < 0 || num_channels < 0 || (num_channels && (!channel_ids || !channels)))
      fail(CG_ERR_INVALID, "cg_tiles_save: bad argument");
    const size_t n = size_t(desc->num_tiles);
    if (n && (!desc->level || !desc->coord || !desc->type || !desc->cell_type))
      fail(CG_ERR_INVALID, "cg_tiles_save: null tree array");
    TileSet t;
    t.h0 = desc->h0;
    t.level.assign(desc->level, desc->level + n);
    t.coord.assign(desc->coord, desc->coord + 3 * n);
    t.type.assign(desc->type, desc->type + n);
    t.ctype.assign(desc->cell_type, desc->cell_type + 512 * n);
    t.channel_id.assign(channel_ids, channel_ids + num_channels);
    t.channel.resize(size_t(num_channels) * 512 * n);
    for (int32_t c = 0; c < num_channels; c++) {
      if (!channels[c]) fail(CG_ERR_INVALID, "cg_tiles_save: null channel");
      std::memcpy(&t.channel[size_t(c) * 512 * n], channels[c], 512 * n * sizeof(float));
    }
    write_tile_file(path, t);
  });
}

cg_status cg_scene_tree_device(const cg_scene_params* sp, int64_t max_tiles, cg_tileset** out) {
  if (!sp || !out) { g_err = "null argument"; return CG_ERR_INVALID; }
  *out = nullptr;
  auto S = std::make_unique<cg_tileset>();
  cg_status st = guard([&] { scene_tree_device(*sp, max_tiles, S->s); });
  if (st == CG_OK) *out = S.release();
  return st;
}

cg_status cg_tree_adapt_device(const cg_tree_desc* tree, const cg_scene_params* sp, int32_t ops, int64_t max_tiles,
                               cg_tileset** out) {
  if (!tree || !sp || !out) { g_err = "null argument"; return CG_ERR_INVALID; }
  *out = nullptr;
  auto S = std::make_unique<cg_tileset>();
  cg_status st = guard([&] { tree_adapt_device(*tree, *sp, ops, max_tiles, S->s); });
  if (st == CG_OK) *out = S.release();
  return st;
}

cg_status cg_tiles_load(const char* path, cg_tileset** out) {
  if (!path || !out) { g_err = "null argument"; return CG_ERR_INVALID; }
  *out = nullptr;
  auto S = std::make_unique<cg_tileset>();
  cg_status st = guard([&] { S->s = read_tile_file(path); });
  if (st == CG_OK) *out = S.release();
  return st;
}

cg_status cg_tiles_import_holder(const char* path, int32_t num_channels, const int32_t* channel_ids, cg_tileset** out) {
  if (!path || !out || num_channels < 0 || (num_channels && !channel_ids)) { g_err = "bad argument"; return CG_ERR_INVALID; }
  *out = nullptr;
  auto S = std::make_unique<cg_tileset>();
  cg_status st = guard([&] {
    S->s = read_holder_dump(path, std::vector<int32_t>(channel_ids, channel_ids + num_channels));
  });
  if (st == CG_OK) *out = S.release();
  return st;
}

cg_status cg_tileset_tree(const cg_tileset* s, cg_tree_desc* desc) {
  if (!s || !desc) { g_err = "null argument"; return CG_ERR_INVALID; }
  desc->h0 = s->s.h0;
  desc->num_tiles = int32_t(s->s.size());
  desc->level = s->s.level.data();
  desc->coord = s->s.coord.data();
  desc->type = s->s.type.data();
  desc->cell_type = s->s.ctype.data();
  return CG_OK;
}

cg_status cg_tileset_channels(const cg_tileset* s, int32_t* num_channels, const int32_t** channel_ids,
                              const float** data) {
  if (!s || !num_channels || !channel_ids || !data) { g_err = "null argument"; return CG_ERR_INVALID; }
  *num_channels = int32_t(s->s.channel_id.size());
  *channel_ids = s->s.channel_id.data();
  *data = s->s.channel.data();
  return CG_OK;
}

cg_status cg_tileset_destroy(cg_tileset* s) {
  delete s;
  return CG_OK;
}

}  // extern "C"
