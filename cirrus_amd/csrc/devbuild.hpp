// Device-side octree build of a synthetic scene (device code and its host
// driver; included by mgpcg.hip inside its anonymous namespace).
//
// The tree follows the reference's refinement and ghost rules (IterativeRefine
// with 2:1 face balance, src/PoissonGrid.h:35-162,317-326; SpawnGhostTiles,
// src/PoissonGrid.cu:30-60) and its cell typing (the walls and the solid
// sphere of the scene on LEAF cells, then CalcCellTypesFromLeafs,
// src/PoissonSolver.cu:7-61), every step a kernel over the tiles:
//   * a hash table of (level, coordinate) keys, filled by compare-and-swap,
//     stands in for the reference's per-level hash tables (src/HAAccessor.h);
//   * each refinement round marks the LEAF tiles whose target level exceeds
//     their level and, for 2:1 face balance, the deepest existing ancestor of
//     every missing level l-1 tile next to a level-l LEAF, then refines the
//     marked tiles (8 LEAF children each); rounds repeat until none marks.
//     Both rules only add tiles, so the rounds reach the same least tree the
//     reference's steps reach, in whatever order they run;
//   * ghosts: an empty face-neighbour position of an active tile of level >= 1
//     whose parent position holds a LEAF;
//   * cell types: LEAF cells by the scene's walls (a cell whose same-level
//     neighbour has no tile is on the wall; NEUMANN wins) and solid sphere,
//     NONLEAF cells from their children (finest level first), GHOST cells from
//     their parents.
// Adaptation of an existing tree (the active tiles of a host tree loaded into
// the table) runs the refinement rounds to the scene's target levels, then
// CoarsenStep (src/PoissonGrid.h:165-314) until a step deletes nothing, as
// the reference's RefineWithPointValue / CoarsenWithPointValue do
// (src/FluidEuler.h:68-111), then spawns ghosts and types the cells.  The
// coarsening runs without ghost tiles: in the reference's step a ghost's flags
// reach no LEAF or NONLEAF tile (steps 1-2 read only children and NONLEAF
// neighbours, step 3 reads a LEAF's NONLEAF parent), and its ghost-only parts
// (step 3 on ghosts, step 4) only delete ghosts whose places the final
// SpawnGhostTiles call decides again, so the ghosts the final spawning gives
// are the reference's.
// The scene's target levels (uniform, sphere band, plane) use only
// comparisons, +, -, *, and a square root, which round the same way on the
// host and the device, so the tree is the host builder's tree exactly.

constexpr unsigned long long kDbEmpty = ~0ull;

struct DevTree {
  int cap = 0;
  uint32_t hmask = 0;
  int* count = nullptr;        // tiles so far
  int* flags = nullptr;        // [0] a tile was marked, [1] capacity exceeded
  int32_t* level = nullptr;
  int32_t* coord = nullptr;    // 3 per tile
  uint8_t* type = nullptr;
  uint8_t* mark = nullptr;
  uint8_t* ctype = nullptr;    // 512 per tile
  unsigned long long* hkeys = nullptr;
  int32_t* hvals = nullptr;
};

__host__ __device__ __forceinline__ bool db_key_ok(int x, int y, int z) {
  return x >= 0 && y >= 0 && z >= 0 && x < (1 << 19) && y < (1 << 19) && z < (1 << 19);
}
__device__ __forceinline__ unsigned long long db_key(int l, int x, int y, int z) {
  return (static_cast<unsigned long long>(l) << 57) | (static_cast<unsigned long long>(x) << 38) |
         (static_cast<unsigned long long>(y) << 19) | static_cast<unsigned long long>(z);
}
__device__ __forceinline__ uint32_t db_hash(unsigned long long k, uint32_t mask) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return uint32_t(k) & mask;
}
// Tile index at (l, x, y, z), -1 if none (bounded probe).
__device__ __forceinline__ int db_find(const DevTree& B, int l, int x, int y, int z) {
  if (l < 0 || !db_key_ok(x, y, z)) return -1;
  const unsigned long long key = db_key(l, x, y, z);
  uint32_t s = db_hash(key, B.hmask);
  for (uint32_t probe = 0; probe <= B.hmask; probe++) {
    const unsigned long long k = B.hkeys[s];
    if (k == key) return B.hvals[s];
    if (k == kDbEmpty) return -1;
    s = (s + 1) & B.hmask;
  }
  return -1;
}
// Insert a tile unless its key is present; its index is visible to lookups in
// later launches.  Capacity overflow raises flags[1].
__device__ __forceinline__ void db_insert(const DevTree& B, int l, int x, int y, int z, uint8_t ty) {
  const unsigned long long key = db_key(l, x, y, z);
  uint32_t s = db_hash(key, B.hmask);
  for (uint32_t probe = 0; probe <= B.hmask; probe++) {
    const unsigned long long prev = atomicCAS(&B.hkeys[s], kDbEmpty, key);
    if (prev == key) return;
    if (prev == kDbEmpty) {
      const int idx = atomicAdd(B.count, 1);
      if (idx >= B.cap) {
        B.flags[1] = 1;
        return;
      }
      B.level[idx] = l;
      B.coord[3 * idx] = x;
      B.coord[3 * idx + 1] = y;
      B.coord[3 * idx + 2] = z;
      B.type[idx] = ty;
      B.mark[idx] = 0;
      B.hvals[s] = idx;
      return;
    }
    s = (s + 1) & B.hmask;
  }
  B.flags[1] = 1;
}

__device__ __forceinline__ bool db_in_domain(const cg_scene_params& p, int l, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0) return false;
  return (x >> l) < p.nroot[0] && (y >> l) < p.nroot[1] && (z >> l) < p.nroot[2];
}
// The scene's target level of a tile (the host builder's rule, same operations).
__device__ __forceinline__ int db_target(const cg_scene_params& p, int level, int bx, int by, int bz) {
  const double h = p.h0 / double(1 << level);
  const int b[3] = {bx, by, bz};
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = b[d] * 8 * h;
    hi[d] = (b[d] + 1) * 8 * h;
  }
  if (p.kind == CG_SCENE_UNIFORM) return p.uniform_level;
  if (p.kind == CG_SCENE_PLANE) return (lo[2] < p.plane_z) ? p.plane_level[0] : p.plane_level[1];
  double dmin2 = 0, dmax2 = 0;
  for (int d = 0; d < 3; d++) {
    const double c = p.band_c[d];
    const double nd = fmax(lo[d] - c, fmax(0.0, c - hi[d]));
    const double fd = fmax(fabs(lo[d] - c), fabs(hi[d] - c));
    dmin2 += nd * nd;
    dmax2 += fd * fd;
  }
  const double dmin = sqrt(dmin2), dmax = sqrt(dmax2);
  for (int w = 0; w < 2; w++)
    if (dmin < p.band_r + p.band_w[w] && dmax > p.band_r - p.band_w[w]) return p.band_level[w];
  return p.band_level[2];
}

__global__ void k_db_roots(DevTree B, cg_scene_params p) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = p.nroot[0] * p.nroot[1] * p.nroot[2];
  if (r >= n) return;
  // the host builder's root order: i outermost
  const int k = r % p.nroot[2], j = (r / p.nroot[2]) % p.nroot[1], i = r / (p.nroot[2] * p.nroot[1]);
  db_insert(B, 0, i, j, k, T_LEAF);
}
// Marks of one refinement round (tiles [0, n)).
__global__ void k_db_mark(DevTree B, cg_scene_params p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || B.type[t] != T_LEAF) return;
  const int l = B.level[t];
  const int x = B.coord[3 * t], y = B.coord[3 * t + 1], z = B.coord[3 * t + 2];
  if (l + 1 < p.num_layers && db_target(p, l, x, y, z) > l) {
    B.mark[t] = 1;
    B.flags[0] = 1;
  }
  if (l < 2) return;
  // 2:1 face balance: the level l-1 tile holding each in-domain face-neighbour position
  for (int f = 0; f < 6; f++) {
    int q[3] = {x, y, z};
    q[f >> 1] += (f & 1) ? 1 : -1;
    if (!db_in_domain(p, l, q[0], q[1], q[2])) continue;
    if (db_find(B, l - 1, q[0] >> 1, q[1] >> 1, q[2] >> 1) >= 0) continue;
    for (int m = l - 2; m >= 0; m--) {
      const int s = l - m;
      const int a = db_find(B, m, q[0] >> s, q[1] >> s, q[2] >> s);
      if (a >= 0) {
        B.mark[a] = 1;
        B.flags[0] = 1;
        break;
      }
    }
  }
}
// Refine the marked LEAF tiles of [0, n): NONLEAF, eight LEAF children.
__global__ void k_db_refine(DevTree B, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || !B.mark[t]) return;
  B.mark[t] = 0;
  if (B.type[t] != T_LEAF) return;
  B.type[t] = T_NONLEAF;
  const int l = B.level[t];
  const int x = B.coord[3 * t], y = B.coord[3 * t + 1], z = B.coord[3 * t + 2];
  for (int o = 0; o < 8; o++) db_insert(B, l + 1, 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + ((o >> 2) & 1), T_LEAF);
}
// Ghost tiles next to the active tiles of [0, n).
__global__ void k_db_ghosts(DevTree B, int n) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = id / 6, f = id % 6;
  if (t >= n || !(B.type[t] & (T_LEAF | T_NONLEAF))) return;
  const int l = B.level[t];
  if (l == 0) return;
  int q[3] = {B.coord[3 * t], B.coord[3 * t + 1], B.coord[3 * t + 2]};
  q[f >> 1] += (f & 1) ? 1 : -1;
  if (!db_key_ok(q[0], q[1], q[2]) || db_find(B, l, q[0], q[1], q[2]) >= 0) return;
  const int pa = db_find(B, l - 1, q[0] >> 1, q[1] >> 1, q[2] >> 1);
  if (pa >= 0 && B.type[pa] == T_LEAF) db_insert(B, l, q[0], q[1], q[2], T_GHOST);
}
// LEAF cell types: the walls, then the solid sphere (one workgroup per tile, a cell per thread).
__global__ __launch_bounds__(512) void k_db_leaf_types(DevTree B, cg_scene_params p, int n) {
  const int t = blockIdx.x, c = threadIdx.x;
  if (t >= n || B.type[t] != T_LEAF) return;
  const int l = B.level[t];
  const double h = p.h0 / double(1 << l);
  const int g[3] = {B.coord[3 * t] * 8 + (c >> 6), B.coord[3 * t + 1] * 8 + ((c >> 3) & 7), B.coord[3 * t + 2] * 8 + (c & 7)};
  bool neu = false, dir = false;
  for (int f = 0; f < 6; f++) {
    int q[3] = {g[0], g[1], g[2]};
    q[f >> 1] += (f & 1) ? 1 : -1;
    if (db_find(B, l, q[0] >> 3, q[1] >> 3, q[2] >> 3) < 0) {
      if (p.face_bc[f] == CG_DIRICHLET) dir = true;
      else neu = true;
    }
  }
  uint8_t ty = C_INT;
  if (neu) ty = C_NEU;
  else if (dir) ty = C_DIR;
  else if (p.has_solid) {
    double d2 = 0;
    for (int d = 0; d < 3; d++) {
      const double xv = (g[d] + 0.5) * h - p.solid_c[d];
      d2 += xv * xv;
    }
    if (d2 <= p.solid_r * p.solid_r) ty = C_NEU;
  }
  B.ctype[size_t(t) * 512 + c] = ty;
}
// NONLEAF cell types of level l from their children (CalcCellTypesFromLeafs).
__global__ __launch_bounds__(512) void k_db_nonleaf_types(DevTree B, int n, int l) {
  const int t = blockIdx.x, c = threadIdx.x;
  if (t >= n || B.level[t] != l || B.type[t] != T_NONLEAF) return;
  const int i = c >> 6, j = (c >> 3) & 7, k = c & 7;
  const int x = B.coord[3 * t], y = B.coord[3 * t + 1], z = B.coord[3 * t + 2];
  int nd = 0, ni = 0;
  for (int s = 0; s < 8; s++) {
    const int fi = 2 * i + (s & 1), fj = 2 * j + ((s >> 1) & 1), fk = 2 * k + ((s >> 2) & 1);
    const int ct = db_find(B, l + 1, 2 * x + (fi >> 3), 2 * y + (fj >> 3), 2 * z + (fk >> 3));
    if (ct < 0 || !(B.type[ct] & (T_LEAF | T_NONLEAF))) continue;
    const uint8_t v = B.ctype[size_t(ct) * 512 + (((fi & 7) << 6) | ((fj & 7) << 3) | (fk & 7))];
    if (v & C_DIR) nd++;
    else if (v & C_INT) ni++;
  }
  B.ctype[size_t(t) * 512 + c] = nd ? C_DIR : (ni ? C_INT : C_NEU);
}
// GHOST cell types from the parent's cells (DIRICHLET without a parent).
__global__ __launch_bounds__(512) void k_db_ghost_types(DevTree B, int n) {
  const int t = blockIdx.x, c = threadIdx.x;
  if (t >= n || B.type[t] != T_GHOST) return;
  const int l = B.level[t];
  const int x = B.coord[3 * t], y = B.coord[3 * t + 1], z = B.coord[3 * t + 2];
  const int pa = db_find(B, l - 1, x >> 1, y >> 1, z >> 1);
  if (pa < 0) {
    B.ctype[size_t(t) * 512 + c] = C_DIR;
    return;
  }
  const int gi = x * 8 + (c >> 6), gj = y * 8 + ((c >> 3) & 7), gk = z * 8 + (c & 7);
  const int pi = (gi >> 1) & 7, pj = (gj >> 1) & 7, pk = (gk >> 1) & 7;
  B.ctype[size_t(t) * 512 + c] = B.ctype[size_t(pa) * 512 + ((pi << 6) | (pj << 3) | pk)];
}

// ---- Adaptation ------------------------------------------------------------
// CoarsenStep's flags, the reference's status bits (src/PoissonTile.h:7).
constexpr uint8_t kDcDelete = 2, kDcCoarsen = 4;

// The active tiles of a host tree (arrays already on the device).
__global__ void k_db_load(DevTree B, const int32_t* __restrict__ level, const int32_t* __restrict__ coord,
                          const uint8_t* __restrict__ type, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || !(type[t] & (T_LEAF | T_NONLEAF))) return;
  db_insert(B, level[t], coord[3 * t], coord[3 * t + 1], coord[3 * t + 2], type[t]);
}
// CoarsenStep 1 on level l (src/PoissonGrid.h:178-201): a NONLEAF tile is
// flagged COARSEN when its 8 children are LEAF tiles flagged DELETE.
__global__ void k_dc_nonleaf(DevTree B, int n, int l) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || B.level[t] != l || B.type[t] != T_NONLEAF) return;
  const int x = B.coord[3 * t], y = B.coord[3 * t + 1], z = B.coord[3 * t + 2];
  bool co = true;
  for (int o = 0; o < 8; o++) {
    const int c = db_find(B, l + 1, 2 * x + (o & 1), 2 * y + ((o >> 1) & 1), 2 * z + ((o >> 2) & 1));
    if (c < 0 || B.type[c] != T_LEAF || !(B.mark[c] & kDcDelete)) co = false;
  }
  B.mark[t] = co ? kDcCoarsen : 0;
}
// CoarsenStep 2 on level l (src/PoissonGrid.h:204-225): a LEAF tile above its
// target level is flagged DELETE unless a same-level NONLEAF face neighbour
// is not flagged COARSEN.
__global__ void k_dc_leaf(DevTree B, cg_scene_params p, int n, int l) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || B.level[t] != l || B.type[t] != T_LEAF) return;
  const int x = B.coord[3 * t], y = B.coord[3 * t + 1], z = B.coord[3 * t + 2];
  bool del = db_target(p, l, x, y, z) < l;
  for (int f = 0; f < 6; f++) {
    int q[3] = {x, y, z};
    q[f >> 1] += (f & 1) ? 1 : -1;
    const int nb = db_find(B, l, q[0], q[1], q[2]);
    if (nb >= 0 && B.type[nb] == T_NONLEAF && !(B.mark[nb] & kDcCoarsen)) del = false;
  }
  B.mark[t] = del ? kDcDelete : 0;
}
// CoarsenStep 3 for LEAF tiles (src/PoissonGrid.h:230-254): a LEAF tile goes
// exactly when its parent is flagged COARSEN.  Parents are NONLEAF tiles,
// which this launch does not write, so all levels go in one launch.
__global__ void k_dc_down(DevTree B, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || B.type[t] != T_LEAF) return;
  const int l = B.level[t];
  const int pa = l > 0 ? db_find(B, l - 1, B.coord[3 * t] >> 1, B.coord[3 * t + 1] >> 1, B.coord[3 * t + 2] >> 1) : -1;
  B.mark[t] = (pa >= 0 && (B.mark[pa] & kDcCoarsen)) ? uint8_t(kDcDelete | kDcCoarsen) : uint8_t(0);
}
// The end of CoarsenStep (src/PoissonGrid.h:278-313): tiles flagged DELETE
// are dropped (counted in D.flags[0]), NONLEAF tiles flagged COARSEN become
// LEAF tiles, the rest keep their type; survivors go into the empty table D.
__global__ void k_dc_compact(DevTree B, DevTree D, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint8_t m = B.mark[t];
  if (m & kDcDelete) {
    atomicAdd(&D.flags[0], 1);
    return;
  }
  const uint8_t ty = (B.type[t] == T_NONLEAF && (m & kDcCoarsen)) ? uint8_t(T_LEAF) : B.type[t];
  db_insert(D, B.level[t], B.coord[3 * t], B.coord[3 * t + 1], B.coord[3 * t + 2], ty);
}
