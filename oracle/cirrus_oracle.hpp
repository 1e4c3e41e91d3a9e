// ============================================================================
// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// A CPU (C++17/OpenMP) restatement of the reference's MGPCG pressure-projection
// path: the HAGrid octree of 8^3 tiles (hash tables, refinement with 2:1 face
// balance, ghost spawning, neighbour tables), cell typing, the divergence RHS,
// the 7-point operator with coarse/fine (ghost) faces, red-black Gauss-Seidel
// and damped-Jacobi smoothing, restriction, prolongation, the V-cycle and the
// MGPCG solve.  Templated on the arithmetic type: T=float is the FP32
// transcription (the GPU path must match it), T=double is the FP64 oracle.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it, and only as the checker or the CPU baseline, never as the product.
//
// Every function cites the reference file:line it restates.  Two deliberate
// deviations, both shared with the GPU path and stated where they occur:
//   * dot products use one fixed, deterministic FP64 summation order
//     (tile_partial / sum_partials below) instead of cub's (cub's order is not
//     pinned by anything in the reference, so that choice stays "parity
//     unpinned");
//   * an INTERIOR cell whose diagonal D is 0 gets no update (the reference
//     divides by zero there: GMGSolver.cu:244-265).
// FP contraction is off (-ffp-contract=off); where the reference's compiler
// would fuse a multiply-add (x += a*b), std::fma is written explicitly so both
// this file and the GPU kernels round identically.
//
// How each part is pinned (tests/test_oracle_pins.py, tests/test_abi_cpu.py),
// since the reference ships no tests or golden vectors and cannot be built here:
//   grid build (IterativeRefine, RefineLeafsOneStep, SpawnGhostTiles,
//     CalculateNeighborTiles)      2:1 face balance, ghost-placement rule and
//                                  neighbour symmetry checked tile by tile;
//                                  closed-form uniform counts
//   CoarsenStep, AdaptScene        tests/test_oracle_adapt.py: uniform
//                                  coarsening against its closed form; the
//                                  adaptation of any tree to a target whose
//                                  value on a tile is its children's maximum
//                                  equals that scene's own build; 2:1 balance,
//                                  the ghost rule and the coarsening's fixed
//                                  point on other targets
//   SceneCellType                  brute-force count of solid / wall cells
//   CalcCellTypesFromLeafs         dense 2x2x2 block reduction per level;
//                                  ghost cells equal their parent cell
//   SceneRHS                       divergence theorem (sum of b = net wall flux)
//   operator (NegLapCoeff, LoadLapTile, NegLapSameLevel, FullNegativeLaplacian)
//                                  an independent leaf-partition finite-volume
//                                  assembly; symmetry and positivity; exactness
//                                  on linear fields across coarse/fine faces;
//                                  dense closed form on uniform grids; A 1 = 0
//                                  without Dirichlet cells
//   GSColor, JacobiSweep           textbook updates with the independent matrix
//   ResidualAndRestrict, ProlongateAndUpdate
//                                  R = P^T; a unit fine value lands in its parent cell
//   VCycle                         symmetric positive preconditioner
//   solve (T=double)               residual of the independent system, a sparse
//                                  direct solve on 32^3, the pure-Neumann form
//   Dot                            value vs an exactly rounded sum
//   ReCenterLeafVoxels             interior mean removed, other cells untouched
//   dampedJacobiSolve              closed-form first iterate, residual decreases
//   bf16_round                     every bfloat16 value and tie case, against
//                                  an exact rational rounding in Python
//   fp16_round, fp16_scale_exp     every float16 value and midpoint and random
//                                  patterns against numpy's float16
//                                  conversion; the scaled mean square's range;
//                                  the float16 solve is exactly equivariant
//                                  under power-of-two scalings of b
//   SolveRefined (no reference counterpart: the mixed-precision form)
//                                  its FP64 residual against the independent
//                                  assembly; the refined x against a direct
//                                  FP64 solve
// Parity unpinned: the exact FP32 bit patterns of the reference's own GPU run
// (they depend on its compiler's contraction choices and on cub's reduction
// order), hence its exact iteration counts; the hash-table slot order
// (hashCoord / tileIdx) is internal and reaches no output, since every export
// uses the canonical (level, Morton) tile order.
// ============================================================================
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace orc {

// The value a bfloat16 store of v leaves (round to nearest, ties to even; the
// top 16 bits of the FP32 pattern), as the GPU's conversion instruction gives
// for finite inputs.  The mixed-precision form stores the V-cycle vectors so.
inline float bf16_round(float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return v;  // inf / nan (never produced by the solver)
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

// The value a float16 store of v leaves: v clamped to the largest finite
// float16 (+-65504, so no store overflows), then rounded to 11 significant
// bits (10 below the 2^-14 subnormal threshold, spacing 2^-24), nearest,
// ties to even.  NaN passes through.
inline float fp16_round(float v) {
  if (v > 65504.f) v = 65504.f;
  if (v < -65504.f) v = -65504.f;
  if (!(v == v) || v == 0.f) return v;
  const double a = std::fabs(double(v));
  int e;
  std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
  const int q = std::max(e - 11, -24);  // exponent of the last kept bit
  const double r = std::nearbyint(std::ldexp(double(v), -q));  // exact scaling, round to nearest even
  return float(std::ldexp(r, q));
}

// Exponent k of the power-of-two scale 2^k that a float16-storage solve
// applies to its right-hand side (and undoes on x and r): the mean square of
// b over the problem's cells, bb / cells = 2^e f with f in [1, 2), gives
// k = -floor(e / 2) - 4, so the scaled root mean square lies near 2^-4 and the
// V-cycle's coarse sums stay far below the float16 maximum.  Scaling by a power
// of two is exact in FP32 and FP64, so only the float16 stores see it.
inline int fp16_scale_exp(double bb, double cells) {
  if (!(bb > 0.0)) return 0;
  const int e = std::ilogb(bb / cells);
  const int half = (e >= 0) ? e / 2 : -((-e + 1) / 2);
  return std::min(120, std::max(-120, -half - 4));
}

// HATileInfo.h:19 / PoissonTile.h:6
enum : uint8_t { LEAF = 1, NONLEAF = 2, GHOST = 4 };
enum : uint8_t { DIRICHLET = 1, INTERIOR = 2, NEUMANN = 4 };
enum : int { REFINE_FLAG = 1, DELETE_FLAG = 2, COARSEN_FLAG = 4 };  // PoissonTile.h:7
// PoissonTile.h:12-24 (node channels and vorticity/dye are not on the path)
enum : int { CH_X = 0, CH_B = 1, CH_R = 1, CH_P = 2, CH_AP = 3, CH_Z = 4, CH_TMP = 5,
             CH_U = 6, CH_V = 7, CH_W = 8, NCH = 9 };

struct Coord {
  int32_t v[3] = {0, 0, 0};
  Coord() = default;
  Coord(int a, int b, int c) { v[0] = a; v[1] = b; v[2] = c; }
  int& operator[](int i) { return v[i]; }
  int operator[](int i) const { return v[i]; }
  bool operator==(const Coord& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
  Coord operator+(const Coord& o) const { return Coord(v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]); }
};

// HAAccessor.h:32-42  (i-major local index i<<6 | j<<3 | k)
inline int localOffset(int i, int j, int k) { return (i << 6) + (j << 3) + k; }
inline int localOffset(const Coord& c) { return localOffset(c[0], c[1], c[2]); }
inline Coord localCoord(int off) { return Coord((off >> 6) & 7, (off >> 3) & 7, off & 7); }
// HAAccessor.h:50-62
inline Coord composeGlobal(const Coord& b, const Coord& l) {
  return Coord((b[0] << 3) + l[0], (b[1] << 3) + l[1], (b[2] << 3) + l[2]);
}
inline void decomposeGlobal(const Coord& g, Coord& b, Coord& l) {
  for (int i = 0; i < 3; i++) { b[i] = g[i] >> 3; l[i] = g[i] - (b[i] << 3); }
}
// HAAccessor.h:141-159
inline Coord childIndexToOffset(int idx) { return Coord(idx & 1, (idx >> 1) & 1, (idx >> 2) & 1); }
inline Coord childCoord(const Coord& c, const Coord& o) { return Coord(c[0] * 2 + o[0], c[1] * 2 + o[1], c[2] * 2 + o[2]); }
inline Coord parentCoord(const Coord& c) { return Coord(c[0] >> 1, c[1] >> 1, c[2] >> 1); }
// HAAccessor.h:162-168  (0..5 = x-, x+, y-, y+, z-, z+)
inline Coord neighborOffset(int nb) { Coord c; c[nb / 2] += (nb % 2 == 0) ? -1 : 1; return c; }

// HAAccessor.h:214-223 (InstantNGP-style XOR-prime hash)
inline uint32_t hashCoord(const Coord& c, uint32_t log2) {
  const uint32_t f[3] = {1u, 2654435761u, 805459861u};
  uint32_t r = 0;
  for (int i = 0; i < 3; i++) r ^= uint32_t(c[i]) * f[i];
  return r & ((1u << log2) - 1u);
}

inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t m = 0;
  for (int b = 0; b < 21; b++) {
    m |= uint64_t((x >> b) & 1u) << (3 * b);
    m |= uint64_t((y >> b) & 1u) << (3 * b + 1);
    m |= uint64_t((z >> b) & 1u) << (3 * b + 2);
  }
  return m;
}

// HATileInfo.h:63-98
struct Info {
  Coord coord;
  int32_t tile = -1;   // storage index, -1 = empty
  int32_t level = 0;
  uint8_t type = 0;
  bool empty() const { return tile < 0; }
  uint8_t subtreeType(int sl) const { return (level == sl && type == NONLEAF) ? uint8_t(LEAF) : type; }
  bool isLeaf(int sl = -1) const { return subtreeType(sl) & LEAF; }
  bool isGhost() const { return type & GHOST; }
  bool isActive() const { return type & (LEAF | NONLEAF); }
};

// PoissonTile.h:51-56 (cell channels only)
template <class T>
struct TileData {
  T ch[NCH][512];
  uint8_t ctype[512];
  Info nbr[6];  // x-, y-, z-, x+, y+, z+  (PoissonTile.h:52)
  int status = 0;
  TileData() { std::memset(ch, 0, sizeof(ch)); std::memset(ctype, 0, sizeof(ctype)); }
};

// ---------------------------------------------------------------------------
// Deterministic FP64 summation order shared with the GPU kernels:
//   per tile, lane l (0..63) owns cells l*8 .. l*8+7 and sums them in order,
//   then a xor-butterfly over masks 32,16,8,4,2,1; the per-tile partials of a
//   problem are summed with lane q taking partials q, q+64, ... in order,
//   followed by the same butterfly.
// ---------------------------------------------------------------------------
inline void butterfly64(double s[64]) {
  for (int m = 32; m >= 1; m >>= 1) {
    double t[64];
    for (int l = 0; l < 64; l++) t[l] = s[l] + s[l ^ m];
    std::memcpy(s, t, sizeof(t));
  }
}
template <class F>
inline double tile_partial(F cell_term) {
  double s[64];
  for (int l = 0; l < 64; l++) {
    double a = 0.0;
    for (int k = 0; k < 8; k++) a += cell_term(l * 8 + k);
    s[l] = a;
  }
  butterfly64(s);
  return s[0];
}
inline double sum_partials(const std::vector<double>& p) {
  double s[64];
  for (int l = 0; l < 64; l++) {
    double a = 0.0;
    for (size_t q = l; q < p.size(); q += 64) a += p[q];
    s[l] = a;
  }
  butterfly64(s);
  return s[0];
}

// ---------------------------------------------------------------------------
// Workload definition (the synthetic scenes of the benchmark configurations).
// ---------------------------------------------------------------------------
enum SceneKind : int { SCENE_UNIFORM = 0, SCENE_SPHERE_BAND = 1, SCENE_PLANE = 2, SCENE_VORTEX = 3 };

struct SceneParams {
  int kind = SCENE_UNIFORM;
  int nroot[3] = {1, 1, 1};          // root tiles per axis at level 0
  double h0 = 0.125;                 // level-0 cell size (FluidEuler.h:415)
  int uniform_level = 3;             // SCENE_UNIFORM target level
  int has_solid = 1;                 // solid sphere cells are NEUMANN
  double solid_c[3] = {0.5, 0.5, 0.5};
  double solid_r = 0.2;
  // SCENE_SPHERE_BAND: level band_level[0] where the tile box meets |d|<band_w[0],
  // band_level[1] where it meets |d|<band_w[1], band_level[2] elsewhere,
  // d = distance to the sphere (band_c, band_r) surface.
  double band_c[3] = {0.5, 0.5, 0.5};
  double band_r = 0.15;
  double band_w[2] = {0.05, 0.15};
  int band_level[3] = {5, 4, 3};
  // SCENE_PLANE: plane_level[0] where the tile box starts below z=plane_z, else plane_level[1]
  double plane_z = 0.5;
  int plane_level[2] = {3, 2};
  // SCENE_VORTEX: sum of vortex tubes; leaf level vortex_level[0] where the tile
  // max |omega| (sampled at tile corners and centre) exceeds vortex_thresh,
  // vortex_level[1] elsewhere.
  int n_tubes = 8;
  double tube_r_lo = 0.02, tube_r_hi = 0.06;
  double vortex_thresh = 0.5;
  int vortex_level[2] = {6, 4};
  // domain-face boundary type, order x-, x+, y-, y+, z-, z+ : NEUMANN or DIRICHLET
  int face_bc[6] = {NEUMANN, NEUMANN, NEUMANN, NEUMANN, NEUMANN, DIRICHLET};
  uint64_t seed = 0;
  int num_layers = 10;               // FluidEuler.h:416
};

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// iid N(0,1) keyed by (seed, level, global ijk, axis), Box-Muller on two
// 53-bit uniforms.  Host code only (libm), shared bit-for-bit with the product.
inline double normal_at(uint64_t seed, int level, int gi, int gj, int gk, int axis) {
  uint64_t h = mix64(seed);
  h = mix64(h ^ uint64_t(uint32_t(level)));
  h = mix64(h ^ uint64_t(uint32_t(gi)));
  h = mix64(h ^ uint64_t(uint32_t(gj)));
  h = mix64(h ^ uint64_t(uint32_t(gk)));
  h = mix64(h ^ uint64_t(uint32_t(axis)));
  uint64_t h2 = mix64(h);
  double u1 = double((h >> 11) + 1) * (1.0 / 9007199254740992.0);
  double u2 = double(h2 >> 11) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Tube { double c[3], a[3], r; };
inline std::vector<Tube> make_tubes(const SceneParams& p) {
  std::vector<Tube> t;
  uint64_t s = mix64(p.seed ^ 0x7475626573ull);
  auto uni = [&]() { s = mix64(s); return double(s >> 11) * (1.0 / 9007199254740992.0); };
  double ext[3] = {p.nroot[0] * 8 * p.h0, p.nroot[1] * 8 * p.h0, p.nroot[2] * 8 * p.h0};
  for (int i = 0; i < p.n_tubes; i++) {
    Tube tb;
    for (int d = 0; d < 3; d++) tb.c[d] = ext[d] * (0.15 + 0.7 * uni());
    double z = 2.0 * uni() - 1.0, ph = 6.283185307179586 * uni();
    double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
    tb.a[0] = rr * std::cos(ph); tb.a[1] = rr * std::sin(ph); tb.a[2] = z;
    tb.r = p.tube_r_lo + (p.tube_r_hi - p.tube_r_lo) * uni();
    t.push_back(tb);
  }
  return t;
}
// |omega| of a sum of Gaussian-core vortex tubes (unit circulation density).
inline double vorticity_mag(const std::vector<Tube>& tubes, const double x[3]) {
  double w[3] = {0, 0, 0};
  for (const auto& t : tubes) {
    double d[3] = {x[0] - t.c[0], x[1] - t.c[1], x[2] - t.c[2]};
    double along = d[0] * t.a[0] + d[1] * t.a[1] + d[2] * t.a[2];
    double q[3] = {d[0] - along * t.a[0], d[1] - along * t.a[1], d[2] - along * t.a[2]};
    double r2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    double m = std::exp(-r2 / (t.r * t.r));
    for (int k = 0; k < 3; k++) w[k] += m * t.a[k];
  }
  return std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
}

// Target level of a tile (level, coord) in the scene; the refinement loop
// refines a LEAF while target > level (PoissonGrid.h:63-65).
inline int scene_target(const SceneParams& p, const std::vector<Tube>& tubes, int level, const Coord& b) {
  double h = p.h0 / double(1 << level);
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) { lo[d] = b[d] * 8 * h; hi[d] = (b[d] + 1) * 8 * h; }
  switch (p.kind) {
    case SCENE_UNIFORM: return p.uniform_level;
    case SCENE_SPHERE_BAND: {
      double dmin2 = 0, dmax2 = 0;
      for (int d = 0; d < 3; d++) {
        double c = p.band_c[d];
        double n = std::max(lo[d] - c, std::max(0.0, c - hi[d]));
        double f = std::max(std::fabs(lo[d] - c), std::fabs(hi[d] - c));
        dmin2 += n * n; dmax2 += f * f;
      }
      double dmin = std::sqrt(dmin2), dmax = std::sqrt(dmax2);
      for (int w = 0; w < 2; w++)
        if (dmin < p.band_r + p.band_w[w] && dmax > p.band_r - p.band_w[w]) return p.band_level[w];
      return p.band_level[2];
    }
    case SCENE_PLANE: return (lo[2] < p.plane_z) ? p.plane_level[0] : p.plane_level[1];
    case SCENE_VORTEX: {
      double m = 0;
      for (int s = 0; s < 27; s++) {
        double x[3] = {lo[0] + (hi[0] - lo[0]) * 0.5 * (s % 3), lo[1] + (hi[1] - lo[1]) * 0.5 * ((s / 3) % 3),
                       lo[2] + (hi[2] - lo[2]) * 0.5 * (s / 9)};
        m = std::max(m, vorticity_mag(tubes, x));
      }
      return (m > p.vortex_thresh) ? p.vortex_level[0] : p.vortex_level[1];
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// The grid: HADeviceGrid (HAGrid.h:205-745) with host and "device" views in
// one address space.
// ---------------------------------------------------------------------------
template <class T>
struct Grid {
  T h0;
  int numLayers;
  int maxLevel = -1;
  std::vector<uint32_t> log2;
  std::vector<std::vector<Info>> hashT;   // hHashTables
  std::vector<std::vector<Info>> arrays;  // hTileArrays (compressed, hash-slot order)
  std::vector<Info> all;                  // dAllTiles (levels 0..N concatenated)
  std::vector<std::unique_ptr<TileData<T>>> store;
  std::vector<size_t> count;              // occupied slots per level

  Grid(T h, const std::vector<uint32_t>& l2) : h0(h), numLayers(int(l2.size())), log2(l2) {
    count.assign(numLayers, 0);
    hashT.resize(numLayers);
    arrays.resize(numLayers);
    for (int i = 0; i < numLayers; i++) hashT[i].assign(size_t(1) << log2[i], Info());
  }
  TileData<T>& tile(const Info& i) { return *store[i.tile]; }
  // HAAccessor.h:72-74
  T voxelSize(int level) const { return h0 / T(1 << level); }

  // HAAccessor.h:225-242 (linear probing)
  uint32_t tileIdx(const std::vector<std::vector<Info>>& tab, uint32_t level, const Coord& b) const {
    if (level >= uint32_t(numLayers)) return uint32_t(-1);
    uint32_t l2 = log2[level], mask = (1u << l2) - 1u;
    uint32_t k0 = hashCoord(b, l2), k1 = k0;
    do {
      const Info& in = tab[level][k1];
      if (in.empty() || in.coord == b) return k1;
      k1 = (k1 + 1) & mask;
    } while (k1 != k0);
    throw std::runtime_error("hash table full");
  }
  // HAAccessor.h:259-267 (checked copy)
  Info tileInfoCopy(const std::vector<std::vector<Info>>& tab, int level, const Coord& b) const {
    if (level < 0 || level >= numLayers) return Info();
    return tab[level][tileIdx(tab, uint32_t(level), b)];
  }
  Info tileInfo(int level, const Coord& b) const { return tileInfoCopy(hashT, level, b); }
  // HAAccessor.h:274-283
  bool findVoxel(int level, const Coord& g, Info& info, Coord& l) const {
    if (level < 0 || level >= numLayers) { info = Info(); return false; }
    Coord b;
    decomposeGlobal(g, b, l);
    info = tileInfoCopy(hashT, level, b);
    return !info.empty();
  }

  // HAGrid.h:387-407 (one allocation per new tile; data copied in).  A level's
  // table doubles when it would become more than half full; the reference's
  // fixed sizes (FluidEuler.h:416) are kept whenever they suffice.
  void setTileHost(int level, const Coord& b, const TileData<T>* data, uint8_t type) {
    uint32_t idx = tileIdx(hashT, uint32_t(level), b);
    if (hashT[level][idx].empty()) {
      if (2 * (count[level] + 1) > hashT[level].size()) { grow(level); idx = tileIdx(hashT, uint32_t(level), b); }
      store.emplace_back(new TileData<T>());
      hashT[level][idx].tile = int32_t(store.size() - 1);
      count[level]++;
    }
    Info& in = hashT[level][idx];
    if (data) *store[in.tile] = *data;
    else *store[in.tile] = TileData<T>();
    in.coord = b; in.level = level; in.type = type;
  }
  // HAGrid.h:409-431: empty the slot, then walk the probe run after it and
  // move back every entry whose home slot does not lie cyclically in
  // (hole, entry] (backward-shift deletion for linear probing), so that
  // later lookups never stop early at the hole.
  void removeTileHost(int level, const Coord& b) {
    uint32_t i = tileIdx(hashT, uint32_t(level), b);
    if (hashT[level][i].empty()) return;
    const uint32_t mask = (1u << log2[level]) - 1u;
    uint32_t j = i;
    while (true) {
      j = (j + 1) & mask;
      if (hashT[level][j].empty()) break;
      uint32_t home = hashCoord(hashT[level][j].coord, log2[level]);
      if ((j > i && (home <= i || home > j)) || (j < i && (home <= i && home > j))) {
        std::swap(hashT[level][i], hashT[level][j]);
        i = j;
      }
    }
    store[hashT[level][i].tile].reset();
    hashT[level][i] = Info();
    count[level]--;
  }
  void grow(int level) {
    std::vector<Info> old = hashT[level];
    log2[level]++;
    hashT[level].assign(size_t(1) << log2[level], Info());
    for (const Info& in : old) if (!in.empty()) hashT[level][tileIdx(hashT, uint32_t(level), in.coord)] = in;
  }
  // HAGrid.h:491-509
  void compressHost() {
    maxLevel = -1;
    for (int i = 0; i < numLayers; i++) {
      arrays[i].clear();
      for (const Info& in : hashT[i]) if (!in.empty()) arrays[i].push_back(in);
      if (!arrays[i].empty()) maxLevel = i;
    }
  }
  // HAGrid.h:510-561 (dAllTiles = concatenation of levels)
  void syncHostAndDevice() {
    all.clear();
    for (int i = 0; i < numLayers; i++) for (const Info& in : arrays[i]) all.push_back(in);
  }
};

// PoissonTile.h:170-190
template <class T>
TileData<T> childTile(const TileData<T>& p, const Coord& off) {
  TileData<T> c;
  c.status = p.status;
  for (int i = 0; i < 8; i++) for (int j = 0; j < 8; j++) for (int k = 0; k < 8; k++) {
    Coord g(off[0] * 8 + i, off[1] * 8 + j, off[2] * 8 + k);
    int pl = localOffset(parentCoord(g)), cl = localOffset(i, j, k);
    for (int ch = 0; ch < NCH; ch++) c.ch[ch][cl] = p.ch[ch][pl];
    c.ctype[cl] = p.ctype[pl];
  }
  return c;
}

// PoissonGrid.cu:30-60
template <class T>
void SpawnGhostTiles(Grid<T>& g) {
  for (int level = g.maxLevel; level > 0; level--) {
    for (size_t i = 0; i < g.arrays[level].size(); i++) {
      const Info info = g.arrays[level][i];
      if (!info.isActive()) continue;
      for (int nb = 0; nb < 6; nb++) {
        Coord n = info.coord + neighborOffset(nb);
        if (g.tileInfo(level, n).empty()) {
          if (g.tileInfo(level - 1, parentCoord(n)).isLeaf()) g.setTileHost(level, n, nullptr, GHOST);
        }
      }
    }
  }
  g.compressHost();
  g.syncHostAndDevice();
}

// PoissonGrid.h:35-162  (device kernels see the tables as they were at the
// start of the step; host edits go to the live hash tables)
template <class T, class F>
int RefineLeafsOneStep(Grid<T>& g, F target) {
  auto devHash = g.hashT;
  auto devArrays = g.arrays;
  int devMax = g.maxLevel;
  int total = 0;
  for (int i = devMax; i >= 0; i--) {
    std::vector<int> flag(devArrays[i].size(), 0);
    for (size_t j = 0; j < devArrays[i].size(); j++) {
      const Info& info = devArrays[i][j];
      if (!(info.type & LEAF)) continue;
      bool to_refine = target(i, info.coord) > info.level;
      if (i < devMax) {
        for (int c = 0; c < 8; c++) {
          Info ci = g.tileInfoCopy(devHash, i + 1, childCoord(info.coord, childIndexToOffset(c)));
          if (ci.isGhost() && (g.tile(ci).status & REFINE_FLAG)) to_refine = true;
        }
      }
      if (i + 1 >= g.numLayers) to_refine = false;
      auto& t = g.tile(info);
      t.status = to_refine ? (t.status | REFINE_FLAG) : (t.status & ~REFINE_FLAG);
      flag[j] = to_refine;
    }
    for (size_t j = 0; j < devArrays[i].size(); j++) {
      const Info& info = devArrays[i][j];
      if (!(info.type & GHOST)) continue;
      bool rf = false;
      for (int nb = 0; nb < 6; nb++) {
        Info ni = g.tileInfoCopy(devHash, i, info.coord + neighborOffset(nb));
        if (ni.isLeaf() && (g.tile(ni).status & REFINE_FLAG)) rf = true;
      }
      auto& t = g.tile(info);
      t.status = rf ? (t.status | REFINE_FLAG) : (t.status & ~REFINE_FLAG);
    }
    for (size_t j = 0; j < flag.size(); j++) {
      if (!flag[j]) continue;
      total++;
      Coord b = devArrays[i][j].coord;
      Info& hin = g.hashT[i][g.tileIdx(g.hashT, uint32_t(i), b)];
      hin.type = NONLEAF;
      TileData<T> parent = g.tile(hin);
      for (int c = 0; c < 8; c++) {
        Coord off = childIndexToOffset(c);
        TileData<T> ct = childTile(parent, off);
        g.setTileHost(i + 1, childCoord(b, off), &ct, LEAF);
      }
    }
  }
  g.compressHost();
  g.syncHostAndDevice();
  return total;
}

// PoissonGrid.h:317-326
template <class T, class F>
void IterativeRefine(Grid<T>& g, F target) {
  while (true) {
    int cnt = RefineLeafsOneStep(g, target);
    SpawnGhostTiles(g);
    if (cnt == 0) break;
  }
}

// PoissonGrid.h:165-314.  Flags live in the tiles' status bits; the "kernel"
// launches of one level read flags that earlier launches wrote and write only
// their own tiles', so the loops below give the launches' results.  Returns
// the number of deleted tiles per level.
template <class T, class F>
std::vector<int> CoarsenStep(Grid<T>& g, F target) {
  auto set = [&](const Info& in, int mask, bool v) {
    int& s = g.tile(in).status;
    s = v ? (s | mask) : (s & ~mask);
  };
  auto status = [&](const Info& in) { return g.tile(in).status; };
  // upstroke, finest level first
  for (int level = g.numLayers - 1; level >= 0; level--) {
    // 1: a NONLEAF tile may coarsen when its 8 children are LEAF tiles flagged DELETE
    for (const Info& info : g.arrays[level]) {
      if (!(info.type & NONLEAF)) continue;
      bool to_coarsen = true;
      for (int c = 0; c < 8; c++) {
        Info ci = g.tileInfo(level + 1, childCoord(info.coord, childIndexToOffset(c)));
        if (ci.isLeaf()) {
          if (!(status(ci) & DELETE_FLAG)) to_coarsen = false;
        } else {
          to_coarsen = false;
        }
      }
      set(info, COARSEN_FLAG, to_coarsen);
      set(info, DELETE_FLAG, false);
    }
    // 2: a LEAF tile above its target level may go unless a same-level NONLEAF
    // neighbour stays refined
    for (const Info& info : g.arrays[level]) {
      if (!(info.type & LEAF)) continue;
      bool to_delete = target(level, info.coord) < info.level;
      for (int nb = 0; nb < 6; nb++) {
        Info ni = g.tileInfo(level, info.coord + neighborOffset(nb));
        if ((ni.type & NONLEAF) && !(status(ni) & COARSEN_FLAG)) to_delete = false;
      }
      set(info, COARSEN_FLAG, false);
      set(info, DELETE_FLAG, to_delete);
    }
  }
  // downstroke, coarsest level first
  for (int level = 0; level < g.numLayers; level++) {
    // 3: LEAF and GHOST tiles go with a coarsening parent (and pass it on)
    for (const Info& info : g.arrays[level]) {
      if (!(info.type & (LEAF | GHOST))) continue;
      Info pi = g.tileInfo(level - 1, parentCoord(info.coord));
      bool to_delete = !pi.empty() && (status(pi) & COARSEN_FLAG);
      set(info, DELETE_FLAG, to_delete);
      if (to_delete) set(info, COARSEN_FLAG, true);
    }
    // 4: a GHOST tile without a surviving same-level LEAF neighbour goes
    for (const Info& info : g.arrays[level]) {
      if (!(info.type & GHOST)) continue;
      bool to_delete = true;
      for (int nb = 0; nb < 6; nb++) {
        Info ni = g.tileInfo(level, info.coord + neighborOffset(nb));
        if ((ni.type & LEAF) && !(status(ni) & DELETE_FLAG)) to_delete = false;
      }
      if (to_delete) set(info, DELETE_FLAG, true);
    }
  }
  // remove DELETE tiles; COARSEN tiles that stay become LEAF tiles
  std::vector<int> deleted(g.numLayers, 0);
  for (int level = 0; level < g.numLayers; level++) {
    const std::vector<Info> arr = g.arrays[level];
    for (const Info& info : arr) {
      const int st = status(info);
      if (st & DELETE_FLAG) {
        g.removeTileHost(level, info.coord);
        deleted[level]++;
      } else if (st & COARSEN_FLAG) {
        g.hashT[level][g.tileIdx(g.hashT, uint32_t(level), info.coord)].type = LEAF;
      }
    }
  }
  g.compressHost();
  g.syncHostAndDevice();
  return deleted;
}

// PoissonGrid.cu:121-134
template <class T>
void CalculateNeighborTiles(Grid<T>& g) {
  for (const Info& info : g.all) {
    auto& t = g.tile(info);
    for (int axis = 0; axis < 3; axis++) {
      Coord nb = info.coord, np = info.coord;
      nb[axis]--; np[axis]++;
      t.nbr[axis] = g.tileInfo(info.level, nb);
      t.nbr[axis + 3] = g.tileInfo(info.level, np);
    }
  }
}

// Boundary and solid typing of LEAF cells, after FluidParams.cu:125-144
// (an empty same-level neighbour voxel marks a boundary cell; NEUMANN wins)
// and FluidParams.cu:241-244 (cells whose centre lies in the solid sphere).
template <class T>
uint8_t SceneCellType(const Grid<T>& g, const SceneParams& p, const Info& info, int vi) {
  Coord l = localCoord(vi), gc = composeGlobal(info.coord, l);
  bool neu = false, dir = false;
  for (int axis = 0; axis < 3; axis++)
    for (int sgn : {-1, 1}) {
      Coord n = gc; n[axis] += sgn;
      Info ni; Coord nl;
      g.findVoxel(info.level, n, ni, nl);
      if (ni.empty()) {
        int face = axis * 2 + (sgn > 0 ? 1 : 0);
        if (p.face_bc[face] == DIRICHLET) dir = true; else neu = true;
      }
    }
  if (neu) return NEUMANN;
  if (dir) return DIRICHLET;
  if (p.has_solid) {
    double h = p.h0 / double(1 << info.level), d2 = 0;
    for (int d = 0; d < 3; d++) { double x = (gc[d] + 0.5) * h - p.solid_c[d]; d2 += x * x; }
    if (d2 <= p.solid_r * p.solid_r) return NEUMANN;
  }
  return INTERIOR;
}

// PoissonSolver.cu:7-61
template <class T>
void CalcCellTypesFromLeafs(Grid<T>& g) {
  for (int level = g.numLayers - 1; level >= 0; level--) {
    const auto& arr = g.arrays[level];
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info info = arr[ti];
      if (info.type != NONLEAF) continue;
      auto& t = g.tile(info);
      for (int vi = 0; vi < 512; vi++) {
        Coord gc = composeGlobal(info.coord, localCoord(vi));
        int nd = 0, ni = 0;
        for (int s = 0; s < 8; s++) {
          Info ci; Coord cl;
          g.findVoxel(level + 1, childCoord(gc, childIndexToOffset(s)), ci, cl);
          if (!ci.empty() && ci.isActive()) {
            uint8_t typ = g.tile(ci).ctype[localOffset(cl)];
            if (typ & DIRICHLET) nd++;
            else if (typ & INTERIOR) ni++;
          }
        }
        t.ctype[vi] = nd > 0 ? DIRICHLET : (ni > 0 ? INTERIOR : NEUMANN);
      }
    }
  }
  const auto& all = g.all;
#pragma omp parallel for schedule(static)
  for (long ti = 0; ti < long(all.size()); ti++) {
    const Info info = all[ti];
    if (!(info.type & GHOST)) continue;
    auto& t = g.tile(info);
    for (int vi = 0; vi < 512; vi++) {
      Coord gc = composeGlobal(info.coord, localCoord(vi));
      Info pi; Coord pl;
      g.findVoxel(info.level - 1, parentCoord(gc), pi, pl);
      t.ctype[vi] = pi.empty() ? DIRICHLET : g.tile(pi).ctype[localOffset(pl)];
    }
  }
}

// The grid, its topology and its FP state for one problem.
template <class T>
struct Problem {
  SceneParams params;
  std::unique_ptr<Grid<T>> grid;
  std::vector<Info> canon;          // all tiles in (level, Morton) order
  std::vector<int> canonOfStore;    // store index -> canonical index
  std::vector<Info> leafCanon;      // LEAF tiles in canonical order
  std::vector<Tube> tubes;
  // AMGSolver's stored coefficients per tile storage index: [0..2] the cell's
  // minus-face off-diagonal along x, y, z, [3] the diagonal (AMGSolver.cu:38-39)
  std::vector<std::array<T, 512>> amg[4];

  Grid<T>& g() { return *grid; }

  void build_canon() {
    Grid<T>& G = g();
    canon = G.all;
    std::sort(canon.begin(), canon.end(), [](const Info& a, const Info& b) {
      if (a.level != b.level) return a.level < b.level;
      return morton3(uint32_t(a.coord[0]), uint32_t(a.coord[1]), uint32_t(a.coord[2])) <
             morton3(uint32_t(b.coord[0]), uint32_t(b.coord[1]), uint32_t(b.coord[2]));
    });
    canonOfStore.assign(G.store.size(), -1);
    for (size_t i = 0; i < canon.size(); i++) canonOfStore[canon[i].tile] = int(i);
    leafCanon.clear();
    for (const Info& in : canon) if (in.type & LEAF) leafCanon.push_back(in);
  }
};

// The scene's octree: FluidEuler.h:415-447 (roots, compress, ghosts), then
// IterativeRefine (PoissonGrid.h:317) to the scene's target levels.
template <class T>
std::unique_ptr<Problem<T>> BuildSceneTree(const SceneParams& p) {
  auto P = std::make_unique<Problem<T>>();
  P->params = p;
  P->tubes = make_tubes(p);
  std::vector<uint32_t> l2(p.num_layers, 16);
  if (p.num_layers > 6) l2[6] = 18;  // FluidEuler.h:416
  P->grid = std::make_unique<Grid<T>>(T(p.h0), l2);
  Grid<T>& G = P->g();
  for (int i = 0; i < p.nroot[0]; i++)
    for (int j = 0; j < p.nroot[1]; j++)
      for (int k = 0; k < p.nroot[2]; k++) G.setTileHost(0, Coord(i, j, k), nullptr, LEAF);
  G.compressHost();
  G.syncHostAndDevice();
  SpawnGhostTiles(G);
  const auto& tubes = P->tubes;
  IterativeRefine(G, [&](int level, const Coord& b) { return scene_target(p, tubes, level, b); });
  return P;
}

// Cell types of the scene on the tree (LEAF cells, then
// CalcCellTypesFromLeafs), neighbour tables and the canonical order.
template <class T>
void FinishScene(Problem<T>& P) {
  Grid<T>& G = P.g();
  const SceneParams& p = P.params;
  for (const Info& info : G.all) {
    if (!(info.type & LEAF)) continue;
    auto& t = G.tile(info);
    for (int vi = 0; vi < 512; vi++) t.ctype[vi] = SceneCellType(G, p, info, vi);
  }
  CalcCellTypesFromLeafs(G);
  CalculateNeighborTiles(G);
  P.build_canon();
}

template <class T>
std::unique_ptr<Problem<T>> BuildScene(const SceneParams& p) {
  auto P = BuildSceneTree<T>(p);
  FinishScene(*P);
  return P;
}

// Adaptation of the tree of scene `from` to the target levels of scene `to`
// (same domain, root tiles and layers), as a simulation step does it
// (FluidEuler.h:820-835): refinement steps with ghost spawning until none
// refines (RefineWithPointValue, FluidEuler.h:68-88), then CoarsenStep until
// none deletes and a final SpawnGhostTiles (CoarsenWithPointValue,
// FluidEuler.h:90-111).  ops: bit 0 refine, bit 1 coarsen.  The result carries
// scene `to`'s cell types.
enum : int { ADAPT_REFINE = 1, ADAPT_COARSEN = 2 };
template <class T>
std::unique_ptr<Problem<T>> AdaptScene(const SceneParams& from, const SceneParams& to, int ops) {
  for (int d = 0; d < 3; d++)
    if (from.nroot[d] != to.nroot[d]) throw std::runtime_error("adapt: root tiles differ");
  if (from.h0 != to.h0 || from.num_layers != to.num_layers) throw std::runtime_error("adapt: domain differs");
  auto P = BuildSceneTree<T>(from);
  P->params = to;
  P->tubes = make_tubes(to);
  Grid<T>& G = P->g();
  const auto& tubes = P->tubes;
  auto target = [&](int level, const Coord& b) { return scene_target(to, tubes, level, b); };
  if (ops & ADAPT_REFINE) IterativeRefine(G, target);
  if (ops & ADAPT_COARSEN) {
    while (true) {
      auto cnt = CoarsenStep(G, target);
      int total = 0;
      for (int c : cnt) total += c;
      if (total == 0) break;
    }
    SpawnGhostTiles(G);
  }
  FinishScene(*P);
  return P;
}

// IterateFaceNeighborCellTypes (PoissonGrid.h:451-489)
template <class T, class F>
void IterateFaceNeighborCellTypes(Grid<T>& g, const Info& info, const Coord& l, int axis, F f) {
  uint8_t type0 = g.tile(info).ctype[localOffset(l)];
  Coord gc = composeGlobal(info.coord, l), ng = gc;
  ng[axis]--;
  Info ni; Coord nl;
  g.findVoxel(info.level, ng, ni, nl);
  if (ni.empty()) return;
  if (!ni.isLeaf()) {
    for (int oj : {0, 1})
      for (int ok : {0, 1}) {
        Coord o; o[axis] = 1; o[(axis + 1) % 3] = oj; o[(axis + 2) % 3] = ok;
        Info ci; Coord cl;
        g.findVoxel(info.level + 1, childCoord(ng, o), ci, cl);
        if (!ci.empty()) f(type0, g.tile(ci).ctype[localOffset(cl)]);
      }
  } else {
    if (ni.isGhost()) g.findVoxel(info.level - 1, parentCoord(ng), ni, nl);
    f(type0, g.tile(ni).ctype[localOffset(nl)]);
  }
}

// The seeded face-velocity field of the scene and the divergence RHS:
// FluidEuler.cu:43-60 (ClearAllNeumannNeighborFaces), FluidEuler.h:583-586
// (PropagateToChildren of u,v,w to ghosts, then VelocityVolumeDivergenceOnLeafs,
// PoissonSolver.cu:66-142, with AccumulateToParents at :141).
template <class T>
void SceneRHS(Problem<T>& P) {
  Grid<T>& G = P.g();
  const SceneParams& p = P.params;
  const auto& all = G.all;
#pragma omp parallel for schedule(static)
  for (long ti = 0; ti < long(all.size()); ti++) {
    const Info info = all[ti];
    if (!(info.type & LEAF)) continue;
    auto& t = G.tile(info);
    for (int vi = 0; vi < 512; vi++) {
      Coord gc = composeGlobal(info.coord, localCoord(vi));
      for (int a = 0; a < 3; a++) t.ch[CH_U + a][vi] = T(float(normal_at(p.seed, info.level, gc[0], gc[1], gc[2], a)));
    }
  }
  // zero faces touching NEUMANN, or between two DIRICHLET cells
#pragma omp parallel for schedule(static)
  for (long ti = 0; ti < long(all.size()); ti++) {
    const Info info = all[ti];
    if (!(info.type & LEAF)) continue;
    auto& t = G.tile(info);
    for (int vi = 0; vi < 512; vi++) {
      Coord l = localCoord(vi);
      for (int axis = 0; axis < 3; axis++) {
        bool to_set = false;
        IterateFaceNeighborCellTypes(G, info, l, axis, [&](uint8_t a, uint8_t b) {
          if ((a & NEUMANN) || (b & NEUMANN) || ((a & DIRICHLET) && (b & DIRICHLET))) to_set = true;
        });
        if (to_set) t.ch[CH_U + axis][vi] = T(0);
      }
    }
  }
  PropagateVelocityToGhosts(G, CH_U);
  VelocityVolumeDivergenceOnLeafs(G, CH_U, CH_B);
}

// FluidEuler.h:582-585: PropagateToChildren(u+a, u+a, -1, GHOST, SUBTREE,
// INTERIOR|DIRICHLET|NEUMANN) (PoissonGrid.cu:635-661), coarse levels first.
template <class T>
void PropagateVelocityToGhosts(Grid<T>& G, int u_ch) {
  const uint8_t ctypes = INTERIOR | DIRICHLET | NEUMANN;
  for (int axis = 0; axis < 3; axis++)
    for (int level = 0; level < G.numLayers; level++) {
      const auto& arr = G.arrays[level];
#pragma omp parallel for schedule(static)
      for (long ti = 0; ti < long(arr.size()); ti++) {
        const Info info = arr[ti];
        if (!(info.type & GHOST)) continue;
        auto& t = G.tile(info);
        for (int vi = 0; vi < 512; vi++) {
          if (!(t.ctype[vi] & ctypes)) continue;
          Coord gc = composeGlobal(info.coord, localCoord(vi));
          Info ci; Coord cl;
          G.findVoxel(level - 1, parentCoord(gc), ci, cl);
          T val = T(0);
          if (!ci.empty()) {
            auto& ct = G.tile(ci);
            val = ct.ch[u_ch + axis][localOffset(cl)];
            if (!(ct.ctype[localOffset(cl)] & ctypes)) val = T(0);
          }
          t.ch[u_ch + axis][vi] = val;
        }
      }
    }
}

// AccumulateToParents(fine, coarse, -1, LEAF, SUBTREE, cell_types, coeff, additive)
// (PoissonGrid.cu:721-755): on LEAF tiles, each cell whose type is in
// cell_types gets coeff times the sum of its 8 children (x-fastest child
// order) whose type is in cell_types; finest levels first.
template <class T>
void AccumulateLeafChildren(Grid<T>& G, int ch, uint8_t cell_types, T coeff) {
  for (int level = G.numLayers - 1; level >= 0; level--) {
    const auto& arr = G.arrays[level];
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info info = arr[ti];
      if (!(info.type & LEAF)) continue;
      auto& t = G.tile(info);
      for (int vi = 0; vi < 512; vi++) {
        if (!(t.ctype[vi] & cell_types)) continue;
        Coord gc = composeGlobal(info.coord, localCoord(vi));
        T sum = T(0);
        for (int c = 0; c < 8; c++) {
          Info fi; Coord fl;
          G.findVoxel(level + 1, childCoord(gc, childIndexToOffset(c)), fi, fl);
          if (!fi.empty()) {
            auto& ft = G.tile(fi);
            if (ft.ctype[localOffset(fl)] & cell_types) sum += ft.ch[ch][localOffset(fl)];
          }
        }
        t.ch[ch][vi] += sum * coeff;
      }
    }
  }
}

// VelocityVolumeDivergenceOnLeafs (PoissonSolver.cu:66-142): on LEAF|GHOST
// tiles, INTERIOR cells sum their face fluxes u h^2 over LEAF-LEAF and
// LEAF-GHOST faces (faces x-, x+, y-, y+, z-, z+; the flux is -u of the cell's
// own minus face, +u of the plus neighbour's), other cells 0; then the ghost
// cells' sums are added to their LEAF parents (INTERIOR cells, weight 1).
template <class T>
void VelocityVolumeDivergenceOnLeafs(Grid<T>& G, int u_ch, int div_ch) {
  const auto& all = G.all;
#pragma omp parallel for schedule(static)
  for (long ti = 0; ti < long(all.size()); ti++) {
    const Info info = all[ti];
    if (!(info.type & (LEAF | GHOST))) continue;
    auto& t = G.tile(info);
    T h = G.voxelSize(info.level);
    uint8_t ttype = info.subtreeType(-1);
    for (int vi = 0; vi < 512; vi++) {
      if (!(t.ctype[vi] & INTERIOR)) { t.ch[div_ch][vi] = T(0); continue; }
      Coord l = localCoord(vi), gc = composeGlobal(info.coord, l);
      T sum = T(0);
      for (int axis = 0; axis < 3; axis++)
        for (int sgn : {-1, 1}) {
          Coord n = gc; n[axis] += sgn;
          Info ni; Coord nl;
          G.findVoxel(info.level, n, ni, nl);
          uint8_t nt;
          T u0 = t.ch[u_ch + axis][vi], u1;
          if (ni.empty()) { nt = ttype; u1 = T(0); }
          else { nt = ni.subtreeType(-1); u1 = G.tile(ni).ch[u_ch + axis][localOffset(nl)]; }
          if (!((ttype == LEAF && nt == LEAF) || (ttype == LEAF && nt == GHOST) || (ttype == GHOST && nt == LEAF))) continue;
          T u = (sgn == -1) ? -u0 : u1;
          sum = std::fma(u * h, h, sum);
        }
      t.ch[div_ch][vi] = sum;
    }
  }
  AccumulateLeafChildren(G, div_ch, INTERIOR, T(1));
}

// AddGradientToFaceCenters (PoissonSolver.cu:144-209): ghost p from the
// parents' INTERIOR values (PropagateValues, PoissonGrid.cu:614-634); on
// LEAF|GHOST tiles (ghost velocities first set to 0) each cell's minus face
// of each axis gets u += (p1 - p0) / dh, p1 / p0 the INTERIOR values (0
// elsewhere) of the cell and its minus neighbour (0 outside the grid), dh = h
// across LEAF-LEAF faces and 1.5 h across LEAF-GHOST faces, nothing across
// other faces or faces touching a NEUMANN cell; then the ghost velocities are
// added to their LEAF parents with weight 1/4 (INTERIOR|DIRICHLET cells).
template <class T>
void AddGradientToFaceCenters(Grid<T>& G, int p_ch, int u_ch) {
  const auto& all = G.all;
  // PropagateValues(p, p, -1, GHOST, SUBTREE): coarse levels first
  for (int level = 0; level < G.numLayers; level++) {
    const auto& arr = G.arrays[level];
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info info = arr[ti];
      if (!(info.type & GHOST)) continue;
      auto& t = G.tile(info);
      for (int vi = 0; vi < 512; vi++) {
        if (!(t.ctype[vi] & INTERIOR)) { t.ch[p_ch][vi] = T(0); continue; }
        Coord gc = composeGlobal(info.coord, localCoord(vi));
        Info ci; Coord cl;
        G.findVoxel(level - 1, parentCoord(gc), ci, cl);
        T v = T(0);
        if (!ci.empty()) {
          auto& ct = G.tile(ci);
          v = (ct.ctype[localOffset(cl)] & INTERIOR) ? ct.ch[p_ch][localOffset(cl)] : T(0);
        }
        t.ch[p_ch][vi] = v;
      }
    }
  }
  std::vector<std::array<T, 3 * 512>> out(all.size());
#pragma omp parallel for schedule(static)
  for (long ti = 0; ti < long(all.size()); ti++) {
    const Info info = all[ti];
    if (!(info.type & (LEAF | GHOST))) continue;
    auto& t = G.tile(info);
    const T h = G.voxelSize(info.level);
    const uint8_t ttype = info.subtreeType(-1);
    for (int vi = 0; vi < 512; vi++) {
      Coord l = localCoord(vi), gc = composeGlobal(info.coord, l);
      const uint8_t ctype = t.ctype[vi];
      const T p1 = (ctype & INTERIOR) ? t.ch[p_ch][vi] : T(0);
      for (int axis = 0; axis < 3; axis++) {
        T u = (ttype == GHOST) ? T(0) : t.ch[u_ch + axis][vi];
        Coord n = gc; n[axis] -= 1;
        Info ni; Coord nl;
        G.findVoxel(info.level, n, ni, nl);
        uint8_t nt, nc;
        T p0;
        if (ni.empty()) { nt = ttype; nc = ctype; p0 = T(0); }
        else {
          auto& nt_ = G.tile(ni);
          nt = ni.subtreeType(-1);
          nc = nt_.ctype[localOffset(nl)];
          p0 = (nc & INTERIOR) ? nt_.ch[p_ch][localOffset(nl)] : T(0);
        }
        T dh;
        bool ok = true;
        if ((ttype & LEAF) && (nt & LEAF)) dh = h;
        else if (((ttype & LEAF) && (nt & GHOST)) || ((ttype & GHOST) && (nt & LEAF))) dh = T(1.5 * double(h));
        else ok = false;
        if (ok && !((ctype & NEUMANN) || (nc & NEUMANN))) u += (p1 - p0) / dh;
        out[ti][axis * 512 + vi] = u;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (long ti = 0; ti < long(all.size()); ti++) {
    const Info info = all[ti];
    if (!(info.type & (LEAF | GHOST))) continue;
    auto& t = G.tile(info);
    for (int axis = 0; axis < 3; axis++) std::memcpy(t.ch[u_ch + axis], &out[ti][axis * 512], 512 * sizeof(T));
  }
  for (int axis = 0; axis < 3; axis++) AccumulateLeafChildren(G, u_ch + axis, INTERIOR | DIRICHLET, T(1.0 / 4.0));
}

// ---------------------------------------------------------------------------
// Solver kernels (GMGSolver.cu)
// ---------------------------------------------------------------------------
// GMGSolver.cu:12-40
template <class T>
inline T NegLapCoeff(T one_over_h, uint8_t t0, uint8_t t1, uint8_t c0, uint8_t c1) {
  int both = (t0 & LEAF) && (t1 & LEAF);
  int lg = ((t0 & LEAF) && (t1 & GHOST)) || ((t0 & GHOST) && (t1 & LEAF));
  T a = one_over_h * T(both);
  T b = (T(2) * one_over_h / T(3)) * T(lg);
  T odh = a + b;
  int hn = (c0 & NEUMANN) || (c1 & NEUMANN);
  return hn ? T(0) : odh * one_over_h;
}

// GMGSolver.cu:42-143  (tile + one-cell face halo)
template <class T>
struct LapTile {
  T x[10][10][10];
  uint8_t tt[10][10][10], ct[10][10][10];
  T& X(int i, int j, int k) { return x[i + 1][j + 1][k + 1]; }
  uint8_t& TT(int i, int j, int k) { return tt[i + 1][j + 1][k + 1]; }
  uint8_t& CT(int i, int j, int k) { return ct[i + 1][j + 1][k + 1]; }
  // GMGSolver.cu:56-76
  T negativeLap(T h, int i, int j, int k, bool diag) {
    uint8_t t0 = TT(i, j, k), c0 = CT(i, j, k);
    T x0 = X(i, j, k), sum = T(0);
    for (int axis = 0; axis < 3; axis++)
      for (int sgn : {-1, 1}) {
        int n[3] = {i, j, k};
        n[axis] += sgn;
        T c = NegLapCoeff(T(1) / h, t0, TT(n[0], n[1], n[2]), c0, CT(n[0], n[1], n[2]));
        T x1 = X(n[0], n[1], n[2]);
        sum = diag ? sum + c : std::fma(c, x0 - x1, sum);
      }
    return (c0 & INTERIOR) ? sum * h * h * h : T(0);
  }
};

template <class T>
void LoadLapTile(Grid<T>& g, const Info& info, LapTile<T>& s, int sl, int xch) {
  auto& t = g.tile(info);
  uint8_t own = info.subtreeType(sl);
  for (int vi = 0; vi < 512; vi++) {
    Coord l = localCoord(vi);
    s.X(l[0], l[1], l[2]) = t.ch[xch][vi];
    s.TT(l[0], l[1], l[2]) = own;
    s.CT(l[0], l[1], l[2]) = t.ctype[vi];
  }
  for (int axis = 0; axis < 3; axis++)
    for (int side = 0; side < 2; side++) {
      const Info& ni = t.nbr[axis + 3 * side];
      for (int f = 0; f < 64; f++) {
        int a = f / 8, b = f % 8;
        int fl[3], nl[3];
        int ax = side ? 8 : -1, nax = side ? 0 : 7;
        if (axis == 0) { fl[0] = ax; fl[1] = a; fl[2] = b; }
        else if (axis == 1) { fl[0] = a; fl[1] = ax; fl[2] = b; }
        else { fl[0] = a; fl[1] = b; fl[2] = ax; }
        for (int d = 0; d < 3; d++) nl[d] = fl[d];
        nl[axis] = nax;
        bool empty = ni.empty();
        s.x[fl[0] + 1][fl[1] + 1][fl[2] + 1] = empty ? T(0) : g.tile(ni).ch[xch][localOffset(nl[0], nl[1], nl[2])];
        s.tt[fl[0] + 1][fl[1] + 1][fl[2] + 1] = empty ? own : ni.subtreeType(sl);
        s.ct[fl[0] + 1][fl[1] + 1][fl[2] + 1] = empty ? uint8_t(DIRICHLET) : g.tile(ni).ctype[localOffset(nl[0], nl[1], nl[2])];
      }
    }
}

enum Smoother : int { SMOOTH_RBGS = 0, SMOOTH_JACOBI = 1 };

template <class T>
struct Solver {
  Problem<T>& P;
  T one_over_alpha = T(1), prolong_coeff = T(1);  // GMGSolver.h:19-22
  // Mixed-precision forms (no reference counterpart): every value the V-cycle
  // stores in its x and rhs channels is rounded to bfloat16 (vround 1) or to
  // float16 (vround 2) (T=float only).
  int vround = 0;
  explicit Solver(Problem<T>& p) : P(p) {}
  T st(T v) const {
    if constexpr (std::is_same_v<T, float>) return vround == 1 ? bf16_round(v) : vround == 2 ? fp16_round(v) : v;
    else return v;
  }
  Grid<T>& G() { return P.g(); }

  // NegativeLaplacianSameLevel128 (GMGSolver.cu:145-200)
  void NegLapSameLevel(const std::vector<Info>& tiles, int sl, uint8_t types, int xch, int axch, bool diag) {
    Grid<T>& g = G();
    std::vector<std::array<T, 512>> out(tiles.size());
#pragma omp parallel for schedule(dynamic, 4)
    for (long ti = 0; ti < long(tiles.size()); ti++) {
      const Info info = tiles[ti];
      if (!(info.subtreeType(sl) & types)) continue;
      LapTile<T> s;
      LoadLapTile(g, info, s, sl, xch);
      T h = g.voxelSize(info.level);
      for (int vi = 0; vi < 512; vi++) { Coord l = localCoord(vi); out[ti][vi] = s.negativeLap(h, l[0], l[1], l[2], diag); }
    }
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(tiles.size()); ti++) {
      const Info info = tiles[ti];
      if (!(info.subtreeType(sl) & types)) continue;
      std::memcpy(g.tile(info).ch[axch], out[ti].data(), sizeof(T) * 512);
    }
  }
  // PoissonGrid.cu:585-605
  void PropagateValuesToGhostTiles(int cch, int fch) {
    Grid<T>& g = G();
    const auto& all = g.all;
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(all.size()); ti++) {
      const Info info = all[ti];
      if (!(info.type & GHOST)) continue;
      auto& t = g.tile(info);
      for (int vi = 0; vi < 512; vi++) {
        if (!(t.ctype[vi] & INTERIOR)) { t.ch[fch][vi] = T(0); continue; }
        Coord gc = composeGlobal(info.coord, localCoord(vi));
        Info ci; Coord cl;
        g.findVoxel(info.level - 1, parentCoord(gc), ci, cl);
        if (!ci.empty()) {
          auto& ct = g.tile(ci);
          int co = localOffset(cl);
          t.ch[fch][vi] = (ct.ctype[co] & INTERIOR) ? ct.ch[cch][co] : T(0);
        } else t.ch[fch][vi] = T(0);
      }
    }
  }
  // AccumulateToParents128 (PoissonGrid.cu:664-719)
  void AccumulateToParents128(int fch, int cch, uint8_t ftypes, T coeff, bool additive, uint8_t ctypes) {
    Grid<T>& g = G();
    const auto& all = g.all;
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(all.size()); ti++) {
      const Info finfo = all[ti];
      if (!(finfo.subtreeType(-1) & ftypes)) continue;
      auto& ft = g.tile(finfo);
      T data[512];
      for (int vi = 0; vi < 512; vi++) data[vi] = (ft.ctype[vi] & ctypes) ? ft.ch[fch][vi] : T(0);
      for (int c = 0; c < 64; c++) {
        Coord fl(c / 16 * 2, (c / 4) % 4 * 2, c % 4 * 2);
        Info ci; Coord cl;
        g.findVoxel(finfo.level - 1, parentCoord(composeGlobal(finfo.coord, fl)), ci, cl);
        if (ci.empty()) continue;
        T sum = T(0);
        for (int ii : {0, 1}) for (int jj : {0, 1}) for (int kk : {0, 1})
          sum += data[localOffset(fl[0] + ii, fl[1] + jj, fl[2] + kk)];
        T& dst = g.tile(ci).ch[cch][localOffset(cl)];
        dst = additive ? dst + sum * coeff : sum * coeff;
      }
    }
  }
  // GMGSolver.cu:203-213
  void FullNegativeLaplacian(int xch, int axch, bool diag = false) {
    PropagateValuesToGhostTiles(xch, xch);
    NegLapSameLevel(G().all, -1, LEAF | GHOST, xch, axch, diag);
    AccumulateToParents128(axch, axch, GHOST, T(1), true, INTERIOR | DIRICHLET | NEUMANN);
  }

  // NegativeLaplacianAndGaussSeidelSameLevel128Kernel (GMGSolver.cu:216-269),
  // one colour over LEAF|NONLEAF tiles of `level` (subtree view at level)
  void GSColor(int level, int xch, int rhsch, int color) {
    Grid<T>& g = G();
    const auto& arr = g.arrays[level];
    std::vector<std::array<T, 512>> out(arr.size());
#pragma omp parallel for schedule(dynamic, 4)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info info = arr[ti];
      if (!(info.subtreeType(level) & (LEAF | NONLEAF))) continue;
      auto& t = g.tile(info);
      LapTile<T> s;
      LoadLapTile(g, info, s, level, xch);
      T h = g.voxelSize(info.level);
      for (int vi = 0; vi < 512; vi++) {
        Coord l = localCoord(vi);
        T xo = s.X(l[0], l[1], l[2]);
        out[ti][vi] = xo;
        if ((l[0] + l[1] + l[2]) % 2 != color) continue;
        uint8_t ct = s.CT(l[0], l[1], l[2]);
        T b = t.ch[rhsch][vi];
        T Ax = s.negativeLap(h, l[0], l[1], l[2], false);
        T D = s.negativeLap(h, l[0], l[1], l[2], true);
        T delta = (ct & INTERIOR) ? ((D != T(0)) ? (b - Ax) / D : T(0)) : T(0);
        out[ti][vi] = st(xo + delta);
      }
    }
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info info = arr[ti];
      if (!(info.subtreeType(level) & (LEAF | NONLEAF))) continue;
      std::memcpy(g.tile(info).ch[xch], out[ti].data(), sizeof(T) * 512);
    }
  }
  // GMGSolver.cu:277-282
  void GaussSeidelGMG(int iters, int order, int level, int xch, int rhsch) {
    for (int i = 0; i < iters; i++) {
      GSColor(level, xch, rhsch, order);
      GSColor(level, xch, rhsch, 1 - order);
    }
  }
  // Damped-Jacobi sweep used as a V-cycle smoother (the update of
  // GMGSolver.cu:508-525 applied with the level's subtree operator).
  void JacobiSweep(int level, int xch, int rhsch, T omega) {
    Grid<T>& g = G();
    const auto& arr = g.arrays[level];
    std::vector<std::array<T, 512>> out(arr.size());
#pragma omp parallel for schedule(dynamic, 4)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info info = arr[ti];
      if (!(info.subtreeType(level) & (LEAF | NONLEAF))) continue;
      auto& t = g.tile(info);
      LapTile<T> s;
      LoadLapTile(g, info, s, level, xch);
      T h = g.voxelSize(info.level);
      for (int vi = 0; vi < 512; vi++) {
        Coord l = localCoord(vi);
        T xo = s.X(l[0], l[1], l[2]);
        uint8_t ct = s.CT(l[0], l[1], l[2]);
        if (!(ct & INTERIOR)) { out[ti][vi] = xo; continue; }
        T Ax = s.negativeLap(h, l[0], l[1], l[2], false);
        T D = s.negativeLap(h, l[0], l[1], l[2], true);
        T r = t.ch[rhsch][vi] - Ax;
        out[ti][vi] = (D == T(0)) ? T(0) : st(xo + (omega * r) / D);
      }
    }
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info info = arr[ti];
      if (!(info.subtreeType(level) & (LEAF | NONLEAF))) continue;
      std::memcpy(g.tile(info).ch[xch], out[ti].data(), sizeof(T) * 512);
    }
  }
  // ResidualAndRestrict128Kernel (GMGSolver.cu:284-343)
  void ResidualAndRestrict(int level, int xch, int rhsch, int cch, uint8_t types) {
    Grid<T>& g = G();
    const auto& arr = g.arrays[level];
#pragma omp parallel for schedule(dynamic, 4)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info finfo = arr[ti];
      if (!(finfo.subtreeType(level) & types)) continue;
      auto& ft = g.tile(finfo);
      LapTile<T> s;
      LoadLapTile(g, finfo, s, level, xch);
      T h = g.voxelSize(finfo.level);
      T res[512];
      for (int vi = 0; vi < 512; vi++) { Coord l = localCoord(vi); res[vi] = ft.ch[rhsch][vi] - s.negativeLap(h, l[0], l[1], l[2], false); }
      for (int c = 0; c < 64; c++) {
        Coord fl(c / 16 * 2, (c / 4) % 4 * 2, c % 4 * 2);
        Info ci; Coord cl;
        g.findVoxel(level - 1, parentCoord(composeGlobal(finfo.coord, fl)), ci, cl);
        if (ci.empty()) continue;
        T sum = T(0);
        for (int ii : {0, 1}) for (int jj : {0, 1}) for (int kk : {0, 1})
          sum += res[localOffset(fl[0] + ii, fl[1] + jj, fl[2] + kk)];
        T& dst = g.tile(ci).ch[cch][localOffset(cl)];
        if (finfo.subtreeType(level) & GHOST) dst = st(dst + sum);
        else dst = st(sum * one_over_alpha);
      }
    }
  }
  // ProlongateAndUpdate128Kernel (GMGSolver.cu:345-380)
  void ProlongateAndUpdate(int level, int cxch, int fxch) {
    Grid<T>& g = G();
    const auto& arr = g.arrays[level];
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info finfo = arr[ti];
      uint8_t ftt = finfo.subtreeType(level);
      if (!(ftt & (LEAF | NONLEAF | GHOST))) continue;
      auto& ft = g.tile(finfo);
      for (int vi = 0; vi < 512; vi++) {
        Coord gc = composeGlobal(finfo.coord, localCoord(vi));
        Info ci; Coord cl;
        g.findVoxel(level - 1, parentCoord(gc), ci, cl);
        T cx = ci.empty() ? T(0) : g.tile(ci).ch[cxch][localOffset(cl)];
        T fx = ft.ch[fxch][vi];
        bool in = ft.ctype[vi] & INTERIOR;
        if (ftt & GHOST) ft.ch[fxch][vi] = in ? cx : T(0);
        else ft.ch[fxch][vi] = in ? st(std::fma(prolong_coeff, cx, fx)) : T(0);
      }
    }
  }
  void Smooth(int smoother, int iters, int order, int level, int xch, int rhsch, T omega) {
    if (smoother == SMOOTH_RBGS) GaussSeidelGMG(iters, order, level, xch, rhsch);
    else for (int i = 0; i < iters; i++) JacobiSweep(level, xch, rhsch, omega);
  }
  // GMGSolver::VCycle (GMGSolver.cu:383-447); the smoother may be damped Jacobi
  void VCycle(int xch, int fch, int rhsch, int level_iters, int coarsest_iters, int smoother = SMOOTH_RBGS, T omega = T(2) / T(3)) {
    Grid<T>& g = G();
    const auto& all = g.all;
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(all.size()); ti++) {
      const Info info = all[ti];
      auto& t = g.tile(info);
      for (int vi = 0; vi < 512; vi++) {
        t.ch[xch][vi] = T(0);
        t.ch[rhsch][vi] = (info.type & GHOST) ? T(0) : st(t.ch[fch][vi]);
      }
    }
    for (int i = g.maxLevel; i > 0; i--) {
      Smooth(smoother, level_iters, 0, i, xch, rhsch, omega);
      ResidualAndRestrict(i, xch, rhsch, rhsch, LEAF | NONLEAF | GHOST);
    }
    GaussSeidelGMG(coarsest_iters / 2, 0, 0, xch, rhsch);
    GaussSeidelGMG(coarsest_iters / 2, 1, 0, xch, rhsch);
    for (int i = 1; i <= g.maxLevel; i++) {
      ProlongateAndUpdate(i, xch, xch);
      Smooth(smoother, level_iters, 1, i, xch, rhsch, omega);
    }
  }

  // Dot over all cells of the tiles whose type is in `types`, in canonical
  // order (PoissonGrid.cu:165-228, with the deterministic order above).
  double Dot(int ach, int bch, uint8_t types) {
    Grid<T>& g = G();
    std::vector<const Info*> tl;
    for (const Info& in : P.canon) if (in.subtreeType(-1) & types) tl.push_back(&in);
    std::vector<double> part(tl.size());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(tl.size()); i++) {
      auto& t = g.tile(*tl[i]);
      part[i] = tile_partial([&](int c) { return double(t.ch[ach][c]) * double(t.ch[bch][c]); });
    }
    return sum_partials(part);
  }
  // PoissonGrid.cu:231-325 MeanAsync + PoissonSolver.cu:356-368 ReCenterLeafVoxels
  void ReCenterLeafVoxels(int ch) {
    Grid<T>& g = G();
    const auto& tl = P.leafCanon;
    std::vector<double> ps(tl.size()), pc(tl.size());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(tl.size()); i++) {
      auto& t = g.tile(tl[i]);
      ps[i] = tile_partial([&](int c) { return double(t.ch[ch][c]); });
      pc[i] = tile_partial([&](int c) { return double(t.ctype[c] & INTERIOR ? 1 : 0); });
    }
    double mean = sum_partials(ps) / sum_partials(pc);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(tl.size()); i++) {
      auto& t = g.tile(tl[i]);
      for (int c = 0; c < 512; c++) if (t.ctype[c] & INTERIOR) t.ch[ch][c] = T(double(t.ch[ch][c]) - mean);
    }
  }
  template <class F>
  void ForLeafCells(F f) {
    Grid<T>& g = G();
    const auto& tl = P.leafCanon;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(tl.size()); i++) {
      auto& t = g.tile(tl[i]);
      for (int c = 0; c < 512; c++) f(t, c);
    }
  }

  // ===== AMGSolver (AMGSolver.cu), the stored-coefficient preconditioner =====
  // AMGSolver.cu:17-36: LEAF-LEAF faces 1/h, LEAF-GHOST faces 1/(1.5 h) (not
  // the GMG operator's (2/3)/h), both over h; 0 across NEUMANN.
  static T AMGCoeff(T h, uint8_t t0, uint8_t t1, uint8_t c0, uint8_t c1) {
    const int both = (t0 & LEAF) && (t1 & LEAF);
    const int lg = ((t0 & LEAF) && (t1 & GHOST)) || ((t0 & GHOST) && (t1 & LEAF));
    const T odh = (T(1) / h) * T(both) + (T(1) / (T(1.5) * h)) * T(lg);
    const int hn = (c0 & NEUMANN) || (c1 & NEUMANN);
    return hn ? T(0) : odh / h;
  }
  std::array<T, 512>& AC(int a, const Info& in) { return P.amg[a][size_t(in.tile)]; }
  // CalculateAMGCoefficients (AMGSolver.cu:40-107) on tiles of `types`
  void CalculateAMGCoefficients(uint8_t types) {
    Grid<T>& g = G();
    const auto& all = g.all;
#pragma omp parallel for schedule(dynamic, 4)
    for (long ti = 0; ti < long(all.size()); ti++) {
      const Info info = all[ti];
      if (!(info.type & types)) continue;
      auto& t = g.tile(info);
      const T h = g.voxelSize(info.level);
      for (int vi = 0; vi < 512; vi++) {
        const Coord gc = composeGlobal(info.coord, localCoord(vi));
        const uint8_t c0 = t.ctype[vi];
        T diag = T(0), off[3] = {T(0), T(0), T(0)};
        for (int axis = 0; axis < 3; axis++)
          for (int sgn : {-1, 1}) {
            Coord ng = gc;
            ng[axis] += sgn;
            Info ni; Coord nl;
            g.findVoxel(info.level, ng, ni, nl);
            const uint8_t t1 = ni.empty() ? info.type : ni.type;
            const uint8_t c1 = ni.empty() ? uint8_t(DIRICHLET) : g.tile(ni).ctype[localOffset(nl)];
            const T coeff = ((AMGCoeff(h, info.type, t1, c0, c1) * h) * h) * h;
            if (sgn == -1 && c1 == INTERIOR) off[axis] = -coeff;
            diag = diag + coeff;
          }
        for (int a = 0; a < 3; a++) AC(a, info)[vi] = (c0 == INTERIOR) ? off[a] : T(0);
        AC(3, info)[vi] = (c0 == INTERIOR) ? diag : T(0);
      }
    }
  }
  // AMGSolver::prepareTypesAndCoeffs = CoarsenTypesAndAMGCoeffs (AMGSolver.cu:110-209):
  // ghost cell types from the parents, LEAF/GHOST coefficients, then finest
  // level first the Galerkin sums of the 8 children into NONLEAF cells (whose
  // cell type becomes INTERIOR if any child is INTERIOR, else DIRICHLET) and
  // into LEAF cells with GHOST children.  Cell types are changed in place.
  void AMGPrepare() {
    Grid<T>& g = G();
    for (int a = 0; a < 4; a++) P.amg[a].assign(g.store.size(), std::array<T, 512>{});
    const auto& all = g.all;
#pragma omp parallel for schedule(static)
    for (long ti = 0; ti < long(all.size()); ti++) {
      const Info info = all[ti];
      if (!(info.type & GHOST)) continue;
      auto& t = g.tile(info);
      for (int vi = 0; vi < 512; vi++) {
        Info pi; Coord pl;
        g.findVoxel(info.level - 1, parentCoord(composeGlobal(info.coord, localCoord(vi))), pi, pl);
        t.ctype[vi] = pi.empty() ? uint8_t(DIRICHLET) : g.tile(pi).ctype[localOffset(pl)];
      }
    }
    CalculateAMGCoefficients(LEAF | GHOST);
    for (int level = g.numLayers - 1; level >= 0; level--) {
      const auto& arr = g.arrays[level];
#pragma omp parallel for schedule(static)
      for (long ti = 0; ti < long(arr.size()); ti++) {
        const Info info = arr[ti];
        if (!(info.type & (NONLEAF | LEAF))) continue;
        auto& t = g.tile(info);
        for (int vi = 0; vi < 512; vi++) {
          const Coord gc = composeGlobal(info.coord, localCoord(vi));
          int interior = 0;
          bool ghost_child = false;
          T cod[3][2][2][2];
          T dsum = T(0);
          for (int ci = 0; ci < 2; ci++)
            for (int cj = 0; cj < 2; cj++)
              for (int ck = 0; ck < 2; ck++) {
                Info cinfo; Coord cl;
                g.findVoxel(level + 1, Coord(gc[0] * 2 + ci, gc[1] * 2 + cj, gc[2] * 2 + ck), cinfo, cl);
                if (!cinfo.empty()) {
                  if (cinfo.type == GHOST) ghost_child = true;
                  const int co = localOffset(cl);
                  dsum = dsum + AC(3, cinfo)[co];
                  interior += g.tile(cinfo).ctype[co] == INTERIOR;
                  for (int a = 0; a < 3; a++) cod[a][ci][cj][ck] = AC(a, cinfo)[co];
                } else {
                  for (int a = 0; a < 3; a++) cod[a][ci][cj][ck] = T(0);
                }
              }
          T offs[3] = {T(0), T(0), T(0)};
          for (int ci = 0; ci < 2; ci++)
            for (int cj = 0; cj < 2; cj++)
              for (int ck = 0; ck < 2; ck++) {
                if (ci == 0) offs[0] = offs[0] + cod[0][ci][cj][ck];
                if (ci == 1) dsum = dsum + T(2) * cod[0][ci][cj][ck];
                if (cj == 0) offs[1] = offs[1] + cod[1][ci][cj][ck];
                if (cj == 1) dsum = dsum + T(2) * cod[1][ci][cj][ck];
                if (ck == 0) offs[2] = offs[2] + cod[2][ci][cj][ck];
                if (ck == 1) dsum = dsum + T(2) * cod[2][ci][cj][ck];
              }
          if (info.type == NONLEAF) {
            t.ctype[vi] = interior > 0 ? uint8_t(INTERIOR) : uint8_t(DIRICHLET);
            for (int a = 0; a < 3; a++) AC(a, info)[vi] = offs[a] * one_over_alpha;
            AC(3, info)[vi] = dsum * one_over_alpha;
          }
          if (info.type == LEAF && ghost_child) {
            for (int a = 0; a < 3; a++) AC(a, info)[vi] = AC(a, info)[vi] + offs[a];
            AC(3, info)[vi] = AC(3, info)[vi] + dsum;
          }
        }
      }
    }
  }
  // A x with the stored coefficients on one tile (AMGLaplacianTileData::
  // negativeLapSameLevel, AMGSolver.cu:232-250, tile + halo as
  // LoadAMGLaplacianTileDataAndTileType, AMGSolver.cu:396-466): the x halo from
  // the face neighbours (0 where there is none), the plus faces' coefficients
  // from the plus neighbours' minus-face entries (0 where there is none).
  void AMGTileRows(const Info& info, int xch, T out[512]) {
    Grid<T>& g = G();
    auto& t = g.tile(info);
    auto X = [&](int i, int j, int k) -> T {
      if (i >= 0 && i < 8 && j >= 0 && j < 8 && k >= 0 && k < 8) return t.ch[xch][localOffset(i, j, k)];
      int f = i < 0 ? 0 : i > 7 ? 3 : j < 0 ? 1 : j > 7 ? 4 : k < 0 ? 2 : 5;
      const Info& ni = t.nbr[f];
      if (ni.empty()) return T(0);
      return g.tile(ni).ch[xch][localOffset((i + 8) & 7, (j + 8) & 7, (k + 8) & 7)];
    };
    auto OFF = [&](int a, int i, int j, int k) -> T {
      if (i < 8 && j < 8 && k < 8) return AC(a, info)[localOffset(i, j, k)];
      const Info& ni = t.nbr[a + 3];
      if (ni.empty()) return T(0);
      return AC(a, ni)[localOffset(i & 7, j & 7, k & 7)];
    };
    for (int vi = 0; vi < 512; vi++) {
      const Coord l = localCoord(vi);
      T sum = X(l[0], l[1], l[2]) * AC(3, info)[vi];
      for (int a = 0; a < 3; a++)
        for (int sgn : {-1, 1}) {
          int n[3] = {l[0], l[1], l[2]}, u[3] = {l[0], l[1], l[2]};
          n[a] += sgn;
          if (sgn == 1) u[a] += 1;
          sum = std::fma(OFF(a, u[0], u[1], u[2]), X(n[0], n[1], n[2]), sum);
        }
      out[vi] = (t.ctype[vi] == INTERIOR) ? sum : T(0);
    }
  }
  // NegativeLaplacianAMG128 (AMGSolver.cu:468-533) on the tiles of `level` whose type is in `types`
  void NegLapAMG(int level, uint8_t types, int xch, int axch) {
    Grid<T>& g = G();
    const auto& arr = g.arrays[level];
    std::vector<std::array<T, 512>> out(arr.size());
#pragma omp parallel for schedule(dynamic, 4)
    for (long ti = 0; ti < long(arr.size()); ti++)
      if (arr[ti].type & types) AMGTileRows(arr[ti], xch, out[ti].data());
    for (long ti = 0; ti < long(arr.size()); ti++)
      if (arr[ti].type & types) std::memcpy(g.tile(arr[ti]).ch[axch], out[ti].data(), sizeof(T) * 512);
  }
  // NegativeLaplacianAndGaussSeidelSameLevel128Kernel (AMGSolver.cu:540-578):
  // one colour on the LEAF|NONLEAF tiles of `level` (D = 0: no update, as in the GMG path)
  void GSColorAMG(int level, int xch, int rhsch, int color) {
    Grid<T>& g = G();
    const auto& arr = g.arrays[level];
    std::vector<std::array<T, 512>> out(arr.size());
#pragma omp parallel for schedule(dynamic, 4)
    for (long ti = 0; ti < long(arr.size()); ti++) {
      const Info info = arr[ti];
      if (!(info.type & (LEAF | NONLEAF))) continue;
      auto& t = g.tile(info);
      T ax[512];
      AMGTileRows(info, xch, ax);
      for (int vi = 0; vi < 512; vi++) {
        const Coord l = localCoord(vi);
        const T x0 = t.ch[xch][vi];
        out[ti][vi] = x0;
        if ((l[0] + l[1] + l[2]) % 2 != color) continue;
        const T D = AC(3, info)[vi];
        // ax holds 0 off INTERIOR cells, where no update happens
        const T delta = (t.ctype[vi] & INTERIOR) ? ((D != T(0)) ? (t.ch[rhsch][vi] - ax[vi]) / D : T(0)) : T(0);
        out[ti][vi] = x0 + delta;
      }
    }
    for (long ti = 0; ti < long(arr.size()); ti++)
      if (arr[ti].type & (LEAF | NONLEAF)) std::memcpy(g.tile(arr[ti]).ch[xch], out[ti].data(), sizeof(T) * 512);
  }
  // GaussSeidelAMG (AMGSolver.cu:598-603)
  void GaussSeidelAMG(int iters, int order, int level, int xch, int rhsch) {
    for (int i = 0; i < iters; i++) {
      GSColorAMG(level, xch, rhsch, order);
      GSColorAMG(level, xch, rhsch, 1 - order);
    }
  }
  // VCycleAMG (AMGSolver.cu:610-715)
  void VCycleAMG(int xch, int fch, int tmpch, int rhsch, int level_iters, int coarsest_iters) {
    Grid<T>& g = G();
    for (const Info& info : g.all) {
      auto& t = g.tile(info);
      for (int vi = 0; vi < 512; vi++) {
        t.ch[xch][vi] = T(0);
        t.ch[rhsch][vi] = (info.type & GHOST) ? T(0) : t.ch[fch][vi];
      }
    }
    for (int i = g.maxLevel; i > 0; i--) {
      GaussSeidelAMG(level_iters, 0, i, xch, rhsch);
      NegLapAMG(i, LEAF | NONLEAF | GHOST, xch, tmpch);
      for (const Info& info : g.arrays[i]) {  // BinaryTransform rhs - tmp on INTERIOR cells
        auto& t = g.tile(info);
        for (int vi = 0; vi < 512; vi++)
          if (t.ctype[vi] & INTERIOR) t.ch[tmpch][vi] = t.ch[rhsch][vi] - t.ch[tmpch][vi];
      }
      for (const Info& info : g.arrays[i - 1]) {
        auto& t = g.tile(info);
        if (info.type == NONLEAF) {  // Restrict (PoissonSolver.cu:292-316)
          for (int vi = 0; vi < 512; vi++) {
            const Coord gc = composeGlobal(info.coord, localCoord(vi));
            T sum = T(0);
            for (int sidx = 0; sidx < 8; sidx++) {
              Info ci; Coord cl;
              g.findVoxel(i, childCoord(gc, childIndexToOffset(sidx)), ci, cl);
              if (!ci.empty()) {
                const auto& ct = g.tile(ci);
                const int co = localOffset(cl);
                sum = sum + ((ct.ctype[co] & INTERIOR) ? ct.ch[tmpch][co] : T(0));
              }
            }
            t.ch[rhsch][vi] = sum * one_over_alpha;
          }
        } else if (info.type == LEAF) {  // Copy f, then AccumulateToParents of the ghost children (PoissonGrid.cu:721-755)
          for (int vi = 0; vi < 512; vi++) {
            const Coord gc = composeGlobal(info.coord, localCoord(vi));
            T sum = T(0);
            for (int sidx = 0; sidx < 8; sidx++) {
              Info ci; Coord cl;
              g.findVoxel(i, childCoord(gc, childIndexToOffset(sidx)), ci, cl);
              if (!ci.empty()) sum = sum + g.tile(ci).ch[tmpch][localOffset(cl)];
            }
            t.ch[rhsch][vi] = std::fma(sum, T(1), t.ch[fch][vi]);
          }
        } else {  // GHOST: Fill 0
          for (int vi = 0; vi < 512; vi++) t.ch[rhsch][vi] = T(0);
        }
      }
    }
    GaussSeidelAMG(coarsest_iters / 2, 0, 0, xch, rhsch);
    GaussSeidelAMG(coarsest_iters / 2, 1, 0, xch, rhsch);
    for (int i = 1; i <= g.maxLevel; i++) {
      ProlongateAndUpdate(i, xch, xch);
      GaussSeidelAMG(level_iters, 1, i, xch, rhsch);
    }
  }

  struct Result { int iters; double rel_res; double bb; double rr; };

  // GMGSolver::solve (GMGSolver.cu:531-666)
  // sync_stride: the residual is tested after iterations k*sync_stride
  // (GMGSolver.cu:611); <= 0 never tests and reports a relative residual of -1
  // (GMGSolver.cu:655-665).  b = 0 returns (0, 0) in every mode (with
  // sync_stride <= 0 the reference would go on and divide 0 by 0).
  // thr_floor (no reference counterpart, used by the refinement's inner
  // solves): the stop threshold is max(tol^2 |b|^2, thr_floor).
  // With float16 V-cycle storage (vround 2) the right-hand side is first
  // scaled by 2^k (fp16_scale_exp) and x and r are scaled back at the end;
  // |b|^2 and |r|^2 are reported unscaled, the threshold floor scaled with b.
  Result solve(int max_iters, double tol, int level_iters, int coarsest_iters, bool pure_neumann,
               int smoother = SMOOTH_RBGS, T omega = T(2) / T(3), int sync_stride = 1, double thr_floor = 0.0) {
    if (vround != 2 || !std::is_same_v<T, float>) return solve_body(max_iters, tol, level_iters, coarsest_iters, pure_neumann, smoother, omega, sync_stride, thr_floor);
    const double bb0 = Dot(CH_R, CH_R, LEAF);
    const int k = fp16_scale_exp(bb0, 512.0 * double(P.leafCanon.size()));
    const T up = T(std::ldexp(1.0, k)), down = T(std::ldexp(1.0, -k));
    ForLeafCells([&](TileData<T>& t, int c) { t.ch[CH_R][c] = t.ch[CH_R][c] * up; });
    Result r = solve_body(max_iters, tol, level_iters, coarsest_iters, pure_neumann, smoother, omega, sync_stride,
                          std::ldexp(thr_floor, 2 * k));
    ForLeafCells([&](TileData<T>& t, int c) {
      t.ch[CH_X][c] = t.ch[CH_X][c] * down;
      t.ch[CH_R][c] = t.ch[CH_R][c] * down;
    });
    r.bb = std::ldexp(r.bb, -2 * k);
    r.rr = std::ldexp(r.rr, -2 * k);
    return r;
  }
  // amg: the preconditioner is VCycleAMG (AMGSolver::solve, AMGSolver.cu:722-855,
  // the same PCG loop as GMGSolver::solve); AMGPrepare must have run.
  bool amg = false;
  void Precondition(int level_iters, int coarsest_iters, int smoother, T omega) {
    if (amg) VCycleAMG(CH_Z, CH_R, CH_AP, CH_TMP, level_iters, coarsest_iters);
    else VCycle(CH_Z, CH_R, CH_AP, level_iters, coarsest_iters, smoother, omega);
  }
  Result solve_body(int max_iters, double tol, int level_iters, int coarsest_iters, bool pure_neumann,
                    int smoother, T omega, int sync_stride, double thr_floor) {
    ForLeafCells([](TileData<T>& t, int c) { t.ch[CH_X][c] = T(0); });
    double bb = Dot(CH_R, CH_R, LEAF);
    if (bb == 0.0) return {0, 0.0, 0.0, 0.0};
    double thr = std::max(std::max(tol * tol * bb, thr_floor), std::numeric_limits<double>::min());
    if (pure_neumann) ReCenterLeafVoxels(CH_R);
    Precondition(level_iters, coarsest_iters, smoother, omega);
    ForLeafCells([](TileData<T>& t, int c) { t.ch[CH_P][c] = t.ch[CH_Z][c]; });
    double gamma = Dot(CH_Z, CH_R, LEAF);
    double rr = bb;
    for (int i = 0; i < max_iters; i++) {
      FullNegativeLaplacian(CH_P, CH_AP);
      double pAp = Dot(CH_P, CH_AP, LEAF);
      T alpha = T(gamma / pAp);
      ForLeafCells([&](TileData<T>& t, int c) {
        if (t.ctype[c] & INTERIOR) {
          t.ch[CH_X][c] = std::fma(alpha, t.ch[CH_P][c], t.ch[CH_X][c]);
          t.ch[CH_R][c] = std::fma(-alpha, t.ch[CH_AP][c], t.ch[CH_R][c]);
        }
      });
      if (sync_stride > 0 && (i + 1) % sync_stride == 0) {
        rr = Dot(CH_R, CH_R, LEAF);
        if (rr < thr) return {i + 1, std::sqrt(rr / bb), bb, rr};
      }
      if (pure_neumann) ReCenterLeafVoxels(CH_R);
      Precondition(level_iters, coarsest_iters, smoother, omega);
      double gamma_old = gamma;
      gamma = Dot(CH_R, CH_Z, LEAF);
      T beta = T(gamma / gamma_old);
      ForLeafCells([&](TileData<T>& t, int c) {
        if (t.ctype[c] & INTERIOR) t.ch[CH_P][c] = std::fma(beta, t.ch[CH_P][c], t.ch[CH_Z][c]);
      });
    }
    rr = Dot(CH_R, CH_R, LEAF);
    return {max_iters, sync_stride > 0 ? std::sqrt(rr / bb) : -1.0, bb, rr};
  }

  // GMGSolver::dampedJacobiSolve (GMGSolver.cu:449-529)
  Result dampedJacobiSolve(int max_iters, double tol, T omega) {
    Grid<T>& g = G();
    for (const Info& info : g.all) { auto& t = g.tile(info); for (int c = 0; c < 512; c++) t.ch[CH_X][c] = T(0); }
    double bnorm = std::sqrt(Dot(CH_B, CH_B, LEAF | NONLEAF | GHOST));
    double thr = std::max(tol * bnorm, std::numeric_limits<double>::min());
    int i = 0;
    double rn = bnorm;
    for (i = 0; i < max_iters; i++) {
      FullNegativeLaplacian(CH_X, CH_P, false);
      ForLeafCells([](TileData<T>& t, int c) { if (t.ctype[c] & INTERIOR) t.ch[CH_P][c] = t.ch[CH_B][c] - t.ch[CH_P][c]; });
      rn = std::sqrt(Dot(CH_P, CH_P, LEAF));
      if (rn < thr) break;
      FullNegativeLaplacian(CH_X, CH_AP, true);
      ForLeafCells([&](TileData<T>& t, int c) {
        if (t.ctype[c] & INTERIOR) {
          T D = t.ch[CH_AP][c];
          if (D == T(0)) t.ch[CH_X][c] = T(0);
          else t.ch[CH_X][c] = t.ch[CH_X][c] + (omega * t.ch[CH_P][c]) / D;
        }
      });
    }
    return {i, rn / bnorm, bnorm * bnorm, rn * rn};
  }
};

// ---------------------------------------------------------------------------
// Mixed-precision form (no reference counterpart)
// ---------------------------------------------------------------------------
// A copy of a problem with another arithmetic type (same tiles, hash tables,
// canonical order and cell types; channel values converted).
template <class B, class A>
std::unique_ptr<Problem<B>> ConvertProblem(Problem<A>& a) {
  auto P = std::make_unique<Problem<B>>();
  P->params = a.params;
  P->tubes = a.tubes;
  Grid<A>& ga = a.g();
  P->grid = std::make_unique<Grid<B>>(B(ga.h0), ga.log2);
  Grid<B>& gb = P->g();
  gb.maxLevel = ga.maxLevel;
  gb.hashT = ga.hashT;
  gb.arrays = ga.arrays;
  gb.all = ga.all;
  gb.count = ga.count;
  gb.store.clear();
  for (const auto& t : ga.store) {
    if (!t) {  // a removed tile's slot
      gb.store.emplace_back();
      continue;
    }
    auto d = std::make_unique<TileData<B>>();
    for (int ch = 0; ch < NCH; ch++)
      for (int c = 0; c < 512; c++) d->ch[ch][c] = B(t->ch[ch][c]);
    std::memcpy(d->ctype, t->ctype, 512);
    for (int f = 0; f < 6; f++) d->nbr[f] = t->nbr[f];
    d->status = t->status;
    gb.store.push_back(std::move(d));
  }
  P->canon = a.canon;
  P->canonOfStore = a.canonOfStore;
  P->leafCanon = a.leafCanon;
  return P;
}

// Iterative refinement: x and r = b - A x in FP64 (Pd, an FP64 copy of Pf whose
// b channel holds the FP32 b); each step solves A d = float(r) with the FP32
// MGPCG on Pf (V-cycle stores rounded to bfloat16 when bf16 is set) until its
// recursive |r_in|^2 < max(inner_tol^2 |float(r)|^2, thr / 4), thr =
// max(tol^2 |b|^2, DBL_MIN), and adds d to x.  Stops when |r|^2 < thr (FP64, LEAF tiles) or after
// max_outer steps.  On return Pd's x channel holds x and Pf's holds float(x).
struct RefineResult { int outer, iters, converged; double rel_res, bb, rr; };
inline RefineResult SolveRefined(Problem<float>& Pf, Problem<double>& Pd, double tol, double inner_tol, int max_outer,
                                 int max_iters, int level_iters, int coarsest_iters, bool pure_neumann, int smoother,
                                 float omega, int sync_stride, int storage) {
  Solver<double> Sd(Pd);
  Solver<float> Sf(Pf);
  Sf.vround = storage;
  Grid<double>& gd = Pd.g();
  Grid<float>& gf = Pf.g();
  for (const Info& in : gd.all) { auto& t = gd.tile(in); for (int c = 0; c < 512; c++) t.ch[CH_X][c] = 0.0; }
  const double bb = Sd.Dot(CH_B, CH_B, LEAF);
  RefineResult res{0, 0, 0, 0.0, bb, 0.0};
  const auto& tl = Pd.leafCanon;
  if (bb == 0.0) {
    res.converged = 1;
    for (const Info& in : tl) { auto& t = gf.tile(in); for (int c = 0; c < 512; c++) t.ch[CH_X][c] = 0.f; }
    return res;
  }
  const double thr = std::max(tol * tol * bb, std::numeric_limits<double>::min());
  while (true) {
    Sd.FullNegativeLaplacian(CH_X, CH_AP);
    for (const Info& in : tl) {
      auto& t = gd.tile(in);
      for (int c = 0; c < 512; c++) t.ch[CH_TMP][c] = t.ch[CH_B][c] - t.ch[CH_AP][c];
    }
    if (pure_neumann) Sd.ReCenterLeafVoxels(CH_TMP);
    res.rr = Sd.Dot(CH_TMP, CH_TMP, LEAF);
    if (res.rr < thr) { res.converged = 1; break; }
    if (res.outer >= max_outer) break;
    for (const Info& in : tl) {
      auto& td = gd.tile(in);
      auto& tf = gf.tile(in);
      for (int c = 0; c < 512; c++) tf.ch[CH_R][c] = float(td.ch[CH_TMP][c]);
    }
    auto r = Sf.solve(max_iters, inner_tol, level_iters, coarsest_iters, pure_neumann, smoother, omega, sync_stride,
                      0.25 * thr);
    res.outer++;
    res.iters += r.iters;
    for (const Info& in : tl) {
      auto& td = gd.tile(in);
      auto& tf = gf.tile(in);
      for (int c = 0; c < 512; c++) td.ch[CH_X][c] = td.ch[CH_X][c] + double(tf.ch[CH_X][c]);
    }
  }
  for (const Info& in : tl) {
    auto& td = gd.tile(in);
    auto& tf = gf.tile(in);
    for (int c = 0; c < 512; c++) tf.ch[CH_X][c] = float(td.ch[CH_X][c]);
  }
  res.rel_res = std::sqrt(res.rr / bb);
  return res;
}

}  // namespace orc
