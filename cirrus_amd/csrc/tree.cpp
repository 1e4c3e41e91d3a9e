// Host-side octree construction, cell typing and the synthetic scenes.
//
// The tree follows the reference's refinement semantics (a LEAF refines when
// its target level exceeds its level, PoissonGrid.h:63-65, or when a
// same-level face neighbour of one of its child positions refines,
// PoissonGrid.h:67-105, i.e. 2:1 face balance) but computes the resulting
// least fixed point directly on a hash map instead of the reference's
// device/host ping-pong per level.  Ghost tiles follow SpawnGhostTiles
// (PoissonGrid.cu:30-60): an empty face-neighbour position of an active tile
// whose parent is a LEAF.
#include "tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace cg {

int Tree::add(int l, int x, int y, int z, uint8_t t) {
  int id = size();
  level.push_back(l);
  coord.push_back({x, y, z});
  type.push_back(t);
  ctype.resize(size_t(id + 1) * 512, 0);
  index[tile_key(l, x, y, z)] = id;
  return id;
}

void Tree::refine(int t) {
  type[t] = NONLEAF;
  int l = level[t];
  auto c = coord[t];
  for (int o = 0; o < 8; o++) {
    int x = 2 * c[0] + (o & 1), y = 2 * c[1] + ((o >> 1) & 1), z = 2 * c[2] + ((o >> 2) & 1);
    int e = find(l + 1, x, y, z);
    if (e >= 0) type[e] = LEAF;  // a ghost at this position becomes a LEAF
    else add(l + 1, x, y, z, LEAF);
  }
}

void Tree::build(const std::function<int(int, const std::array<int32_t, 3>&)>& target) {
  for (int i = 0; i < nroot[0]; i++)
    for (int j = 0; j < nroot[1]; j++)
      for (int k = 0; k < nroot[2]; k++) add(0, i, j, k, LEAF);
  bool changed = true;
  while (changed) {
    changed = false;
    // target rule
    std::vector<int> stack(size());
    std::iota(stack.begin(), stack.end(), 0);
    while (!stack.empty()) {
      int t = stack.back();
      stack.pop_back();
      if (type[t] != LEAF || level[t] + 1 >= num_layers) continue;
      if (target(level[t], coord[t]) > level[t]) {
        int first = size();
        refine(t);
        changed = true;
        for (int c = first; c < size(); c++) stack.push_back(c);
      }
    }
    // 2:1 face balance: a LEAF at level l needs the level l-1 tile that holds
    // each in-domain face neighbour position
    int maxl = *std::max_element(level.begin(), level.end());
    for (int l = maxl; l >= 2; l--) {
      std::vector<int> lv;
      for (int t = 0; t < size(); t++) if (level[t] == l && type[t] == LEAF) lv.push_back(t);
      for (int t : lv) {
        for (int f = 0; f < 6; f++) {
          std::array<int32_t, 3> n = coord[t];
          n[f / 2] += (f & 1) ? 1 : -1;
          if (!in_domain(l, n[0], n[1], n[2])) continue;
          if (find(l - 1, n[0] >> 1, n[1] >> 1, n[2] >> 1) >= 0) continue;
          int m = l - 2;
          int a = find(m, n[0] >> 2, n[1] >> 2, n[2] >> 2);
          while (a < 0) { m--; a = find(m, n[0] >> (l - m), n[1] >> (l - m), n[2] >> (l - m)); }
          while (level[a] < l - 1) {
            refine(a);
            changed = true;
            int cl = level[a] + 1, s = l - cl;
            a = find(cl, n[0] >> s, n[1] >> s, n[2] >> s);
          }
        }
      }
    }
  }
  spawn_ghosts();
}

void Tree::spawn_ghosts() {
  int n0 = size();
  for (int t = 0; t < n0; t++) {
    if (!(type[t] & (LEAF | NONLEAF)) || level[t] == 0) continue;
    int l = level[t];
    for (int f = 0; f < 6; f++) {
      std::array<int32_t, 3> n = coord[t];
      n[f / 2] += (f & 1) ? 1 : -1;
      if (!key_ok(n[0], n[1], n[2]) || find(l, n[0], n[1], n[2]) >= 0) continue;
      int p = find(l - 1, n[0] >> 1, n[1] >> 1, n[2] >> 1);
      if (p >= 0 && type[p] == LEAF) add(l, n[0], n[1], n[2], GHOST);
    }
  }
}

void Tree::sort_canonical() {
  int n = size();
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<uint64_t> mk(n);
  for (int t = 0; t < n; t++) mk[t] = morton3(uint32_t(coord[t][0]), uint32_t(coord[t][1]), uint32_t(coord[t][2]));
  std::sort(perm.begin(), perm.end(), [&](int a, int b) {
    if (level[a] != level[b]) return level[a] < level[b];
    return mk[a] < mk[b];
  });
  std::vector<int32_t> lv(n);
  std::vector<std::array<int32_t, 3>> cd(n);
  std::vector<uint8_t> ty(n), ct(size_t(n) * 512);
  for (int i = 0; i < n; i++) {
    lv[i] = level[perm[i]];
    cd[i] = coord[perm[i]];
    ty[i] = type[perm[i]];
    std::memcpy(&ct[size_t(i) * 512], &ctype[size_t(perm[i]) * 512], 512);
  }
  level.swap(lv); coord.swap(cd); type.swap(ty); ctype.swap(ct);
  index.clear();
  for (int i = 0; i < n; i++) index[tile_key(level[i], coord[i][0], coord[i][1], coord[i][2])] = i;
}

// CalcCellTypesFromLeafs (PoissonSolver.cu:7-61): a NONLEAF cell is DIRICHLET
// if any child cell is, else INTERIOR if any child cell is, else NEUMANN
// (children taken fine-to-coarse); a GHOST cell takes its parent cell's type.
void Tree::cell_types_from_leafs() {
  int maxl = 0;
  for (int t = 0; t < size(); t++) maxl = std::max(maxl, level[t]);
  for (int l = maxl; l >= 0; l--) {
    for (int t = 0; t < size(); t++) {
      if (level[t] != l || type[t] != NONLEAF) continue;
      int ch[8];
      for (int o = 0; o < 8; o++)
        ch[o] = find(l + 1, 2 * coord[t][0] + (o & 1), 2 * coord[t][1] + ((o >> 1) & 1), 2 * coord[t][2] + ((o >> 2) & 1));
      for (int c = 0; c < 512; c++) {
        int i = c >> 6, j = (c >> 3) & 7, k = c & 7;
        int nd = 0, ni = 0;
        for (int s = 0; s < 8; s++) {
          int fi = 2 * i + (s & 1), fj = 2 * j + ((s >> 1) & 1), fk = 2 * k + ((s >> 2) & 1);
          int o = (fi >> 3) | ((fj >> 3) << 1) | ((fk >> 3) << 2);
          int ct = ch[o];
          if (ct < 0 || !(type[ct] & (LEAF | NONLEAF))) continue;
          uint8_t v = ctype[size_t(ct) * 512 + (((fi & 7) << 6) | ((fj & 7) << 3) | (fk & 7))];
          if (v & DIRICHLET) nd++;
          else if (v & INTERIOR) ni++;
        }
        ctype[size_t(t) * 512 + c] = nd ? DIRICHLET : (ni ? INTERIOR : NEUMANN);
      }
    }
  }
  for (int t = 0; t < size(); t++) {
    if (type[t] != GHOST) continue;
    int l = level[t];
    int p = find(l - 1, coord[t][0] >> 1, coord[t][1] >> 1, coord[t][2] >> 1);
    for (int c = 0; c < 512; c++) {
      if (p < 0) { ctype[size_t(t) * 512 + c] = DIRICHLET; continue; }
      int gi = coord[t][0] * 8 + (c >> 6), gj = coord[t][1] * 8 + ((c >> 3) & 7), gk = coord[t][2] * 8 + (c & 7);
      int pi = (gi >> 1) & 7, pj = (gj >> 1) & 7, pk = (gk >> 1) & 7;
      ctype[size_t(t) * 512 + c] = ctype[size_t(p) * 512 + ((pi << 6) | (pj << 3) | pk)];
    }
  }
}

// ---------------------------------------------------------------------------
// Scenes
// ---------------------------------------------------------------------------
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// iid N(0,1) keyed by (seed, level, global cell, axis): Box-Muller on two
// 53-bit uniforms from a splitmix64 chain.
static double normal_at(uint64_t seed, int level, int gi, int gj, int gk, int axis) {
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
static std::vector<Tube> make_tubes(const cg_scene_params& p) {
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
static double vorticity_mag(const std::vector<Tube>& tubes, const double x[3]) {
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

static int scene_target(const cg_scene_params& p, const std::vector<Tube>& tubes, int level, const std::array<int32_t, 3>& b) {
  double h = p.h0 / double(1 << level), lo[3], hi[3];
  for (int d = 0; d < 3; d++) { lo[d] = b[d] * 8 * h; hi[d] = (b[d] + 1) * 8 * h; }
  switch (p.kind) {
    case CG_SCENE_UNIFORM: return p.uniform_level;
    case CG_SCENE_SPHERE_BAND: {
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
    case CG_SCENE_PLANE: return (lo[2] < p.plane_z) ? p.plane_level[0] : p.plane_level[1];
    case CG_SCENE_VORTEX: {
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

void scene_params_default(cg_scene_params& p) {
  std::memset(&p, 0, sizeof(p));
  p.kind = CG_SCENE_UNIFORM;
  p.nroot[0] = p.nroot[1] = p.nroot[2] = 1;
  p.h0 = 0.125;
  p.uniform_level = 3;
  p.has_solid = 1;
  p.solid_c[0] = p.solid_c[1] = p.solid_c[2] = 0.5;
  p.solid_r = 0.2;
  p.band_c[0] = p.band_c[1] = p.band_c[2] = 0.5;
  p.band_r = 0.15;
  p.band_w[0] = 0.05; p.band_w[1] = 0.15;
  p.band_level[0] = 5; p.band_level[1] = 4; p.band_level[2] = 3;
  p.plane_z = 0.5;
  p.plane_level[0] = 3; p.plane_level[1] = 2;
  p.n_tubes = 8;
  p.tube_r_lo = 0.02; p.tube_r_hi = 0.06;
  p.vortex_thresh = 0.5;
  p.vortex_level[0] = 6; p.vortex_level[1] = 4;
  for (int f = 0; f < 5; f++) p.face_bc[f] = CG_NEUMANN;
  p.face_bc[5] = CG_DIRICHLET;
  p.seed = 0;
  p.num_layers = 10;
}

// Global cell -> (tile, local offset) at a level; -1 if no tile there.
static inline int cell_at(const Tree& T, int l, int gi, int gj, int gk, int& off) {
  int t = T.find(l, gi >> 3, gj >> 3, gk >> 3);
  off = ((gi & 7) << 6) | ((gj & 7) << 3) | (gk & 7);
  return t;
}

void scene_build(const cg_scene_params& p, Scene& s) {
  s.params = p;
  Tree& T = s.tree;
  T = Tree();
  T.h0 = p.h0;
  for (int d = 0; d < 3; d++) T.nroot[d] = p.nroot[d];
  T.num_layers = p.num_layers;
  auto tubes = make_tubes(p);
  T.build([&](int l, const std::array<int32_t, 3>& c) { return scene_target(p, tubes, l, c); });
  T.sort_canonical();
  const int n = T.size();

  // Leaf cell types: a cell with an empty same-level neighbour lies on the
  // domain boundary (NEUMANN wins over DIRICHLET, FluidParams.cu:125-144);
  // otherwise the solid sphere is NEUMANN (FluidParams.cu:241-244).
  for (int t = 0; t < n; t++) {
    if (T.type[t] != LEAF) continue;
    int l = T.level[t];
    double h = p.h0 / double(1 << l);
    for (int c = 0; c < 512; c++) {
      int g[3] = {T.coord[t][0] * 8 + (c >> 6), T.coord[t][1] * 8 + ((c >> 3) & 7), T.coord[t][2] * 8 + (c & 7)};
      bool neu = false, dir = false;
      for (int f = 0; f < 6; f++) {
        int q[3] = {g[0], g[1], g[2]};
        q[f / 2] += (f & 1) ? 1 : -1;
        int off;
        if (cell_at(T, l, q[0], q[1], q[2], off) < 0) {
          if (p.face_bc[f] == CG_DIRICHLET) dir = true; else neu = true;
        }
      }
      uint8_t ty = INTERIOR;
      if (neu) ty = NEUMANN;
      else if (dir) ty = DIRICHLET;
      else if (p.has_solid) {
        double d2 = 0;
        for (int d = 0; d < 3; d++) { double x = (g[d] + 0.5) * h - p.solid_c[d]; d2 += x * x; }
        if (d2 <= p.solid_r * p.solid_r) ty = NEUMANN;
      }
      T.ctype[size_t(t) * 512 + c] = ty;
    }
  }
  T.cell_types_from_leafs();

  // Face velocities on LEAF cells (minus faces), zeroed where a face touches a
  // NEUMANN cell or joins two DIRICHLET cells (FluidEuler.cu:43-60, with the
  // neighbour found as in PoissonGrid.h:451-489).
  s.uvw.assign(size_t(n) * 3 * 512, 0.f);
  auto U = [&](int t, int a, int c) -> float& { return s.uvw[(size_t(t) * 3 + a) * 512 + c]; };
  auto CT = [&](int t, int c) { return T.ctype[size_t(t) * 512 + c]; };
  for (int t = 0; t < n; t++) {
    if (T.type[t] != LEAF) continue;
    int l = T.level[t];
    for (int c = 0; c < 512; c++) {
      int g[3] = {T.coord[t][0] * 8 + (c >> 6), T.coord[t][1] * 8 + ((c >> 3) & 7), T.coord[t][2] * 8 + (c & 7)};
      for (int a = 0; a < 3; a++) U(t, a, c) = float(normal_at(p.seed, l, g[0], g[1], g[2], a));
    }
  }
  for (int t = 0; t < n; t++) {
    if (T.type[t] != LEAF) continue;
    int l = T.level[t];
    for (int c = 0; c < 512; c++) {
      uint8_t t0 = CT(t, c);
      int g[3] = {T.coord[t][0] * 8 + (c >> 6), T.coord[t][1] * 8 + ((c >> 3) & 7), T.coord[t][2] * 8 + (c & 7)};
      for (int a = 0; a < 3; a++) {
        int q[3] = {g[0], g[1], g[2]};
        q[a] -= 1;
        auto hit = [&](uint8_t t1) { return (t0 & NEUMANN) || (t1 & NEUMANN) || ((t0 & DIRICHLET) && (t1 & DIRICHLET)); };
        bool zero = false;
        int off, nt = cell_at(T, l, q[0], q[1], q[2], off);
        if (nt < 0) continue;
        if (T.type[nt] != LEAF) {
          // Not a LEAF: look at the finer cells across this face.  For a GHOST
          // neighbour there are none (a ghost's parent is a LEAF), so faces
          // towards coarser cells are left as they are (PoissonGrid.h:462 tests
          // isLeaf(), which is false for ghosts).
          for (int oj = 0; oj < 2; oj++)
            for (int ok = 0; ok < 2; ok++) {
              int f[3];
              f[a] = 2 * q[a] + 1;
              f[(a + 1) % 3] = 2 * q[(a + 1) % 3] + oj;
              f[(a + 2) % 3] = 2 * q[(a + 2) % 3] + ok;
              int fo, ft = cell_at(T, l + 1, f[0], f[1], f[2], fo);
              if (ft >= 0 && hit(CT(ft, fo))) zero = true;
            }
        } else if (hit(CT(nt, off))) {
          zero = true;
        }
        if (zero) U(t, a, c) = 0.f;
      }
    }
  }
  // ghost velocities from parents (PoissonGrid.cu:635-661; all cells are typed)
  for (int t = 0; t < n; t++) {
    if (T.type[t] != GHOST) continue;
    int l = T.level[t];
    for (int c = 0; c < 512; c++) {
      int g[3] = {T.coord[t][0] * 8 + (c >> 6), T.coord[t][1] * 8 + ((c >> 3) & 7), T.coord[t][2] * 8 + (c & 7)};
      int off, pt = cell_at(T, l - 1, g[0] >> 1, g[1] >> 1, g[2] >> 1, off);
      for (int a = 0; a < 3; a++) U(t, a, c) = (pt >= 0) ? U(pt, a, off) : 0.f;
    }
  }
  // Divergence b = sum of face fluxes * h^2 on INTERIOR cells of LEAF|GHOST
  // tiles (PoissonSolver.cu:72-137), ghost rows then added to their LEAF
  // parents (PoissonSolver.cu:141).
  s.rhs.assign(size_t(n) * 512, 0.f);
  for (int t = 0; t < n; t++) {
    if (!(T.type[t] & (LEAF | GHOST))) continue;
    int l = T.level[t];
    float h = float(p.h0 / double(1 << l));
    uint8_t tt = T.type[t];
    for (int c = 0; c < 512; c++) {
      if (!(CT(t, c) & INTERIOR)) continue;
      int g[3] = {T.coord[t][0] * 8 + (c >> 6), T.coord[t][1] * 8 + ((c >> 3) & 7), T.coord[t][2] * 8 + (c & 7)};
      float sum = 0.f;
      for (int a = 0; a < 3; a++)
        for (int sg = -1; sg <= 1; sg += 2) {
          int q[3] = {g[0], g[1], g[2]};
          q[a] += sg;
          int off, nt = cell_at(T, l, q[0], q[1], q[2], off);
          uint8_t ntt = nt < 0 ? tt : T.type[nt];
          bool ok = (tt == LEAF && (ntt == LEAF || ntt == GHOST)) || (tt == GHOST && ntt == LEAF);
          if (!ok) continue;
          float u = (sg < 0) ? -U(t, a, c) : (nt < 0 ? 0.f : U(nt, a, off));
          sum = std::fma(u * h, h, sum);
        }
      s.rhs[size_t(t) * 512 + c] = sum;
    }
  }
  for (int t = 0; t < n; t++) {
    if (T.type[t] != LEAF) continue;
    int l = T.level[t];
    for (int c = 0; c < 512; c++) {
      if (!(CT(t, c) & INTERIOR)) continue;
      int g[3] = {T.coord[t][0] * 8 + (c >> 6), T.coord[t][1] * 8 + ((c >> 3) & 7), T.coord[t][2] * 8 + (c & 7)};
      float sum = 0.f;
      for (int o = 0; o < 8; o++) {
        int off, ft = cell_at(T, l + 1, 2 * g[0] + (o & 1), 2 * g[1] + ((o >> 1) & 1), 2 * g[2] + ((o >> 2) & 1), off);
        if (ft >= 0 && (CT(ft, off) & INTERIOR)) sum += s.rhs[size_t(ft) * 512 + off];
      }
      s.rhs[size_t(t) * 512 + c] += sum;
    }
  }
}

}  // namespace cg
