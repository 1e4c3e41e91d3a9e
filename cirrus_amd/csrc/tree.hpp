// Host-side octree of 8^3 tiles: construction (refinement with 2:1 face
// balance and ghost tiles), cell typing and the synthetic benchmark scenes.
#pragma once
#include <array>
#include <cstdint>
#include <functional>
#include <string>
#include <unordered_map>
#include <vector>

#include "cirrus_mgpcg.h"

namespace cg {

constexpr uint8_t LEAF = CG_LEAF, NONLEAF = CG_NONLEAF, GHOST = CG_GHOST;
constexpr uint8_t DIRICHLET = CG_DIRICHLET, INTERIOR = CG_INTERIOR, NEUMANN = CG_NEUMANN;

inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t m = 0;
  for (int b = 0; b < 21; b++) {
    m |= uint64_t((x >> b) & 1u) << (3 * b);
    m |= uint64_t((y >> b) & 1u) << (3 * b + 1);
    m |= uint64_t((z >> b) & 1u) << (3 * b + 2);
  }
  return m;
}

// (level, coord) -> 64-bit key; coordinates must lie in [0, 2^19).
inline bool key_ok(int x, int y, int z) { return x >= 0 && y >= 0 && z >= 0 && x < (1 << 19) && y < (1 << 19) && z < (1 << 19); }
inline uint64_t tile_key(int l, int x, int y, int z) {
  return (uint64_t(l) << 57) | (uint64_t(x) << 38) | (uint64_t(y) << 19) | uint64_t(z);
}

struct Tree {
  double h0 = 0.125;
  int nroot[3] = {1, 1, 1};
  int num_layers = 10;
  std::vector<int32_t> level;
  std::vector<std::array<int32_t, 3>> coord;
  std::vector<uint8_t> type;
  std::vector<uint8_t> ctype;  // 512 per tile
  std::unordered_map<uint64_t, int32_t> index;

  int size() const { return int(level.size()); }
  int find(int l, int x, int y, int z) const {
    if (l < 0 || !key_ok(x, y, z)) return -1;
    auto it = index.find(tile_key(l, x, y, z));
    return it == index.end() ? -1 : it->second;
  }
  bool in_domain(int l, int x, int y, int z) const {
    if (x < 0 || y < 0 || z < 0) return false;
    return (x >> l) < nroot[0] && (y >> l) < nroot[1] && (z >> l) < nroot[2];
  }
  int add(int l, int x, int y, int z, uint8_t t);
  void refine(int t);
  // Refine every LEAF whose target exceeds its level, then enforce 2:1 face
  // balance, until neither rule refines anything; then spawn ghosts.
  void build(const std::function<int(int, const std::array<int32_t, 3>&)>& target);
  void spawn_ghosts();
  // Sort tiles by (level, Morton code) and rebuild the index.
  void sort_canonical();
  // NONLEAF cell types from their children, GHOST cell types from parents.
  void cell_types_from_leafs();
};

// Scene (benchmark workload) construction.
struct Scene {
  cg_scene_params params;
  Tree tree;
  std::vector<float> rhs;  // 512 per tile
  std::vector<float> uvw;  // 3*512 per tile
};
void scene_build(const cg_scene_params& p, Scene& s);
void scene_params_default(cg_scene_params& p);

}  // namespace cg
