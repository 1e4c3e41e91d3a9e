// Host-side checks of the library under AddressSanitizer / UndefinedBehavior-
// Sanitizer or ThreadSanitizer (built and run by tests/test_host_sanitizers.py
// on a machine without a GPU; nothing here launches a kernel).
//
//   host_check SCENES.bin THREADS TILEDIR
//
// SCENES.bin holds an array of cg_scene_params.  The scenes are built on
// THREADS threads at once (as the benchmark builds them); each tree's cell types
// are recomputed from its leaves and compared; a few trees go through a tile
// file and back; then the whole set is handed to cg_problem_create_batch, whose
// host side (trees built on up to 16 threads, canonical order, records,
// worklists, coarse plans and their range checks) runs before its first HIP
// call: without a device that call fails with CG_ERR_HIP, with one the pool is
// created and destroyed.  Last, the serving engine's range allocator is run
// against a bitmap model.  Exit status 0 when every check passes.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "cirrus_mgpcg.h"
#include "range_alloc.hpp"

static int g_fail = 0;
#define CHECK(cond, ...)                         \
  do {                                           \
    if (!(cond)) {                               \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);         \
      std::fprintf(stderr, "\n");                \
      g_fail = 1;                                \
    }                                            \
  } while (0)

static void check_range_alloc() {
  const int64_t cap = 4096;
  cg::RangeAlloc a;
  a.init(cap);
  std::vector<uint8_t> used(cap, 0);
  std::vector<std::pair<int64_t, int64_t>> live;
  std::mt19937_64 rng(1234);
  for (int step = 0; step < 20000; step++) {
    if (live.empty() || rng() % 3 != 0) {
      const int64_t n = 1 + int64_t(rng() % 200);
      const int64_t at = a.alloc(n);
      // the model: lowest start of n free slots in a row
      int64_t want = -1;
      for (int64_t s = 0, run = 0; s < cap; s++) {
        run = used[s] ? 0 : run + 1;
        if (run == n) { want = s - n + 1; break; }
      }
      CHECK(at == want, "alloc(%lld) = %lld, model %lld", (long long)n, (long long)at, (long long)want);
      if (at >= 0) {
        for (int64_t s = at; s < at + n; s++) used[s] = 1;
        live.emplace_back(at, n);
      }
    } else {
      const size_t k = size_t(rng() % live.size());
      const auto r = live[k];
      live.erase(live.begin() + long(k));
      CHECK(a.release(r.first, r.second), "release of a live range refused");
      for (int64_t s = r.first; s < r.first + r.second; s++) used[s] = 0;
      CHECK(!a.release(r.first, r.second), "second release of a range accepted");
    }
  }
  CHECK(!a.release(-1, 4), "release of a negative start accepted");
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: host_check SCENES.bin THREADS TILEDIR\n");
    return 2;
  }
  std::ifstream f(argv[1], std::ios::binary);
  std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const size_t ns = raw.size() / sizeof(cg_scene_params);
  CHECK(ns * sizeof(cg_scene_params) == raw.size() && ns > 0, "scene file size %zu", raw.size());
  std::vector<cg_scene_params> sp(ns);
  std::memcpy(sp.data(), raw.data(), ns * sizeof(cg_scene_params));
  const int nth = std::max(1, std::atoi(argv[2]));
  const std::string dir = argv[3];

  // scenes on several threads at once
  std::vector<cg_scene*> scenes(ns, nullptr);
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; t++)
    pool.emplace_back([&]() {
      for (size_t q = next++; q < ns; q = next++) {
        const cg_status st = cg_scene_create(&sp[q], &scenes[q]);
        CHECK(st == CG_OK, "scene %zu: %s", q, cg_last_error());
      }
    });
  for (auto& th : pool) th.join();

  std::vector<cg_tree_desc> descs(ns);
  size_t tiles = 0;
  for (size_t q = 0; q < ns; q++) {
    if (!scenes[q]) continue;
    const float* rhs = nullptr;
    CHECK(cg_scene_tree(scenes[q], &descs[q], &rhs) == CG_OK && rhs, "scene %zu tree", q);
    const cg_tree_desc& d = descs[q];
    tiles += size_t(d.num_tiles);
    // cell types of NONLEAF and GHOST tiles from the leaves, as built
    std::vector<uint8_t> ct(d.cell_type, d.cell_type + size_t(d.num_tiles) * 512);
    for (int t = 0; t < d.num_tiles; t++)
      if (d.type[t] != CG_LEAF) std::memset(&ct[size_t(t) * 512], 0, 512);
    CHECK(cg_calc_cell_types_from_leafs(&d, ct.data()) == CG_OK, "scene %zu cell types: %s", q, cg_last_error());
    CHECK(std::memcmp(ct.data(), d.cell_type, ct.size()) == 0, "scene %zu: cell types differ", q);
    // a tile file round trip for every eighth scene
    if (q % 8 == 0) {
      const std::string path = dir + "/scene" + std::to_string(q) + ".ctf";
      const int32_t id = 1;
      const float* chans[1] = {rhs};
      CHECK(cg_tiles_save(path.c_str(), &d, 1, &id, chans) == CG_OK, "save %zu: %s", q, cg_last_error());
      cg_tileset* ts = nullptr;
      CHECK(cg_tiles_load(path.c_str(), &ts) == CG_OK, "load %zu: %s", q, cg_last_error());
      if (ts) {
        cg_tree_desc e;
        int32_t m = 0;
        const int32_t* ids = nullptr;
        const float* data = nullptr;
        CHECK(cg_tileset_tree(ts, &e) == CG_OK && e.num_tiles == d.num_tiles, "tileset %zu tree", q);
        CHECK(cg_tileset_channels(ts, &m, &ids, &data) == CG_OK && m == 1 && ids[0] == 1, "tileset %zu channels", q);
        CHECK(std::memcmp(e.cell_type, d.cell_type, size_t(d.num_tiles) * 512) == 0, "tileset %zu cell types", q);
        CHECK(std::memcmp(data, rhs, size_t(d.num_tiles) * 512 * sizeof(float)) == 0, "tileset %zu rhs", q);
        cg_tileset_destroy(ts);
      }
      // a corrupted file is refused
      std::fstream g(path, std::ios::in | std::ios::out | std::ios::binary);
      g.seekp(100);
      g.put(char(0x5a));
      g.close();
      ts = nullptr;
      CHECK(cg_tiles_load(path.c_str(), &ts) != CG_OK, "corrupted tile file %zu accepted", q);
    }
  }

  // the pool's host planning, then (without a device) the first HIP call fails
  cg_problem* p = nullptr;
  const cg_status st = cg_problem_create_batch(descs.data(), int32_t(ns), &p);
  CHECK(st == CG_OK || st == CG_ERR_HIP, "pool of %zu problems: status %d: %s", ns, int(st), cg_last_error());
  if (p) cg_problem_destroy(p);

  check_range_alloc();
  for (cg_scene* s : scenes)
    if (s) cg_scene_destroy(s);
  std::printf("host_check: %zu scenes, %zu tiles, pool status %d, %s\n", ns, tiles, int(st), g_fail ? "FAILED" : "ok");
  return g_fail;
}
