// First-fit allocator of index ranges (host side of the serving engine: tile,
// leaf-slot and ghost-slot ranges of resident instances).
#pragma once
#include <cstdint>
#include <iterator>
#include <map>

namespace cg {

// First-fit allocator of index ranges [0, cap): alloc returns the start of the
// lowest free range of length >= n (-1 if none); release merges neighbours.
struct RangeAlloc {
  std::map<int64_t, int64_t> free_;  // start -> length
  void init(int64_t cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  int64_t alloc(int64_t n) {
    if (n <= 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const int64_t at = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[at + n] = len - n;
      return at;
    }
    return -1;
  }
  // false (and no change) if [at, at + n) overlaps a free range: a range
  // released twice or never allocated
  bool release(int64_t at, int64_t n) {
    if (n <= 0) return true;
    if (at < 0) return false;
    auto after = free_.lower_bound(at);
    if (after != free_.end() && after->first < at + n) return false;
    if (after != free_.begin()) {
      auto before = std::prev(after);
      if (before->first + before->second > at) return false;
    }
    auto it = free_.emplace(at, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
    return true;
  }
};

}  // namespace cg
