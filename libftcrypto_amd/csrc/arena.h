// Device-memory arena of the serving loop: first-fit over a free list keyed
// by offset, 256-byte granules, neighbours coalesced on release.  Plain C++
// (no device calls): the server hands it the base of one hipMalloc of the
// HBM budget, and a request that creates a vector is admitted only when
// alloc succeeds.  Not thread-safe; the server calls it under its lock.
#pragma once
#include <cstddef>
#include <iterator>
#include <map>

namespace ft {

constexpr size_t ALIGN = 256;

struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
  std::map<size_t, size_t> free_;     // offset -> length
  void init(char *b, size_t s) { base = b; size = s; free_.clear(); free_[0] = s; used = 0; }
  bool alloc(size_t bytes, size_t *off) {
    bytes = (bytes + ALIGN - 1) / ALIGN * ALIGN;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes, o = it->first;
        free_.erase(it);
        if (rest) free_[o + bytes] = rest;
        used += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + ALIGN - 1) / ALIGN * ALIGN;
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
    }
  }
};

}  // namespace ft
