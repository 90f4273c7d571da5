// Host build of the serving loop's device-memory arena (arena.h) for the CPU
// tests: the same header the library compiles, driven through a C interface.
#include <cstddef>
#include <iterator>

#include "../../libftcrypto_amd/csrc/arena.h"

extern "C" {

void *arena_new(size_t size) {
  auto *a = new ft::Arena;
  a->init(nullptr, size);
  return a;
}
void arena_delete(void *a) { delete static_cast<ft::Arena *>(a); }
int arena_alloc(void *a, size_t bytes, size_t *off) { return static_cast<ft::Arena *>(a)->alloc(bytes, off) ? 1 : 0; }
void arena_release(void *a, size_t off, size_t bytes) { static_cast<ft::Arena *>(a)->release(off, bytes); }
size_t arena_used(void *a) { return static_cast<ft::Arena *>(a)->used; }
size_t arena_nfree(void *a) { return static_cast<ft::Arena *>(a)->free_.size(); }
void arena_free_block(void *a, size_t i, size_t *off, size_t *len) {
  auto it = static_cast<ft::Arena *>(a)->free_.begin();
  std::advance(it, i);
  *off = it->first;
  *len = it->second;
}
}
