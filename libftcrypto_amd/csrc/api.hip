// C ABI of the library: the reference's 38 functions (include/ftcrypto.h)
// and the extensions (include/ftcrypto_ext.h), as C++ host code over the
// gfx950 kernels.
//
// State: one device (current at ft_crypto_init), one library stream at
// default priority, the base point's comb table, and a grow-only device
// staging buffer for host inputs.  Every reference-compatible call enqueues
// on the library stream in issue order (as the reference's legacy default
// stream does) and has consumed its host inputs when it returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <thread>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ftcrypto.h"
#include "../../include/ftcrypto_ext.h"
#include "kernels.h"
#include "misc.h"

struct ft_ge25519_array_s {
  uint32_t *d;      // 4 planes x cap x 32 bytes
  size_t n;
  size_t cap;
};

struct ftx_key_s {
  uint32_t *table;    // device comb table of H
  uint32_t *table32;  // device radix-32 comb table of H
};

struct ftx_graph_s {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace {

using namespace ft;

struct Lib {
  bool init = false;
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t *btab = nullptr;        // comb table of B
  uint32_t *btab32 = nullptr;      // radix-32 comb table of B
  int8_t *lnaf = nullptr;          // width-3 NAF digits of the group order l
  int lnaf_len = 0;
  unsigned long long *flag = nullptr;
  void *stage = nullptr;           // host-input staging
  size_t stage_bytes = 0;
  std::mutex mu;                   // initialisation
  std::mutex io;                   // the staging buffer and the flag word
};
Lib G;

// Calls that use the shared staging buffer or flag word hold G.io from start
// to end, so host threads calling the library at once do not reuse either
// while another call's work still reads it (growing the staging buffer frees
// the old one).
#define FT_IO_LOCK std::lock_guard<std::mutex> io_lock(G.io)

const char *ERR_STR[FT_ERR_MAX_ERR + 1] = {
    "No error",
    "Out of memory",
    "Size mismatch",
    "Bad size",
    "Index out of bounds",
    "Invalid point",
    "Failed to initialise sodium library",
    "Insufficient space",
    "Bad ciphertext",
    "Internal library error",
    "Insufficient device or driver",
};

inline PtView view(const_ft_ge25519_array a) { return PtView{a->d, a->cap}; }
inline hipStream_t S(void *s) { return s ? (hipStream_t)s : G.stream; }
inline ft_error herr(hipError_t e) {
  if (e == hipSuccess) return FT_ERR_NO_ERROR;
  if (e == hipErrorOutOfMemory) return FT_ERR_OUT_OF_MEMORY;
  return FT_ERR_INTERNAL_LIBRARY_ERROR;
}
#define CK(expr)                                   \
  do {                                             \
    hipError_t e_ = (expr);                        \
    if (e_ != hipSuccess) return herr(e_);         \
  } while (0)
#define CKE(expr)                                  \
  do {                                             \
    ft_error f_ = (expr);                          \
    if (f_ != FT_ERR_NO_ERROR) return f_;          \
  } while (0)

ft_error alloc_array(ft_ge25519_array *out, size_t n) {
  auto *a = new (std::nothrow) ft_ge25519_array_s;
  if (!a) return FT_ERR_OUT_OF_MEMORY;
  a->n = n;
  a->cap = n ? n : 1;
  if (hipMalloc(&a->d, a->cap * 128) != hipSuccess) {
    (void)hipGetLastError();
    delete a;
    return FT_ERR_OUT_OF_MEMORY;
  }
  *out = a;
  return FT_ERR_NO_ERROR;
}

// device staging of host input (stream-ordered; the host buffer is free on return)
ft_error stage(const void *h, size_t bytes, void **d) {
  if (bytes > G.stage_bytes) {
    CK(hipStreamSynchronize(G.stream));
    if (G.stage) (void)hipFree(G.stage);
    G.stage = nullptr;
    size_t want = std::max<size_t>(bytes, 1 << 20);
    if (hipMalloc(&G.stage, want) != hipSuccess) {
      (void)hipGetLastError();
      G.stage_bytes = 0;
      return FT_ERR_OUT_OF_MEMORY;
    }
    G.stage_bytes = want;
  }
  if (bytes) {
    CK(hipMemcpyAsync(G.stage, h, bytes, hipMemcpyHostToDevice, G.stream));
    CK(hipStreamSynchronize(G.stream));
  }
  *d = G.stage;
  return FT_ERR_NO_ERROR;
}

// a temporary device buffer released by the destructor after a stream sync
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) {
      (void)hipStreamSynchronize(G.stream);
      (void)hipFree(p);
    }
  }
  ft_error alloc(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return FT_ERR_OUT_OF_MEMORY;
    }
    return FT_ERR_NO_ERROR;
  }
  ft_error upload(const void *h, size_t bytes) {
    CKE(alloc(bytes));
    if (bytes) {
      CK(hipMemcpyAsync(p, h, bytes, hipMemcpyHostToDevice, G.stream));
      CK(hipStreamSynchronize(G.stream));
    }
    return FT_ERR_NO_ERROR;
  }
};

struct ArrGuard {
  ft_ge25519_array a = nullptr;
  ~ArrGuard() { if (a) ft_array_free(a); }
  ft_ge25519_array release() { auto r = a; a = nullptr; return r; }
};

// out[0] = sum of src; part (reduce_scratch(n) points) holds the block sums
size_t reduce_scratch(size_t n) { return reduce_grid(n, REDUCE_BLOCKS_MAX); }

ft_error reduce_run_blocks(ft_ge25519_array out, const_ft_ge25519_array src, PtView part, unsigned max_blocks,
                           hipStream_t s) {
  const size_t n = src->n;
  if (n == 0) { CK(launch_set_identity(view(out), 0, 1, s)); return FT_ERR_NO_ERROR; }
  const unsigned g1 = reduce_grid(n, max_blocks);
  CK(launch_reduce_blocks(part, view(src), n, g1, s));
  CK(launch_reduce_blocks(view(out), part, g1, 1, s));
  return FT_ERR_NO_ERROR;
}

ft_error reduce_run(ft_ge25519_array out, const_ft_ge25519_array src, PtView part, hipStream_t s) {
  return reduce_run_blocks(out, src, part, (unsigned)tune_for(src->n).reduce, s);
}

ft_error reduce_into(ft_ge25519_array out, const_ft_ge25519_array src, hipStream_t s) {
  ArrGuard part;
  CKE(alloc_array(&part.a, reduce_scratch(src->n)));
  CKE(reduce_run(out, src, view(part.a), s));
  CK(hipStreamSynchronize(s));
  return FT_ERR_NO_ERROR;
}

ft_error scan_rec(PtView v, size_t n, bool exclusive, hipStream_t s) {
  if (n == 0) return FT_ERR_NO_ERROR;
  ArrGuard scr;
  CKE(alloc_array(&scr.a, scan_scratch_points(n)));
  CK(scan_launch(v, n, exclusive, view(scr.a), s));
  CK(hipStreamSynchronize(s));
  return FT_ERR_NO_ERROR;
}

size_t seg_scratch_points(size_t n) { return segsum_scratch_points(n); }

ft_error segsum_run(ft_ge25519_array out, const int64_t *d_off, size_t nseg, size_t n_elems,
                    const int32_t *d_gather, const_ft_ge25519_array src, PtView scratch, hipStream_t s) {
  CK(segsum_launch(view(out), d_off, nseg, n_elems, d_gather, view(src), src->n, scratch, s));
  return FT_ERR_NO_ERROR;
}

// 32-bit-word helpers for the scalar generator / reducer
const uint32_t L_W[8] = {0x5cf5d3edu, 0x5812631au, 0xa2f79cd6u, 0x14def9deu, 0, 0, 0, 0x10000000u};

// Width-3 NAF of l, lowest digit first: digits in {0, +-1, +-3}, every
// nonzero digit followed by at least two zeros; returns the digit count.
// l is public, so the validation kernel may branch on these digits.
static int naf3_of_l(int8_t d[256]) {
  uint32_t k[9] = {0};
  for (int q = 0; q < 8; ++q) k[q] = L_W[q];
  int len = 0;
  auto nonzero = [&k] {
    for (int q = 0; q < 9; ++q)
      if (k[q]) return true;
    return false;
  };
  while (nonzero()) {
    int v = 0;
    if (k[0] & 1u) {
      v = (int)(k[0] & 7u);
      if (v > 4) v -= 8;
      // k -= v (v odd, |v| <= 3): borrow or carry through the words
      uint64_t c = (uint64_t)(v < 0 ? -v : v);
      for (int q = 0; q < 9 && c; ++q) {
        if (v > 0) {
          const uint64_t t = (uint64_t)k[q] - c;
          k[q] = (uint32_t)t;
          c = (t >> 63) & 1;
        } else {
          const uint64_t t = (uint64_t)k[q] + c;
          k[q] = (uint32_t)t;
          c = t >> 32;
        }
      }
    }
    d[len++] = (int8_t)v;
    for (int q = 0; q < 8; ++q) k[q] = (k[q] >> 1) | (k[q + 1] << 31);
    k[8] >>= 1;
  }
  return len;
}

struct Barrett {
  uint32_t mu[9];   // floor(2^512 / l), 261 bits
  Barrett() {
    // long division of 2^512 by l, bit by bit
    uint32_t r[9] = {0}, q[17] = {0};
    for (int bit = 512; bit >= 0; --bit) {
      uint32_t c = (bit == 512) ? 1u : 0u;
      for (int k = 0; k < 9; ++k) { uint32_t nc = r[k] >> 31; r[k] = (r[k] << 1) | c; c = nc; }
      int ge = 1;
      for (int k = 8; k >= 0; --k) {
        uint32_t lk = k < 8 ? L_W[k] : 0;
        if (r[k] != lk) { ge = r[k] > lk; break; }
      }
      if (ge) {
        uint64_t br = 0;
        for (int k = 0; k < 9; ++k) {
          uint64_t lk = k < 8 ? L_W[k] : 0;
          uint64_t d = (uint64_t)r[k] - lk - br;
          r[k] = (uint32_t)d;
          br = (d >> 63) & 1;
        }
        q[bit / 32] |= 1u << (bit % 32);
      }
    }
    for (int k = 0; k < 9; ++k) mu[k] = q[k];
  }
  // x: 16 words (512 bits) -> x mod l in 8 words
  void reduce(uint32_t out[8], const uint32_t x[16]) const {
    // q1 = x >> 252 (9 words)
    uint32_t q1[9];
    for (int k = 0; k < 9; ++k) {
      const int bit = 252 + 32 * k, j = bit / 32, sft = bit % 32;
      uint64_t v = (uint64_t)x[j] >> sft;
      if (j + 1 < 16) v |= (uint64_t)x[j + 1] << (32 - sft);
      q1[k] = (uint32_t)v;
    }
    // q2 = q1 * mu (18 words); q3 = q2 >> 260
    uint32_t q2[19] = {0};
    for (int i = 0; i < 9; ++i) {
      uint64_t c = 0;
      for (int j = 0; j < 9; ++j) {
        uint64_t t = (uint64_t)q1[i] * mu[j] + q2[i + j] + c;
        q2[i + j] = (uint32_t)t;
        c = t >> 32;
      }
      q2[i + 9] = (uint32_t)c;
    }
    uint32_t q3[9];
    for (int k = 0; k < 9; ++k) {
      const int bit = 260 + 32 * k, j = bit / 32, sft = bit % 32;
      uint64_t v = (uint64_t)q2[j] >> sft;
      if (j + 1 < 19) v |= (uint64_t)q2[j + 1] << (32 - sft);
      q3[k] = (uint32_t)v;
    }
    // r = x - q3 * l  (mod 2^288); q - q3 <= 3 (x / l - q1 mu / 2^260 < 2)
    uint32_t ql[18] = {0};
    for (int i = 0; i < 9; ++i) {
      uint64_t c = 0;
      for (int j = 0; j < 8; ++j) {
        uint64_t t = (uint64_t)q3[i] * L_W[j] + ql[i + j] + c;
        ql[i + j] = (uint32_t)t;
        c = t >> 32;
      }
      ql[i + 8] = (uint32_t)c;
    }
    uint32_t r[9];
    uint64_t br = 0;
    for (int k = 0; k < 9; ++k) {
      uint64_t d = (uint64_t)x[k] - ql[k] - br;
      r[k] = (uint32_t)d;
      br = (d >> 63) & 1;
    }
    for (int it = 0; it < 3; ++it) {
      int ge = 1;
      for (int k = 8; k >= 0; --k) {
        uint32_t lk = k < 8 ? L_W[k] : 0;
        if (r[k] != lk) { ge = r[k] > lk; break; }
      }
      if (!ge) break;
      uint64_t b2 = 0;
      for (int k = 0; k < 9; ++k) {
        uint64_t lk = k < 8 ? L_W[k] : 0;
        uint64_t d = (uint64_t)r[k] - lk - b2;
        r[k] = (uint32_t)d;
        b2 = (d >> 63) & 1;
      }
    }
    for (int k = 0; k < 8; ++k) out[k] = r[k];
  }
};

const Barrett &barrett() {
  static Barrett b;
  return b;
}

inline uint64_t splitmix_at(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void words_to_bytes(uint8_t *o, const uint32_t *w, int nw) {
  for (int k = 0; k < nw; ++k)
    for (int j = 0; j < 4; ++j) o[4 * k + j] = (uint8_t)(w[k] >> (8 * j));
}

// radix-16 message digits for a public bound of `bits` bits (0 or 64: any
// uint64, 17 digits); the top digit keeps the final carry
int msg_digits(int bits) { return bits <= 0 || bits >= 64 ? 17 : (bits + 3) / 4 + 1; }

}  // namespace

namespace ft {
const uint32_t *base_table() { return G.btab; }
const uint32_t *base_table32() { return G.btab32; }
const uint32_t *key_table(const ftx_key_s *key) { return key->table; }
const uint32_t *key_table32(const ftx_key_s *key) { return key->table32; }
}  // namespace ft

extern "C" {

// ------------------------------------------------------------------ init

const char *ft_error_str(ft_error err) {
  if (err < 0 || err > FT_ERR_MAX_ERR) return "Unknown error";
  return ERR_STR[err];
}

ft_error ft_crypto_init(void) {
  std::lock_guard<std::mutex> lk(G.mu);
  if (G.init) return FT_ERR_NO_ERROR;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return FT_ERR_INSUFFICIENT_DEVICE_OR_DRIVER;
  }
  CK(hipGetDevice(&G.device));
  hipDeviceProp_t p;
  CK(hipGetDeviceProperties(&p, G.device));
  if (std::string(p.gcnArchName).rfind("gfx950", 0) != 0) return FT_ERR_INSUFFICIENT_DEVICE_OR_DRIVER;
  CK(hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking));
  CK(hipMalloc(&G.btab, COMB_TABLE_WORDS * 4));
  CK(hipMalloc(&G.btab32, COMB32_TABLE_WORDS * 4));
  static int8_t naf[256];
  G.lnaf_len = naf3_of_l(naf);
  CK(hipMalloc(&G.lnaf, sizeof naf));
  CK(hipMalloc(&G.flag, 8));
  CK(hipMemcpyAsync(G.lnaf, naf, sizeof naf, hipMemcpyHostToDevice, G.stream));
  CK(launch_build_comb(G.btab, PtView{nullptr, 0}, 0, true, G.stream));
  CK(launch_build_comb32(G.btab32, PtView{nullptr, 0}, 0, true, G.stream));
  CK(hipStreamSynchronize(G.stream));
  CK(noalloc_resolve(G.device));
  tune_init_from_env();
  G.init = true;
  return FT_ERR_NO_ERROR;
}

ft_error ft_crypto_device_memory(size_t *free_mem, size_t *total_mem) {
  CK(hipMemGetInfo(free_mem, total_mem));
  return FT_ERR_NO_ERROR;
}

// ------------------------------------------------------------ allocation

ft_error ftx_array_alloc(ft_ge25519_array *array, size_t n) { return alloc_array(array, n); }

ft_error ftx_array_device_view(const_ft_ge25519_array a, void **planes, size_t *capacity) {
  *planes = a->d;
  *capacity = a->cap;
  return FT_ERR_NO_ERROR;
}

void *ftx_default_stream(void) { return G.stream; }

ft_error ftx_synchronize(void *stream) {
  CK(hipStreamSynchronize(S(stream)));
  return FT_ERR_NO_ERROR;
}

void ft_array_free(ft_ge25519_array array) {
  if (!array) return;
  (void)hipStreamSynchronize(G.stream);
  (void)hipFree(array->d);
  delete array;
}

size_t ft_array_get_length(const_ft_ge25519_array array) { return array->n; }

ft_error ft_array_set_length(ft_ge25519_array a, size_t newlen) {
  // ftcrypto.h:296-308: shrink in place unless newlen <= len / 2; grow with zeros
  const size_t len = a->n;
  if (newlen <= len && newlen > len / 2) { a->n = newlen; return FT_ERR_NO_ERROR; }
  if (newlen > len && newlen <= a->cap) {
    CK(launch_set_identity(view(a), len, newlen, G.stream));
    a->n = newlen;
    return FT_ERR_NO_ERROR;
  }
  ft_ge25519_array b;
  CKE(alloc_array(&b, newlen));
  const size_t keep = std::min(len, newlen);
  CK(launch_copy(view(b), view(a), keep, G.stream));
  CK(launch_set_identity(view(b), keep, newlen, G.stream));
  CK(hipStreamSynchronize(G.stream));
  (void)hipFree(a->d);
  a->d = b->d;
  a->cap = b->cap;
  a->n = newlen;
  delete b;
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_init_scalar(ft_ge25519_array *array, size_t n, const uint8_t *scalar) {
  FT_IO_LOCK;
  ArrGuard g;
  CKE(alloc_array(&g.a, n));
  if (n) {
    if (!scalar) {
      CK(launch_set_identity(view(g.a), 0, n, G.stream));
    } else {
      void *d;
      CKE(stage(scalar, 32, &d));
      CK(launch_fixed_base(view(g.a), d, false, 1, G.btab32, G.stream));
      CK(launch_broadcast(view(g.a), n, view(g.a), 0, G.stream));
    }
  }
  CK(hipStreamSynchronize(G.stream));
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_init_point(ft_ge25519_array *array, size_t n, const_ft_ge25519_array src, size_t idx) {
  if (idx >= src->n) return FT_ERR_OUT_OF_BOUNDS;
  ArrGuard g;
  CKE(alloc_array(&g.a, n));
  CK(launch_broadcast(view(g.a), n, view(src), idx, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

// ------------------------------------------------------------ data movement

static ft_error check_idxs(const size_t *idxs, size_t n, size_t bound) {
  for (size_t i = 0; i < n; ++i)
    if (idxs[i] >= bound) return FT_ERR_OUT_OF_BOUNDS;
  return FT_ERR_NO_ERROR;
}

// Copy n host indices to the device and range-check them there (one pass
// over the caller's buffer): FT_ERR_OUT_OF_BOUNDS when one is >= bound.
// Uses the flag word: the caller holds FT_IO_LOCK.
static ft_error upload_idxs(DevBuf &di, const size_t *idxs, size_t n, size_t bound) {
  CKE(di.upload(idxs, n * sizeof(size_t)));
  if (!n) return FT_ERR_NO_ERROR;
  CK(keys_out_of_range(G.flag, static_cast<const uint64_t *>(di.p), n, bound, G.stream));
  unsigned long long bad = 0;
  CK(hipMemcpyAsync(&bad, G.flag, 8, hipMemcpyDeviceToHost, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return bad ? FT_ERR_OUT_OF_BOUNDS : FT_ERR_NO_ERROR;
}

ft_error ft_array_get_subset(ft_ge25519_array *array, const size_t *idxs, size_t n_idxs,
                             const_ft_ge25519_array src) {
  FT_IO_LOCK;
  DevBuf di;
  CKE(upload_idxs(di, idxs, n_idxs, src->n));
  ArrGuard g;
  CKE(alloc_array(&g.a, n_idxs));
  CK(launch_gather(view(g.a), view(src), (const uint64_t *)di.p, 0, 0, n_idxs, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

static size_t range_count(size_t start, size_t stop, size_t step) {
  if (stop <= start || step == 0) return 0;
  return (stop - start + step - 1) / step;
}

ft_error ft_array_get_range(ft_ge25519_array *array, size_t start, size_t stop, size_t step,
                            const_ft_ge25519_array src) {
  if (step == 0) return FT_ERR_BAD_SIZE;
  if (stop > src->n || start > stop) return FT_ERR_OUT_OF_BOUNDS;
  const size_t n = range_count(start, stop, step);
  ArrGuard g;
  CKE(alloc_array(&g.a, n));
  CK(launch_gather(view(g.a), view(src), nullptr, start, step, n, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_assign_to_subset(ft_ge25519_array array, const size_t *idxs, size_t n_idxs,
                                   const_ft_ge25519_array src) {
  if (n_idxs != src->n) return FT_ERR_SIZE_MISMATCH;
  FT_IO_LOCK;
  DevBuf di;
  CKE(upload_idxs(di, idxs, n_idxs, array->n));
  CK(launch_scatter(view(array), view(src), (const uint64_t *)di.p, 0, 0, n_idxs, G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_assign_zero_to_subset(ft_ge25519_array array, const size_t *idxs, size_t n_idxs) {
  FT_IO_LOCK;
  DevBuf di;
  CKE(upload_idxs(di, idxs, n_idxs, array->n));
  CK(launch_scatter(view(array), PtView{nullptr, 0}, (const uint64_t *)di.p, 0, 0, n_idxs, G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_assign_to_range(ft_ge25519_array array, size_t start, size_t stop, size_t step,
                                  const_ft_ge25519_array src) {
  if (step == 0) return FT_ERR_BAD_SIZE;
  if (stop > array->n || start > stop) return FT_ERR_OUT_OF_BOUNDS;
  const size_t n = range_count(start, stop, step);
  if (n != src->n) return FT_ERR_SIZE_MISMATCH;
  CK(launch_scatter(view(array), view(src), nullptr, start, step, n, G.stream));
  return FT_ERR_NO_ERROR;
}

static ft_error keep_only(ft_ge25519_array array, const std::vector<uint8_t> &drop) {
  std::vector<size_t> keep;
  keep.reserve(array->n);
  for (size_t i = 0; i < array->n; ++i)
    if (!drop[i]) keep.push_back(i);
  ft_ge25519_array b;
  CKE(alloc_array(&b, keep.size()));
  ArrGuard g;
  g.a = b;
  DevBuf di;
  CKE(di.upload(keep.data(), keep.size() * sizeof(size_t)));
  CK(launch_gather(view(b), view(array), (const uint64_t *)di.p, 0, 0, keep.size(), G.stream));
  CK(hipStreamSynchronize(G.stream));
  std::swap(array->d, b->d);
  std::swap(array->cap, b->cap);
  array->n = keep.size();
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_delete_subset(ft_ge25519_array array, const size_t *idxs, size_t n_idxs) {
  CKE(check_idxs(idxs, n_idxs, array->n));
  std::vector<uint8_t> drop(array->n, 0);
  for (size_t i = 0; i < n_idxs; ++i) drop[idxs[i]] = 1;
  return keep_only(array, drop);
}

ft_error ft_array_delete_range(ft_ge25519_array array, size_t start, size_t stop, size_t step) {
  if (step == 0) return FT_ERR_BAD_SIZE;
  if (stop > array->n || start > stop) return FT_ERR_OUT_OF_BOUNDS;
  std::vector<uint8_t> drop(array->n, 0);
  for (size_t i = start; i < stop; i += step) drop[i] = 1;
  return keep_only(array, drop);
}

ft_error ft_array_mux(ft_ge25519_array iftrue, const int64_t *cond, const_ft_ge25519_array iffalse) {
  if (iftrue->n != iffalse->n) return FT_ERR_SIZE_MISMATCH;
  DevBuf dc;
  CKE(dc.upload(cond, iftrue->n * sizeof(int64_t)));
  CK(launch_mux(view(iftrue), (const int64_t *)dc.p, view(iffalse), iftrue->n, G.stream));
  return FT_ERR_NO_ERROR;
}

// ------------------------------------------------------------ serialisation

ft_error ftx_encode_dev(void *d_out, const_ft_ge25519_array array, void *stream) {
  CK(launch_encode((uint32_t *)d_out, view(array), array->n, S(stream)));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_decode_dev(ft_ge25519_array out, const void *d_in, void *d_fail, void *stream) {
  // ftcrypto.cu:718-790 with device input: *d_fail = 1 + largest failing index
  if (!d_fail) return FT_ERR_BAD_SIZE;
  CK(launch_decode(view(out), (const uint32_t *)d_in, out->n, (unsigned long long *)d_fail, S(stream)));
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_to_bytes_folded(uint8_t *mem, size_t n_bytes, const_ft_ge25519_array array) {
  if (n_bytes / FT_FOLDED_POINT_BYTES < array->n) return FT_ERR_INSUFFICIENT_SPACE;
  DevBuf out;
  CKE(out.alloc(array->n * 32));
  CK(launch_encode((uint32_t *)out.p, view(array), array->n, G.stream));
  CK(hipMemcpyAsync(mem, out.p, array->n * 32, hipMemcpyDeviceToHost, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_to_bytes_affine(uint8_t *mem, size_t n_bytes, const_ft_ge25519_array array) {
  if (n_bytes / FT_AFFINE_POINT_BYTES < array->n) return FT_ERR_INSUFFICIENT_SPACE;
  DevBuf out;
  CKE(out.alloc(array->n * 64));
  CK(launch_to_affine((uint32_t *)out.p, view(array), array->n, G.stream));
  CK(hipMemcpyAsync(mem, out.p, array->n * 64, hipMemcpyDeviceToHost, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_to_bytes(uint8_t *mem, size_t n_bytes, const_ft_ge25519_array array) {
  if (n_bytes / FT_EXTENDED_POINT_BYTES < array->n) return FT_ERR_INSUFFICIENT_SPACE;
  DevBuf out;
  CKE(out.alloc(array->n * 160));
  CK(launch_to_raw((uint32_t *)out.p, view(array), array->n, G.stream));
  CK(hipMemcpyAsync(mem, out.p, array->n * 160, hipMemcpyDeviceToHost, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_from_bytes_folded(int64_t *res, ft_ge25519_array *array, const uint8_t *mem, size_t n_bytes) {
  FT_IO_LOCK;
  if (n_bytes % FT_FOLDED_POINT_BYTES) return FT_ERR_BAD_SIZE;
  const size_t n = n_bytes / 32;
  ArrGuard g;
  CKE(alloc_array(&g.a, n));
  DevBuf in;
  CKE(in.upload(mem, n_bytes));
  CK(launch_decode(view(g.a), (const uint32_t *)in.p, n, G.flag, G.stream));
  unsigned long long bad = 0;
  CK(hipMemcpyAsync(&bad, G.flag, 8, hipMemcpyDeviceToHost, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *res = (int64_t)bad - 1;
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_from_bytes_affine(ft_ge25519_array *array, const uint8_t *mem, size_t n_bytes) {
  if (n_bytes % FT_AFFINE_POINT_BYTES) return FT_ERR_BAD_SIZE;
  const size_t n = n_bytes / 64;
  ArrGuard g;
  CKE(alloc_array(&g.a, n));
  DevBuf in;
  CKE(in.upload(mem, n_bytes));
  CK(launch_from_affine(view(g.a), (const uint32_t *)in.p, n, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_from_bytes(ft_ge25519_array *array, const uint8_t *mem, size_t n_bytes) {
  if (n_bytes % FT_EXTENDED_POINT_BYTES) return FT_ERR_BAD_SIZE;
  const size_t n = n_bytes / 160;
  ArrGuard g;
  CKE(alloc_array(&g.a, n));
  DevBuf in;
  CKE(in.upload(mem, n_bytes));
  CK(launch_from_raw(view(g.a), (const uint32_t *)in.p, n, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_validate(int64_t *res, const_ft_ge25519_array array) {
  FT_IO_LOCK;
  CK(launch_decode(PtView{nullptr, 0}, nullptr, 0, G.flag, G.stream));    // clears the flag
  CK(launch_validate(G.flag, view(array), array->n, G.lnaf, G.lnaf_len, G.stream));
  unsigned long long bad = 0;
  CK(hipMemcpyAsync(&bad, G.flag, 8, hipMemcpyDeviceToHost, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *res = (int64_t)bad - 1;
  return FT_ERR_NO_ERROR;
}

// ------------------------------------------------------------ scalar mults

ft_error ft_array_from_scalars(ft_ge25519_array *array, const uint8_t *scalars, size_t n) {
  FT_IO_LOCK;
  ArrGuard g;
  CKE(alloc_array(&g.a, n));
  void *d;
  CKE(stage(scalars, n * 32, &d));
  CK(launch_fixed_base(view(g.a), d, false, n, G.btab32, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_from_small_scalars(ft_ge25519_array *array, const uint64_t *small, size_t n) {
  FT_IO_LOCK;
  ArrGuard g;
  CKE(alloc_array(&g.a, n));
  void *d;
  CKE(stage(small, n * 8, &d));
  CK(launch_fixed_base(view(g.a), d, true, n, G.btab32, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *array = g.release();
  return FT_ERR_NO_ERROR;
}

void ft_scale(ft_ge25519_array array, const uint8_t *scalar) {
  FT_IO_LOCK;
  void *d;
  if (stage(scalar, 32, &d) != FT_ERR_NO_ERROR) return;
  (void)launch_varbase(view(array), (const uint32_t *)d, 0, array->n, G.stream);
  (void)hipStreamSynchronize(G.stream);
}

ft_error ft_mul(ft_ge25519_array array, const uint8_t *scalars, size_t n) {
  FT_IO_LOCK;
  if (n != array->n) return FT_ERR_SIZE_MISMATCH;
  void *d;
  CKE(stage(scalars, n * 32, &d));
  CK(launch_varbase(view(array), (const uint32_t *)d, 8, n, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return FT_ERR_NO_ERROR;
}

// ------------------------------------------------------------ group ops

ft_error ft_add(ft_ge25519_array res, const_ft_ge25519_array summand) {
  if (res->n != summand->n) return FT_ERR_SIZE_MISMATCH;
  CK(launch_add(view(res), view(summand), res->n, false, G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ft_sub(ft_ge25519_array res, const_ft_ge25519_array summand) {
  if (res->n != summand->n) return FT_ERR_SIZE_MISMATCH;
  CK(launch_add(view(res), view(summand), res->n, true, G.stream));
  return FT_ERR_NO_ERROR;
}

void ft_neg(ft_ge25519_array array) { (void)launch_neg(view(array), array->n, G.stream); }

// res[keys[i]] += summand[i] with device keys, stream-ordered.  Keys sort on
// the device into m + 1 CSR buckets (bucket m takes every key >= m); the
// gather form of the segmented sum adds each bucket's elements in increasing
// i, bucket m's sum is dropped, and the others are added into res.
static size_t isum_work_bytes(size_t n, size_t m) { return 8 * (m + 2) + csr_sort_bytes(n); }
static size_t isum_scratch_points(size_t n, size_t m) { return segsum_gather_scratch_points(n, n) + m + 1; }

static ft_error isum_run(ft_ge25519_array res, const uint64_t *d_keys, const_ft_ge25519_array summand,
                         PtView scratch, void *work, hipStream_t s) {
  const size_t n = summand->n, m = res->n;
  int64_t *off = static_cast<int64_t *>(work);
  const int32_t *order = nullptr;
  CK(csr_sort_keys(off, &order, d_keys, n, m, off + (m + 2), s));
  const size_t tail = segsum_gather_scratch_points(n, n);
  const PtView sums{scratch.w + tail * 8, scratch.cap};      // element slots [tail, tail + m + 1)
  CK(segsum_launch(sums, off, m + 1, n, order, view(summand), n, scratch, s));
  CK(launch_add(view(res), sums, m, false, s));
  return FT_ERR_NO_ERROR;
}

static ft_error isum_checked(ft_ge25519_array res, const void *d_keys, const_ft_ge25519_array summand,
                             ft_ge25519_array scratch, void *d_work, hipStream_t s) {
  const size_t n = summand->n, m = res->n;
  if (n == 0 || m == 0) return FT_ERR_NO_ERROR;
  if (n >= ((size_t)1 << 31) || m >= ((size_t)1 << 31) - 1) return FT_ERR_BAD_SIZE;
  if (scratch && scratch->cap < isum_scratch_points(n, m)) return FT_ERR_INSUFFICIENT_SPACE;
  if (d_work && (reinterpret_cast<uintptr_t>(d_work) & 7)) return FT_ERR_BAD_SIZE;
  ArrGuard own;
  DevBuf work;
  if (!scratch) {
    CKE(alloc_array(&own.a, isum_scratch_points(n, m)));
    scratch = own.a;
  }
  if (!d_work) {
    CKE(work.alloc(isum_work_bytes(n, m)));
    d_work = work.p;
  }
  CKE(isum_run(res, static_cast<const uint64_t *>(d_keys), summand, view(scratch), d_work, s));
  if (own.a || work.p) CK(hipStreamSynchronize(s));
  return FT_ERR_NO_ERROR;
}

ft_error ft_reduce_isum(ft_ge25519_array res, const size_t *idxs, size_t n_idxs,
                        const_ft_ge25519_array summand) {
  // ftcrypto.cu:1150-1217: res[idxs[i]] += summand[i], duplicate keys accumulate;
  // the keys are range-checked on the device after their copy (one pass over
  // the caller's buffer instead of two), before anything is added
  if (n_idxs != summand->n) return FT_ERR_SIZE_MISMATCH;
  if (n_idxs == 0) return FT_ERR_NO_ERROR;
  FT_IO_LOCK;
  DevBuf dkeys;
  CKE(dkeys.upload(idxs, n_idxs * sizeof(size_t)));
  CK(keys_out_of_range(G.flag, static_cast<const uint64_t *>(dkeys.p), n_idxs, res->n, G.stream));
  unsigned long long bad = 0;
  CK(hipMemcpyAsync(&bad, G.flag, 8, hipMemcpyDeviceToHost, G.stream));
  CK(hipStreamSynchronize(G.stream));
  if (bad) return FT_ERR_OUT_OF_BOUNDS;
  CKE(isum_checked(res, dkeys.p, summand, nullptr, nullptr, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_reduce(ft_ge25519_array array) {
  ArrGuard out;
  CKE(alloc_array(&out.a, 1));
  CKE(reduce_into(out.a, array, G.stream));
  std::swap(array->d, out.a->d);
  std::swap(array->cap, out.a->cap);
  array->n = 1;
  return FT_ERR_NO_ERROR;
}

ft_error ft_array_scan(ft_ge25519_array array) { return scan_rec(view(array), array->n, false, G.stream); }
ft_error ft_array_prescan(ft_ge25519_array array) { return scan_rec(view(array), array->n, true, G.stream); }

// ------------------------------------------------------------ predicates

static ft_error download_i64(int64_t *res, const void *d, size_t n) {
  CK(hipMemcpyAsync(res, d, n * 8, hipMemcpyDeviceToHost, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return FT_ERR_NO_ERROR;
}

void ft_index(int64_t *res, const_ft_ge25519_array array) {
  DevBuf out;
  if (out.alloc(array->n * 8) != FT_ERR_NO_ERROR) return;
  if (launch_index((int64_t *)out.p, view(array), array->n, G.stream) != hipSuccess) return;
  (void)download_i64(res, out.p, array->n);
}

static ft_error equal_impl(int64_t *res, const_ft_ge25519_array a, const_ft_ge25519_array b, bool neq) {
  if (a->n != b->n) return FT_ERR_SIZE_MISMATCH;
  DevBuf out;
  CKE(out.alloc(a->n * 8));
  CK(launch_equal((int64_t *)out.p, view(a), view(b), a->n, neq, G.stream));
  return download_i64(res, out.p, a->n);
}

ft_error ft_equal(int64_t *res, const_ft_ge25519_array a, const_ft_ge25519_array b) {
  return equal_impl(res, a, b, false);
}
ft_error ft_not_equal(int64_t *res, const_ft_ge25519_array a, const_ft_ge25519_array b) {
  return equal_impl(res, a, b, true);
}

ft_error ft_array_contains(int64_t *res, const_ft_ge25519_array a1, const_ft_ge25519_array a2) {
  // ftcrypto.cu:1046-1086: res[i] = 1 when a1[i] equals some element of a2;
  // canonical encodings of both, then a device hash set of a2's
  const size_t n1 = a1->n, n2 = a2->n;
  if (n1 == 0) return FT_ERR_NO_ERROR;
  DevBuf e1, e2, slots, out;
  CKE(e1.alloc(n1 * 32));
  CKE(e2.alloc(n2 * 32));
  CKE(slots.alloc(set_slots(n2) * 4));
  CKE(out.alloc(n1 * 8));
  CK(launch_encode((uint32_t *)e1.p, view(a1), n1, G.stream));
  CK(launch_encode((uint32_t *)e2.p, view(a2), n2, G.stream));
  CK(launch_set_contains((int64_t *)out.p, (const uint32_t *)e1.p, n1, (const uint32_t *)e2.p, n2,
                         (uint32_t *)slots.p, G.stream));
  return download_i64(res, out.p, n1);
}

// ------------------------------------------------------------ extensions

ft_error ftx_key_create(ftx_key *key, const_ft_ge25519_array pk, size_t pk_idx) {
  if (pk_idx >= pk->n) return FT_ERR_OUT_OF_BOUNDS;
  auto *k = new (std::nothrow) ftx_key_s;
  if (!k) return FT_ERR_OUT_OF_MEMORY;
  k->table = nullptr;
  k->table32 = nullptr;
  if (hipMalloc(&k->table, COMB_TABLE_WORDS * 4) != hipSuccess ||
      hipMalloc(&k->table32, COMB32_TABLE_WORDS * 4) != hipSuccess) {
    (void)hipGetLastError();
    if (k->table) (void)hipFree(k->table);
    delete k;
    return FT_ERR_OUT_OF_MEMORY;
  }
  hipError_t e = launch_build_comb(k->table, view(pk), pk_idx, false, G.stream);
  if (e == hipSuccess) e = launch_build_comb32(k->table32, view(pk), pk_idx, false, G.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
  if (e != hipSuccess) {
    (void)hipFree(k->table);
    (void)hipFree(k->table32);
    delete k;
    return herr(e);
  }
  *key = k;
  return FT_ERR_NO_ERROR;
}

void ftx_key_free(ftx_key key) {
  if (!key) return;
  (void)hipStreamSynchronize(G.stream);
  (void)hipFree(key->table);
  (void)hipFree(key->table32);
  delete key;
}

// A bound 0 < msg_bits < 64 is checked on the device before the encryption:
// the check writes the fail word, and the encryption launch is gated by it
// (writes nothing when it is nonzero).  Without a bound (0 or 64) the fail
// word, when given, is cleared.
static bool msg_bounded(int msg_bits) { return msg_bits > 0 && msg_bits < 64; }

ft_error ftx_encrypt_dev(ft_ge25519_array c1, ft_ge25519_array c2, ftx_key key, const void *d_m,
                         const void *d_r, int msg_bits, void *d_fail, void *stream) {
  if (c1->n != c2->n) return FT_ERR_SIZE_MISMATCH;
  if (c1->d == c2->d) return FT_ERR_BAD_SIZE;
  if (msg_bits < 0 || msg_bits > 64) return FT_ERR_BAD_SIZE;
  const bool bounded = msg_bounded(msg_bits);
  if (bounded && !d_fail) return FT_ERR_BAD_SIZE;
  auto *fail = static_cast<unsigned long long *>(d_fail);
  if (fail) CK(launch_msg_bound((const uint64_t *)d_m, bounded ? c1->n : 0, msg_bits, fail, S(stream)));
  CK(launch_encrypt(view(c1), view(c2), G.btab, key->table, (const uint64_t *)d_m, msg_digits(msg_bits),
                    (const uint32_t *)d_r, c1->n, false, S(stream), bounded ? fail : nullptr, G.btab32,
                    key->table32));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_encrypt_batch_dev(const ftx_enc_job *jobs, size_t n_jobs, ftx_key key, int msg_bits, void *d_fail,
                               void *stream) {
  if (n_jobs > (size_t)FTX_ENC_MAX_JOBS) return FT_ERR_BAD_SIZE;
  if (msg_bits < 0 || msg_bits > 64) return FT_ERR_BAD_SIZE;
  const bool bounded = msg_bounded(msg_bits);
  if (bounded && !d_fail) return FT_ERR_BAD_SIZE;
  EncBatch B;
  std::memset(&B, 0, sizeof B);
  B.J = (int)n_jobs;
  for (size_t q = 0; q < n_jobs; ++q) {
    const ftx_enc_job &a = jobs[q];
    if (!a.c1 || !a.c2 || a.c1->d == a.c2->d) return FT_ERR_BAD_SIZE;
    if (a.c1->n != a.c2->n) return FT_ERR_SIZE_MISMATCH;
    for (size_t o = 0; o < q; ++o)      // jobs write concurrently: outputs pairwise distinct
      if (jobs[o].c1->d == a.c1->d || jobs[o].c1->d == a.c2->d || jobs[o].c2->d == a.c1->d ||
          jobs[o].c2->d == a.c2->d)
        return FT_ERR_BAD_SIZE;
    B.job[q] = EncJob{view(a.c1), view(a.c2), (const uint64_t *)a.d_m, (const uint32_t *)a.d_r, 0, a.c1->n};
  }
  auto *fail = static_cast<unsigned long long *>(d_fail);
  if (fail) {
    EncBatch C = B;                       // no message is read without a bound
    if (!bounded)
      for (int q = 0; q < C.J; ++q) C.job[q].n = 0;
    CK(launch_msg_bound_batch(C, bounded ? msg_bits : 1, fail, S(stream)));
  }
  CK(launch_encrypt_batch(B, G.btab32, key->table32, msg_digits(msg_bits), S(stream), bounded ? fail : nullptr));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_encrypt(ft_ge25519_array *c1, ft_ge25519_array *c2, ftx_key key, const uint64_t *m,
                     const uint8_t *r, size_t n, int msg_bits) {
  if (msg_bits < 0 || msg_bits > 64) return FT_ERR_BAD_SIZE;
  // the bound is public and checked here, before anything is allocated
  if (msg_bounded(msg_bits))
    for (size_t i = 0; i < n; ++i)
      if (m[i] >> msg_bits) return FT_ERR_OUT_OF_BOUNDS;
  ArrGuard a1, a2;
  CKE(alloc_array(&a1.a, n));
  CKE(alloc_array(&a2.a, n));
  DevBuf dm, dr;
  CKE(dm.upload(m, n * 8));
  CKE(dr.upload(r, n * 32));
  CK(launch_encrypt(view(a1.a), view(a2.a), G.btab, key->table, (const uint64_t *)dm.p, msg_digits(msg_bits),
                    (const uint32_t *)dr.p, n, false, G.stream, nullptr, G.btab32, key->table32));
  CK(hipStreamSynchronize(G.stream));
  *c1 = a1.release();
  *c2 = a2.release();
  return FT_ERR_NO_ERROR;
}

ft_error ftx_rerandomize_dev(ft_ge25519_array c1, ft_ge25519_array c2, ftx_key key, const void *d_r,
                             void *stream) {
  if (c1->n != c2->n) return FT_ERR_SIZE_MISMATCH;
  if (c1->d == c2->d) return FT_ERR_BAD_SIZE;
  CK(launch_encrypt(view(c1), view(c2), G.btab, key->table, nullptr, 0, (const uint32_t *)d_r, c1->n, true,
                    S(stream), nullptr, G.btab32, key->table32));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_rerandomize(ft_ge25519_array c1, ft_ge25519_array c2, ftx_key key, const uint8_t *r) {
  DevBuf dr;
  CKE(dr.upload(r, c1->n * 32));
  CKE(ftx_rerandomize_dev(c1, c2, key, dr.p, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_mul_u32_dev(ft_ge25519_array c1, ft_ge25519_array c2, const void *d_k, void *stream) {
  if (c2 && c2->n != c1->n) return FT_ERR_SIZE_MISMATCH;
  if (c2 && c2->d == c1->d) return FT_ERR_BAD_SIZE;   // the halves' lanes run at once
  CK(launch_mul_u32(view(c1), c2 ? view(c2) : view(c1), c2 != nullptr, (const uint32_t *)d_k, c1->n,
                    S(stream)));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_mul_u32(ft_ge25519_array c1, ft_ge25519_array c2, const uint32_t *k, size_t n) {
  if (n != c1->n || (c2 && c2->n != n)) return FT_ERR_SIZE_MISMATCH;
  DevBuf dk;
  CKE(dk.upload(k, n * 4));
  CKE(ftx_mul_u32_dev(c1, c2, dk.p, G.stream));
  CK(hipStreamSynchronize(G.stream));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_from_scalars_dev(ft_ge25519_array out, const void *d_scalars, void *stream) {
  CK(launch_fixed_base(view(out), d_scalars, false, out->n, G.btab32, S(stream)));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_mul_dev(ft_ge25519_array array, const void *d_scalars, void *stream) {
  CK(launch_varbase(view(array), (const uint32_t *)d_scalars, 8, array->n, S(stream)));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_add_dev(ft_ge25519_array res, const_ft_ge25519_array summand, void *stream) {
  if (res->n != summand->n) return FT_ERR_SIZE_MISMATCH;
  CK(launch_add(view(res), view(summand), res->n, false, S(stream)));
  return FT_ERR_NO_ERROR;
}

// ------------------------------------------------------------ streams and graphs

ft_error ftx_stream_create(void **stream) {
  hipStream_t st;
  CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));   // default priority
  *stream = st;
  return FT_ERR_NO_ERROR;
}

void ftx_stream_destroy(void *stream) {
  if (stream) (void)hipStreamDestroy((hipStream_t)stream);
}

ft_error ftx_capture_begin(void *stream) {
  if (!stream) return FT_ERR_BAD_SIZE;            // the library stream is never captured
  CK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
  return FT_ERR_NO_ERROR;
}

ft_error ftx_capture_end(void *stream, ftx_graph *graph) {
  auto *g = new (std::nothrow) ftx_graph_s;
  if (!g) return FT_ERR_OUT_OF_MEMORY;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    delete g;
    return herr(e);
  }
  *graph = g;
  return FT_ERR_NO_ERROR;
}

ft_error ftx_graph_launch(ftx_graph graph, void *stream) {
  CK(hipGraphLaunch(graph->exec, S(stream)));
  return FT_ERR_NO_ERROR;
}

void ftx_graph_free(ftx_graph graph) {
  if (!graph) return;
  (void)hipGraphExecDestroy(graph->exec);
  (void)hipGraphDestroy(graph->graph);
  delete graph;
}

size_t ftx_segsum_scratch_points(size_t n_elems) { return seg_scratch_points(n_elems); }
size_t ftx_reduce_isum_scratch_points(size_t n_keys, size_t n_res) { return isum_scratch_points(n_keys, n_res); }
size_t ftx_reduce_isum_work_bytes(size_t n_keys, size_t n_res) { return isum_work_bytes(n_keys, n_res); }

ft_error ftx_reduce_isum_dev(ft_ge25519_array res, const void *d_keys, const_ft_ge25519_array summand,
                             ft_ge25519_array scratch, void *d_work, void *stream) {
  return isum_checked(res, d_keys, summand, scratch, d_work, S(stream));
}
size_t ftx_gather_scratch_points(size_t n_elems, size_t n_src) {
  return segsum_gather_scratch_points(n_elems, n_src);
}

ft_error ftx_segmented_sum_dev(ft_ge25519_array out, const void *d_offsets, size_t n_seg, size_t n_elems,
                               const void *d_gather, const_ft_ge25519_array src, ft_ge25519_array scratch,
                               void *stream) {
  if (out->n < n_seg) return FT_ERR_INSUFFICIENT_SPACE;
  if (out->d == src->d) return FT_ERR_BAD_SIZE;   // sums are written while other lanes read src
  if (!d_gather && n_elems > src->n) return FT_ERR_OUT_OF_BOUNDS;
  const size_t need = seg_scratch_points(n_elems);
  ArrGuard own;
  if (!scratch) {
    CKE(alloc_array(&own.a, d_gather ? segsum_gather_scratch_points(n_elems, src->n) : need));
    scratch = own.a;
  } else if (scratch->cap < need) {
    return FT_ERR_INSUFFICIENT_SPACE;
  }
  CKE(segsum_run(out, (const int64_t *)d_offsets, n_seg, n_elems, (const int32_t *)d_gather, src,
                 view(scratch), S(stream)));
  if (own.a) CK(hipStreamSynchronize(S(stream)));
  return FT_ERR_NO_ERROR;
}

// job table of ftx_aggregate_batch_dev -> kernel jobs; FT_ERR_NO_ERROR or
// the first size error
static ft_error agg_jobs(const ftx_agg_job *jobs, size_t n_jobs, AggJob *out) {
  if (n_jobs > (size_t)FTX_AGG_MAX_JOBS) return FT_ERR_BAD_SIZE;
  for (size_t i = 0; i < n_jobs; ++i) {
    const ftx_agg_job &a = jobs[i];
    AggJob &b = out[i];
    std::memset(&b, 0, sizeof b);
    b.kind = a.kind;
    if (!a.out) return FT_ERR_BAD_SIZE;
    b.out = view(a.out);
    switch (a.kind) {
      case FTX_AGG_SEGSUM:
        if (!a.src || a.src->d == a.out->d) return FT_ERR_BAD_SIZE;
        if (a.out->n < a.n_seg) return FT_ERR_INSUFFICIENT_SPACE;
        if (a.n_elems > a.src->n) return FT_ERR_OUT_OF_BOUNDS;
        if (!a.d_offsets) {
          if (a.seg_len == 0 || a.n_seg != (a.n_elems + a.seg_len - 1) / a.seg_len) return FT_ERR_BAD_SIZE;
        }
        b.src = view(a.src);
        b.off = (const int64_t *)a.d_offsets;
        b.len = a.seg_len;
        b.nseg = a.n_seg;
        b.n = a.n_elems;
        break;
      case FTX_AGG_REDUCE:
        if (!a.src) return FT_ERR_BAD_SIZE;
        if (a.out->n < 1) return FT_ERR_INSUFFICIENT_SPACE;
        b.src = view(a.src);
        b.n = a.src->n;
        break;
      case FTX_AGG_SCAN:
      case FTX_AGG_XSCAN:
        b.n = a.out->n;
        break;
      default:
        return FT_ERR_BAD_SIZE;
    }
  }
  // jobs run concurrently: a job's output is no other job's output or input
  for (size_t i = 0; i < n_jobs; ++i)
    for (size_t j = 0; j < n_jobs; ++j) {
      if (i == j) continue;
      const uint32_t *oi = jobs[i].out->d;
      const bool scan_j = jobs[j].kind == FTX_AGG_SCAN || jobs[j].kind == FTX_AGG_XSCAN;
      if (oi == jobs[j].out->d || (!scan_j && oi == jobs[j].src->d)) return FT_ERR_BAD_SIZE;
    }
  return FT_ERR_NO_ERROR;
}

size_t ftx_aggregate_batch_scratch_points(const ftx_agg_job *jobs, size_t n_jobs) {
  AggJob kj[FTX_AGG_MAX_JOBS];
  if (agg_jobs(jobs, n_jobs, kj) != FT_ERR_NO_ERROR) return 1;
  return agg_batch_scratch_points(kj, (int)n_jobs);
}

ft_error ftx_aggregate_batch_dev(const ftx_agg_job *jobs, size_t n_jobs, ft_ge25519_array scratch,
                                 void *stream) {
  AggJob kj[FTX_AGG_MAX_JOBS];
  CKE(agg_jobs(jobs, n_jobs, kj));
  if (!n_jobs) return FT_ERR_NO_ERROR;
  const size_t need = agg_batch_scratch_points(kj, (int)n_jobs);
  ArrGuard own;
  if (!scratch) {
    CKE(alloc_array(&own.a, need));
    scratch = own.a;
  } else if (scratch->cap < need) {
    return FT_ERR_INSUFFICIENT_SPACE;
  }
  CK(agg_batch_launch(kj, (int)n_jobs, view(scratch), S(stream)));
  if (own.a) CK(hipStreamSynchronize(S(stream)));
  return FT_ERR_NO_ERROR;
}

static ft_error check_csr(const int64_t *off, size_t nrows, size_t limit) {
  if (off[0] < 0) return FT_ERR_OUT_OF_BOUNDS;
  for (size_t s = 0; s < nrows; ++s)
    if (off[s + 1] < off[s]) return FT_ERR_BAD_SIZE;
  if ((size_t)off[nrows] > limit) return FT_ERR_OUT_OF_BOUNDS;
  return FT_ERR_NO_ERROR;
}

ft_error ftx_segmented_sum(ft_ge25519_array *res, const int64_t *offsets, size_t n_seg,
                           const_ft_ge25519_array src) {
  CKE(check_csr(offsets, n_seg, src->n));
  ArrGuard out;
  CKE(alloc_array(&out.a, n_seg));
  DevBuf doff;
  CKE(doff.upload(offsets, (n_seg + 1) * 8));
  CKE(ftx_segmented_sum_dev(out.a, doff.p, n_seg, (size_t)offsets[n_seg], nullptr, src, nullptr, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *res = out.release();
  return FT_ERR_NO_ERROR;
}

ft_error ftx_gather_sum(ft_ge25519_array *res, const int64_t *row_off, const int32_t *cols, size_t n_rows,
                        const_ft_ge25519_array src) {
  CKE(check_csr(row_off, n_rows, (size_t)INT64_MAX));
  const size_t ne = (size_t)row_off[n_rows];
  for (size_t j = 0; j < ne; ++j)
    if (cols[j] < 0 || (size_t)cols[j] >= src->n) return FT_ERR_OUT_OF_BOUNDS;
  ArrGuard out;
  CKE(alloc_array(&out.a, n_rows));
  DevBuf doff, dcol;
  CKE(doff.upload(row_off, (n_rows + 1) * 8));
  CKE(dcol.upload(cols, ne * 4));
  CKE(ftx_segmented_sum_dev(out.a, doff.p, n_rows, ne, dcol.p, src, nullptr, G.stream));
  CK(hipStreamSynchronize(G.stream));
  *res = out.release();
  return FT_ERR_NO_ERROR;
}

size_t ftx_reduce_scratch_points(size_t n) { return reduce_scratch(n); }
size_t ftx_scan_scratch_points(size_t n) { return scan_scratch_points(n); }

ft_error ftx_reduce_dev(ft_ge25519_array out, const_ft_ge25519_array src, ft_ge25519_array scratch, void *stream) {
  if (out->n < 1) return FT_ERR_INSUFFICIENT_SPACE;
  if (!scratch) return reduce_into(out, src, S(stream));
  if (scratch->cap < reduce_scratch(src->n)) return FT_ERR_INSUFFICIENT_SPACE;
  return reduce_run(out, src, view(scratch), S(stream));
}

static_assert(FTX_REDUCE_NOALLOC_POINTS == NOALLOC_POINTS, "static reduce workspace size");

// ftx_reduce_dev with the library's static block-sum array as its scratch:
// two kernel launches, nothing allocated, nothing synchronised
ft_error ftx_reduce_noalloc_dev(ft_ge25519_array out, const_ft_ge25519_array src, void *stream) {
  if (out->n < 1) return FT_ERR_INSUFFICIENT_SPACE;
  if (reduce_scratch(src->n) > NOALLOC_POINTS) return FT_ERR_BAD_SIZE;   // beyond the static array
  PtView part;
  CK(noalloc_scratch(&part));
  return reduce_run(out, src, part, S(stream));
}

ft_error ftx_scan_dev(ft_ge25519_array array, int exclusive, ft_ge25519_array scratch, void *stream) {
  if (!scratch) return scan_rec(view(array), array->n, exclusive != 0, S(stream));
  if (scratch->cap < scan_scratch_points(array->n)) return FT_ERR_INSUFFICIENT_SPACE;
  CK(scan_launch(view(array), array->n, exclusive != 0, view(scratch), S(stream)));
  return FT_ERR_NO_ERROR;
}

// ------------------------------------------------------------ inputs (host)

static void parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t)> &fn) {
  unsigned hw = std::thread::hardware_concurrency();
  size_t nt = std::min<size_t>(hw ? std::min(hw, 16u) : 1, (n + grain - 1) / grain);
  if (nt <= 1) { fn(0, n); return; }
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t)
    th.emplace_back([&, t] { fn(n * t / nt, n * (t + 1) / nt); });
  for (auto &x : th) x.join();
}

void ftx_scalar_reduce64(uint8_t *out, const uint8_t *in, size_t n) {
  const Barrett &b = barrett();
  parallel_for(n, 4096, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      uint32_t x[16], r[8];
      for (int k = 0; k < 16; ++k)
        x[k] = (uint32_t)in[64 * i + 4 * k] | ((uint32_t)in[64 * i + 4 * k + 1] << 8) |
               ((uint32_t)in[64 * i + 4 * k + 2] << 16) | ((uint32_t)in[64 * i + 4 * k + 3] << 24);
      b.reduce(r, x);
      words_to_bytes(out + 32 * i, r, 8);
    }
  });
}

void ftx_random_scalars(uint8_t *out, size_t n, uint64_t seed, uint64_t stream) {
  const Barrett &b = barrett();
  const uint64_t base = stream << 40;
  parallel_for(n, 4096, [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      uint32_t x[16], r[8];
      for (int j = 0; j < 8; ++j) {
        const uint64_t z = splitmix_at(seed, base + 8 * (uint64_t)i + (uint64_t)j);
        x[2 * j] = (uint32_t)z;
        x[2 * j + 1] = (uint32_t)(z >> 32);
      }
      b.reduce(r, x);
      words_to_bytes(out + 32 * i, r, 8);
    }
  });
}

}  // extern "C"
