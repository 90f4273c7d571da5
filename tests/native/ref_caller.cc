// A C++ caller written the way the reference's own test driver is
// (run_tests.cc: the header wrapped in extern "C", nothing but ft_ calls),
// compiled with g++ against include/ftcrypto.h and linked against
// libftcrypto.so: the drop-in boundary exercised from compiled code.  It
// runs the driver's flow and asserts what the driver checks and what it only
// prints (run_tests.cc:49-189).  Exit status 0 when every check holds.
#include <cstdint>
#include <cstdio>
#include <initializer_list>
#include <vector>

extern "C" {
#include "ftcrypto.h"
}

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAIL: %s (ref_caller.cc:%d)\n", #c, __LINE__);  \
      ++fails;                                                     \
    }                                                              \
  } while (0)

static std::vector<uint8_t> folded(const_ft_ge25519_array a) {
  std::vector<uint8_t> b(32 * ft_array_get_length(a));
  CHECK(ft_array_to_bytes_folded(b.data(), b.size(), a) == FT_ERR_NO_ERROR);
  return b;
}

// k_i B for small k_i
static ft_ge25519_array small(std::initializer_list<uint64_t> v) {
  std::vector<uint64_t> s(v);
  ft_ge25519_array a = nullptr;
  CHECK(ft_array_from_small_scalars(&a, s.data(), s.size()) == FT_ERR_NO_ERROR);
  return a;
}

int main() {
  CHECK(ft_crypto_init() == FT_ERR_NO_ERROR);                              // :49-50

  // a 2^26-element array and its length (:61-69)
  ft_ge25519_array big = nullptr;
  CHECK(ft_array_init_scalar(&big, (size_t)1 << 26, nullptr) == FT_ERR_NO_ERROR);
  CHECK(big && ft_array_get_length(big) == ((size_t)1 << 26));
  ft_array_free(big);

  // init_scalar with no scalar: identities, encoded 01 00 .. 00 (:67-71)
  ft_ge25519_array z = nullptr;
  CHECK(ft_array_init_scalar(&z, 16, nullptr) == FT_ERR_NO_ERROR);
  const std::vector<uint8_t> zb = folded(z);
  for (size_t i = 0; i < zb.size(); ++i) CHECK(zb[i] == (i % 32 == 0 ? 1 : 0));

  // self-aliased ft_add doubles every element (:87)
  ft_ge25519_array d = small({1, 2, 3}), d2 = small({2, 4, 6});
  CHECK(ft_add(d, d) == FT_ERR_NO_ERROR);
  CHECK(folded(d) == folded(d2));

  // affine round trip, then every point validates (:92-103)
  std::vector<uint8_t> aff(64 * 3);
  CHECK(ft_array_to_bytes_affine(aff.data(), aff.size(), d) == FT_ERR_NO_ERROR);
  ft_ge25519_array back = nullptr;
  CHECK(ft_array_from_bytes_affine(&back, aff.data(), aff.size()) == FT_ERR_NO_ERROR);
  int64_t bad = 0;
  CHECK(ft_array_validate(&bad, back) == FT_ERR_NO_ERROR && bad == -1);
  CHECK(folded(back) == folded(d));

  // reduce of {0..8} B is 36 B (:121-126)
  ft_ge25519_array r = small({0, 1, 2, 3, 4, 5, 6, 7, 8}), r36 = small({36});
  CHECK(ft_array_reduce(r) == FT_ERR_NO_ERROR && ft_array_get_length(r) == 1);
  CHECK(folded(r) == folded(r36));

  // inclusive scan of {0..6} B (:128-134)
  ft_ge25519_array s = small({0, 1, 2, 3, 4, 5, 6}), sw = small({0, 1, 3, 6, 10, 15, 21});
  CHECK(ft_array_scan(s) == FT_ERR_NO_ERROR && folded(s) == folded(sw));

  // ft_equal of 5B and 5B, then of 5B and 10B (:140-152)
  ft_ge25519_array e1 = small({5}), e2 = small({5});
  int64_t eq = -1;
  CHECK(ft_equal(&eq, e1, e2) == FT_ERR_NO_ERROR && eq == 1);
  CHECK(ft_add(e2, e1) == FT_ERR_NO_ERROR);
  CHECK(ft_equal(&eq, e1, e2) == FT_ERR_NO_ERROR && eq == 0);

  // {3, 5, 8, 2} B in {1..7} B (:157-170)
  ft_ge25519_array q = small({3, 5, 8, 2}), set = small({1, 2, 3, 4, 5, 6, 7});
  int64_t has[4] = {-1, -1, -1, -1};
  CHECK(ft_array_contains(has, q, set) == FT_ERR_NO_ERROR);
  CHECK(has[0] == 1 && has[1] == 1 && has[2] == 0 && has[3] == 1);

  // ft_reduce_isum: {10, 20, 30} B at keys [2, 2, 3] into {5, 8, 20, 30, 50} B (:175-189)
  ft_ge25519_array res = small({5, 8, 20, 30, 50}), add = small({10, 20, 30}), rw = small({5, 8, 50, 60, 50});
  const size_t keys[3] = {2, 2, 3};
  CHECK(ft_reduce_isum(res, keys, 3, add) == FT_ERR_NO_ERROR);
  CHECK(folded(res) == folded(rw));

  for (ft_ge25519_array a : {z, d, d2, back, r, r36, s, sw, e1, e2, q, set, res, add, rw}) ft_array_free(a);
  std::printf(fails ? "ref_caller: %d checks failed\n" : "ref_caller: all checks passed\n", fails);
  return fails ? 1 : 0;
}
