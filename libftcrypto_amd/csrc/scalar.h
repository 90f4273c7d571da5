// Scalar arithmetic modulo the group order l = 2^252 + delta (delta =
// 27742317777372353535851937790883648493) and the signed radix-32 digits of
// the fixed-base comb.  Branch-free 32-bit integer code, shared by the
// kernels and the host tests (tests/native/fe_host.cc).
#pragma once
#include "fe25519.h"

namespace ft {

// word i (low first) of l
FT_HD uint32_t l_word(int i) {
  return i == 0 ? 0x5cf5d3edu : i == 1 ? 0x5812631au : i == 2 ? 0xa2f79cd6u : i == 3 ? 0x14def9deu
       : i == 7 ? 0x10000000u : 0u;
}

// m <- m mod l for any m < 2^256.  With m = q 2^252 + r (q < 16, r < 2^252)
// and 2^252 = l - delta, m = q l + (r - q delta) where r - q delta lies in
// (-2^129, 2^252): l is added back when it is negative.
FT_HD void scalar_mod_l(uint32_t m[8]) {
  const uint32_t q = m[7] >> 28;
  m[7] &= 0x0fffffffu;
  uint32_t qd[5];
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c += (uint64_t)q * l_word(i);
    qd[i] = (uint32_t)c;
    c >>= 32;
  }
  qd[4] = (uint32_t)c;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint64_t s = (uint64_t)m[i] - (i < 5 ? qd[i] : 0u) - borrow;
    m[i] = (uint32_t)s;
    borrow = (uint32_t)(s >> 63);
  }
  const uint32_t mask = 0u - borrow;
  c = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    c += (uint64_t)m[i] + (l_word(i) & mask);
    m[i] = (uint32_t)c;
    c >>= 32;
  }
}

// m <- l - m when neg, for 0 <= m < l (m = 0 gives l, which is 0 mod l)
FT_HD void scalar_cneg_l(uint32_t m[8], uint32_t neg) {
  const uint32_t mask = 0u - neg;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint64_t s = (uint64_t)l_word(i) - m[i] - borrow;
    borrow = (uint32_t)(s >> 63);
    m[i] = ((uint32_t)s & mask) | (m[i] & ~mask);
  }
}

// ND signed radix-32 digits of m, lowest first, as bytes (digit v at byte
// v & 3 of e[v >> 2]).  Digits 0 .. ND-2 lie in [-16, 15]; the top digit
// takes the rest: at most 8 for m <= l with ND = 51, at most 16 for
// m < 2^64 with ND = 13.
template <int ND>
FT_HD void recode32(uint32_t e[(ND + 3) / 4], const uint32_t m[8]) {
  int carry = 0;
#pragma unroll
  for (int q = 0; q < (ND + 3) / 4; ++q) e[q] = 0;
#pragma unroll
  for (int v = 0; v < ND; ++v) {
    const int bit = 5 * v, w = bit >> 5, sh = bit & 31;
    uint32_t x = m[w] >> sh;
    if (sh > 27 && w + 1 < 8) x |= m[w + 1] << (32 - sh);
    int d = (int)(x & 31u) + carry;
    if (v < ND - 1) {
      carry = (d + 16) >> 5;
      d -= carry * 32;
    }
    e[v >> 2] |= ((uint32_t)d & 0xffu) << (8 * (v & 3));
  }
}

// signed radix-32 digits of a message x < 2^bits over nd = (bits + 5) / 5
// digits (nd <= 13), the top one keeping the last carry: at most 16 (its
// field holds at most 4 bits of x), as bytes like recode32; digits from nd
// on are zero
FT_HD void recode32_msg(uint32_t e[4], uint64_t x, int nd) {
  int carry = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) e[q] = 0;
#pragma unroll
  for (int v = 0; v < 13; ++v) {
    const int d = (int)((x >> (5 * v)) & 31u) + carry;
    const int c = (d + 16) >> 5;
    const bool top = v == nd - 1;
    const int out = top ? d : d - c * 32;
    carry = top ? 0 : c;
    const uint32_t db = v < nd ? ((uint32_t)out & 0xffu) : 0u;
    e[v >> 2] |= db << (8 * (v & 3));
  }
}

}  // namespace ft
