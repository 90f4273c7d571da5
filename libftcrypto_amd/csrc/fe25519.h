// GF(2^255 - 19) arithmetic for gfx950, in nine unsigned 29-bit limbs.
//
// Replaces the reference's fe25519 layer (sodium_fe.cu:10-1479: ten signed
// limbs of 25.5 bits, 100 int32 x int32 -> int64 products per multiply).
// Here a field element is v[0..8] with value sum v[i] * 2^(29 i) (a 261-bit
// span), reduced modulo p only lazily.  A multiply is 81
// v_mad_u64_u32 (full rate on gfx950) into 17 uint64 column sums, a fold of
// the high columns by 2^261 = 1216 (mod p) done on their 32-bit halves (16
// more mads), and one carry pass over the low columns.  No branches, no data-dependent addresses: every routine is
// constant-time.
//
// Bound classes (largest value of any limb):
//   N  = 2^29 + 2^17      output of mul, sq, carry and unpack
//   A1 = 2 N              sum of two N
//   A2 = 3 N              sum of three N, or x + K - y with x, y <= N
// fe_mul is exact (no uint64 column overflows) for the operand classes
// (N,N) (N,A1) (N,A2) (A1,A1) (A1,A2); fe_sq for N and A1.  The worst case
// is checked numerically in tests/test_bounds.py.  K = 128 p written with
// limbs 2^30 - 2432, 2^30 - 2, ..., 2^30 - 2 so that x + K - y never
// underflows for y <= N.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define FT_HD __host__ __device__ __forceinline__
#else
#define FT_HD static inline
#endif

namespace ft {

struct fe { uint32_t v[9]; };

constexpr uint32_t M29 = (1u << 29) - 1;
constexpr uint32_t K0 = (1u << 30) - 2432;   // 128 p, limb 0
constexpr uint32_t KI = (1u << 30) - 2;      // 128 p, limbs 1..8
constexpr uint32_t FOLD261 = 1216;           // 2^261 mod p = 64 * 19

FT_HD void fe_0(fe &r) {
#pragma unroll
  for (int i = 0; i < 9; ++i) r.v[i] = 0;
}
FT_HD void fe_1(fe &r) { fe_0(r); r.v[0] = 1; }

FT_HD void fe_add(fe &r, const fe &a, const fe &b) {
#pragma unroll
  for (int i = 0; i < 9; ++i) r.v[i] = a.v[i] + b.v[i];
}

// r = a - b + 128 p; b <= N, result <= a + 2^30
FT_HD void fe_sub(fe &r, const fe &a, const fe &b) {
  r.v[0] = a.v[0] + (K0 - b.v[0]);
#pragma unroll
  for (int i = 1; i < 9; ++i) r.v[i] = a.v[i] + (KI - b.v[i]);
}

// r = 128 p - a; a <= N, result <= 2^30
FT_HD void fe_neg(fe &r, const fe &a) {
  r.v[0] = K0 - a.v[0];
#pragma unroll
  for (int i = 1; i < 9; ++i) r.v[i] = KI - a.v[i];
}

// r = neg ? K - b : b for a 0 / all-ones mask (x ^ ~0 = -1 - x, so adding
// K + 1 gives K - x modulo 2^32); b of class N
FT_HD void fe_cneg_mask(fe &r, const fe &b, uint32_t mask) {
  r.v[0] = (b.v[0] ^ mask) + (mask & (K0 + 1u));
#pragma unroll
  for (int l = 1; l < 9; ++l) r.v[l] = (b.v[l] ^ mask) + (mask & (KI + 1u));
}

// carry to class N; limbs up to 2^32 - 16 (callers pass at most 2^31.4)
FT_HD void fe_carry(fe &r) {
#pragma unroll
  for (int i = 0; i < 8; ++i) { r.v[i + 1] += r.v[i] >> 29; r.v[i] &= M29; }
  uint32_t c = r.v[8] >> 29;
  r.v[8] &= M29;
  r.v[0] += c * FOLD261;
}

// Reduce 17 column sums (weights 2^(29k)) to class N.
// A high column k >= 9 is folded without carrying it first: its low word
// lo weighs 2^(29k) = 2^261 2^(29(k-9)) -> 1216 lo into column k-9, its high
// word hi weighs 2^32 2^(29k) = 8 2^(29(k+1)) -> 9728 hi into column k-8.
// Each low column gains < 2^45.5, so the sums stay below 2^64 for every
// allowed operand class (9 A B <= 2^63.76).  One carry pass then remains.
FT_HD void fe_reduce_cols(fe &r, uint64_t c[17]) {
#pragma unroll
  for (int k = 9; k < 17; ++k) {
    c[k - 9] += (uint64_t)(uint32_t)c[k] * FOLD261;
    c[k - 8] += (uint64_t)(uint32_t)(c[k] >> 32) * (8u * FOLD261);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) { c[k + 1] += c[k] >> 29; r.v[k] = (uint32_t)c[k] & M29; }
  r.v[8] = (uint32_t)c[8] & M29;
  uint64_t t = c[8] >> 29;                       // bits >= 2^261, < 2^35
  uint64_t r0 = (uint64_t)r.v[0] + t * FOLD261;  // < 2^46
  r.v[0] = (uint32_t)r0 & M29;
  r.v[1] += (uint32_t)(r0 >> 29);                // v1 < 2^29 + 2^17
}

FT_HD void fe_mul(fe &r, const fe &a, const fe &b) {
  uint64_t c[17];
#pragma unroll
  for (int k = 0; k < 17; ++k) c[k] = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = 0; j < 9; ++j) c[i + j] += (uint64_t)a.v[i] * b.v[j];
  fe_reduce_cols(r, c);
}

FT_HD void fe_sq(fe &r, const fe &a) {
  uint64_t c[17];
  uint32_t d[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = a.v[i] << 1;
#pragma unroll
  for (int k = 0; k < 17; ++k) c[k] = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    c[2 * i] += (uint64_t)a.v[i] * a.v[i];
#pragma unroll
    for (int j = i + 1; j < 9; ++j) c[i + j] += (uint64_t)a.v[i] * d[j];
  }
  fe_reduce_cols(r, c);
}

FT_HD void fe_sqn(fe &r, const fe &a, int n) {
  fe_sq(r, a);
  for (int i = 1; i < n; ++i) fe_sq(r, r);
}

// r = a if bit else r (bit in {0,1}); branch-free select
FT_HD void fe_cmov(fe &r, const fe &a, uint32_t bit) {
  uint32_t m = 0u - bit;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.v[i] ^= m & (r.v[i] ^ a.v[i]);
}

FT_HD void fe_cswap(fe &a, fe &b, uint32_t bit) {
  uint32_t m = 0u - bit;
#pragma unroll
  for (int i = 0; i < 9; ++i) { uint32_t t = m & (a.v[i] ^ b.v[i]); a.v[i] ^= t; b.v[i] ^= t; }
}

// z^(2^250 - 1) and z^11 (shared prefix of the invert and pow22523 chains)
FT_HD void fe_pow250(fe &z250, fe &z11, const fe &z) {
  fe z2, z9, t, z5, z10, z20, z50, z100;
  fe_sq(z2, z);
  fe_sqn(t, z2, 2);
  fe_mul(z9, t, z);
  fe_mul(z11, z9, z2);
  fe_sq(t, z11);
  fe_mul(z5, t, z9);
  fe_sqn(t, z5, 5);
  fe_mul(z10, t, z5);
  fe_sqn(t, z10, 10);
  fe_mul(z20, t, z10);
  fe_sqn(t, z20, 20);
  fe_mul(t, t, z20);
  fe_sqn(t, t, 10);
  fe_mul(z50, t, z10);
  fe_sqn(t, z50, 50);
  fe_mul(z100, t, z50);
  fe_sqn(t, z100, 100);
  fe_mul(t, t, z100);
  fe_sqn(t, t, 50);
  fe_mul(z250, t, z50);
}

// z^(p-2); invert(0) = 0 (as sodium_fe.cu:1352-1408)
FT_HD void fe_invert(fe &r, const fe &z) {
  fe z250, z11, t;
  fe_pow250(z250, z11, z);
  fe_sqn(t, z250, 5);
  fe_mul(r, t, z11);
}

// z^((p-5)/8) (as sodium_fe.cu:1410-1463)
FT_HD void fe_pow22523(fe &r, const fe &z) {
  fe z250, z11, t;
  fe_pow250(z250, z11, z);
  fe_sqn(t, z250, 2);
  fe_mul(r, t, z);
}

// Input of class N (or any limbs <= 2^29 + 2^17): fold the bits at and above
// 2^255 so that limbs 0..7 < 2^29 and limb 8 <= 2^23 (value < 2^255 + 2^232).
FT_HD void fe_fold255(fe &t) {
  uint32_t c = t.v[8] >> 23;
  t.v[8] &= (1u << 23) - 1;
  t.v[0] += 19u * c;
#pragma unroll
  for (int i = 0; i < 8; ++i) { t.v[i + 1] += t.v[i] >> 29; t.v[i] &= M29; }
}

// 255-bit limbs (limb 8 < 2^24) -> 8 little-endian 32-bit words
FT_HD void fe_pack_words(uint32_t w[8], const fe &t) {
  uint64_t acc = 0;
  int bits = 0, k = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    acc |= (uint64_t)t.v[i] << bits;
    bits += 29;
    if (bits >= 32) { w[k++] = (uint32_t)acc; acc >>= 32; bits -= 32; }
  }
  if (k < 8) w[k] = (uint32_t)acc;
}

// canonical residue in [0, p) as 8 words (sodium_fe.cu:168-289 semantics)
FT_HD void fe_canonical_words(uint32_t w[8], const fe &a) {
  fe t = a;
  fe_carry(t);
  fe_fold255(t);
  uint32_t q = (t.v[0] + 19u) >> 29;
#pragma unroll
  for (int i = 1; i < 8; ++i) q = (t.v[i] + q) >> 29;
  q = (t.v[8] + q) >> 23;
  t.v[0] += 19u * q;
#pragma unroll
  for (int i = 0; i < 8; ++i) { t.v[i + 1] += t.v[i] >> 29; t.v[i] &= M29; }
  t.v[8] &= (1u << 23) - 1;
  fe_pack_words(w, t);
}

// class N -> value < 2^256 (not canonical) as 8 words: the resident format
FT_HD void fe_store_words(uint32_t w[8], const fe &a) {
  fe t = a;
  fe_fold255(t);
  fe_pack_words(w, t);
}

// 8 little-endian words (value < 2^256) -> class N (limb 8 < 2^24)
FT_HD void fe_from_words(fe &r, const uint32_t w[8]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int bit = 29 * i, j = bit >> 5, s = bit & 31;
    uint32_t x = w[j] >> s;
    if (s > 3 && j + 1 < 8) x |= w[j + 1] << (32 - s);
    r.v[i] = (i < 8) ? (x & M29) : x;
  }
}

FT_HD uint32_t fe_isnegative(const fe &a) {
  uint32_t w[8];
  fe_canonical_words(w, a);
  return w[0] & 1u;
}

FT_HD uint32_t fe_iszero(const fe &a) {
  uint32_t w[8];
  fe_canonical_words(w, a);
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) acc |= w[i];
  return (uint32_t)(((uint64_t)acc - 1) >> 63);
}

// Constants in radix-2^29 limbs (class N), derived from their definitions and
// checked against them by tests/test_bounds.py:
//   d = -121665/121666 mod p, 2d, sqrt(-1) = 2^((p-1)/4) mod p,
//   1/2 = (p + 1)/2 = 2^254 - 9, 1/d = -121666/121665 mod p, 1/(2d).
constexpr fe FE_D = {{0x135978a3u, 0x0f5a6e50u, 0x10762addu, 0x00149a82u, 0x1e898007u,
                      0x003cbbbcu, 0x19ce331du, 0x1dc56dffu, 0x0052036cu}};
constexpr fe FE_D2 = {{0x06b2f159u, 0x1eb4dca1u, 0x00ec55bau, 0x00293505u, 0x1d13000eu,
                       0x00797779u, 0x139c663au, 0x1b8adbffu, 0x002406d9u}};
constexpr fe FE_SQRTM1 = {{0x0a0ea0b0u, 0x0770d93au, 0x0bf91e31u, 0x06300d5au, 0x1d7a72f4u,
                           0x004c9efdu, 0x1c2cad34u, 0x1009f83bu, 0x002b8324u}};
constexpr fe FE_INV2 = {{0x1ffffff7u, 0x1fffffffu, 0x1fffffffu, 0x1fffffffu, 0x1fffffffu,
                         0x1fffffffu, 0x1fffffffu, 0x1fffffffu, 0x003fffffu}};
constexpr fe FE_INV_D = {{0x0dc9f843u, 0x0f0793b6u, 0x1e550b89u, 0x1bad3084u, 0x1cf660b5u,
                          0x108a66dcu, 0x190cac58u, 0x1a429ab9u, 0x0040907eu}};
constexpr fe FE_INV_D2 = {{0x06e4fc18u, 0x1783c9dbu, 0x0f2a85c4u, 0x1dd69842u, 0x0e7b305au,
                           0x0845336eu, 0x1c86562cu, 0x0d214d5cu, 0x0060483fu}};

}  // namespace ft
