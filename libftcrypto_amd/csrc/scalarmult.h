// Scalar-multiplication building blocks shared by the kernels: signed
// radix-16 digit streams, the fixed-base comb pass and the constant-time
// variable-base window loop.
#pragma once
#include "layout.h"
#include "scalar.h"

namespace ft {

constexpr int TPB = 256;
// Encryption kernels hold two 28 KB comb tables in LDS per block.  For large
// batches, blocks of 512 lanes limited to 128 registers (4 spilled words) run
// two blocks per CU, four waves per SIMD: 6 % faster than 256-lane blocks at
// two waves (384-lane blocks, three waves uneven over the SIMDs, were slower
// than both).  Below ENC_WIDE_MIN elements there are too few 512-lane blocks
// to fill 256 CUs, and 256-lane blocks are used.
constexpr int ETPB = 512;
constexpr size_t ENC_WIDE_MIN = (size_t)ETPB * 512;
// Small batches leave few waves per SIMD with one lane per ciphertext; two
// lanes per ciphertext (one per half) double that.  Which of the two a size
// class uses is a tuned schedule (AggTune::encl).  Two-lane launches of at
// least one 512-lane block per CU use 512-lane blocks.
constexpr size_t ENC_SPLIT_WIDE_MIN = (size_t)ETPB * 256;

// ------------------------------------------------------------ digits
//
// Signed radix-16 digits (sodium_ge.cu:693-708) are packed four to a word
// as bytes and consumed through a register shift stream: the byte taken in
// a window depends only on the window counter, so neither a branch nor an
// address ever depends on a secret digit.

// 64 digits of k: stream words e[0..15], digit i at byte i&3 of e[i>>2]
__device__ __forceinline__ void recode16(uint32_t e[16], const uint32_t k[8]) {
  int carry = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) e[i] = 0;
#pragma unroll
  for (int i = 0; i < 64; ++i) {
    int d = (int)((k[i >> 3] >> (4 * (i & 7))) & 15u) + carry;
    if (i < 63) {
      carry = (d + 8) >> 4;
      d -= carry << 4;
    }
    e[i >> 2] |= ((uint32_t)d & 0xffu) << (8 * (i & 3));
  }
}

// split into the odd-digit and even-digit streams of the comb
__device__ __forceinline__ void split_odd_even(uint32_t odd[8], uint32_t even[8], const uint32_t e[16]) {
#pragma unroll
  for (int q = 0; q < 8; ++q) { odd[q] = 0; even[q] = 0; }
#pragma unroll
  for (int i = 0; i < 64; ++i) {
    const uint32_t d = (e[i >> 2] >> (8 * (i & 3))) & 0xffu;
    const int q = i >> 1;
    if (i & 1) odd[q >> 2] |= d << (8 * (q & 3));
    else even[q >> 2] |= d << (8 * (q & 3));
  }
}

// pop the lowest digit of a low-first stream of NW words (q = pops so far)
template <int NW>
__device__ __forceinline__ int pop_lo(uint32_t s[NW], int q) {
  const int b = (int)(int8_t)(uint8_t)(s[0] & 0xffu);
  s[0] >>= 8;
  if ((q & 3) == 3) {
#pragma unroll
    for (int k = 0; k < NW - 1; ++k) s[k] = s[k + 1];
    s[NW - 1] = 0;
  }
  return b;
}

// pop the highest digit of a 16-word high-first stream (v = its digit index)
__device__ __forceinline__ int pop_hi(uint32_t s[16], int v) {
  const int b = (int)(int8_t)(uint8_t)(s[15] >> 24);
  s[15] <<= 8;
  if ((v & 3) == 0) {
#pragma unroll
    for (int k = 15; k > 0; --k) s[k] = s[k - 1];
    s[0] = 0;
  }
  return b;
}

__device__ __forceinline__ void load_scalar(uint32_t k[8], const uint32_t *sc, size_t i) {
  ld_words(k, sc + i * 8);
}

// comb over one table: h += sum over NWIN windows of d * row(q), rows in LDS,
// digits from a low-first stream (q-th digit uses row q of the table)
template <int NW>
__device__ __forceinline__ void comb_pass(ge_p3 &h, const uint32_t *tab, uint32_t s[NW], int nwin,
                                          int row_step = 1) {
  ge_precomp t;
  ge_p1p1 r;
#pragma unroll 1
  for (int q = 0; q < nwin; ++q) {
    const int b = pop_lo<NW>(s, q);
    comb_select(t, tab + q * row_step * 8 * COMB_ENTRY_WORDS, b);
    ge_madd(r, h, t);
    ge_p1p1_to_p3(h, r);
  }
}

__device__ __forceinline__ void dbl4(ge_p3 &h);

// Comb in one loop: n0 windows of stream s, four doublings, n1 windows of
// stream s1.  Window w of a phase (w counted from 0 in each phase) uses row
// w of ta for w < 32 and row w - 32 of tb after that (the message digits of
// an encryption, over the table of B).  One loop instead of a loop per
// phase lets the compiler emit the window body once.
template <int NW>
__device__ __forceinline__ void comb_run(ge_p3 &h, const uint32_t *ta, const uint32_t *tb, uint32_t s[NW],
                                         const uint32_t s1[NW], int n0, int n1) {
  ge_precomp t;
  ge_p1p1 r;
  // keep the trip count opaque: with a known count the compiler reshapes
  // the loop and emits a slower body (an extra partial product and about
  // 170 register moves per window)
  asm volatile("" : "+s"(n0), "+s"(n1));
#pragma unroll 1
  for (int q = 0; q < n0 + n1; ++q) {
    if (q == n0) {
      dbl4(h);
#pragma unroll
      for (int k = 0; k < NW; ++k) s[k] = s1[k];
    }
    const int w = q < n0 ? q : q - n0;
    const int b = pop_lo<NW>(s, w);
    const uint32_t *row = w < 32 ? ta + w * 8 * COMB_ENTRY_WORDS : tb + (w - 32) * 8 * COMB_ENTRY_WORDS;
    comb_select(t, row, b);
    ge_madd(r, h, t);
    ge_p1p1_to_p3(h, r);
  }
  pin_p3(h);
}

__device__ __forceinline__ void dbl4(ge_p3 &h) {
  ge_p1p1 r;
  ge_p2 s;
  ge_p3_dbl(r, h);
  ge_p1p1_to_p2(s, r);
  ge_p2_dbl(r, s);
  ge_p1p1_to_p2(s, r);
  ge_p2_dbl(r, s);
  ge_p1p1_to_p2(s, r);
  ge_p2_dbl(r, s);
  ge_p1p1_to_p3(h, r);
}

// Radix-32 comb of the base point: h = sum over rows r < nrows of d_r *
// row r, digits from a low-first byte stream of NW words, rows read from
// global memory by scalar loads.  Row 0's entry gives h directly
// (ge_precomp_to_p3); every later row costs a select over its 16 entries
// and one mixed add, with no doubling.
template <int NW>
__device__ __forceinline__ void comb32_run(ge_p3 &h, const uint32_t *__restrict__ tab, uint32_t s[NW],
                                           int nrows) {
  ge_precomp t;
  ge_p1p1 r;
  comb_select_uniform<16>(t, tab, pop_lo<NW>(s, 0));
  ge_precomp_to_p3(h, t);
  asm volatile("" : "+s"(nrows));
#pragma unroll 1
  for (int q = 1; q < nrows; ++q) {
    const int b = pop_lo<NW>(s, q);
    comb_select_uniform<16>(t, tab + q * COMB32_ROW_WORDS, b);
    ge_madd(r, h, t);
    ge_p1p1_to_p3(h, r);
  }
  pin_p3(h);
}

// h = sum over rows q in [q0, q1) of d_q * row q of a radix-32 table (q0 <
// q1, both wave-uniform): the stream's first q0 digits are popped unused
template <int NW>
__device__ __forceinline__ void comb32_run_from(ge_p3 &h, const uint32_t *__restrict__ tab, uint32_t s[NW], int q0,
                                                int q1) {
  ge_precomp t;
  ge_p1p1 r;
  asm volatile("" : "+s"(q0), "+s"(q1));
#pragma unroll 1
  for (int q = 0; q < q0; ++q) (void)pop_lo<NW>(s, q);
  comb_select_uniform<16>(t, tab + q0 * COMB32_ROW_WORDS, pop_lo<NW>(s, q0));
  ge_precomp_to_p3(h, t);
#pragma unroll 1
  for (int q = q0 + 1; q < q1; ++q) {
    const int b = pop_lo<NW>(s, q);
    comb_select_uniform<16>(t, tab + q * COMB32_ROW_WORDS, b);
    ge_madd(r, h, t);
    ge_p1p1_to_p3(h, r);
  }
  pin_p3(h);
}

// h += sum over rows q < nrows of d_q * row q of a radix-32 table, digits
// from a low-first byte stream of NW words
template <int NW>
__device__ __forceinline__ void comb32_add_rows(ge_p3 &h, const uint32_t *__restrict__ tab, uint32_t s[NW],
                                                int nrows) {
  ge_precomp t;
  ge_p1p1 r;
  asm volatile("" : "+s"(nrows));
#pragma unroll 1
  for (int q = 0; q < nrows; ++q) {
    const int b = pop_lo<NW>(s, q);
    comb_select_uniform<16>(t, tab + q * COMB32_ROW_WORDS, b);
    ge_madd(r, h, t);
    ge_p1p1_to_p3(h, r);
  }
  pin_p3(h);
}

__device__ __forceinline__ void lds_load_table(uint32_t *lds, const uint32_t *g) {
  for (int i = threadIdx.x; i < COMB_TABLE_WORDS / 4; i += blockDim.x)
    reinterpret_cast<uint4 *>(lds)[i] = reinterpret_cast<const uint4 *>(g)[i];
}

// cached multiples 1P..8P (sodium_ge.cu:664-689)
__device__ __forceinline__ void build_cached8(ge_cached tab[8], const ge_p3 &p) {
  ge_p3 m[8];
  ge_p1p1 t;
  m[0] = p;
  ge_p3_dbl(t, p);  ge_p1p1_to_p3(m[1], t);
  ge_p3_to_cached(tab[0], m[0]);
  ge_p3_to_cached(tab[1], m[1]);
  ge_add(t, m[0], tab[1]); ge_p1p1_to_p3(m[2], t);     // 3P
  ge_p3_dbl(t, m[1]); ge_p1p1_to_p3(m[3], t);          // 4P
  ge_p3_to_cached(tab[2], m[2]);
  ge_p3_to_cached(tab[3], m[3]);
  ge_add(t, m[0], tab[3]); ge_p1p1_to_p3(m[4], t);     // 5P
  ge_p3_dbl(t, m[2]); ge_p1p1_to_p3(m[5], t);          // 6P
  ge_p3_to_cached(tab[4], m[4]);
  ge_p3_to_cached(tab[5], m[5]);
  ge_add(t, m[0], tab[5]); ge_p1p1_to_p3(m[6], t);     // 7P
  ge_p3_dbl(t, m[3]); ge_p1p1_to_p3(m[7], t);          // 8P
  ge_p3_to_cached(tab[6], m[6]);
  ge_p3_to_cached(tab[7], m[7]);
}

// |b| * P from the register table (all eight entries merged), b < 0 negates.
__device__ __forceinline__ void cached_select(ge_cached &t, const ge_cached tab[8], int b) {
  const uint32_t neg = (uint32_t)b >> 31;
  const uint32_t babs = (uint32_t)(b < 0 ? -b : b);
  ge_cached_0(t);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t m = ct_eq(babs, (uint32_t)(j + 1));
#pragma unroll
    for (int l = 0; l < 9; ++l) {
      t.YpX.v[l] = m ? tab[j].YpX.v[l] : t.YpX.v[l];
      t.YmX.v[l] = m ? tab[j].YmX.v[l] : t.YmX.v[l];
      t.Z2.v[l] = m ? tab[j].Z2.v[l] : t.Z2.v[l];
      t.T2d.v[l] = m ? tab[j].T2d.v[l] : t.T2d.v[l];
    }
  }
  ge_cached_cneg(t, neg);
}

// h = k * p over digits vlo..63 of a high-first stream (sodium_ge.cu:650-731:
// per window one table add and four doublings)
__device__ __forceinline__ void varbase_windows(ge_p3 &h, const ge_p3 &p, uint32_t e[16], int vlo) {
  ge_cached tab[8];
  build_cached8(tab, p);
  ge_p3_0(h);
  ge_cached c;
  ge_p1p1 r;
  ge_p2 s;
#pragma unroll 1
  for (int v = 63; v > vlo; --v) {
    cached_select(c, tab, pop_hi(e, v));
    ge_add(r, h, c);
    ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s); ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s); ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s); ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s);
    ge_p1p1_to_p3(h, r);
  }
  pin_p3(h);
  cached_select(c, tab, pop_hi(e, vlo));
  ge_add(r, h, c);
  ge_p1p1_to_p3(h, r);
}


// Radix-8 window loop for public short multipliers: h = k * p over digits
// vlo..63 of a high-first stream of signed radix-8 digits in [-4, 4], with a
// 4-entry cached table (1P..4P) and three doublings per window.  Only the
// 32-bit-multiplier kernels use it; 256-bit scalars keep radix 16, whose
// top-digit rule the reference defines.
__device__ __forceinline__ void build_cached4(ge_cached tab[4], const ge_p3 &p) {
  ge_p3 m1, m2, m3;
  ge_p1p1 t;
  ge_p3_to_cached(tab[0], p);
  ge_p3_dbl(t, p);  ge_p1p1_to_p3(m1, t);            // 2P
  ge_p3_to_cached(tab[1], m1);
  ge_add(t, p, tab[1]); ge_p1p1_to_p3(m2, t);        // 3P
  ge_p3_dbl(t, m1); ge_p1p1_to_p3(m3, t);            // 4P
  ge_p3_to_cached(tab[2], m2);
  ge_p3_to_cached(tab[3], m3);
}

// The same four multiples scaled to one shared Z (lambda_j = prod of the
// other entries' Z): each entry then needs three field elements, the select
// merges 27 words per entry instead of 36 and the table takes 117 registers
// instead of 144, which shortens the 256-bit window loop's spills.  It
// costs 19 multiplies per scalar mult: 1.6 % faster over 85 windows, slower
// over the 11 windows of a 32-bit multiplier (which keeps the plain table).
struct ge_cached3 { fe YpX, YmX, T2d; };

// cached fields of the extended point m scaled by l: (Y+X) l, (Y-X) l, 2d T l
__device__ __forceinline__ void scaled_cached3(ge_cached3 &t, const ge_p3 &m, const fe &l) {
  fe u, w;
  fe_add(u, m.Y, m.X);                 // A1
  fe_mul(t.YpX, u, l);
  fe_sub(u, m.Y, m.X);                 // A2
  fe_mul(t.YmX, u, l);
  fe_mul(w, m.T, l);
  fe_mul(t.T2d, w, FE_D2);
}

// 1P..4P, each scaled to the shared Z = Z1 Z2 Z3 Z4; the multiples are
// converted one at a time so that few full points are live together
__device__ __forceinline__ void build_cached4_coz(ge_cached3 tab[4], fe &zc, const ge_p3 &p) {
  ge_p3 m1, m2, m3;
  ge_p1p1 t;
  ge_cached c1;
  ge_p3_dbl(t, p);  ge_p1p1_to_p3(m1, t);            // 2P
  ge_p3_to_cached(c1, m1);
  ge_add(t, p, c1); ge_p1p1_to_p3(m2, t);            // 3P
  ge_p3_dbl(t, m1); ge_p1p1_to_p3(m3, t);            // 4P
  fe a, b, l;
  fe_mul(a, p.Z, m1.Z);
  fe_mul(b, m2.Z, m3.Z);
  fe_mul(zc, a, b);
  fe_mul(l, m1.Z, b);
  scaled_cached3(tab[0], p, l);
  fe_mul(l, p.Z, b);
  scaled_cached3(tab[1], m1, l);
  fe_mul(l, m3.Z, a);
  scaled_cached3(tab[2], m2, l);
  fe_mul(l, m2.Z, a);
  scaled_cached3(tab[3], m3, l);
}

// Entries of the shared-Z table parked in LDS: 27 words of an entry in
// seven 16-byte chunks, chunk q at park[q * PARK_STRIDE] (one slot per lane
// and entry, the address independent of any digit)
constexpr int PARK_STRIDE = 256;
constexpr int PARK_CHUNKS = 7;

__device__ __forceinline__ void park_cached3(uint4 *park, const ge_cached3 &e) {
  uint32_t w[28];
#pragma unroll
  for (int l = 0; l < 9; ++l) { w[l] = e.YpX.v[l]; w[9 + l] = e.YmX.v[l]; w[18 + l] = e.T2d.v[l]; }
  w[27] = 0;
#pragma unroll
  for (int q = 0; q < PARK_CHUNKS; ++q) park[q * PARK_STRIDE] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

// |b| * P from the shared-Z table with entries 1P..NR P in registers and
// the others parked in LDS (park + (j - NR) * PARK_CHUNKS * PARK_STRIDE for
// entry j, all read every window, whatever b); b = 0 gives the identity
template <int NR>
__device__ __forceinline__ void cached_select4_coz_park(ge_cached &t, const ge_cached3 tab[NR], const uint4 *park,
                                                        const fe &zc, int b) {
  const uint32_t neg = (uint32_t)b >> 31;
  const uint32_t babs = (uint32_t)(b < 0 ? -b : b);
  t.YpX = zc;
  t.YmX = zc;
  fe_0(t.T2d);
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const uint32_t m = ct_eq(babs, (uint32_t)(j + 1));
#pragma unroll
    for (int l = 0; l < 9; ++l) {
      t.YpX.v[l] = m ? tab[j].YpX.v[l] : t.YpX.v[l];
      t.YmX.v[l] = m ? tab[j].YmX.v[l] : t.YmX.v[l];
      t.T2d.v[l] = m ? tab[j].T2d.v[l] : t.T2d.v[l];
    }
  }
#pragma unroll
  for (int j = NR; j < 4; ++j) {
    const uint4 *e = park + (j - NR) * PARK_CHUNKS * PARK_STRIDE;
    uint32_t w[28];
#pragma unroll
    for (int q = 0; q < PARK_CHUNKS; ++q) {
      const uint4 x = e[q * PARK_STRIDE];
      w[4 * q] = x.x; w[4 * q + 1] = x.y; w[4 * q + 2] = x.z; w[4 * q + 3] = x.w;
    }
    const uint32_t m = ct_eq(babs, (uint32_t)(j + 1));
#pragma unroll
    for (int l = 0; l < 9; ++l) {
      t.YpX.v[l] = m ? w[l] : t.YpX.v[l];
      t.YmX.v[l] = m ? w[9 + l] : t.YmX.v[l];
      t.T2d.v[l] = m ? w[18 + l] : t.T2d.v[l];
    }
  }
  fe_add(t.Z2, zc, zc);
  ge_cached_cneg(t, neg);
}

// entries NR..3 of the lane's table in its LDS slots, as read above
template <int NR>
__device__ __forceinline__ void park_entries(uint4 *park, const ge_cached3 tab[4]) {
#pragma unroll
  for (int j = NR; j < 4; ++j) park_cached3(park + (j - NR) * PARK_CHUNKS * PARK_STRIDE, tab[j]);
}

__device__ __forceinline__ void cached_select4(ge_cached &t, const ge_cached tab[4], int b) {
  const uint32_t neg = (uint32_t)b >> 31;
  const uint32_t babs = (uint32_t)(b < 0 ? -b : b);
  ge_cached_0(t);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t m = ct_eq(babs, (uint32_t)(j + 1));
#pragma unroll
    for (int l = 0; l < 9; ++l) {
      t.YpX.v[l] = m ? tab[j].YpX.v[l] : t.YpX.v[l];
      t.YmX.v[l] = m ? tab[j].YmX.v[l] : t.YmX.v[l];
      t.Z2.v[l] = m ? tab[j].Z2.v[l] : t.Z2.v[l];
      t.T2d.v[l] = m ? tab[j].T2d.v[l] : t.T2d.v[l];
    }
  }
  ge_cached_cneg(t, neg);
}

__device__ __forceinline__ void varbase_windows8(ge_p3 &h, const ge_p3 &p, uint32_t e[16], int vlo) {
  ge_cached tab[4];
  build_cached4(tab, p);
  ge_p3_0(h);
  ge_cached c;
  ge_p1p1 r;
  ge_p2 s;
  asm volatile("" : "+s"(vlo));
#pragma unroll 1
  for (int v = 63; v > vlo; --v) {
    cached_select4(c, tab, pop_hi(e, v));
    ge_add(r, h, c);
    ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s); ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s); ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s);
    ge_p1p1_to_p3(h, r);
  }
  pin_p3(h);
  cached_select4(c, tab, pop_hi(e, vlo));
  ge_add(r, h, c);
  ge_p1p1_to_p3(h, r);
}

// ------------------------------------------ radix 8 for 256-bit scalars
//
// The reference's radix-16 recoding (sodium_ge.cu:693-708) gives a top digit
// d63 in [0, 16]; a top digit above 8 selects no table entry, so the
// reference computes (k - d63 2^252) P for such k (only k >= 2^255, outside
// the documented precondition).  scalar_effective rewrites k to exactly that
// multiple, as a magnitude and a sign, without a branch; the radix-8 loop
// then gives the reference's result for every 32-byte input.

// |k - [d63 > 8] d63 2^252| into m, its sign (1 = negative) returned
__device__ __forceinline__ uint32_t scalar_effective(uint32_t m[8], const uint32_t k[8]) {
  uint32_t e[16];
  recode16(e, k);
  const int d63 = (int)(int8_t)(uint8_t)(e[15] >> 24);
  const uint32_t big = (uint32_t)(8 - d63) >> 31;              // d63 > 8
  const uint32_t sub = (0u - big) & (uint32_t)d63;             // 0..16
#pragma unroll
  for (int i = 0; i < 7; ++i) m[i] = k[i];
  const uint64_t t = (uint64_t)k[7] - ((uint64_t)sub << 28);   // 2^252 = bit 28 of word 7
  m[7] = (uint32_t)t;
  const uint32_t neg = (uint32_t)(t >> 63);                    // borrow out of the top word
  const uint32_t mask = 0u - neg;
  uint64_t c = neg;                                            // two's complement negate when negative
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint64_t x = (uint64_t)(m[i] ^ mask) + c;
    m[i] = (uint32_t)x;
    c = x >> 32;
  }
  return neg;
}

// signed radix-8 digits of m < 2^256: 86 digits in [-4, 4] (the top one
// covers bit 255 plus the carry, at most 2) as 4-bit two's-complement
// nibbles; digit v sits at nibble position v + 2 (word (v+2) >> 3) so that
// digit 85 is the top nibble of word 10
__device__ __forceinline__ void recode8(uint32_t e[11], const uint32_t m[8]) {
  int carry = 0;
#pragma unroll
  for (int q = 0; q < 11; ++q) e[q] = 0;
#pragma unroll
  for (int v = 0; v < 86; ++v) {
    const int bit = 3 * v, w = bit >> 5, sh = bit & 31;
    uint32_t x = m[w] >> sh;
    if (sh > 29 && w + 1 < 8) x |= m[w + 1] << (32 - sh);
    int d = (int)(x & 7u) + carry;
    if (v < 85) {
      carry = (d + 4) >> 3;
      d -= carry << 3;
    }
    const int p = v + 2;
    e[p >> 3] |= ((uint32_t)d & 0xfu) << (4 * (p & 7));
  }
}

// pop the top nibble digit (sign-extended) of an 11-word high-first stream;
// p = its position
__device__ __forceinline__ int pop_hi4(uint32_t s[11], int p) {
  const int b = ((int)s[10]) >> 28;
  s[10] <<= 4;
  if ((p & 7) == 0) {
#pragma unroll
    for (int k = 10; k > 0; --k) s[k] = s[k - 1];
    s[0] = 0;
  }
  return b;
}

// h = k p for a 256-bit scalar in radix 8 (85 windows of an add and three
// doublings, then the last add), with the reference's result for k >= 2^255;
// the table's entries VB8_REG + 1 .. 4 (times P) live in the lane's LDS
// slots from `park` (see park_cached3), which cuts the 256-register window
// loop's spills from 92 to 12 bytes per lane (and its time by about 1 %)
constexpr int VB8_REG = 2;
constexpr int VB8_PARK_WORDS = (4 - VB8_REG) * PARK_CHUNKS * 4;     // per lane

__device__ __forceinline__ void varbase_mul8(ge_p3 &h, const ge_p3 &p_in, const uint32_t k[8], uint4 *park) {
  uint32_t m[8], e[11];
  const uint32_t neg = scalar_effective(m, k);
  recode8(e, m);
  // finish the digit stream here: left alone, the compiler keeps the 86
  // digits apart (spilled) until the loop and packs them there
  asm volatile("" : "+v"(e[0]), "+v"(e[1]), "+v"(e[2]), "+v"(e[3]), "+v"(e[4]), "+v"(e[5]), "+v"(e[6]),
                    "+v"(e[7]), "+v"(e[8]), "+v"(e[9]), "+v"(e[10]));
  ge_p3 p = p_in, np;
  ge_p3_neg(np, p);
  fe_cmov(p.X, np.X, neg);
  fe_cmov(p.T, np.T, neg);
  ge_cached3 tab[4];
  fe zc;
  build_cached4_coz(tab, zc, p);
  park_entries<VB8_REG>(park, tab);
  ge_p3_0(h);
  ge_cached c;
  ge_p1p1 r;
  ge_p2 s;
  int plo = 2;
  asm volatile("" : "+s"(plo));
#pragma unroll 1
  for (int q = 87; q > plo; --q) {
    cached_select4_coz_park<VB8_REG>(c, tab, park, zc, pop_hi4(e, q));
    ge_add(r, h, c);
    ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s); ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s); ge_p1p1_to_p2(s, r);
    ge_p2_dbl(r, s);
    ge_p1p1_to_p3(h, r);
  }
  pin_p3(h);
  cached_select4_coz_park<VB8_REG>(c, tab, park, zc, pop_hi4(e, plo));
  ge_add(r, h, c);
  ge_p1p1_to_p3(h, r);
}

__device__ __forceinline__ void recode16_words(uint32_t e[16], const uint32_t k[8]) { recode16(e, k); }

__device__ __forceinline__ void varbase_mul(ge_p3 &h, const ge_p3 &p, uint32_t e[16]) {
  varbase_windows(h, p, e, 0);
}

}  // namespace ft
