// Per-element bodies of the elementwise ciphertext kernels, shared by the
// single-array kernels (kernels.hip) and the batched multi-request kernels of
// the serving loop (serve.hip).
#pragma once
#include "scalarmult.h"
#include "quad.h"

namespace ft {

// a[i] = a[i] +/- b[j]
template <bool SUB>
__device__ __forceinline__ void add_one(PtView a, size_t i, PtView b, size_t j) {
  ge_p3 p;
  ge_cached c;
  load_p3(p, a, i);
  load_cached(c, b, j);
  if (SUB) ge_cached_cneg(c, 1u);
  ge_p1p1 t;
  ge_add(t, p, c);
  ge_p1p1_to_p3(p, t);
  store_p3(a, i, p);
}

// 32-bit multipliers: 9 signed digits (the bit length is public), placed at
// stream positions 55..63 so the window loop runs over nine windows only
__device__ __forceinline__ void recode_u32(uint32_t e[16], uint32_t k) {
  int carry = 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) e[q] = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    int d = (i < 8 ? (int)((k >> (4 * i)) & 15u) : 0) + carry;
    if (i < 8) {
      carry = (d + 8) >> 4;
      d -= carry << 4;
    }
    const int v = i + 55;
    e[v >> 2] |= ((uint32_t)d & 0xffu) << (8 * (v & 3));
  }
}

// 32-bit multipliers in radix 8: 11 signed digits in [-4, 4] (the top one
// covers bits 30..32 plus the carry, at most 4), at stream positions 53..63
__device__ __forceinline__ void recode_u32_r8(uint32_t e[16], uint32_t k) {
  int carry = 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) e[q] = 0;
#pragma unroll
  for (int i = 0; i < 11; ++i) {
    int d = (int)((i < 10 ? (k >> (3 * i)) : (k >> 30)) & 7u) + carry;
    if (i < 10) {
      carry = (d + 4) >> 3;
      d -= carry << 3;
    }
    const int v = i + 53;
    e[v >> 2] |= ((uint32_t)d & 0xffu) << (8 * (v & 3));
  }
}

// v[i] = k v[i] for a 32-bit k; W = 4 (radix 16, 8-entry table, 9 windows)
// or W = 3 (radix 8, 4-entry table, 11 windows).  The kernels give each point
// of a ciphertext its own lane.
template <int W>
__device__ __forceinline__ void mul_u32_point(PtView v, size_t i, uint32_t k) {
  uint32_t e[16];
  ge_p3 p, h;
  if (W == 4) recode_u32(e, k);
  else recode_u32_r8(e, k);
  // the stream words are finished here, not next to their first use
  asm volatile("" : "+v"(e[13]), "+v"(e[14]), "+v"(e[15]));
  load_p3(p, v, i);
  if (W == 4) varbase_windows(h, p, e, 55);
  else varbase_windows8(h, p, e, 53);
  store_p3(v, i, h);
}

// c1[i] = r B, c2[i] = m B + r H (ACC: c1[i] += r B, c2[i] += r H).  The two
// combs share the four middle doublings of r H with m B; the message uses nd
// digits (nd = 0 for rerandomisation), a public bound on its bit length.
// tb, th: LDS comb tables of B and H; r: the lane's 8-word scalar.
template <bool ACC>
__device__ __forceinline__ void encrypt_one(PtView c1, PtView c2, size_t i, const uint32_t *tb,
                                            const uint32_t *th, uint64_t x, int nd, const uint32_t *r) {
  uint32_t k[8];
  ld_words(k, r);
  // C1 = r B, finished and stored before C2 starts so that only one comb's
  // digit streams and accumulator are live at a time
  {
    uint32_t e[16], o1[8], e1[8];
    recode16(e, k);
    split_odd_even(o1, e1, e);
    ge_p3 a1;
    ge_p3_0(a1);
    comb_run<8>(a1, tb, tb, o1, e1, 32, 32);
    if (ACC) {
      add_stored(a1, c1, i);
    }
    store_p3(c1, i, a1);
  }
  // C2 = r H + m B, sharing the four doublings: rows of H for r's digits
  // (stream words 0..7), rows of B for the message's (at most 17 digits:
  // 8 odd, 9 even, in words 8..10)
  uint32_t e[16], odd[11], even[11];
  recode16(e, k);
#pragma unroll
  for (int w = 8; w < 11; ++w) { odd[w] = 0; even[w] = 0; }
  if (!ACC) {
    int carry = 0;
#pragma unroll
    for (int q = 0; q < 17; ++q) {
      int d = (q < 16 ? (int)((x >> (4 * q)) & 15u) : 0) + carry;
      carry = (d + 8) >> 4;
      d -= carry << 4;
      if (q == nd - 1) d += carry << 4;       // top digit keeps the final carry
      const uint32_t db = q < nd ? ((uint32_t)d & 0xffu) : 0u;
      const int h = q >> 1;
      if (q & 1) odd[8 + (h >> 2)] |= db << (8 * (h & 3));
      else even[8 + (h >> 2)] |= db << (8 * (h & 3));
      if (q >= nd - 1) carry = 0;
    }
  }
  const int mo = ACC ? 0 : nd >> 1, me = ACC ? 0 : (nd + 1) >> 1;
  split_odd_even(odd, even, e);
  ge_p3 a2;
  ge_p3_0(a2);
  comb_run<11>(a2, th, tb, odd, even, 32 + mo, 32 + me);
  if (ACC) {
    add_stored(a2, c2, i);
  }
  store_p3(c2, i, a2);
}

// One half of an encryption over radix-32 tables: half 0 C1 = r B, half 1
// C2 = r H + m B (ACC: C1 += r B, C2 += r H).  The half must be the same on
// every lane of the wave: the rows are read by scalar loads from the table
// the half selects.  r is rewritten by the radix-16 top-digit rule and
// reduced mod l as for the fixed base (the same multiple of B and H, both
// of order l); the message takes nd32 rows of B's table.
template <bool ACC>
__device__ __forceinline__ void encrypt_half32(PtView c1, PtView c2, size_t i, uint32_t half,
                                               const uint32_t *__restrict__ tb32, const uint32_t *__restrict__ th32,
                                               uint64_t x, int nd32, const uint32_t *r) {
  uint32_t k[8], mm[8], e[13];
  ld_words(k, r);
  const uint32_t neg = scalar_effective(mm, k);
  scalar_mod_l(mm);
  scalar_cneg_l(mm, neg);
  recode32<COMB32_ROWS>(e, mm);
  ge_p3 a;
  comb32_run<13>(a, half ? th32 : tb32, e, COMB32_ROWS);
  if (!ACC && half) {
    uint32_t d[4];
    recode32_msg(d, x, nd32);
    comb32_add_rows<4>(a, tb32, d, nd32);
  }
  const PtView dst = half ? c2 : c1;
  if (ACC) add_stored(a, dst, i);
  store_p3(dst, i, a);
}

// One wave's part of a ciphertext half in the four-wave schedule: the rows of
// the radix-32 comb of B (half 0, C1) or H (half 1, C2) are split at row
// `split` (wave-uniform), part 0 summing rows [0, split) and part 1 rows
// [split, 51), plus, for C2, the message rows over B's table.  The two parts'
// sums add up to encrypt_half32's point.
template <bool ACC>
__device__ __forceinline__ void encrypt_part32(ge_p3 &a, uint32_t half, uint32_t part, int split,
                                               const uint32_t *__restrict__ tb32, const uint32_t *__restrict__ th32,
                                               uint64_t x, int nd32, const uint32_t *r) {
  uint32_t k[8], mm[8], e[13];
  ld_words(k, r);
  const uint32_t neg = scalar_effective(mm, k);
  scalar_mod_l(mm);
  scalar_cneg_l(mm, neg);
  recode32<COMB32_ROWS>(e, mm);
  const uint32_t *tab = half ? th32 : tb32;
  if (!part) {
    comb32_run<13>(a, tab, e, split);
  } else {
    comb32_run_from<13>(a, tab, e, split, COMB32_ROWS);
    if (!ACC && half) {
      uint32_t d[4];
      recode32_msg(d, x, nd32);
      comb32_add_rows<4>(a, tb32, d, nd32);
    }
  }
}

}  // namespace ft
