// Resident layout of point arrays in HBM and the comb tables in LDS.
//
// A point array of capacity `cap` is four coordinate planes X | Y | Z | T,
// each `cap` x 32 bytes: element i's coordinate c is the 256-bit value (not
// reduced mod p) at word offset (c * cap + i) * 8.  A wavefront reading one
// coordinate of 64 consecutive elements touches one contiguous 2 KiB run, in
// two 16-byte loads per lane.  (The reference keeps an array-of-structures of
// 160-byte ge25519_p3 in managed memory, ftcrypto.cu:21-27, 215-242.)
//
// A comb table holds 32 rows x 8 affine entries ((y+x)/2, (y-x)/2, dxy) in
// the 9-limb form, each entry padded to 28 words (112 bytes): 28,672 bytes.
#pragma once
#include <hip/hip_runtime.h>
#include "ge25519.h"

namespace ft {

struct PtView {
  uint32_t *w;    // base of plane X
  size_t cap;     // elements per plane
};

constexpr int COMB_ROWS = 32;
constexpr int COMB_ENTRY_WORDS = 28;
constexpr int COMB_TABLE_WORDS = COMB_ROWS * 8 * COMB_ENTRY_WORDS;   // 7168 words
// radix-32 comb of the base point: 51 rows of 16 entries, (j+1) 32^r B
constexpr int COMB32_ROWS = 51;
constexpr int COMB32_ROW_WORDS = 16 * COMB_ENTRY_WORDS;
constexpr int COMB32_TABLE_WORDS = COMB32_ROWS * COMB32_ROW_WORDS;  // 22848 words, 89.25 KiB

__device__ __forceinline__ void ld_words(uint32_t w[8], const uint32_t *p) {
  const uint4 a = *reinterpret_cast<const uint4 *>(p);
  const uint4 b = *reinterpret_cast<const uint4 *>(p + 4);
  w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
  w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
}

__device__ __forceinline__ void st_words(uint32_t *p, const uint32_t w[8]) {
  *reinterpret_cast<uint4 *>(p) = make_uint4(w[0], w[1], w[2], w[3]);
  *reinterpret_cast<uint4 *>(p + 4) = make_uint4(w[4], w[5], w[6], w[7]);
}

__device__ __forceinline__ void load_fe(fe &r, const uint32_t *p) {
  uint32_t w[8];
  ld_words(w, p);
  fe_from_words(r, w);
}

__device__ __forceinline__ void store_fe(uint32_t *p, const fe &a) {
  uint32_t w[8];
  fe_store_words(w, a);
  st_words(p, w);
}

// Resident point format: planes X | Y | Z | 2d T.  The fourth plane holds T
// premultiplied by 2d, the form in which the unified add takes its second
// operand (ge_cached::T2d): the aggregation walks then add a loaded point
// without a product of its own (add_pw), for one product per point
// written.  load_p3 / un_pw give the extended point (T = plane 3 / 2d),
// store_p3 / pk_pw write one.
__device__ __forceinline__ void load_p3(ge_p3 &r, const PtView v, size_t i) {
  fe t2d;
  load_fe(r.X, v.w + (0 * v.cap + i) * 8);
  load_fe(r.Y, v.w + (1 * v.cap + i) * 8);
  load_fe(r.Z, v.w + (2 * v.cap + i) * 8);
  load_fe(t2d, v.w + (3 * v.cap + i) * 8);
  fe_mul(r.T, t2d, FE_INV_D2);
}

__device__ __forceinline__ void store_p3(const PtView v, size_t i, const ge_p3 &r) {
  fe t2d;
  fe_mul(t2d, r.T, FE_D2);
  store_fe(v.w + (0 * v.cap + i) * 8, r.X);
  store_fe(v.w + (1 * v.cap + i) * 8, r.Y);
  store_fe(v.w + (2 * v.cap + i) * 8, r.Z);
  store_fe(v.w + (3 * v.cap + i) * 8, t2d);
}

// the cached form (Y + X, Y - X, 2 Z, 2d T) of the point at v[i], no product
__device__ __forceinline__ void load_cached(ge_cached &r, const PtView v, size_t i) {
  fe x, y, z;
  load_fe(x, v.w + (0 * v.cap + i) * 8);
  load_fe(y, v.w + (1 * v.cap + i) * 8);
  load_fe(z, v.w + (2 * v.cap + i) * 8);
  load_fe(r.T2d, v.w + (3 * v.cap + i) * 8);
  fe_add(r.YpX, y, x);
  fe_sub(r.YmX, y, x);
  fe_carry(r.YmX);
  fe_add(r.Z2, z, z);
}

// identity (0, 1, 1, 0) written without arithmetic
__device__ __forceinline__ void store_identity(const PtView v, size_t i) {
  const uint32_t z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t o[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  st_words(v.w + (0 * v.cap + i) * 8, z);
  st_words(v.w + (1 * v.cap + i) * 8, o);
  st_words(v.w + (2 * v.cap + i) * 8, o);
  st_words(v.w + (3 * v.cap + i) * 8, z);
}

// copy of the 32 packed bytes of each coordinate, no arithmetic.  All four
// coordinates are loaded before any is stored: the two arrays may alias, so
// the compiler would otherwise keep each load behind the previous store (four
// memory round trips in a row instead of one).
__device__ __forceinline__ void copy_point(const PtView d, size_t i, const PtView s, size_t j) {
  uint32_t w[4][8];
#pragma unroll
  for (int c = 0; c < 4; ++c) ld_words(w[c], s.w + (c * s.cap + j) * 8);
#pragma unroll
  for (int c = 0; c < 4; ++c) st_words(d.w + (c * d.cap + i) * 8, w[c]);
}

// A point as its packed 32-byte coordinates.  A lane walking consecutive
// points loads four of them (i0 a multiple of 4: per coordinate one whole
// 128-byte line) before using any, so no line is fetched twice; `pin_pw` is an
// empty asm that takes the loaded words as register operands, which makes the
// compiler issue all 32 loads up front instead of sinking them to their uses.
struct PW { uint32_t x[8], y[8], z[8], t[8]; };

__device__ __forceinline__ void ld_pw(PW &p, const PtView v, size_t i) {
  ld_words(p.x, v.w + (0 * v.cap + i) * 8);
  ld_words(p.y, v.w + (1 * v.cap + i) * 8);
  ld_words(p.z, v.w + (2 * v.cap + i) * 8);
  ld_words(p.t, v.w + (3 * v.cap + i) * 8);
}

__device__ __forceinline__ void un_pw(ge_p3 &r, const PW &p) {
  fe t2d;
  fe_from_words(r.X, p.x);
  fe_from_words(r.Y, p.y);
  fe_from_words(r.Z, p.z);
  fe_from_words(t2d, p.t);
  fe_mul(r.T, t2d, FE_INV_D2);
}

// the cached form of a point from its resident words (no product)
__device__ __forceinline__ void un_pw_cached(ge_cached &r, const PW &p) {
  fe x, y, z;
  fe_from_words(x, p.x);
  fe_from_words(y, p.y);
  fe_from_words(z, p.z);
  fe_from_words(r.T2d, p.t);
  fe_add(r.YpX, y, x);
  fe_sub(r.YmX, y, x);
  fe_carry(r.YmX);
  fe_add(r.Z2, z, z);
}

// acc += v[i] (8 products: the stored point's 2d T enters the add as is)
__device__ __forceinline__ void add_stored(ge_p3 &acc, const PtView v, size_t i) {
  ge_cached c;
  ge_p1p1 t;
  load_cached(c, v, i);
  ge_add(t, acc, c);
  ge_p1p1_to_p3(acc, t);
}

// acc += the point whose resident words are q: 8 products
__device__ __forceinline__ void add_pw(ge_p3 &acc, const PW &q) {
  ge_cached c;
  ge_p1p1 t;
  un_pw_cached(c, q);
  ge_add(t, acc, c);
  ge_p1p1_to_p3(acc, t);
}

__device__ __forceinline__ void st_pw(const PtView v, size_t i, const PW &p) {
  st_words(v.w + (0 * v.cap + i) * 8, p.x);
  st_words(v.w + (1 * v.cap + i) * 8, p.y);
  st_words(v.w + (2 * v.cap + i) * 8, p.z);
  st_words(v.w + (3 * v.cap + i) * 8, p.t);
}

// points g .. g + cnt - 1 (cnt <= 4) stored coordinate by coordinate, so that
// each coordinate's line (four consecutive points, g a multiple of 4) is
// written by consecutive stores
__device__ __forceinline__ void st_pw4(const PtView v, size_t g, int cnt, const PW &a, const PW &b, const PW &c,
                                       const PW &d) {
  uint32_t *px = v.w + (0 * v.cap + g) * 8, *py = v.w + (1 * v.cap + g) * 8;
  uint32_t *pz = v.w + (2 * v.cap + g) * 8, *pt = v.w + (3 * v.cap + g) * 8;
  st_words(px, a.x);
  if (cnt > 1) st_words(px + 8, b.x);
  if (cnt > 2) st_words(px + 16, c.x);
  if (cnt > 3) st_words(px + 24, d.x);
  st_words(py, a.y);
  if (cnt > 1) st_words(py + 8, b.y);
  if (cnt > 2) st_words(py + 16, c.y);
  if (cnt > 3) st_words(py + 24, d.y);
  st_words(pz, a.z);
  if (cnt > 1) st_words(pz + 8, b.z);
  if (cnt > 2) st_words(pz + 16, c.z);
  if (cnt > 3) st_words(pz + 24, d.z);
  st_words(pt, a.t);
  if (cnt > 1) st_words(pt + 8, b.t);
  if (cnt > 2) st_words(pt + 16, c.t);
  if (cnt > 3) st_words(pt + 24, d.t);
}

__device__ __forceinline__ void pk_pw(PW &p, const ge_p3 &r) {
  fe t2d;
  fe_mul(t2d, r.T, FE_D2);
  fe_store_words(p.x, r.X);
  fe_store_words(p.y, r.Y);
  fe_store_words(p.z, r.Z);
  fe_store_words(p.t, t2d);
}

__device__ __forceinline__ void pin_pw(PW &p) {
  asm volatile("" : "+v"(p.x[0]), "+v"(p.x[1]), "+v"(p.x[2]), "+v"(p.x[3]), "+v"(p.x[4]), "+v"(p.x[5]),
                    "+v"(p.x[6]), "+v"(p.x[7]), "+v"(p.y[0]), "+v"(p.y[1]), "+v"(p.y[2]), "+v"(p.y[3]),
                    "+v"(p.y[4]), "+v"(p.y[5]), "+v"(p.y[6]), "+v"(p.y[7]));
  asm volatile("" : "+v"(p.z[0]), "+v"(p.z[1]), "+v"(p.z[2]), "+v"(p.z[3]), "+v"(p.z[4]), "+v"(p.z[5]),
                    "+v"(p.z[6]), "+v"(p.z[7]), "+v"(p.t[0]), "+v"(p.t[1]), "+v"(p.t[2]), "+v"(p.t[3]),
                    "+v"(p.t[4]), "+v"(p.t[5]), "+v"(p.t[6]), "+v"(p.t[7]));
}

// an empty asm taking a field element's limbs as register operands: a
// scheduling fence that keeps work on the limbs from moving across it
__device__ __forceinline__ void pin_fe(fe &a) {
  asm volatile("" : "+v"(a.v[0]), "+v"(a.v[1]), "+v"(a.v[2]), "+v"(a.v[3]), "+v"(a.v[4]), "+v"(a.v[5]),
                    "+v"(a.v[6]), "+v"(a.v[7]), "+v"(a.v[8]));
}

__device__ __forceinline__ void pin_p3(ge_p3 &p) {
  pin_fe(p.X);
  pin_fe(p.Y);
  pin_fe(p.Z);
  pin_fe(p.T);
}

__device__ __forceinline__ void pin_cached(ge_cached &p) {
  pin_fe(p.YpX);
  pin_fe(p.YmX);
  pin_fe(p.Z2);
  pin_fe(p.T2d);
}

// Point-contiguous copy kept in the tail slots of a scratch array's planes
// (from slot `base` on): point j's 128 bytes (X, Y, Z, T) are the four
// consecutive slots base + 4 (j >> 2) .. +3 of plane j & 3.
__device__ __forceinline__ uint32_t *aos_point(const PtView v, size_t base, size_t j) {
  return v.w + ((j & 3) * v.cap + base + 4 * (j >> 2)) * 8;
}

__device__ __forceinline__ void ld_pw_aos(PW &p, const PtView v, size_t base, size_t j) {
  const uint32_t *s = aos_point(v, base, j);
  ld_words(p.x, s);
  ld_words(p.y, s + 8);
  ld_words(p.z, s + 16);
  ld_words(p.t, s + 24);
}

// points g .. g+3 (indices at or past `hi` re-read point g; their values are unused)
#define FT_LOAD4(q0, q1, q2, q3, v, g, hi)                   \
  do {                                                       \
    ld_pw(q0, v, g);                                         \
    ld_pw(q1, v, ((g) + 1 < (hi)) ? (g) + 1 : (g));          \
    ld_pw(q2, v, ((g) + 2 < (hi)) ? (g) + 2 : (g));          \
    ld_pw(q3, v, ((g) + 3 < (hi)) ? (g) + 3 : (g));          \
    pin_pw(q0); pin_pw(q1); pin_pw(q2); pin_pw(q3);           \
  } while (0)

// ---------------------------------------------------------- comb entries

__device__ __forceinline__ void comb_entry_store(uint32_t *dst, const ge_precomp &e) {
#pragma unroll
  for (int i = 0; i < 9; ++i) { dst[i] = e.ypx_h.v[i]; dst[9 + i] = e.ymx_h.v[i]; dst[18 + i] = e.xyd.v[i]; }
  dst[27] = 0;
}

// Constant-time selection of |b| * row entry from an LDS-resident row of NE
// entries: every entry is read (the LDS address never depends on b) and
// merged with v_cndmask; b < 0 negates the result.  b outside [-NE, NE]
// selects the identity (for NE = 8 the reference's cmov8 behaviour,
// sodium_ge.cu:458-488).
template <int NE = 8>
__device__ __forceinline__ void comb_select(ge_precomp &t, const uint32_t *row, int b) {
  const uint32_t neg = (uint32_t)b >> 31;
  const uint32_t babs = (uint32_t)(b < 0 ? -b : b);
  ge_precomp_0(t);
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const uint32_t m = ct_eq(babs, (uint32_t)(j + 1));
    const uint32_t *e = row + j * COMB_ENTRY_WORDS;
    uint32_t w[28];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const uint4 x = reinterpret_cast<const uint4 *>(e)[q];
      w[4 * q] = x.x; w[4 * q + 1] = x.y; w[4 * q + 2] = x.z; w[4 * q + 3] = x.w;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      t.ypx_h.v[i] = m ? w[i] : t.ypx_h.v[i];
      t.ymx_h.v[i] = m ? w[9 + i] : t.ymx_h.v[i];
      t.xyd.v[i] = m ? w[18 + i] : t.xyd.v[i];
    }
  }
  ge_precomp_cneg(t, neg);
}

// The same selection from a row in global memory at a wave-uniform address:
// the words come in through scalar loads and are merged with v_bfi_b32
// under a per-lane mask (a v_cndmask cannot take a scalar operand besides
// its lane mask), so a row costs no LDS traffic.  Constant time as above:
// every entry is read, at an address that never depends on b.
template <int NE>
__device__ __forceinline__ void comb_select_uniform(ge_precomp &t, const uint32_t *__restrict__ row, int b) {
  const uint32_t neg = (uint32_t)b >> 31;
  const uint32_t babs = (uint32_t)(b < 0 ? -b : b);
  ge_precomp_0(t);
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    uint32_t mk = 0u - (uint32_t)(babs == (uint32_t)(j + 1));
    // keep the mask opaque: known to be 0 or ~0, the merge below would be
    // turned back into a select, which needs each scalar word moved to a
    // vector register first
    asm volatile("" : "+v"(mk));
    const uint32_t *e = row + j * COMB_ENTRY_WORDS;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      t.ypx_h.v[i] = (e[i] & mk) | (t.ypx_h.v[i] & ~mk);
      t.ymx_h.v[i] = (e[9 + i] & mk) | (t.ymx_h.v[i] & ~mk);
      t.xyd.v[i] = (e[18 + i] & mk) | (t.xyd.v[i] & ~mk);
    }
  }
  ge_precomp_cneg(t, neg);
}

}  // namespace ft
