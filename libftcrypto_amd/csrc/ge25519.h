// Edwards25519 group arithmetic on the 9x29 field layer.
//
// Replaces the reference's ge25519 layer (sodium_ge.cu:10-899).  Formulas are
// the extended twisted-Edwards ones of Hisil, Wong, Carter and Dawson (2008)
// with a = -1, as ref10 uses them:
//   unified add (p3 + cached -> completed) ......... sodium_ge.cu:76-92
//   mixed add with an affine (y+x, y-x, 2dxy) entry . sodium_ge.cu:231-246
//   doubling of a projective point .................. sodium_ge.cu:306-320
//   completed -> projective / extended .............. sodium_ge.cu:273-292
// The completed form is kept as (E, F, G, H) with X = E F, Y = G H,
// Z = F G, T = E H.  Bound contracts (classes of fe25519.h):
//   ge_p3 / ge_p2 ......... every coordinate N
//   ge_cached ............. YpX, YmX <= A1, Z2 <= A1, T2d <= A1
//   ge_precomp ............ every field N
//   ge_p1p1 (E,F,G,H) ..... from dbl: E <= A2, F N, G <= A2, H <= A1;
//                           from add and madd: E, G <= A1, F, H <= A2;
//                           from sub: E, G <= A2, F, H <= A1
// which keeps every fe_mul operand pair inside the exact set.
// ge_cached holds 2 Z (Z2) rather than Z: the add's D = 2 Z1 Z2 is then one
// product, and with the sums arranged as below no field of the completed
// form needs a carry.
// ge_precomp is the affine form of a table entry with its sums halved:
// ((y + x)/2, (y - x)/2, d x y).  Every product of a mixed add then comes
// out halved, so its completed form (all four coordinates half of the usual
// ones: the same point) needs neither D = 2 Z nor a carry of F = D - C.
#pragma once
#include "fe25519.h"

namespace ft {

struct ge_p3 { fe X, Y, Z, T; };
struct ge_p2 { fe X, Y, Z; };
struct ge_p1p1 { fe E, F, G, H; };
struct ge_cached { fe YpX, YmX, Z2, T2d; };   // Y + X, Y - X, 2 Z, 2d T
struct ge_precomp { fe ypx_h, ymx_h, xyd; };   // (y + x)/2, (y - x)/2, d x y

FT_HD void ge_p3_0(ge_p3 &h) { fe_0(h.X); fe_1(h.Y); fe_1(h.Z); fe_0(h.T); }
FT_HD void ge_cached_0(ge_cached &h) { fe_1(h.YpX); fe_1(h.YmX); fe_0(h.Z2); h.Z2.v[0] = 2; fe_0(h.T2d); }
FT_HD void ge_precomp_0(ge_precomp &h) { h.ypx_h = FE_INV2; h.ymx_h = FE_INV2; fe_0(h.xyd); }

FT_HD void ge_p1p1_to_p3(ge_p3 &r, const ge_p1p1 &p) {
  fe_mul(r.X, p.E, p.F);
  fe_mul(r.Y, p.G, p.H);
  fe_mul(r.Z, p.G, p.F);
  fe_mul(r.T, p.E, p.H);
}

FT_HD void ge_p1p1_to_p2(ge_p2 &r, const ge_p1p1 &p) {
  fe_mul(r.X, p.E, p.F);
  fe_mul(r.Y, p.G, p.H);
  fe_mul(r.Z, p.G, p.F);
}

FT_HD void ge_p3_to_cached(ge_cached &r, const ge_p3 &p) {
  fe_add(r.YpX, p.Y, p.X);
  fe_sub(r.YmX, p.Y, p.X);
  fe_carry(r.YmX);
  fe_add(r.Z2, p.Z, p.Z);
  fe_mul(r.T2d, p.T, FE_D2);
}

// r = p + q (unified: also correct for p == q and for the identity).  With
// A = (Y1 - X1)(Y2 - X2), B = (Y1 + X1)(Y2 + X2), C = 2d T1 T2, D = 2 Z1 Z2:
// E = B - A, F = D - C, G = D + C, H = B + A, computed as E = B + (-A) from
// (X1 - Y1)(Y2 - X2) = -A, so that E and G are sums (<= A1) and F and H
// differences (<= A2): every product of the p1p1 conversions pairs A1 with A2.
FT_HD void ge_add(ge_p1p1 &r, const ge_p3 &p, const ge_cached &q) {
  fe a, b, nA, B, C, D;
  fe_sub(a, p.X, p.Y);        // A2
  fe_add(b, p.Y, p.X);        // A1
  fe_mul(nA, a, q.YmX);
  fe_mul(B, b, q.YpX);
  fe_mul(C, p.T, q.T2d);
  fe_mul(D, p.Z, q.Z2);
  fe_add(r.E, B, nA);         // A1
  fe_sub(r.H, B, nA);         // A2
  fe_add(r.G, D, C);          // A1
  fe_sub(r.F, D, C);          // A2
}

// r = p - q = p + (-q), -q = (YmX, YpX, Z2, -T2d) read in place: A, B and D
// as for the add of -q, C = 2d T1 T2 (so -q's C is -C); E = B - A, F = D + C,
// G = D - C, H = B + A: E, G <= A2 and F, H <= A1.
FT_HD void ge_sub(ge_p1p1 &r, const ge_p3 &p, const ge_cached &q) {
  fe a, b, A, B, C, D;
  fe_sub(a, p.Y, p.X);        // A2
  fe_add(b, p.Y, p.X);        // A1
  fe_mul(A, a, q.YpX);
  fe_mul(B, b, q.YmX);
  fe_mul(C, p.T, q.T2d);
  fe_mul(D, p.Z, q.Z2);
  fe_sub(r.E, B, A);          // A2
  fe_add(r.H, B, A);          // A1
  fe_add(r.F, D, C);          // A1
  fe_sub(r.G, D, C);          // A2
}

// r = p + q for an affine precomputed q (halved sums, see above).  With
// A = (Y - X)(y - x), B = (Y + X)(y + x), C = 2d T x y, D = 2 Z the usual
// completed form is E = B - A, F = D - C, G = D + C, H = B + A; here each is
// produced halved, from (X - Y) (y - x)/2 = -A/2, (Y + X) (y + x)/2 = B/2 and
// T d x y = C/2: E/2 = B/2 + (-A/2) <= A1, H/2 = B/2 - (-A/2) <= A2,
// G/2 = Z + C/2 <= A1, F/2 = Z - C/2 <= A2.
FT_HD void ge_madd(ge_p1p1 &r, const ge_p3 &p, const ge_precomp &q) {
  fe a, b, nA, B, C;
  fe_sub(a, p.X, p.Y);        // A2
  fe_add(b, p.Y, p.X);        // A1
  fe_mul(nA, a, q.ymx_h);
  fe_mul(B, b, q.ypx_h);
  fe_mul(C, p.T, q.xyd);
  fe_add(r.E, B, nA);         // A1
  fe_sub(r.H, B, nA);         // A2
  fe_add(r.G, p.Z, C);        // A1
  fe_sub(r.F, p.Z, C);        // A2
}

// r = 2 p: E = 2XY, G = Y^2 - X^2, H = X^2 + Y^2, F = 2Z^2 - G
// (the dbl-2008-hwcd values with F and H negated, which leaves the point unchanged)
FT_HD void ge_dbl(ge_p1p1 &r, const fe &X, const fe &Y, const fe &Z) {
  fe XX, YY, ZZ, t;
  fe_sq(XX, X);
  fe_sq(YY, Y);
  fe_sq(ZZ, Z);
  fe_add(t, Y, Y);
  fe_mul(r.E, X, t);
  fe_sub(r.G, YY, XX);
  fe_add(r.H, XX, YY);
  fe_add(t, ZZ, ZZ);
  fe_add(t, t, XX);
  fe_sub(r.F, t, YY);         // <= 3N + 2^30
  fe_carry(r.F);
}

FT_HD void ge_p3_dbl(ge_p1p1 &r, const ge_p3 &p) { ge_dbl(r, p.X, p.Y, p.Z); }
FT_HD void ge_p2_dbl(ge_p1p1 &r, const ge_p2 &p) { ge_dbl(r, p.X, p.Y, p.Z); }

// One step of a double-and-add over PUBLIC signed digits: s = 2 s + d P for
// d in {0, +-1, +-3}, c1 = P and c3 = 3P in cached form.  Branches on d, so
// only for scalars that are not secret.
FT_HD void ge_dbl_add_public(ge_p2 &s, int d, const ge_cached &c1, const ge_cached &c3) {
  ge_p1p1 r;
  ge_p2_dbl(r, s);
  if (d != 0) {
    ge_p3 h;
    ge_p1p1_to_p3(h, r);
    if (d == 1) ge_add(r, h, c1);
    else if (d == -1) ge_sub(r, h, c1);
    else if (d == 3) ge_add(r, h, c3);
    else ge_sub(r, h, c3);
  }
  ge_p1p1_to_p2(s, r);
}

// -q for a cached point when neg = 1: swap Y+X and Y-X (two selects per
// limb), negate 2dT (T2d of class N) as (x ^ mask) + (mask & (K + 1)), one
// instruction per limb; the values are those of fe_cswap and fe_neg
FT_HD void ge_cached_cneg(ge_cached &q, uint32_t neg) {
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    const uint32_t a = q.YpX.v[l], b = q.YmX.v[l];
    q.YpX.v[l] = neg ? b : a;
    q.YmX.v[l] = neg ? a : b;
  }
  fe_cneg_mask(q.T2d, q.T2d, 0u - neg);
}

// -q for neg = 1: (-x, y) swaps the halved sums and negates d x y
FT_HD void ge_precomp_cneg(ge_precomp &q, uint32_t neg) {
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    const uint32_t a = q.ypx_h.v[l], b = q.ymx_h.v[l];
    q.ypx_h.v[l] = neg ? b : a;
    q.ymx_h.v[l] = neg ? a : b;
  }
  fe_cneg_mask(q.xyd, q.xyd, 0u - neg);
}

// full p3 + p3 -> p3 (9M)
FT_HD void ge_p3_add(ge_p3 &r, const ge_p3 &p, const ge_p3 &q) {
  ge_cached c;
  ge_p1p1 t;
  ge_p3_to_cached(c, q);
  ge_add(t, p, c);
  ge_p1p1_to_p3(r, t);
}

FT_HD void ge_p3_sub(ge_p3 &r, const ge_p3 &p, const ge_p3 &q) {
  ge_cached c;
  ge_p1p1 t;
  ge_p3_to_cached(c, q);
  ge_cached_cneg(c, 1u);
  ge_add(t, p, c);
  ge_p1p1_to_p3(r, t);
}

// -p = (-x, y): negate X and T (the reference's k_ft_neg negates Y and T,
// ftcrypto.cu:163-167, which yields -p plus the point of order two)
FT_HD void ge_p3_neg(ge_p3 &r, const ge_p3 &p) {
  fe_neg(r.X, p.X); fe_carry(r.X);
  r.Y = p.Y; r.Z = p.Z;
  fe_neg(r.T, p.T); fe_carry(r.T);
}

// affine precomputed form of a point with halved sums (one inversion):
// ((y + x)/2, (y - x)/2, d x y)
FT_HD void ge_p3_to_precomp(ge_precomp &r, const ge_p3 &p) {
  fe zi, x, y, xh, yh, xy;
  fe_invert(zi, p.Z);
  fe_mul(x, p.X, zi);
  fe_mul(y, p.Y, zi);
  fe_mul(xh, x, FE_INV2);
  fe_mul(yh, y, FE_INV2);
  fe_add(r.ypx_h, yh, xh); fe_carry(r.ypx_h);
  fe_sub(r.ymx_h, yh, xh); fe_carry(r.ymx_h);
  fe_mul(xy, x, y);
  fe_mul(r.xyd, xy, FE_D);
}

// the extended point of an entry in that form (one product): x = ypx_h -
// ymx_h, y = ypx_h + ymx_h, Z = 1, T = (d x y) / d; the fields may be
// swapped and d x y negated (ge_precomp_cneg)
FT_HD void ge_precomp_to_p3(ge_p3 &r, const ge_precomp &q) {
  fe_sub(r.X, q.ypx_h, q.ymx_h); fe_carry(r.X);
  fe_add(r.Y, q.ypx_h, q.ymx_h); fe_carry(r.Y);
  fe_1(r.Z);
  fe_mul(r.T, q.xyd, FE_INV_D);
}

// ---------------------------------------------------------------- encoding

// compressed encoding words with affine x, y known (sodium_ge.cu:382-394)
FT_HD void ge_encode_xy(uint32_t w[8], const fe &x, const fe &y) {
  uint32_t xw[8];
  fe_canonical_words(w, y);
  fe_canonical_words(xw, x);
  w[7] |= (xw[0] & 1u) << 31;
}

FT_HD void ge_p3_encode(uint32_t w[8], const ge_p3 &p) {
  fe zi, x, y;
  fe_invert(zi, p.Z);
  fe_mul(x, p.X, zi);
  fe_mul(y, p.Y, zi);
  ge_encode_xy(w, x, y);
}

// Decompression with the reference's semantics (sodium_ge.cu:137-180):
// y from bits 0..254 (values >= p accepted), x = sqrt((y^2-1)/(dy^2+1)) with
// the sign of bit 255; x = 0 with the sign bit set is accepted.  Returns 1 on
// success, 0 when no square root exists (the point is then unspecified).
FT_HD uint32_t ge_decode(ge_p3 &h, const uint32_t s[8]) {
  uint32_t yw[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) yw[i] = s[i];
  yw[7] &= 0x7fffffffu;
  const uint32_t sign = s[7] >> 31;
  fe u, v, v3, vxx, chk, xs, one;
  fe_from_words(h.Y, yw);
  fe_1(one);
  h.Z = one;
  fe_sq(u, h.Y);
  fe_mul(v, u, FE_D);
  fe_sub(u, u, one); fe_carry(u);        // y^2 - 1
  fe_add(v, v, one);                     // d y^2 + 1
  fe_sq(v3, v);
  fe_mul(v3, v3, v);                     // v^3
  fe_sq(h.X, v3);
  fe_mul(h.X, h.X, v);
  fe_mul(h.X, h.X, u);                   // u v^7
  fe_pow22523(h.X, h.X);
  fe_mul(h.X, h.X, v3);
  fe_mul(h.X, h.X, u);                   // u v^3 (u v^7)^((p-5)/8)
  fe_sq(vxx, h.X);
  fe_mul(vxx, vxx, v);
  fe_sub(chk, vxx, u);
  const uint32_t m_root = fe_iszero(chk);
  fe_add(chk, vxx, u);
  const uint32_t p_root = fe_iszero(chk);
  fe_mul(xs, h.X, FE_SQRTM1);
  fe_cmov(h.X, xs, 1u - m_root);
  fe_neg(xs, h.X); fe_carry(xs);
  fe_cmov(h.X, xs, fe_isnegative(h.X) ^ sign);
  fe_mul(h.T, h.X, h.Y);
  return m_root | p_root;
}

// ------------------------------------------------------------ helpers

FT_HD uint32_t ct_eq(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)(a ^ b) - 1) >> 63); }

}  // namespace ft
