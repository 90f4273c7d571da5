// Host build of the library's field and group headers, for CPU tests of the
// exact arithmetic the gfx950 kernels run (bounds, reduction, encoding).
#include <cstring>
#include "../../libftcrypto_amd/csrc/fe25519.h"
#include "../../libftcrypto_amd/csrc/ge25519.h"
#include "../../libftcrypto_amd/csrc/scalar.h"

using namespace ft;

static fe ld(const uint32_t *a) { fe x; std::memcpy(x.v, a, 36); return x; }
static void st(uint32_t *r, const fe &x) { std::memcpy(r, x.v, 36); }

extern "C" {
void h_fe_mul(uint32_t *r, const uint32_t *a, const uint32_t *b) { fe z; fe_mul(z, ld(a), ld(b)); st(r, z); }
void h_fe_sq(uint32_t *r, const uint32_t *a) { fe z; fe_sq(z, ld(a)); st(r, z); }
void h_fe_carry(uint32_t *r, const uint32_t *a) { fe z = ld(a); fe_carry(z); st(r, z); }
void h_fe_invert(uint32_t *r, const uint32_t *a) { fe z; fe_invert(z, ld(a)); st(r, z); }
void h_fe_pow22523(uint32_t *r, const uint32_t *a) { fe z; fe_pow22523(z, ld(a)); st(r, z); }
void h_fe_canonical(uint32_t *w, const uint32_t *a) { fe_canonical_words(w, ld(a)); }
void h_fe_store(uint32_t *w, const uint32_t *a) { fe_store_words(w, ld(a)); }
void h_fe_from_words(uint32_t *r, const uint32_t *w) { fe z; fe_from_words(z, w); st(r, z); }
void h_fe_sub(uint32_t *r, const uint32_t *a, const uint32_t *b) { fe z; fe_sub(z, ld(a), ld(b)); st(r, z); }
void h_constants(uint32_t *d, uint32_t *d2, uint32_t *s) { st(d, FE_D); st(d2, FE_D2); st(s, FE_SQRTM1); }
void h_inv2(uint32_t *r) { st(r, FE_INV2); }
void h_inv_d(uint32_t *r) { st(r, FE_INV_D); }
void h_inv_d2(uint32_t *r) { st(r, FE_INV_D2); }
void h_scalar_mod_l(uint32_t *m) { scalar_mod_l(m); }
void h_scalar_cneg_l(uint32_t *m, uint32_t neg) { scalar_cneg_l(m, neg); }
void h_recode32_51(uint32_t *e, const uint32_t *m) { recode32<51>(e, m); }
void h_recode32_13(uint32_t *e, const uint32_t *m) { recode32<13>(e, m); }
// (+-b) + a where +-b starts as the extended point of b's table entry form
// (negated when neg; b = NULL: the identity entry) and a is mixed-added
void h_precomp_start_encoded(uint32_t *out, const uint32_t *a, const uint32_t *b, int neg) {
  ge_p3 p, q, h, r;
  ge_precomp ea, eb;
  ge_p1p1 t;
  ge_decode(p, a);
  ge_p3_to_precomp(ea, p);
  if (b) {
    ge_decode(q, b);
    ge_p3_to_precomp(eb, q);
  } else {
    ge_precomp_0(eb);
  }
  ge_precomp_cneg(eb, (uint32_t)neg);
  ge_precomp_to_p3(h, eb);
  ge_madd(t, h, ea);
  ge_p1p1_to_p3(r, t);
  ge_p3_encode(out, r);
}
// a + q (q negated when neg) by the mixed add over q's affine precomputed form
// (the comb tables' entry); q = NULL adds the identity entry
void h_madd_encoded(uint32_t *out, const uint32_t *a, const uint32_t *b, int neg) {
  ge_p3 p, q, r;
  ge_precomp e;
  ge_p1p1 t;
  ge_decode(p, a);
  if (b) {
    ge_decode(q, b);
    ge_p3_to_precomp(e, q);
  } else {
    ge_precomp_0(e);
  }
  ge_precomp_cneg(e, (uint32_t)neg);
  ge_madd(t, p, e);
  ge_p1p1_to_p3(r, t);
  ge_p3_encode(out, r);
}
int h_decode_encode(uint32_t *out, const uint32_t *in) {
  ge_p3 p;
  uint32_t ok = ge_decode(p, in);
  ge_p3_encode(out, p);
  return (int)ok;
}
void h_add_encoded(uint32_t *out, const uint32_t *a, const uint32_t *b) {
  ge_p3 p, q;
  ge_decode(p, a);
  ge_decode(q, b);
  ge_p3_add(p, p, q);
  ge_p3_encode(out, p);
}
void h_sub_encoded(uint32_t *out, const uint32_t *a, const uint32_t *b) {
  ge_p3 p, q;
  ge_decode(p, a);
  ge_decode(q, b);
  ge_p3_sub(p, p, q);
  ge_p3_encode(out, p);
}
void h_dbl_encoded(uint32_t *out, const uint32_t *a) {
  ge_p3 p;
  ge_p1p1 t;
  ge_decode(p, a);
  ge_p3_dbl(t, p);
  ge_p1p1_to_p3(p, t);
  ge_p3_encode(out, p);
}
// l-style public multiple: digits (lowest first, top digit positive, in
// {0, +-1, +-3}) applied to the decoded point with ge_dbl_add_public
void h_mul_public_encoded(uint32_t *out, const uint32_t *a, const int8_t *digits, int len) {
  ge_p3 p, q;
  ge_p1p1 r;
  ge_cached c1, c3;
  ge_decode(p, a);
  ge_p3_to_cached(c1, p);
  ge_p3_dbl(r, p);
  ge_p1p1_to_p3(q, r);
  ge_add(r, q, c1);
  ge_p1p1_to_p3(q, r);
  ge_p3_to_cached(c3, q);
  ge_p2 s;
  if (digits[len - 1] == 3) { s.X = q.X; s.Y = q.Y; s.Z = q.Z; }
  else { s.X = p.X; s.Y = p.Y; s.Z = p.Z; }
  for (int b = len - 2; b >= 0; --b) ge_dbl_add_public(s, digits[b], c1, c3);
  ge_p3 h;
  h.X = s.X; h.Y = s.Y; h.Z = s.Z;
  fe_0(h.T);                      // the encoding reads X, Y, Z only
  ge_p3_encode(out, h);
}
void h_neg_encoded(uint32_t *out, const uint32_t *a) {
  ge_p3 p;
  ge_decode(p, a);
  ge_p3_neg(p, p);
  ge_p3_encode(out, p);
}
}
