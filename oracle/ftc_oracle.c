/*
 * ftc_oracle.c -- CPU oracle for the batched Ed25519 ElGamal path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the algorithm of
 * the reference's CPU path (ftcrypto.c over libsodium's ref10 group code) so
 * that the HIP library can be checked byte-for-byte against it.  It is never
 * linked into, loaded by, or called from the product library; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Representation: GF(2^255-19) in 5 unsigned 51-bit limbs with unsigned
 * __int128 products (the limb layout libsodium uses on 64-bit hosts,
 * selected by HAVE_TI_MODE at ftcrypto.c:12).  This is deliberately a
 * different representation from the GPU library's 9x29-bit limbs so that a
 * limb-level mistake cannot hide in both.
 *
 * Reference semantics followed (file:line in the reference tree, src/):
 *   field decode ignores bit 255, accepts non-canonical y .. sodium_fe.cu:70-141
 *   canonical encode ..................................... sodium_fe.cu:168-289
 *   invert / pow22523 exponents .......................... sodium_fe.cu:1352-1463
 *   point decode (sqrt, sign bit, failure code) ........... sodium_ge.cu:137-180
 *   point encode (y canonical, bit 255 = parity of x) ..... sodium_ge.cu:382-394
 *   unified extended add (p3 + cached) / sub .............. sodium_ge.cu:72-92, 517-533
 *   madd with affine precomp .............................. sodium_ge.cu:231-246
 *   doubling .............................................. sodium_ge.cu:306-320
 *   variable-base scalar mult, signed 4-bit windows,
 *     digit |e| outside 1..8 selects the identity ......... sodium_ge.cu:650-731, 458-511
 *   fixed-base comb, table[i][j] = (j+1)*256^i*B ........... sodium_ge.cu:743-792, 480
 *   identity encoding 01 00..00, base point y = 4/5 ....... ftcrypto.cu:95-125, ftcrypto.c:99-104
 *   array add / sub loops ................................. ftcrypto.c:143-160, 363-412
 *   array mul / scale ..................................... ftcrypto.c:340-360
 *   sparse gather-sum (empty row -> identity) ............. ftcrypto.c:435-474
 *   reduce_isum (duplicate keys accumulate) ............... ftcrypto.cu:1150-1217
 *   reduce / inclusive scan / exclusive scan definitions .. ftcrypto.h:424-435
 *   small-scalar packing (uint64 -> 32 B little-endian) ... ftcrypto.cu:937-963
 *
 * The ElGamal composition Enc(m; r) = (r*B, m*B + r*H) has no reference
 * function; it is the composition FinTracer makes from from_scalars,
 * init_point, mul and add (README.md:4, ftcrypto.h:257-384).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef struct { uint64_t v[5]; } fe;
#define MASK51 ((1ULL << 51) - 1)

/* ------------------------------------------------------------------ field */

static void fe_0(fe *h) { memset(h, 0, sizeof *h); }
static void fe_1(fe *h) { fe_0(h); h->v[0] = 1; }

static void fe_carry(fe *h)
{
    uint64_t c;
    c = h->v[0] >> 51; h->v[0] &= MASK51; h->v[1] += c;
    c = h->v[1] >> 51; h->v[1] &= MASK51; h->v[2] += c;
    c = h->v[2] >> 51; h->v[2] &= MASK51; h->v[3] += c;
    c = h->v[3] >> 51; h->v[3] &= MASK51; h->v[4] += c;
    c = h->v[4] >> 51; h->v[4] &= MASK51; h->v[0] += 19 * c;
}

static void fe_add(fe *h, const fe *f, const fe *g)
{
    for (int i = 0; i < 5; ++i) h->v[i] = f->v[i] + g->v[i];
    fe_carry(h);
}

/* f - g + 4p, limbs of 4p are 2^53-76 and 2^53-4; inputs are carried. */
static void fe_sub(fe *h, const fe *f, const fe *g)
{
    h->v[0] = f->v[0] + ((1ULL << 53) - 76) - g->v[0];
    for (int i = 1; i < 5; ++i) h->v[i] = f->v[i] + ((1ULL << 53) - 4) - g->v[i];
    fe_carry(h);
}

static void fe_neg(fe *h, const fe *f) { fe z; fe_0(&z); fe_sub(h, &z, f); }

static void fe_mul(fe *h, const fe *f, const fe *g)
{
    const uint64_t *a = f->v, *b = g->v;
    uint64_t b1 = 19 * b[1], b2 = 19 * b[2], b3 = 19 * b[3], b4 = 19 * b[4];
    u128 r0 = (u128)a[0] * b[0] + (u128)a[1] * b4 + (u128)a[2] * b3 + (u128)a[3] * b2 + (u128)a[4] * b1;
    u128 r1 = (u128)a[0] * b[1] + (u128)a[1] * b[0] + (u128)a[2] * b4 + (u128)a[3] * b3 + (u128)a[4] * b2;
    u128 r2 = (u128)a[0] * b[2] + (u128)a[1] * b[1] + (u128)a[2] * b[0] + (u128)a[3] * b4 + (u128)a[4] * b3;
    u128 r3 = (u128)a[0] * b[3] + (u128)a[1] * b[2] + (u128)a[2] * b[1] + (u128)a[3] * b[0] + (u128)a[4] * b4;
    u128 r4 = (u128)a[0] * b[4] + (u128)a[1] * b[3] + (u128)a[2] * b[2] + (u128)a[3] * b[1] + (u128)a[4] * b[0];
    r1 += (uint64_t)(r0 >> 51); uint64_t h0 = (uint64_t)r0 & MASK51;
    r2 += (uint64_t)(r1 >> 51); uint64_t h1 = (uint64_t)r1 & MASK51;
    r3 += (uint64_t)(r2 >> 51); uint64_t h2 = (uint64_t)r2 & MASK51;
    r4 += (uint64_t)(r3 >> 51); uint64_t h3 = (uint64_t)r3 & MASK51;
    uint64_t h4 = (uint64_t)r4 & MASK51;
    u128 t0 = (u128)h0 + (u128)(uint64_t)(r4 >> 51) * 19;
    h->v[0] = (uint64_t)t0 & MASK51;
    h->v[1] = h1 + (uint64_t)(t0 >> 51);
    h->v[2] = h2; h->v[3] = h3; h->v[4] = h4;
}

static void fe_sq(fe *h, const fe *f) { fe_mul(h, f, f); }

static void fe_sqn(fe *h, const fe *f, int n)
{
    fe_sq(h, f);
    for (int i = 1; i < n; ++i) fe_sq(h, h);
}

/* z^(2^250 - 1), shared by invert and pow22523; also returns z^11. */
static void fe_pow250(fe *z250, fe *z11, const fe *z)
{
    fe z2, z9, t, z5, z10, z20, z50, z100;
    fe_sq(&z2, z);                 /* 2 */
    fe_sqn(&t, &z2, 2);            /* 8 */
    fe_mul(&z9, &t, z);            /* 9 */
    fe_mul(z11, &z9, &z2);         /* 11 */
    fe_sq(&t, z11);                /* 22 */
    fe_mul(&z5, &t, &z9);          /* 2^5 - 1 */
    fe_sqn(&t, &z5, 5);
    fe_mul(&z10, &t, &z5);         /* 2^10 - 1 */
    fe_sqn(&t, &z10, 10);
    fe_mul(&z20, &t, &z10);        /* 2^20 - 1 */
    fe_sqn(&t, &z20, 20);
    fe_mul(&t, &t, &z20);          /* 2^40 - 1 */
    fe_sqn(&t, &t, 10);
    fe_mul(&z50, &t, &z10);        /* 2^50 - 1 */
    fe_sqn(&t, &z50, 50);
    fe_mul(&z100, &t, &z50);       /* 2^100 - 1 */
    fe_sqn(&t, &z100, 100);
    fe_mul(&t, &t, &z100);         /* 2^200 - 1 */
    fe_sqn(&t, &t, 50);
    fe_mul(z250, &t, &z50);        /* 2^250 - 1 */
}

/* z^(p-2) = z^(2^255 - 21) */
static void fe_invert(fe *out, const fe *z)
{
    fe z250, z11, t;
    fe_pow250(&z250, &z11, z);
    fe_sqn(&t, &z250, 5);          /* 2^255 - 32 */
    fe_mul(out, &t, &z11);         /* 2^255 - 21 */
}

/* z^((p-5)/8) = z^(2^252 - 3) */
static void fe_pow22523(fe *out, const fe *z)
{
    fe z250, z11, t;
    fe_pow250(&z250, &z11, z);
    fe_sqn(&t, &z250, 2);          /* 2^252 - 4 */
    fe_mul(out, &t, z);            /* 2^252 - 3 */
}

/* canonical little-endian encoding of the residue in [0, p) */
static void fe_tobytes(uint8_t s[32], const fe *f)
{
    fe h = *f;
    fe_carry(&h); fe_carry(&h);
    uint64_t q = (h.v[0] + 19) >> 51;
    q = (h.v[1] + q) >> 51; q = (h.v[2] + q) >> 51;
    q = (h.v[3] + q) >> 51; q = (h.v[4] + q) >> 51;
    h.v[0] += 19 * q;
    uint64_t c;
    c = h.v[0] >> 51; h.v[0] &= MASK51; h.v[1] += c;
    c = h.v[1] >> 51; h.v[1] &= MASK51; h.v[2] += c;
    c = h.v[2] >> 51; h.v[2] &= MASK51; h.v[3] += c;
    c = h.v[3] >> 51; h.v[3] &= MASK51; h.v[4] += c;
    h.v[4] &= MASK51;
    uint64_t w[4];
    w[0] = h.v[0] | (h.v[1] << 51);
    w[1] = (h.v[1] >> 13) | (h.v[2] << 38);
    w[2] = (h.v[2] >> 26) | (h.v[3] << 25);
    w[3] = (h.v[3] >> 39) | (h.v[4] << 12);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) s[8 * i + j] = (uint8_t)(w[i] >> (8 * j));
}

/* bits 0..254 of s; bit 255 ignored, values >= p accepted */
static void fe_frombytes(fe *h, const uint8_t s[32])
{
    uint64_t w[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) w[i] |= (uint64_t)s[8 * i + j] << (8 * j);
    h->v[0] = w[0] & MASK51;
    h->v[1] = ((w[0] >> 51) | (w[1] << 13)) & MASK51;
    h->v[2] = ((w[1] >> 38) | (w[2] << 26)) & MASK51;
    h->v[3] = ((w[2] >> 25) | (w[3] << 39)) & MASK51;
    h->v[4] = (w[3] >> 12) & MASK51;
}

static int fe_isnegative(const fe *f) { uint8_t s[32]; fe_tobytes(s, f); return s[0] & 1; }

static int fe_iszero(const fe *f)
{
    uint8_t s[32]; fe_tobytes(s, f);
    uint8_t acc = 0;
    for (int i = 0; i < 32; ++i) acc |= s[i];
    return acc == 0;
}

static void fe_cmov(fe *f, const fe *g, unsigned b)
{
    uint64_t mask = (uint64_t)0 - (uint64_t)(b & 1);
    for (int i = 0; i < 5; ++i) f->v[i] ^= mask & (f->v[i] ^ g->v[i]);
}

static fe FE_D, FE_D2, FE_SQRTM1;

/* ------------------------------------------------------------------ group */

typedef struct { fe X, Y, Z, T; } ge_p3;
typedef struct { fe X, Y, Z; } ge_p2;
typedef struct { fe X, Y, Z, T; } ge_p1p1;        /* X=E, Y=H, Z=G, T=F */
typedef struct { fe YplusX, YminusX, Z, T2d; } ge_cached;
typedef struct { fe yplusx, yminusx, xy2d; } ge_precomp;

static void ge_p3_0(ge_p3 *h) { fe_0(&h->X); fe_1(&h->Y); fe_1(&h->Z); fe_0(&h->T); }
static void ge_cached_0(ge_cached *h) { fe_1(&h->YplusX); fe_1(&h->YminusX); fe_1(&h->Z); fe_0(&h->T2d); }
static void ge_precomp_0(ge_precomp *h) { fe_1(&h->yplusx); fe_1(&h->yminusx); fe_0(&h->xy2d); }

static void ge_p3_to_cached(ge_cached *r, const ge_p3 *p)
{
    fe_add(&r->YplusX, &p->Y, &p->X);
    fe_sub(&r->YminusX, &p->Y, &p->X);
    r->Z = p->Z;
    fe_mul(&r->T2d, &p->T, &FE_D2);
}

static void ge_p1p1_to_p3(ge_p3 *r, const ge_p1p1 *p)
{
    fe_mul(&r->X, &p->X, &p->T);
    fe_mul(&r->Y, &p->Y, &p->Z);
    fe_mul(&r->Z, &p->Z, &p->T);
    fe_mul(&r->T, &p->X, &p->Y);
}

static void ge_p1p1_to_p2(ge_p2 *r, const ge_p1p1 *p)
{
    fe_mul(&r->X, &p->X, &p->T);
    fe_mul(&r->Y, &p->Y, &p->Z);
    fe_mul(&r->Z, &p->Z, &p->T);
}

/* unified addition (Hisil-Wong-Carter-Dawson 2008, a = -1) */
static void ge_add(ge_p1p1 *r, const ge_p3 *p, const ge_cached *q)
{
    fe a, b, A, B, C, D;
    fe_sub(&a, &p->Y, &p->X);
    fe_add(&b, &p->Y, &p->X);
    fe_mul(&A, &a, &q->YminusX);
    fe_mul(&B, &b, &q->YplusX);
    fe_mul(&C, &q->T2d, &p->T);
    fe_mul(&D, &p->Z, &q->Z);
    fe_add(&D, &D, &D);
    fe_sub(&r->X, &B, &A);
    fe_add(&r->Y, &B, &A);
    fe_add(&r->Z, &D, &C);
    fe_sub(&r->T, &D, &C);
}

static void ge_sub(ge_p1p1 *r, const ge_p3 *p, const ge_cached *q)
{
    fe a, b, A, B, C, D;
    fe_sub(&a, &p->Y, &p->X);
    fe_add(&b, &p->Y, &p->X);
    fe_mul(&A, &a, &q->YplusX);
    fe_mul(&B, &b, &q->YminusX);
    fe_mul(&C, &q->T2d, &p->T);
    fe_mul(&D, &p->Z, &q->Z);
    fe_add(&D, &D, &D);
    fe_sub(&r->X, &B, &A);
    fe_add(&r->Y, &B, &A);
    fe_sub(&r->Z, &D, &C);
    fe_add(&r->T, &D, &C);
}

static void ge_madd(ge_p1p1 *r, const ge_p3 *p, const ge_precomp *q)
{
    fe a, b, A, B, C, D;
    fe_sub(&a, &p->Y, &p->X);
    fe_add(&b, &p->Y, &p->X);
    fe_mul(&A, &a, &q->yminusx);
    fe_mul(&B, &b, &q->yplusx);
    fe_mul(&C, &q->xy2d, &p->T);
    fe_add(&D, &p->Z, &p->Z);
    fe_sub(&r->X, &B, &A);
    fe_add(&r->Y, &B, &A);
    fe_add(&r->Z, &D, &C);
    fe_sub(&r->T, &D, &C);
}

/* doubling of a projective point (dbl-2008-hwcd, a = -1) */
static void ge_p2_dbl(ge_p1p1 *r, const ge_p2 *p)
{
    fe XX, YY, ZZ2, s, ss;
    fe_sq(&XX, &p->X);
    fe_sq(&YY, &p->Y);
    fe_sq(&ZZ2, &p->Z); fe_add(&ZZ2, &ZZ2, &ZZ2);
    fe_add(&s, &p->X, &p->Y);
    fe_sq(&ss, &s);
    fe_add(&r->Y, &YY, &XX);        /* H' = X^2 + Y^2 */
    fe_sub(&r->Z, &YY, &XX);        /* G  = Y^2 - X^2 */
    fe_sub(&r->X, &ss, &r->Y);      /* E  = 2XY */
    fe_sub(&r->T, &ZZ2, &r->Z);     /* F' = 2Z^2 - G */
}

static void ge_p3_dbl(ge_p1p1 *r, const ge_p3 *p)
{
    ge_p2 q; q.X = p->X; q.Y = p->Y; q.Z = p->Z;
    ge_p2_dbl(r, &q);
}

static void ge_p3_add(ge_p3 *r, const ge_p3 *p, const ge_p3 *q)
{
    ge_cached c; ge_p1p1 t;
    ge_p3_to_cached(&c, q);
    ge_add(&t, p, &c);
    ge_p1p1_to_p3(r, &t);
}

static void ge_p3_sub(ge_p3 *r, const ge_p3 *p, const ge_p3 *q)
{
    ge_cached c; ge_p1p1 t;
    ge_p3_to_cached(&c, q);
    ge_sub(&t, p, &c);
    ge_p1p1_to_p3(r, &t);
}

static void ge_p3_neg(ge_p3 *r, const ge_p3 *p)
{
    fe_neg(&r->X, &p->X);
    r->Y = p->Y; r->Z = p->Z;
    fe_neg(&r->T, &p->T);
}

static void ge_p3_tobytes(uint8_t s[32], const ge_p3 *h)
{
    fe recip, x, y;
    fe_invert(&recip, &h->Z);
    fe_mul(&x, &h->X, &recip);
    fe_mul(&y, &h->Y, &recip);
    fe_tobytes(s, &y);
    s[31] ^= (uint8_t)(fe_isnegative(&x) << 7);
}

/* returns 0 on success, -1 when (y^2-1)/(dy^2+1) has no square root */
static int ge_frombytes(ge_p3 *h, const uint8_t s[32])
{
    fe u, v, v3, vxx, chk, x_sqrtm1, negx;
    fe_frombytes(&h->Y, s);
    fe_1(&h->Z);
    fe_sq(&u, &h->Y);
    fe_mul(&v, &u, &FE_D);
    fe_sub(&u, &u, &h->Z);                 /* u = y^2 - 1 */
    fe_add(&v, &v, &h->Z);                 /* v = d y^2 + 1 */
    fe_sq(&v3, &v); fe_mul(&v3, &v3, &v);  /* v^3 */
    fe_sq(&h->X, &v3); fe_mul(&h->X, &h->X, &v); fe_mul(&h->X, &h->X, &u); /* u v^7 */
    fe_pow22523(&h->X, &h->X);
    fe_mul(&h->X, &h->X, &v3);
    fe_mul(&h->X, &h->X, &u);              /* u v^3 (u v^7)^((p-5)/8) */
    fe_sq(&vxx, &h->X); fe_mul(&vxx, &vxx, &v);
    fe_sub(&chk, &vxx, &u);
    int has_m_root = fe_iszero(&chk);
    fe_add(&chk, &vxx, &u);
    int has_p_root = fe_iszero(&chk);
    fe_mul(&x_sqrtm1, &h->X, &FE_SQRTM1);
    fe_cmov(&h->X, &x_sqrtm1, (unsigned)(1 - has_m_root));
    fe_neg(&negx, &h->X);
    fe_cmov(&h->X, &negx, (unsigned)(fe_isnegative(&h->X) ^ (s[31] >> 7)));
    fe_mul(&h->T, &h->X, &h->Y);
    return (has_m_root | has_p_root) - 1;
}

/* signed radix-16 digits: e[i] in [-8, 8) for i < 63, e[63] = top nibble + carry */
static void recode_nibbles(signed char e[64], const uint8_t a[32])
{
    for (int i = 0; i < 32; ++i) {
        e[2 * i] = (signed char)(a[i] & 15);
        e[2 * i + 1] = (signed char)((a[i] >> 4) & 15);
    }
    signed char carry = 0;
    for (int i = 0; i < 63; ++i) {
        e[i] = (signed char)(e[i] + carry);
        carry = (signed char)((e[i] + 8) >> 4);
        e[i] = (signed char)(e[i] - carry * 16);
    }
    e[63] = (signed char)(e[63] + carry);
}

static unsigned eq_u8(unsigned a, unsigned b) { return ((a ^ b) - 1U) >> 31; }

static void select_cached(ge_cached *t, const ge_cached tab[8], signed char b)
{
    unsigned neg = ((unsigned char)b) >> 7;
    unsigned babs = (unsigned)(b < 0 ? -b : b) & 0xff;
    ge_cached_0(t);
    for (unsigned j = 0; j < 8; ++j) {
        unsigned m = eq_u8(babs, j + 1);
        fe_cmov(&t->YplusX, &tab[j].YplusX, m);
        fe_cmov(&t->YminusX, &tab[j].YminusX, m);
        fe_cmov(&t->Z, &tab[j].Z, m);
        fe_cmov(&t->T2d, &tab[j].T2d, m);
    }
    ge_cached mt;
    mt.YplusX = t->YminusX; mt.YminusX = t->YplusX; mt.Z = t->Z;
    fe_neg(&mt.T2d, &t->T2d);
    fe_cmov(&t->YplusX, &mt.YplusX, neg);
    fe_cmov(&t->YminusX, &mt.YminusX, neg);
    fe_cmov(&t->T2d, &mt.T2d, neg);
}

/* h = a * p, a = 32-byte little-endian integer (constant-time in a) */
static void ge_scalarmult(ge_p3 *h, const uint8_t a[32], const ge_p3 *p)
{
    ge_cached tab[8];
    ge_p3 m[9];
    ge_p1p1 t;
    m[1] = *p;
    ge_p3_dbl(&t, p); ge_p1p1_to_p3(&m[2], &t);
    for (int k = 3; k <= 8; ++k) {
        if (k % 2 == 0) { ge_p3_dbl(&t, &m[k / 2]); ge_p1p1_to_p3(&m[k], &t); }
        else ge_p3_add(&m[k], &m[k - 1], p);
    }
    for (int k = 0; k < 8; ++k) ge_p3_to_cached(&tab[k], &m[k + 1]);

    signed char e[64];
    recode_nibbles(e, a);
    ge_p3_0(h);
    ge_cached c;
    ge_p2 s;
    for (int i = 63; i > 0; --i) {
        select_cached(&c, tab, e[i]);
        ge_add(&t, h, &c);
        ge_p1p1_to_p2(&s, &t);
        ge_p2_dbl(&t, &s); ge_p1p1_to_p2(&s, &t);
        ge_p2_dbl(&t, &s); ge_p1p1_to_p2(&s, &t);
        ge_p2_dbl(&t, &s); ge_p1p1_to_p2(&s, &t);
        ge_p2_dbl(&t, &s);
        ge_p1p1_to_p3(h, &t);
    }
    select_cached(&c, tab, e[0]);
    ge_add(&t, h, &c);
    ge_p1p1_to_p3(h, &t);
}

/* comb table for a fixed point P: tab[i][j] = (j+1) * 256^i * P, affine */
typedef struct { ge_precomp t[32][8]; } comb_table;

static void ge_p3_to_precomp(ge_precomp *r, const ge_p3 *p)
{
    fe recip, x, y, xy;
    fe_invert(&recip, &p->Z);
    fe_mul(&x, &p->X, &recip);
    fe_mul(&y, &p->Y, &recip);
    fe_add(&r->yplusx, &y, &x);
    fe_sub(&r->yminusx, &y, &x);
    fe_mul(&xy, &x, &y);
    fe_mul(&r->xy2d, &xy, &FE_D2);
}

static void comb_table_build(comb_table *ct, const ge_p3 *P)
{
    ge_p3 row = *P;                           /* 256^i * P */
    for (int i = 0; i < 32; ++i) {
        ge_p3 acc = row;
        for (int j = 0; j < 8; ++j) {
            ge_p3_to_precomp(&ct->t[i][j], &acc);
            ge_p3_add(&acc, &acc, &row);
        }
        ge_p1p1 t;
        for (int k = 0; k < 8; ++k) { ge_p3_dbl(&t, &row); ge_p1p1_to_p3(&row, &t); }
    }
}

static void select_precomp(ge_precomp *t, const ge_precomp row[8], signed char b)
{
    unsigned neg = ((unsigned char)b) >> 7;
    unsigned babs = (unsigned)(b < 0 ? -b : b) & 0xff;
    ge_precomp_0(t);
    for (unsigned j = 0; j < 8; ++j) {
        unsigned m = eq_u8(babs, j + 1);
        fe_cmov(&t->yplusx, &row[j].yplusx, m);
        fe_cmov(&t->yminusx, &row[j].yminusx, m);
        fe_cmov(&t->xy2d, &row[j].xy2d, m);
    }
    ge_precomp mt;
    mt.yplusx = t->yminusx; mt.yminusx = t->yplusx;
    fe_neg(&mt.xy2d, &t->xy2d);
    fe_cmov(&t->yplusx, &mt.yplusx, neg);
    fe_cmov(&t->yminusx, &mt.yminusx, neg);
    fe_cmov(&t->xy2d, &mt.xy2d, neg);
}

/* h = a * P using P's comb table (odd digits, x16, even digits) */
static void ge_scalarmult_comb(ge_p3 *h, const uint8_t a[32], const comb_table *ct)
{
    signed char e[64];
    recode_nibbles(e, a);
    ge_p3_0(h);
    ge_precomp t;
    ge_p1p1 r;
    ge_p2 s;
    for (int i = 1; i < 64; i += 2) {
        select_precomp(&t, ct->t[i / 2], e[i]);
        ge_madd(&r, h, &t); ge_p1p1_to_p3(h, &r);
    }
    ge_p3_dbl(&r, h); ge_p1p1_to_p2(&s, &r);
    ge_p2_dbl(&r, &s); ge_p1p1_to_p2(&s, &r);
    ge_p2_dbl(&r, &s); ge_p1p1_to_p2(&s, &r);
    ge_p2_dbl(&r, &s); ge_p1p1_to_p3(h, &r);
    for (int i = 0; i < 64; i += 2) {
        select_precomp(&t, ct->t[i / 2], e[i]);
        ge_madd(&r, h, &t); ge_p1p1_to_p3(h, &r);
    }
}

/* ------------------------------------------------------------ constants */

static ge_p3 GE_B;
static comb_table *B_TABLE;
static int INITIALISED;

static void fe_from_u64(fe *h, uint64_t x) { fe_0(h); h->v[0] = x & MASK51; h->v[1] = x >> 51; }

int ftco_init(void)
{
    if (INITIALISED) return 0;
    /* d = -121665/121666 */
    fe n, dd, inv;
    fe_from_u64(&n, 121665); fe_neg(&n, &n);
    fe_from_u64(&dd, 121666); fe_invert(&inv, &dd);
    fe_mul(&FE_D, &n, &inv);
    fe_add(&FE_D2, &FE_D, &FE_D);
    /* sqrt(-1) = 2^((p-1)/4): (p-1)/4 = 2^253 - 5 = 4*(2^251 - 2) + 3; computed by square-and-multiply */
    {
        fe two, r; fe_from_u64(&two, 2); fe_1(&r);
        /* exponent bits of 2^253 - 5: binary 1...1011 (253 bits) */
        uint8_t ebits[32];
        memset(ebits, 0xff, 32);
        ebits[0] = 0xfb;                   /* ...11111011 = 2^256 - 5 */
        ebits[31] = 0x1f;                  /* clear bits 253..255 -> 2^253 - 5 */
        for (int i = 255; i >= 0; --i) {
            fe_sq(&r, &r);
            if ((ebits[i / 8] >> (i % 8)) & 1) fe_mul(&r, &r, &two);
        }
        FE_SQRTM1 = r;
    }
    /* B = (x, 4/5) with x even */
    {
        fe four, five, y; uint8_t s[32];
        fe_from_u64(&four, 4); fe_from_u64(&five, 5);
        fe_invert(&inv, &five); fe_mul(&y, &four, &inv);
        fe_tobytes(s, &y);
        if (ge_frombytes(&GE_B, s) != 0) return -1;
    }
    B_TABLE = (comb_table *)malloc(sizeof(comb_table));
    if (!B_TABLE) return -1;
    comb_table_build(B_TABLE, &GE_B);
    INITIALISED = 1;
    return 0;
}

/* ------------------------------------------------------- field test hooks */

void ftco_fe_mul_bytes(uint8_t *out, const uint8_t *a, const uint8_t *b)
{ fe x, y, z; fe_frombytes(&x, a); fe_frombytes(&y, b); fe_mul(&z, &x, &y); fe_tobytes(out, &z); }
void ftco_fe_add_bytes(uint8_t *out, const uint8_t *a, const uint8_t *b)
{ fe x, y, z; fe_frombytes(&x, a); fe_frombytes(&y, b); fe_add(&z, &x, &y); fe_tobytes(out, &z); }
void ftco_fe_sub_bytes(uint8_t *out, const uint8_t *a, const uint8_t *b)
{ fe x, y, z; fe_frombytes(&x, a); fe_frombytes(&y, b); fe_sub(&z, &x, &y); fe_tobytes(out, &z); }
void ftco_fe_invert_bytes(uint8_t *out, const uint8_t *a)
{ fe x, z; fe_frombytes(&x, a); fe_invert(&z, &x); fe_tobytes(out, &z); }
void ftco_fe_pow22523_bytes(uint8_t *out, const uint8_t *a)
{ fe x, z; fe_frombytes(&x, a); fe_pow22523(&z, &x); fe_tobytes(out, &z); }
void ftco_fe_tobytes_roundtrip(uint8_t *out, const uint8_t *a)
{ fe x; fe_frombytes(&x, a); fe_tobytes(out, &x); }
void ftco_constants(uint8_t *d, uint8_t *d2, uint8_t *sqrtm1, uint8_t *base)
{ fe_tobytes(d, &FE_D); fe_tobytes(d2, &FE_D2); fe_tobytes(sqrtm1, &FE_SQRTM1); ge_p3_tobytes(base, &GE_B); }

/* -------------------------------------------------------------- arrays */

typedef struct { ge_p3 *p; size_t n; } ftco_array;

static int nthreads_or_default(int nt) { return nt > 0 ? nt : 1; }

ftco_array *ftco_array_new(size_t n)
{
    ftco_array *a = (ftco_array *)malloc(sizeof *a);
    if (!a) return NULL;
    a->p = (ge_p3 *)malloc((n ? n : 1) * sizeof(ge_p3));
    if (!a->p) { free(a); return NULL; }
    a->n = n;
    for (size_t i = 0; i < n; ++i) ge_p3_0(&a->p[i]);
    return a;
}

void ftco_array_free(ftco_array *a) { if (a) { free(a->p); free(a); } }
size_t ftco_array_len(const ftco_array *a) { return a->n; }

ftco_array *ftco_array_copy(const ftco_array *src)
{
    ftco_array *a = ftco_array_new(src->n);
    if (a) memcpy(a->p, src->p, src->n * sizeof(ge_p3));
    return a;
}

/* decode n folded encodings; returns the largest failing index or -1 */
int64_t ftco_array_from_bytes(ftco_array *a, const uint8_t *mem, int nt)
{
    int64_t bad = -1;
    #pragma omp parallel for num_threads(nthreads_or_default(nt)) reduction(max:bad)
    for (int64_t i = 0; i < (int64_t)a->n; ++i)
        if (ge_frombytes(&a->p[i], mem + 32 * i) != 0 && i > bad) bad = i;
    return bad;
}

void ftco_array_to_bytes(uint8_t *mem, const ftco_array *a, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)a->n; ++i) ge_p3_tobytes(mem + 32 * i, &a->p[i]);
}

/* a[i] = scalars[i] * B (comb) */
void ftco_from_scalars(ftco_array *a, const uint8_t *scalars, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)a->n; ++i) ge_scalarmult_comb(&a->p[i], scalars + 32 * i, B_TABLE);
}

/* a[i] = scalars[i] * B through the variable-base routine (as ftcrypto.cu:895-913 does) */
void ftco_from_scalars_varbase(ftco_array *a, const uint8_t *scalars, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)a->n; ++i) ge_scalarmult(&a->p[i], scalars + 32 * i, &GE_B);
}

static void u64_to_scalar(uint8_t s[32], uint64_t x)
{
    memset(s, 0, 32);
    for (int j = 0; j < 8; ++j) s[j] = (uint8_t)(x >> (8 * j));
}

void ftco_from_small_scalars(ftco_array *a, const uint64_t *small, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)a->n; ++i) {
        uint8_t s[32]; u64_to_scalar(s, small[i]);
        ge_scalarmult_comb(&a->p[i], s, B_TABLE);
    }
}

void ftco_add(ftco_array *res, const ftco_array *summand, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)res->n; ++i) ge_p3_add(&res->p[i], &res->p[i], &summand->p[i]);
}

void ftco_sub(ftco_array *res, const ftco_array *sub, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)res->n; ++i) ge_p3_sub(&res->p[i], &res->p[i], &sub->p[i]);
}

void ftco_neg(ftco_array *a, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)a->n; ++i) ge_p3_neg(&a->p[i], &a->p[i]);
}

/* a[i] = scalars[i] * a[i] (variable base, constant time) */
void ftco_mul(ftco_array *a, const uint8_t *scalars, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)a->n; ++i) ge_scalarmult(&a->p[i], scalars + 32 * i, &a->p[i]);
}

void ftco_scale(ftco_array *a, const uint8_t *scalar, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)a->n; ++i) ge_scalarmult(&a->p[i], scalar, &a->p[i]);
}

/* a[i] = k[i] * a[i] for 32-bit unsigned multipliers */
void ftco_mul_u32(ftco_array *a, const uint32_t *k, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)a->n; ++i) {
        uint8_t s[32]; u64_to_scalar(s, k[i]);
        ge_scalarmult(&a->p[i], s, &a->p[i]);
    }
}

/* sum of all elements, written as one encoded point */
void ftco_reduce(uint8_t out[32], const ftco_array *a, int nt)
{
    int T = nthreads_or_default(nt);
    ge_p3 *part = (ge_p3 *)malloc((size_t)T * sizeof(ge_p3));
    for (int t = 0; t < T; ++t) ge_p3_0(&part[t]);
    #pragma omp parallel num_threads(T)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        size_t lo = a->n * (size_t)t / (size_t)T, hi = a->n * (size_t)(t + 1) / (size_t)T;
        for (size_t i = lo; i < hi; ++i) ge_p3_add(&part[t], &part[t], &a->p[i]);
    }
    ge_p3 s; ge_p3_0(&s);
    for (int t = 0; t < T; ++t) ge_p3_add(&s, &s, &part[t]);
    ge_p3_tobytes(out, &s);
    free(part);
}

/* in-place inclusive (exclusive = 0) or exclusive (exclusive = 1) scan */
void ftco_scan(ftco_array *a, int exclusive, int nt)
{
    int T = nthreads_or_default(nt);
    size_t n = a->n;
    ge_p3 *tot = (ge_p3 *)malloc((size_t)(T + 1) * sizeof(ge_p3));
    #pragma omp parallel num_threads(T)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        size_t lo = n * (size_t)t / (size_t)T, hi = n * (size_t)(t + 1) / (size_t)T;
        ge_p3 s; ge_p3_0(&s);
        for (size_t i = lo; i < hi; ++i) ge_p3_add(&s, &s, &a->p[i]);
        tot[t + 1] = s;
        #pragma omp barrier
        #pragma omp single
        {
            ge_p3_0(&tot[0]);
            for (int k = 1; k <= T; ++k) ge_p3_add(&tot[k], &tot[k], &tot[k - 1]);
        }
        ge_p3 run = tot[t];
        for (size_t i = lo; i < hi; ++i) {
            ge_p3 x = a->p[i];
            ge_p3_add(&run, &run, &x);
            if (exclusive) { ge_p3_sub(&a->p[i], &run, &x); }
            else a->p[i] = run;
        }
    }
    free(tot);
}

/* out[s] = sum_{i in [off[s], off[s+1])} src[idx ? idx[i] : i]; empty -> identity.
 * Returns 0, or -1 when an offset or index is out of bounds. */
int ftco_segmented_sum(ftco_array *out, const int64_t *off, size_t nseg,
                       const int32_t *idx, const ftco_array *src, int nt)
{
    int err = 0;
    for (size_t s = 0; s < nseg; ++s)
        if (off[s] < 0 || off[s + 1] < off[s]) return -1;
    #pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads_or_default(nt)) reduction(|:err)
    for (int64_t s = 0; s < (int64_t)nseg; ++s) {
        ge_p3 acc; ge_p3_0(&acc);
        for (int64_t i = off[s]; i < off[s + 1]; ++i) {
            int64_t j = idx ? (int64_t)idx[i] : i;
            if (j < 0 || (size_t)j >= src->n) { err |= 1; break; }
            if (i == off[s]) acc = src->p[j];
            else ge_p3_add(&acc, &acc, &src->p[j]);
        }
        out->p[s] = acc;
    }
    return err ? -1 : 0;
}

/* res[idxs[i]] += src[i]; returns -1 on an out-of-range key */
int ftco_reduce_isum(ftco_array *res, const uint64_t *idxs, size_t n, const ftco_array *src)
{
    for (size_t i = 0; i < n; ++i)
        if (idxs[i] >= res->n) return -1;
    for (size_t i = 0; i < n; ++i) ge_p3_add(&res->p[idxs[i]], &res->p[idxs[i]], &src->p[i]);
    return 0;
}

/* ElGamal: c1[i] = r[i]*B, c2[i] = m[i]*B + r[i]*H, H given as an encoding */
int ftco_encrypt(ftco_array *c1, ftco_array *c2, const uint8_t H_enc[32],
                 const uint64_t *m, const uint8_t *r, int nt)
{
    ge_p3 H;
    if (ge_frombytes(&H, H_enc) != 0) return -1;
    comb_table *HT = (comb_table *)malloc(sizeof(comb_table));
    if (!HT) return -1;
    comb_table_build(HT, &H);
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)c1->n; ++i) {
        uint8_t ms[32]; u64_to_scalar(ms, m[i]);
        ge_p3 mB, rH;
        ge_scalarmult_comb(&c1->p[i], r + 32 * i, B_TABLE);
        ge_scalarmult_comb(&mB, ms, B_TABLE);
        ge_scalarmult_comb(&rH, r + 32 * i, HT);
        ge_p3_add(&c2->p[i], &mB, &rH);
    }
    free(HT);
    return 0;
}

/* c1[i] += r[i]*B, c2[i] += r[i]*H */
int ftco_rerandomize(ftco_array *c1, ftco_array *c2, const uint8_t H_enc[32],
                     const uint8_t *r, int nt)
{
    ge_p3 H;
    if (ge_frombytes(&H, H_enc) != 0) return -1;
    comb_table *HT = (comb_table *)malloc(sizeof(comb_table));
    if (!HT) return -1;
    comb_table_build(HT, &H);
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)c1->n; ++i) {
        ge_p3 rB, rH;
        ge_scalarmult_comb(&rB, r + 32 * i, B_TABLE);
        ge_scalarmult_comb(&rH, r + 32 * i, HT);
        ge_p3_add(&c1->p[i], &c1->p[i], &rB);
        ge_p3_add(&c2->p[i], &c2->p[i], &rH);
    }
    free(HT);
    return 0;
}

/* single-point helpers on encodings (for known-answer tests) */
int ftco_scalarmult_base_bytes(uint8_t out[32], const uint8_t k[32])
{ ge_p3 h; ge_scalarmult_comb(&h, k, B_TABLE); ge_p3_tobytes(out, &h); return 0; }

int ftco_scalarmult_bytes(uint8_t out[32], const uint8_t k[32], const uint8_t P[32])
{ ge_p3 p, h; if (ge_frombytes(&p, P)) return -1; ge_scalarmult(&h, k, &p); ge_p3_tobytes(out, &h); return 0; }

int ftco_add_bytes(uint8_t out[32], const uint8_t P[32], const uint8_t Q[32])
{ ge_p3 p, q; if (ge_frombytes(&p, P) || ge_frombytes(&q, Q)) return -1; ge_p3_add(&p, &p, &q); ge_p3_tobytes(out, &p); return 0; }

int ftco_decode_roundtrip(uint8_t out[32], const uint8_t P[32])
{ ge_p3 p; int r = ge_frombytes(&p, P); if (r) return r; ge_p3_tobytes(out, &p); return 0; }

/* ------------------------------------------------ scalar reduction mod l */

/* l = 2^252 + 27742317777372353535851937790883648493, little-endian 32-bit words */
static const uint32_t L_WORDS[8] = {
    0x5cf5d3ed, 0x5812631a, 0xa2f79cd6, 0x14def9de, 0, 0, 0, 0x10000000 };

/* x (16 words) mod l by shift-and-subtract; slow but obviously correct */
static void mod_l_512(uint32_t out[8], const uint32_t in[16])
{
    uint32_t r[9] = {0};
    for (int bit = 511; bit >= 0; --bit) {
        /* r = 2r + bit */
        uint32_t c = (in[bit / 32] >> (bit % 32)) & 1;
        for (int k = 0; k < 9; ++k) { uint32_t nc = r[k] >> 31; r[k] = (r[k] << 1) | c; c = nc; }
        /* if r >= l: r -= l */
        int ge = 1;
        for (int k = 8; k >= 0; --k) {
            uint32_t lk = k < 8 ? L_WORDS[k] : 0;
            if (r[k] != lk) { ge = r[k] > lk; break; }
        }
        if (ge) {
            uint64_t br = 0;
            for (int k = 0; k < 9; ++k) {
                uint64_t lk = k < 8 ? L_WORDS[k] : 0;
                uint64_t d = (uint64_t)r[k] - lk - br;
                r[k] = (uint32_t)d; br = (d >> 63) & 1;
            }
        }
    }
    memcpy(out, r, 32);
}

void ftco_scalar_reduce64(uint8_t *out, const uint8_t *in, size_t n, int nt)
{
    #pragma omp parallel for num_threads(nthreads_or_default(nt))
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        uint32_t w[16], o[8];
        for (int k = 0; k < 16; ++k)
            w[k] = (uint32_t)in[64 * i + 4 * k] | ((uint32_t)in[64 * i + 4 * k + 1] << 8) |
                   ((uint32_t)in[64 * i + 4 * k + 2] << 16) | ((uint32_t)in[64 * i + 4 * k + 3] << 24);
        mod_l_512(o, w);
        for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 4; ++j) out[32 * i + 4 * k + j] = (uint8_t)(o[k] >> (8 * j));
    }
}
