"""CPU checks of the library's 9x29-bit field arithmetic (fe25519.h) and group
formulas (ge25519.h), compiled for the host from the same headers the gfx950
kernels use.

The field multiply is exact only when its operands stay inside the bound
classes documented in fe25519.h (N = 2^29 + 2^17, A1 = 2N, A2 = 3N).  The
worst case of every uint64 column is reached with every limb at its class
maximum, so each allowed operand pair is checked there (and at random points
inside the class) against Python integers mod p; a deliberately disallowed
pair (A2 x A2) is shown to overflow, which pins that the bound check has
teeth.
"""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

import edref

P = edref.P
HERE = os.path.dirname(os.path.abspath(__file__))
N = 2**29 + 2**17
CLS = {"N": N, "A1": 2 * N, "A2": 3 * N}


@pytest.fixture(scope="module")
def H(tmp_path_factory):
    out = tmp_path_factory.mktemp("fehost") / "fe_host.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(out),
                    os.path.join(HERE, "native", "fe_host.cc")], check=True)
    L = ctypes.CDLL(str(out))
    return L


def arr(v):
    return (ctypes.c_uint32 * len(v))(*v)


def val(l):
    return sum(int(x) << (29 * i) for i, x in enumerate(l))


def words_val(w):
    return sum(int(x) << (32 * i) for i, x in enumerate(w))


def limbs_of(x):
    return [(x >> (29 * i)) & (2**29 - 1) for i in range(9)]


def to_words(x):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)]


def mul(H, a, b):
    r = (ctypes.c_uint32 * 9)()
    H.h_fe_mul(r, arr(a), arr(b))
    return list(r)


@pytest.mark.parametrize("ca,cb", [("N", "N"), ("N", "A1"), ("N", "A2"), ("A1", "A1"), ("A1", "A2")])
def test_mul_exact_at_class_bounds(H, ca, cb):
    rng = random.Random(hash((ca, cb)) & 0xFFFF)
    cases = [([CLS[ca]] * 9, [CLS[cb]] * 9)]
    for _ in range(400):
        cases.append(([rng.randrange(CLS[ca] + 1) for _ in range(9)], [rng.randrange(CLS[cb] + 1) for _ in range(9)]))
    for a, b in cases:
        r = mul(H, a, b)
        assert val(r) % P == val(a) * val(b) % P
        assert max(r) <= N


def test_disallowed_pair_overflows(H):
    a = [CLS["A2"]] * 9
    r = mul(H, a, a)
    assert val(r) % P != val(a) * val(a) % P


def test_square_and_carry(H):
    rng = random.Random(3)
    for cls in ("N", "A1"):
        for t in range(300):
            a = [CLS[cls]] * 9 if t == 0 else [rng.randrange(CLS[cls] + 1) for _ in range(9)]
            r = (ctypes.c_uint32 * 9)()
            H.h_fe_sq(r, arr(a))
            assert val(r) % P == val(a) ** 2 % P
            assert max(r) <= N
    for t in range(300):
        a = [2**32 - 16] * 9 if t == 0 else [rng.randrange(2**32 - 15) for _ in range(9)]
        r = (ctypes.c_uint32 * 9)()
        H.h_fe_carry(r, arr(a))
        assert val(r) % P == val(a) % P
        assert max(r) <= N


def test_sub_never_underflows(H):
    rng = random.Random(4)
    for t in range(300):
        a = [0] * 9 if t == 0 else [rng.randrange(N + 1) for _ in range(9)]
        b = [N] * 9 if t < 2 else [rng.randrange(N + 1) for _ in range(9)]
        r = (ctypes.c_uint32 * 9)()
        H.h_fe_sub(r, arr(a), arr(b))
        assert val(r) % P == (val(a) - val(b)) % P
        assert max(r) <= CLS["A2"]


def test_canonical_store_and_unpack(H):
    rng = random.Random(5)
    specials = [0, 1, P - 1, P, P + 1, 2**255 - 1, 2**255 + 5, 2**256 - 1]
    for t in range(400):
        if t < len(specials):
            x = specials[t]
            a = limbs_of(x) if x < 2**261 else None
        elif t == len(specials):                   # every limb at the class-N maximum
            a = [N] * 9
            x = val(a)
        else:
            a = [rng.randrange(N + 1) for _ in range(9)]
            x = val(a)
        w = (ctypes.c_uint32 * 8)()
        H.h_fe_canonical(w, arr(a))
        assert words_val(w) == x % P
        H.h_fe_store(w, arr(a))
        assert words_val(w) < 2**256 and words_val(w) % P == x % P
        # the resident format: limb 8's bits from 2^255 on fold into limb 0
        # (times 19), the rest is the plain integer sum of the limbs
        v = list(a)
        c = v[8] >> 23
        v[8] &= (1 << 23) - 1
        v[0] += 19 * c
        assert words_val(w) == sum(vi << (29 * i) for i, vi in enumerate(v))
        r = (ctypes.c_uint32 * 9)()
        H.h_fe_from_words(r, w)
        assert val(r) == words_val(w) and max(r) <= N


def test_invert_pow_constants(H):
    rng = random.Random(6)
    for _ in range(20):
        x = rng.randrange(1, P)
        r = (ctypes.c_uint32 * 9)()
        H.h_fe_invert(r, arr(limbs_of(x)))
        assert val(r) % P == pow(x, P - 2, P)
        H.h_fe_pow22523(r, arr(limbs_of(x)))
        assert val(r) % P == pow(x, (P - 5) // 8, P)
    d, d2, s = (ctypes.c_uint32 * 9)(), (ctypes.c_uint32 * 9)(), (ctypes.c_uint32 * 9)()
    H.h_constants(d, d2, s)
    assert val(d) == edref.D and val(d2) == 2 * edref.D % P
    assert val(s) == edref.SQRTM1
    h = (ctypes.c_uint32 * 9)()
    H.h_inv2(h)
    assert val(h) == (P + 1) // 2 and max(h) <= N
    H.h_inv_d(h)
    assert val(h) * edref.D % P == 1 and max(h) <= N
    H.h_inv_d2(h)
    assert val(h) * 2 * edref.D % P == 1 and max(h) <= N


def enc_words(pt):
    return to_words(int.from_bytes(edref.encode(pt), "little"))


def test_group_formulas(H):
    rng = random.Random(7)
    out = (ctypes.c_uint32 * 8)()
    for _ in range(25):
        a, b = rng.randrange(edref.L), rng.randrange(edref.L)
        pa, pb = edref.mul(a, edref.G), edref.mul(b, edref.G)
        H.h_add_encoded(out, arr(enc_words(pa)), arr(enc_words(pb)))
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(pa, pb)), "little")
        H.h_sub_encoded(out, arr(enc_words(pa)), arr(enc_words(pb)))
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(pa, edref.neg(pb))), "little")
        H.h_dbl_encoded(out, arr(enc_words(pa)))
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(pa, pa)), "little")
        H.h_neg_encoded(out, arr(enc_words(pa)))
        assert words_val(out) == int.from_bytes(edref.encode(edref.neg(pa)), "little")
        # mixed add over the comb tables' entry form (halved sums, d x y),
        # with and without the conditional negation, and the identity entry
        H.h_madd_encoded(out, arr(enc_words(pa)), arr(enc_words(pb)), 0)
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(pa, pb)), "little")
        H.h_madd_encoded(out, arr(enc_words(pa)), arr(enc_words(pb)), 1)
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(pa, edref.neg(pb))), "little")
        H.h_madd_encoded(out, arr(enc_words(pa)), None, 0)
        assert words_val(out) == int.from_bytes(edref.encode(pa), "little")
        H.h_madd_encoded(out, arr(enc_words(pa)), arr(enc_words(pa)), 0)
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(pa, pa)), "little")
        # the radix-32 comb's start: row 0's entry as an extended point (also
        # negated, and the identity entry), then a mixed add
        H.h_precomp_start_encoded(out, arr(enc_words(pa)), arr(enc_words(pb)), 0)
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(pb, pa)), "little")
        H.h_precomp_start_encoded(out, arr(enc_words(pa)), arr(enc_words(pb)), 1)
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(edref.neg(pb), pa)), "little")
        H.h_precomp_start_encoded(out, arr(enc_words(pa)), None, 1)
        assert words_val(out) == int.from_bytes(edref.encode(pa), "little")
        H.h_precomp_start_encoded(out, arr(enc_words(pa)), arr(enc_words(edref.neg(pa))), 1)
        assert words_val(out) == int.from_bytes(edref.encode(edref.add(pa, pa)), "little")


def naf3(k):
    """Width-3 NAF digits of k > 0, lowest first (digits in {0, +-1, +-3})."""
    d = []
    while k:
        v = 0
        if k & 1:
            v = k & 7
            if v > 4:
                v -= 8
            k -= v
        d.append(v)
        k >>= 1
    return d


def test_public_double_and_add(H):
    """ge_dbl_add_public (and ge_sub inside it), the step of the validation
    kernel's l P, against Python: l and random public scalars times random
    points and points with an order-8 component."""
    H.h_mul_public_encoded.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    rng = random.Random(9)
    t8 = None
    y = 2
    while t8 is None:
        pt = edref.decode(int.to_bytes(y, 32, "little"))
        if pt is not None:
            t = edref.mul(edref.L, pt)
            if edref.mul(4, t) != edref.IDENT:
                t8 = t
        y += 1
    out = (ctypes.c_uint32 * 8)()
    for trial in range(12):
        k = edref.L if trial < 4 else rng.randrange(1, 2**253)
        p = edref.mul(rng.randrange(1, edref.L), edref.G)
        if trial % 2:
            p = edref.add(p, edref.mul(trial, t8))
        d = naf3(k)
        dig = (ctypes.c_int8 * len(d))(*d)
        H.h_mul_public_encoded(out, arr(enc_words(p)), dig, len(d))
        assert words_val(out) == int.from_bytes(edref.encode(edref.mul(k, p)), "little"), trial


def _digits(e, nd):
    b = np.frombuffer(bytes(e), dtype=np.int8)[:nd]
    return [int(x) for x in b]


def test_scalar_mod_l_negate_and_radix32_digits(H):
    """scalar.h: m mod l for any 256-bit m (both outcomes of the q delta
    subtraction), l - m, and the signed radix-32 digits of the base-point comb
    (value, digit ranges, the top digit's bound that keeps it in a 16-entry row)"""
    L = edref.L
    rng = random.Random(11)
    ms = [0, 1, L - 1, L, L + 1, 2 * L - 1, 2 * L, 2**252 - 1, 2**252, 2**253 - 1, 2**255, 15 * 2**252,
          2**256 - 1, 15 * L + 7] + [rng.getrandbits(256) for _ in range(3000)]
    for m in ms:
        w = arr(to_words(m))
        H.h_scalar_mod_l(w)
        assert words_val(w) == m % L, hex(m)
    small = [0, 1, 2, L - 1, L - 2] + [rng.randrange(L) for _ in range(500)]
    for m in small:
        for neg in (0, 1):
            w = arr(to_words(m))
            H.h_scalar_cneg_l(w, neg)
            assert words_val(w) == ((L - m) if neg else m)
    e = (ctypes.c_uint32 * 13)()
    tops = []
    for m in small + [L, 2**253 - 1, 2**252 + 2**251] + [rng.getrandbits(253) for _ in range(500)]:
        H.h_recode32_51(e, arr(to_words(m)))
        d = _digits(e, 51)
        assert sum(x * 32**v for v, x in enumerate(d)) == m
        assert all(-16 <= x <= 15 for x in d[:50]) and 0 <= d[50] <= 8
        tops.append(d[50])
    assert max(tops) == 8
    e4 = (ctypes.c_uint32 * 4)()
    for m in [0, 1, 2**63, 2**64 - 1, 2**60 - 1] + [rng.getrandbits(64) for _ in range(500)]:
        H.h_recode32_13(e4, arr(to_words(m)))
        d = _digits(e4, 13)
        assert sum(x * 32**v for v, x in enumerate(d)) == m
        assert all(-16 <= x <= 15 for x in d[:12]) and 0 <= d[12] <= 16
    H.h_recode32_13(e4, arr(to_words(2**64 - 1)))
    assert _digits(e4, 13)[12] == 16


def test_decode_semantics(H):
    rng = random.Random(8)
    out = (ctypes.c_uint32 * 8)()
    bad = 0
    for t in range(200):
        s = bytes(rng.getrandbits(8) for _ in range(32))
        ok = H.h_decode_encode(out, arr(to_words(int.from_bytes(s, "little"))))
        want = edref.decode(s)
        if want is None:
            assert ok == 0
            bad += 1
        else:
            assert ok == 1 and words_val(out) == int.from_bytes(edref.encode(want), "little")
    assert bad > 20
