"""CPU tests that pin the oracle before it is trusted as the parity checker.

Pinning of each oracle function (none is left unpinned):
  field mul/add/sub/invert/pow22523/encode ... Python big integers mod p,
                                               special values 0, p-1, p..2^255-1, bit 255
  constants d, 2d, sqrt(-1), base point ...... defining equations; B's encoding 58 66..66
                                               (ftcrypto.cu:120-125)
  fixed-base comb / variable-base mult ....... RFC 8032 section 7.1 key-generation answers,
                                               libsodium public API, Python double-and-add,
                                               l*B = identity, (l-1)*B = -B
  point add / sub / neg ...................... libsodium crypto_core_ed25519_add, Python
  point decode / encode ...................... Python decoder (incl. non-canonical y,
                                               x = 0 with sign bit, non-squares)
  reduce / scan / prescan .................... closed forms (sum of scalars)*B;
                                               run_tests.cc:121-134 expectations
  segmented / gather sum ..................... Python loops over the same CSR
  reduce_isum ................................ run_tests.cc:175-189 expectation
  encrypt / rerandomize ...................... closed form (m + r x)*B with x the secret key
  mul_u32 / scale ............................ closed forms k*s*B mod l
  scalar reduce mod l ........................ Python int % l, libsodium scalar_reduce
"""
import ctypes
import hashlib
import os
import random

import numpy as np
import pytest

import edref
from oracle import oracle as O

P, L = edref.P, edref.L


def rnd_bytes(rng, n):
    return bytes(rng.getrandbits(8) for _ in range(n))


# --------------------------------------------------------------------- field

def test_field_ops_against_python():
    rng = random.Random(11)
    specials = [0, 1, 2, P - 1, P, P + 1, 2**255 - 1, 2**255 - 20, 19, 2**254]
    vals = specials + [rng.getrandbits(255) for _ in range(300)]
    for i, a in enumerate(vals):
        b = vals[(i * 7 + 3) % len(vals)]
        ab, bb = int.to_bytes(a, 32, "little"), int.to_bytes(b, 32, "little")
        # bit 255 must be ignored by decoding
        ab_hi = bytearray(ab)
        ab_hi[31] |= 0x80
        assert O.fe_op("mul_bytes", bytes(ab_hi), bb) == edref.fe_bytes(a * b)
        assert O.fe_op("add_bytes", ab, bb) == edref.fe_bytes(a + b)
        assert O.fe_op("sub_bytes", ab, bb) == edref.fe_bytes(a - b)
        assert O.fe_op("tobytes_roundtrip", ab) == edref.fe_bytes(a)
        if a % P:
            assert O.fe_op("invert_bytes", ab) == edref.fe_bytes(pow(a, P - 2, P))
        assert O.fe_op("pow22523_bytes", ab) == edref.fe_bytes(pow(a, (P - 5) // 8, P))


def test_constants():
    d, d2, sqrtm1, base = O.constants()
    dv = int.from_bytes(d, "little")
    assert (dv * 121666 + 121665) % P == 0
    assert int.from_bytes(d2, "little") == 2 * dv % P
    s = int.from_bytes(sqrtm1, "little")
    assert s * s % P == P - 1
    assert base == bytes([0x58] + [0x66] * 31)
    assert base == edref.encode(edref.G)


# --------------------------------------------------------------------- group

RFC8032 = [
    ("9d61b19deffd5a60ba844af492ec2cc44449c5697b326919703bac031cae7f60",
     "d75a980182b10ab7d54bfed3c964073a0ee172f3daa62325af021a68f707511a"),
    ("4ccd089b28ff96da9db6c346ec114e0f5b8a319f35aba624da8cf6ed4fb8a6fb",
     "3d4017c3e843895a92b70aa74d1b7ebc9c982ccf2ec4968cc0cd55f12af4660c"),
    ("c5aa8df43f9f837bedb7442f31dcb7b166d38535076f094b85ce3a2e0b4458f7",
     "fc51cd8e6218a1a38da47ed00230f0580816ed13ba3303ac5deb911548908025"),
]


@pytest.mark.parametrize("sk,pk", RFC8032)
def test_rfc8032_keygen(sk, pk):
    pub, a = edref.rfc8032_public_key(bytes.fromhex(sk))
    assert pub.hex() == pk
    assert O.scalarmult_base(a).hex() == pk
    B = bytes([0x58] + [0x66] * 31)
    assert O.scalarmult(a, B).hex() == pk


def test_group_order_and_small_multiples():
    B = edref.encode(edref.G)
    ident = bytes([1] + [0] * 31)
    assert O.scalarmult_base(edref.scalar_bytes(0)) == ident
    assert O.scalarmult_base(edref.scalar_bytes(L)) == ident
    assert O.scalarmult(edref.scalar_bytes(L), B) == ident
    negB = edref.encode(edref.neg(edref.G))
    assert negB[31] == 0xE6
    assert O.scalarmult_base(edref.scalar_bytes(L - 1)) == negB
    for k in list(range(0, 40)) + [255, 256, 257, 2**64 - 1, 2**128 + 5]:
        want = edref.encode(edref.mul(k, edref.G))
        assert O.scalarmult_base(edref.scalar_bytes(k)) == want, k
        assert O.scalarmult(edref.scalar_bytes(k), B) == want, k


def test_scalarmult_random_against_python():
    rng = random.Random(5)
    for _ in range(12):
        k = rng.getrandbits(255)
        s = rng.randrange(L)
        Pt = edref.mul(s, edref.G)
        want = edref.encode(edref.mul(k, Pt))
        assert O.scalarmult(edref.scalar_bytes(k), edref.encode(Pt)) == want
        assert O.scalarmult_base(edref.scalar_bytes(k)) == edref.encode(edref.mul(k, edref.G))


def test_top_digit_semantics_beyond_precondition():
    """Scalars with byte 31 > 127 break the reference precondition
    (ftcrypto.h:361-362); its recoding then selects the identity for a top digit
    above 8.  The oracle follows the same digit rule for both routines."""
    rng = random.Random(9)
    B = edref.encode(edref.G)
    for _ in range(4):
        k = rng.getrandbits(256) | (1 << 255)
        e = []
        a = k
        for i in range(64):
            e.append((a >> (4 * i)) & 15)
        carry = 0
        for i in range(63):
            e[i] += carry
            carry = (e[i] + 8) >> 4
            e[i] -= carry * 16
        e[63] += carry
        eff = sum((d if abs(d) <= 8 else 0) * 16**i for i, d in enumerate(e))
        want = edref.encode(edref.mul(eff % L, edref.G))
        assert O.scalarmult_base(edref.scalar_bytes(k)) == want
        assert O.scalarmult(edref.scalar_bytes(k), B) == want


def test_point_add_against_python():
    rng = random.Random(7)
    for _ in range(20):
        a, b = rng.randrange(L), rng.randrange(L)
        pa, pb = edref.mul(a, edref.G), edref.mul(b, edref.G)
        assert O.point_add(edref.encode(pa), edref.encode(pb)) == edref.encode(edref.add(pa, pb))
    # doubling and adding the identity / the inverse through the unified formula
    pa = edref.mul(12345, edref.G)
    assert O.point_add(edref.encode(pa), edref.encode(pa)) == edref.encode(edref.mul(24690, edref.G))
    assert O.point_add(edref.encode(pa), edref.encode(edref.neg(pa))) == bytes([1] + [0] * 31)


def test_decode_semantics():
    rng = random.Random(3)
    n_bad = 0
    for _ in range(200):
        s = rnd_bytes(rng, 32)
        rc, out = O.decode_roundtrip(s)
        want = edref.decode(s)
        if want is None:
            assert rc == -1
            n_bad += 1
        else:
            assert rc == 0
            assert out == edref.encode(want)
    assert 40 < n_bad < 160
    # non-canonical y (y >= p) decodes as y mod p; x = 0 with the sign bit set is accepted
    one_noncanon = int.to_bytes(P + 1, 32, "little")
    assert O.decode_roundtrip(one_noncanon) == (0, bytes([1] + [0] * 31))
    ident_signed = bytearray([1] + [0] * 31)
    ident_signed[31] |= 0x80
    assert O.decode_roundtrip(bytes(ident_signed)) == (0, bytes([1] + [0] * 31))


# ---------------------------------------------------------------- libsodium

def _sodium():
    for path in ("/opt/conda/lib/libsodium.so.23", "/opt/conda/lib/libsodium.so"):
        if os.path.exists(path):
            s = ctypes.CDLL(path)
            if s.sodium_init() < 0:
                return None
            return s
    return None


@pytest.fixture(scope="module")
def sodium():
    s = _sodium()
    if s is None:
        pytest.skip("libsodium not present")
    return s


def test_against_libsodium(sodium):
    rng = random.Random(21)
    out = ctypes.create_string_buffer(32)
    for _ in range(40):
        k = rng.randrange(1, L)
        kb = edref.scalar_bytes(k)
        assert sodium.crypto_scalarmult_ed25519_base_noclamp(out, kb) == 0
        assert O.scalarmult_base(kb) == out.raw
        pb = out.raw
        k2 = rng.randrange(1, L)
        assert sodium.crypto_scalarmult_ed25519_noclamp(out, edref.scalar_bytes(k2), pb) == 0
        assert O.scalarmult(edref.scalar_bytes(k2), pb) == out.raw
        qb = O.scalarmult_base(edref.scalar_bytes(rng.randrange(1, L)))
        assert sodium.crypto_core_ed25519_add(out, pb, qb) == 0
        assert O.point_add(pb, qb) == out.raw
    wide = np.frombuffer(rnd_bytes(rng, 64 * 50), dtype=np.uint8)
    red = O.scalar_reduce64(wide)
    for i in range(50):
        sodium.crypto_core_ed25519_scalar_reduce(out, bytes(wide[64 * i: 64 * i + 64]))
        assert bytes(red[i]) == out.raw
        assert int.from_bytes(bytes(red[i]), "little") == int.from_bytes(bytes(wide[64 * i: 64 * i + 64]), "little") % L


# -------------------------------------------------------------------- arrays

def scal(ks):
    return np.frombuffer(b"".join(edref.scalar_bytes(k) for k in ks), dtype=np.uint8)


def enc_mul(k):
    return O.scalarmult_base(edref.scalar_bytes(k % L))


def test_run_tests_expectations():
    """The relations run_tests.cc prints (run_tests.cc:121-134, :175-189, :67-71)."""
    a = O.Array.from_scalars(scal(range(9)))
    assert bytes(a.reduce()) == enc_mul(36)
    s = O.Array.from_scalars(scal(range(7)))
    s.scan()
    out = s.to_bytes().reshape(-1, 32)
    for i, v in enumerate([0, 1, 3, 6, 10, 15, 21]):
        assert bytes(out[i]) == enc_mul(v)
    res = O.Array.from_small_scalars([5, 8, 20, 30, 50])
    src = O.Array.from_small_scalars([10, 20, 30])
    assert res.reduce_isum([2, 2, 3], src) == 0
    out = res.to_bytes().reshape(-1, 32)
    for i, v in enumerate([5, 8, 50, 60, 50]):
        assert bytes(out[i]) == enc_mul(v)
    z = O.Array(16)
    assert bytes(z.to_bytes()) == bytes([1] + [0] * 31) * 16


def test_array_ops_closed_forms():
    rng = random.Random(4)
    n = 300
    ks = [rng.randrange(L) for _ in range(n)]
    js = [rng.randrange(L) for _ in range(n)]
    A = O.Array.from_scalars(scal(ks))
    V = O.Array.from_scalars(scal(ks), varbase=True)
    assert np.array_equal(A.to_bytes(), V.to_bytes())
    Bv = O.Array.from_scalars(scal(js))
    C = A.copy()
    C.add(Bv)
    got = C.to_bytes().reshape(-1, 32)
    for i in range(0, n, 37):
        assert bytes(got[i]) == enc_mul(ks[i] + js[i])
    C.sub(Bv)
    assert np.array_equal(C.to_bytes(), A.to_bytes())
    C.neg()
    got = C.to_bytes().reshape(-1, 32)
    assert bytes(got[3]) == enc_mul(L - ks[3])
    assert bytes(A.reduce()) == enc_mul(sum(ks))
    for excl in (False, True):
        S = A.copy()
        S.scan(exclusive=excl)
        got = S.to_bytes().reshape(-1, 32)
        for i in (0, 1, 2, 150, n - 1):
            want = sum(ks[:i] if excl else ks[: i + 1])
            assert bytes(got[i]) == enc_mul(want)
    # variable-base and u32 multiples
    M = A.copy()
    M.mul(scal(js))
    got = M.to_bytes().reshape(-1, 32)
    for i in range(0, n, 50):
        assert bytes(got[i]) == enc_mul(ks[i] * js[i])
    u = np.array([rng.getrandbits(32) for _ in range(n)], dtype=np.uint32)
    M = A.copy()
    M.mul_u32(u)
    got = M.to_bytes().reshape(-1, 32)
    for i in range(0, n, 50):
        assert bytes(got[i]) == enc_mul(ks[i] * int(u[i]))
    M = A.copy()
    M.scale(edref.scalar_bytes(12345))
    got = M.to_bytes().reshape(-1, 32)
    assert bytes(got[7]) == enc_mul(ks[7] * 12345)


def test_segmented_and_gather_sum():
    rng = random.Random(8)
    n = 200
    ks = [rng.randrange(L) for _ in range(n)]
    A = O.Array.from_scalars(scal(ks))
    lens = [0, 1, 5, 0, 0, 30, 2, 100, 0, 62]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    out = O.segmented_sum(off, A).to_bytes().reshape(-1, 32)
    for s in range(len(lens)):
        assert bytes(out[s]) == enc_mul(sum(ks[off[s]:off[s + 1]]))
    idx = np.array([rng.randrange(n) for _ in range(int(off[-1]))], dtype=np.int32)
    out = O.segmented_sum(off, A, idx=idx).to_bytes().reshape(-1, 32)
    for s in range(len(lens)):
        assert bytes(out[s]) == enc_mul(sum(ks[j] for j in idx[off[s]:off[s + 1]]))
    with pytest.raises(IndexError):
        O.segmented_sum(off, A, idx=np.full(int(off[-1]), n, dtype=np.int32))


def test_encrypt_closed_form():
    rng = random.Random(12)
    x = rng.randrange(L)
    H = enc_mul(x)
    n = 64
    m = np.array([rng.getrandbits(1) if i % 2 else rng.getrandbits(8) for i in range(n)], dtype=np.uint64)
    rs = [rng.randrange(L) for _ in range(n)]
    c1, c2 = O.encrypt(H, m, scal(rs))
    b1, b2 = c1.to_bytes().reshape(-1, 32), c2.to_bytes().reshape(-1, 32)
    for i in range(n):
        assert bytes(b1[i]) == enc_mul(rs[i])
        assert bytes(b2[i]) == enc_mul(int(m[i]) + rs[i] * x)
    rho = [rng.randrange(L) for _ in range(n)]
    O.rerandomize(c1, c2, H, scal(rho))
    b1, b2 = c1.to_bytes().reshape(-1, 32), c2.to_bytes().reshape(-1, 32)
    for i in range(0, n, 9):
        assert bytes(b1[i]) == enc_mul(rs[i] + rho[i])
        assert bytes(b2[i]) == enc_mul(int(m[i]) + (rs[i] + rho[i]) * x)


def test_scalar_reduce_against_python():
    rng = random.Random(2)
    wide = [rng.getrandbits(512) for _ in range(100)] + [0, L, L - 1, 2**512 - 1]
    buf = np.frombuffer(b"".join(int.to_bytes(w, 64, "little") for w in wide), dtype=np.uint8)
    red = O.scalar_reduce64(buf)
    for i, w in enumerate(wide):
        assert int.from_bytes(bytes(red[i]), "little") == w % L
