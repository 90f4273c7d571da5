"""Parity of the gfx950 library (through its C ABI) with the CPU oracle.

Every comparison is byte-exact on the canonical 32-byte compressed encodings
(or on the exact int64 / index outputs), on identical seeded inputs.  Sizes
keep the oracle to a few seconds; the full-size runs are in
tests/test_gpu_configs.py.
"""
import numpy as np
import pytest

import edref
import libftcrypto_amd as ft
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L = edref.L
IDENT = bytes([1] + [0] * 31)


@pytest.fixture(scope="module", autouse=True)
def _init():
    ft.init()


def scalars(n, seed, stream=0):
    return ft.random_scalars(n, seed, stream)


def sb(k):
    return np.frombuffer(int.to_bytes(k, 32, "little"), dtype=np.uint8)


def edge_scalars():
    ks = [0, 1, 2, 7, 8, 9, 15, 16, 255, 256, L - 1, L, L + 1, 2**252, 2**253 - 1, 2**255 - 1,
          2**254 + 2**251, int("8" * 63, 16), int("7" * 63, 16)]
    return np.stack([sb(k) for k in ks])


def same(gpu_arr, oracle_arr):
    a = gpu_arr.to_bytes_folded()
    b = oracle_arr.to_bytes()
    assert len(a) == len(b)
    if not np.array_equal(a, b):
        bad = np.nonzero((a.reshape(-1, 32) != b.reshape(-1, 32)).any(axis=1))[0]
        raise AssertionError("%d of %d points differ, first at %d" % (len(bad), len(a) // 32, bad[0]))


# ------------------------------------------------------------ scalar mults

def test_fixed_base_edges_and_random():
    s = np.concatenate([edge_scalars(), scalars(3000, 101)])
    same(ft.GeArray.from_scalars(s), O.Array.from_scalars(s))


def test_fixed_base_top_digit_beyond_precondition():
    # byte 31 > 127 breaks the reference precondition; the digit rule is kept
    rng = np.random.default_rng(5)
    s = rng.integers(0, 256, size=(64, 32), dtype=np.uint8)
    s[:, 31] |= 0x80
    same(ft.GeArray.from_scalars(s), O.Array.from_scalars(s))


B_ENC = bytes([0x58] + [0x66] * 31)


def _times_base_variable(sc32):
    """k_i B through the variable-base kernel (B decoded from its encoding)"""
    p, bad = ft.GeArray.from_bytes_folded(np.frombuffer(B_ENC * len(sc32), dtype=np.uint8))
    assert bad == -1                                   # every point decoded
    p.mul(sc32)
    return p.to_bytes_folded()


def test_fixed_base_every_element_matches_variable_base():
    """the radix-32 comb of B (scalars reduced mod l after the radix-16
    top-digit rule) equals the variable-base product of B on every element
    of a launch of 2^18 + 3001, and the oracle on the edge scalars, top
    digits above 8 and a sample"""
    n = (1 << 18) + 3001
    s = scalars(n, 303)
    rng = np.random.default_rng(8)
    top = rng.integers(0, 256, size=(64, 32), dtype=np.uint8)
    top[:, 31] |= 0x80
    full = np.full((4, 32), 255, dtype=np.uint8)
    full[1, 31] = 0x8f                                 # top nibble 8 with a carry in
    full[2, 31] = 0x87
    full[3, :31] = 0
    special = np.concatenate([edge_scalars(), top, full])
    pos = np.arange(len(special)) * 4099 % n
    pos[-1] = n - 1
    s[pos] = special
    g = ft.GeArray.from_scalars(s).to_bytes_folded()
    assert np.array_equal(g, _times_base_variable(s))
    idx = np.unique(np.concatenate([pos, rng.choice(n, 1500, replace=False)]))
    assert np.array_equal(g.reshape(-1, 32)[idx], O.Array.from_scalars(s[idx]).to_bytes().reshape(-1, 32))


def test_fixed_base_small_scalars_every_element():
    n = (1 << 18) + 777
    rng = np.random.default_rng(9)
    m = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)
    edges = np.array([0, 1, 2, 15, 16, 17, 2**32, 2**60 - 1, 2**63, 2**64 - 1, 2**64 - 2], dtype=np.uint64)
    pos = np.arange(len(edges)) * 7919 % n
    m[pos] = edges
    g = ft.GeArray.from_small_scalars(m).to_bytes_folded()
    sc32 = np.zeros((n, 32), dtype=np.uint8)
    sc32[:, :8] = m.view(np.uint8).reshape(n, 8)
    assert np.array_equal(g, _times_base_variable(sc32))
    idx = np.unique(np.concatenate([pos, rng.choice(n, 800, replace=False)]))
    assert np.array_equal(g.reshape(-1, 32)[idx], O.Array.from_small_scalars(m[idx]).to_bytes().reshape(-1, 32))


def test_small_scalars():
    rng = np.random.default_rng(6)
    m = np.concatenate([np.array([0, 1, 2, 5, 2**32, 2**63, 2**64 - 1], dtype=np.uint64),
                        rng.integers(0, 2**63, size=500, dtype=np.uint64)])
    same(ft.GeArray.from_small_scalars(m), O.Array.from_small_scalars(m))


def test_variable_base_mul_and_scale():
    base = scalars(1500, 102)
    k = np.concatenate([edge_scalars(), scalars(1500 - len(edge_scalars()), 103)])
    g = ft.GeArray.from_scalars(base)
    o = O.Array.from_scalars(base)
    g.mul(k)
    o.mul(k)
    same(g, o)
    one = bytes(scalars(1, 104)[0])
    g.scale(one)
    o.scale(one)
    same(g, o)
    with pytest.raises(ft.FtError) as e:
        g.mul(k[:10])
    assert e.value.code == ft.FT_ERR_SIZE_MISMATCH


def test_mul_u32_pair():
    n = 1200
    a, b = scalars(n, 105), scalars(n, 106)
    k = np.random.default_rng(7).integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    k[:4] = [0, 1, 2**32 - 1, 2**31]
    g1, g2 = ft.GeArray.from_scalars(a), ft.GeArray.from_scalars(b)
    o1, o2 = O.Array.from_scalars(a), O.Array.from_scalars(b)
    g1.mul_u32(k, g2)
    o1.mul_u32(k)
    o2.mul_u32(k)
    same(g1, o1)
    same(g2, o2)


# ------------------------------------------------------------ group ops

def test_add_sub_neg_and_aliasing():
    n = 2000
    a, b = scalars(n, 107), scalars(n, 108)
    ga, gb = ft.GeArray.from_scalars(a), ft.GeArray.from_scalars(b)
    oa, ob = O.Array.from_scalars(a), O.Array.from_scalars(b)
    ga.add(gb)
    oa.add(ob)
    same(ga, oa)
    ga.sub(gb)
    oa.sub(ob)
    same(ga, oa)
    ga.add(ga)              # ft_add(arr, arr): doubling through the unified formula
    oa.add(oa.copy())
    same(ga, oa)
    ga.neg()
    oa.neg()
    same(ga, oa)
    z = ft.GeArray.init_scalar(n, None)
    z.add(gb)
    same(z, ob)


def test_init_scalar_and_point():
    z = ft.GeArray.init_scalar(16)
    assert bytes(z.to_bytes_folded()) == IDENT * 16
    k = scalars(1, 109)[0]
    g = ft.GeArray.init_scalar(40, k)
    want = O.scalarmult_base(bytes(k))
    assert bytes(g.to_bytes_folded()) == want * 40
    src = ft.GeArray.from_scalars(scalars(5, 110))
    bc = ft.GeArray.init_point(33, src, 3)
    assert bytes(bc.to_bytes_folded()) == bytes(src.to_bytes_folded()[96:128]) * 33
    with pytest.raises(ft.FtError):
        ft.GeArray.init_point(3, src, 5)


# ------------------------------------------------------------ I/O

def test_encode_decode_roundtrip_and_failures():
    s = scalars(700, 111)
    enc = O.Array.from_scalars(s).to_bytes()
    g, bad = ft.GeArray.from_bytes_folded(enc)
    assert bad == -1
    assert np.array_equal(g.to_bytes_folded(), enc)
    rng = np.random.default_rng(8)
    junk = rng.integers(0, 256, size=600 * 32, dtype=np.uint8)
    g, bad = ft.GeArray.from_bytes_folded(junk)
    o, obad = O.Array.from_bytes(junk)
    assert bad == obad and bad > 0
    # points that decode must match the oracle's decoding
    ge = g.to_bytes_folded().reshape(-1, 32)
    oe = o.to_bytes().reshape(-1, 32)
    ok = [i for i in range(600) if O.decode_roundtrip(bytes(junk[32 * i:32 * i + 32]))[0] == 0]
    assert len(ok) > 100
    for i in ok:
        assert bytes(ge[i]) == bytes(oe[i])
    with pytest.raises(ft.FtError) as e:
        ft.GeArray.from_bytes_folded(junk[:33])
    assert e.value.code == ft.FT_ERR_BAD_SIZE


def test_decode_dev_matches_host_decode():
    """ftx_decode_dev (device bytes, stream-ordered, failure index in device
    memory) against ft_array_from_bytes_folded on valid and random bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    L = ft.lib()
    rng = np.random.default_rng(9)
    valid = O.Array.from_scalars(scalars(3000, 113)).to_bytes()
    junk = rng.integers(0, 256, size=2500 * 32, dtype=np.uint8)
    junk[-32:] = valid[:32]                     # last one decodes: the failure index is an earlier one
    st = ft.Stream()
    for data in (valid, junk):
        n = len(data) // 32
        d_in, d_fail = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(d_in), len(data)) == 0
        assert hip.hipMalloc(ctypes.byref(d_fail), 8) == 0
        assert hip.hipMemcpy(d_in, data.ctypes.data_as(ctypes.c_void_p), len(data), 1) == 0
        out = ft.GeArray.alloc(n)
        ft.check(L.ftx_decode_dev(out.h, d_in, d_fail, st.h))
        st.synchronize()
        fail = np.zeros(1, dtype=np.uint64)
        assert hip.hipMemcpy(fail.ctypes.data_as(ctypes.c_void_p), d_fail, 8, 2) == 0
        hip.hipFree(d_in)
        hip.hipFree(d_fail)
        g, bad = ft.GeArray.from_bytes_folded(data)
        assert int(fail[0]) - 1 == bad
        assert np.array_equal(out.to_bytes_folded(), g.to_bytes_folded())
    assert bad > 0


def special_encodings():
    """Encodings the reference's decoder treats specially (sodium_fe.cu:70-141,
    sodium_ge.cu:137-180): y >= p (bit 255 ignored, y not reduced first), x = 0
    with the sign bit set, y = 0, y = -1."""
    P = 2**255 - 19

    def enc(y, sign=0):
        return (y | (sign << 255)).to_bytes(32, "little")

    cases = {
        "y=p+1 (identity)": enc(P + 1),
        "y=p+1, sign set": enc(P + 1, 1),
        "identity, bit 255 set": enc(1, 1),
        "y=p-1=2^255-20 (order 2)": enc(2**255 - 20),
        "y=p-1, sign set": enc(2**255 - 20, 1),
        "y=p (y=0: order 4)": enc(P),
        "y=p, sign set": enc(P, 1),
        "y=0": enc(0),
        "y=0, sign set": enc(0, 1),
        "y=2^255-1": enc(2**255 - 1),
        "y=2^255-1, sign set": enc(2**255 - 1, 1),
        "y=p+2": enc(P + 2),
        "y=p+3, sign set": enc(P + 3, 1),
        "B": bytes.fromhex("58" + "66" * 31),
        "B with y + p": enc(int.from_bytes(bytes.fromhex("58" + "66" * 31), "little") + P),
    }
    return cases


def test_decode_special_encodings():
    """Each special encoding decodes on the GPU (host and device entry points)
    exactly as in the oracle: the same success or failure, the same canonical
    re-encoding, and the same point (its sum with 3B and its double agree)."""
    import ctypes
    cases = special_encodings()
    names = list(cases)
    data = np.frombuffer(b"".join(cases[k] for k in names), dtype=np.uint8).copy()
    three_b = O.Array.from_small_scalars(np.full(1, 3, dtype=np.uint64)).to_bytes()
    L = ft.lib()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    d_in, d_fail = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_in), len(data)) == 0 and hip.hipMalloc(ctypes.byref(d_fail), 8) == 0
    for i, name in enumerate(names):
        one = data[32 * i:32 * i + 32].copy()
        rc, _ = O.decode_roundtrip(bytes(one))
        g, bad = ft.GeArray.from_bytes_folded(one)
        assert (bad == -1) == (rc == 0), name
        assert hip.hipMemcpy(d_in, one.ctypes.data_as(ctypes.c_void_p), 32, 1) == 0
        dv = ft.GeArray.alloc(1)
        ft.check(L.ftx_decode_dev(dv.h, d_in, d_fail, None))
        ft.check(L.ftx_synchronize(None))
        fail = np.zeros(1, dtype=np.uint64)
        assert hip.hipMemcpy(fail.ctypes.data_as(ctypes.c_void_p), d_fail, 8, 2) == 0
        assert int(fail[0]) - 1 == bad, name
        if rc != 0:
            continue
        o, obad = O.Array.from_bytes(one)
        assert obad == -1, name
        assert np.array_equal(g.to_bytes_folded(), o.to_bytes()), name
        assert np.array_equal(dv.to_bytes_folded(), o.to_bytes()), name
        # the decoded coordinates, not only the encoding: P + 3B and 2P
        g.add(ft.GeArray.from_small_scalars(np.full(1, 3, dtype=np.uint64)))
        o3 = O.Array.from_bytes(three_b)[0]
        o.add(o3)
        assert np.array_equal(g.to_bytes_folded(), o.to_bytes()), name
        dv.add(dv)
        o2 = O.Array.from_bytes(one)[0]
        o2.add(O.Array.from_bytes(one)[0])
        assert np.array_equal(dv.to_bytes_folded(), o2.to_bytes()), name
    # all of them at once through the device entry point: the largest failing index
    assert hip.hipMemcpy(d_in, data.ctypes.data_as(ctypes.c_void_p), len(data), 1) == 0
    allv = ft.GeArray.alloc(len(names))
    ft.check(L.ftx_decode_dev(allv.h, d_in, d_fail, None))
    ft.check(L.ftx_synchronize(None))
    fail = np.zeros(1, dtype=np.uint64)
    assert hip.hipMemcpy(fail.ctypes.data_as(ctypes.c_void_p), d_fail, 8, 2) == 0
    _, obad = O.Array.from_bytes(data)
    assert int(fail[0]) - 1 == obad
    hip.hipFree(d_in)
    hip.hipFree(d_fail)


def test_affine_and_extended_roundtrip_and_validate():
    s = scalars(300, 112)
    g = ft.GeArray.from_scalars(s)
    want = g.to_bytes_folded()
    aff = g.to_bytes_affine()
    for i in (0, 7, 299):
        pt = edref.mul(int.from_bytes(bytes(s[i]), "little"), edref.G)
        assert bytes(aff[64 * i:64 * i + 32]) == edref.fe_bytes(pt[0])
        assert bytes(aff[64 * i + 32:64 * i + 64]) == edref.fe_bytes(pt[1])
    ga = ft.GeArray.from_bytes_affine(aff)
    assert np.array_equal(ga.to_bytes_folded(), want)
    assert ga.validate() == -1
    raw = g.to_bytes()
    assert len(raw) == 300 * 160
    gr = ft.GeArray.from_bytes(raw)
    assert np.array_equal(gr.to_bytes_folded(), want)
    bad = aff.copy()
    bad[64 * 41 + 32] ^= 1           # y changed: off the curve
    bad[64 * 7] ^= 4
    assert ft.GeArray.from_bytes_affine(bad).validate() == 41


# ------------------------------------------------------------ small-order components

def order8_torsion():
    """E[8] = {k T : k = 0..7} for a point T of order 8 ([l] P of a decoded P)."""
    y = 2
    while True:
        pt = edref.decode(int.to_bytes(y, 32, "little"))
        if pt is not None:
            t = edref.mul(L, pt)
            if edref.mul(4, t) != edref.IDENT:
                return [edref.mul(k, t) for k in range(8)]
        y += 1


def test_points_with_torsion_components():
    """Points s B + k T (T of order 8): decode, variable-base mult (edge
    scalars included), scale, x u32, add / sub / neg, reduce and contains
    against the oracle; validate follows the reference's main-subgroup test
    X(l P) = 0 (sodium_ge.cu:795-882), which passes a component of order 1
    or 2 and fails one of order 4 or 8."""
    e8 = order8_torsion()
    rng = np.random.default_rng(31)
    n = 128
    sv = [int(v) for v in rng.integers(1, 1 << 30, size=n)]
    pts = [edref.add(edref.mul(sv[i], edref.G), e8[i % 8]) for i in range(n)]
    enc = np.frombuffer(b"".join(edref.encode(q) for q in pts), dtype=np.uint8)
    g, bad = ft.GeArray.from_bytes_folded(enc)
    o, obad = O.Array.from_bytes(enc)
    assert bad == -1 and obad == -1
    assert np.array_equal(g.to_bytes_folded(), o.to_bytes())
    # variable-base mult, random and edge scalars
    ks = np.concatenate([edge_scalars(), scalars(n - len(edge_scalars()), 130)])
    gm, om = ft.GeArray.from_bytes_folded(enc)[0], O.Array.from_bytes(enc)[0]
    gm.mul(ks)
    om.mul(ks)
    same(gm, om)
    gs, osc = ft.GeArray.from_bytes_folded(enc)[0], O.Array.from_bytes(enc)[0]
    gs.scale(ks[5])
    osc.scale(bytes(ks[5]))
    same(gs, osc)
    k32 = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    k32[:4] = [0, 1, 8, 2**32 - 1]
    gu, ou = ft.GeArray.from_bytes_folded(enc)[0], O.Array.from_bytes(enc)[0]
    gu.mul_u32(k32)
    ou.mul_u32(k32)
    same(gu, ou)
    # add / sub / neg against arrays shifted by other torsion elements
    sh = [edref.add(edref.mul(sv[i] + 7, edref.G), e8[(3 * i + 1) % 8]) for i in range(n)]
    enc2 = np.frombuffer(b"".join(edref.encode(q) for q in sh), dtype=np.uint8)
    for op in ("add", "sub"):
        ga, oa = ft.GeArray.from_bytes_folded(enc)[0], O.Array.from_bytes(enc)[0]
        getattr(ga, op)(ft.GeArray.from_bytes_folded(enc2)[0])
        getattr(oa, op)(O.Array.from_bytes(enc2)[0])
        same(ga, oa)
    gn, on = ft.GeArray.from_bytes_folded(enc)[0], O.Array.from_bytes(enc)[0]
    gn.neg()
    on.neg()
    same(gn, on)
    gr, orr = ft.GeArray.from_bytes_folded(enc)[0], O.Array.from_bytes(enc)[0]
    gr.reduce()
    assert bytes(gr.to_bytes_folded()) == bytes(orr.reduce())
    # contains: points differing only by a torsion component are different
    probe = [edref.add(pts[i], e8[4]) if i % 2 else pts[i] for i in range(n)]
    encp = np.frombuffer(b"".join(edref.encode(q) for q in probe), dtype=np.uint8)
    got = ft.GeArray.from_bytes_folded(encp)[0].contains(g)
    want = [1 if encp[32 * i:32 * i + 32].tobytes() in {enc[32 * j:32 * j + 32].tobytes() for j in range(n)}
            else 0 for i in range(n)]
    assert list(got) == want and sum(want) >= n // 2
    # validate: components of order 1 or 2 pass (k = 0, 4), order 4 or 8 fail
    for k in range(8):
        q = edref.add(edref.mul(sv[0], edref.G), e8[k])
        one = ft.GeArray.from_bytes_folded(np.frombuffer(edref.encode(q), dtype=np.uint8))[0]
        clean = ft.GeArray.from_small_scalars(np.arange(1, 17, dtype=np.uint64))
        clean.assign_to_subset([9], one)
        assert clean.validate() == (-1 if k in (0, 4) else 9), k


# ------------------------------------------------------------ aggregation

@pytest.mark.parametrize("n", [1, 9, 256, 257, 4097, 70000])
def test_reduce(n):
    s = scalars(n, 113 + n)
    g = ft.GeArray.from_scalars(s)
    g.reduce()
    assert len(g) == 1
    assert bytes(g.to_bytes_folded()) == bytes(O.Array.from_scalars(s).reduce())


def test_reduce_run_tests_case():
    g = ft.GeArray.from_small_scalars(np.arange(9, dtype=np.uint64))
    g.reduce()
    assert bytes(g.to_bytes_folded()) == O.scalarmult_base(bytes(sb(36)))


@pytest.mark.parametrize("n", [1, 7, 16, 17, 300, 4500])
def test_scan_and_prescan(n):
    s = scalars(n, 200 + n)
    for excl in (False, True):
        g = ft.GeArray.from_scalars(s)
        o = O.Array.from_scalars(s)
        if excl:
            g.prescan()
        else:
            g.scan()
        o.scan(exclusive=excl)
        same(g, o)


def test_scan_run_tests_case():
    g = ft.GeArray.from_small_scalars(np.arange(7, dtype=np.uint64))
    g.scan()
    want = b"".join(O.scalarmult_base(bytes(sb(v))) for v in [0, 1, 3, 6, 10, 15, 21])
    assert bytes(g.to_bytes_folded()) == want


def seg_lengths(rng, nseg):
    lens = np.clip(np.round(rng.lognormal(np.log(16), 1.5, size=nseg)), 1, 6000).astype(np.int64)
    lens[rng.integers(0, nseg, size=nseg // 10)] = 0
    lens[:3] = [0, 0, 4100]
    lens[-1] = 0
    return lens


def test_segmented_sum():
    rng = np.random.default_rng(9)
    lens = seg_lengths(rng, 400)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    s = scalars(int(off[-1]), 114)
    g = ft.GeArray.from_scalars(s)
    o = O.Array.from_scalars(s)
    same(ft.segmented_sum(off, g), O.segmented_sum(off, o))


def test_segmented_sum_chunk_aligned_and_single():
    s = scalars(16 * 40, 115)
    g, o = ft.GeArray.from_scalars(s), O.Array.from_scalars(s)
    for lens in ([16] * 40, [640], [1] * 640, [0, 640, 0], [15, 17] * 20):
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        same(ft.segmented_sum(off, g), O.segmented_sum(off, o))


def test_gather_sum():
    rng = np.random.default_rng(10)
    nsrc, nrows = 3000, 500
    deg = rng.poisson(8, size=nrows)
    deg[5] = 2000
    deg[6] = 0
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    cols = rng.integers(0, nsrc, size=int(off[-1]), dtype=np.int32)
    s = scalars(nsrc, 116)
    g, o = ft.GeArray.from_scalars(s), O.Array.from_scalars(s)
    same(ft.gather_sum(off, cols, g), O.segmented_sum(off, o, idx=cols))
    with pytest.raises(ft.FtError) as e:
        ft.gather_sum(off, np.full_like(cols, nsrc), g)
    assert e.value.code == ft.FT_ERR_OUT_OF_BOUNDS


def test_reduce_isum():
    res = ft.GeArray.from_small_scalars(np.array([5, 8, 20, 30, 50], dtype=np.uint64))
    src = ft.GeArray.from_small_scalars(np.array([10, 20, 30], dtype=np.uint64))
    res.reduce_isum([2, 2, 3], src)
    want = b"".join(O.scalarmult_base(bytes(sb(v))) for v in [5, 8, 50, 60, 50])
    assert bytes(res.to_bytes_folded()) == want
    rng = np.random.default_rng(11)
    r, s = scalars(300, 117), scalars(2000, 118)
    keys = rng.integers(0, 300, size=2000).astype(np.uint64)
    g, o = ft.GeArray.from_scalars(r), O.Array.from_scalars(r)
    g.reduce_isum(keys, ft.GeArray.from_scalars(s))
    assert o.reduce_isum(keys, O.Array.from_scalars(s)) == 0
    same(g, o)
    with pytest.raises(ft.FtError):
        g.reduce_isum(np.array([300] * 2000, dtype=np.uint64), ft.GeArray.from_scalars(s))


# ------------------------------------------------------------ ElGamal

@pytest.mark.parametrize("bits", [1, 8, 64])
def test_encrypt_and_rerandomize(bits):
    n = 1500
    x = scalars(1, 119)
    pk = ft.GeArray.from_scalars(x)
    H = bytes(pk.to_bytes_folded())
    key = ft.Key(pk, 0)
    rng = np.random.default_rng(12 + bits)
    hi = 2**bits
    m = rng.integers(0, hi if bits < 64 else 2**63, size=n, dtype=np.uint64)
    if bits == 64:
        m[0] = 2**64 - 1
    m[1] = 0
    r = scalars(n, 120)
    c1, c2 = ft.encrypt(key, m, r)
    o1, o2 = O.encrypt(H, m, r)
    same(c1, o1)
    same(c2, o2)
    rho = scalars(n, 121)
    ft.rerandomize(c1, c2, key, rho)
    O.rerandomize(o1, o2, H, rho)
    same(c1, o1)
    same(c2, o2)


# ------------------------------------------------------------ data movement / predicates

def test_subset_range_assign_delete():
    s = scalars(100, 122)
    g = ft.GeArray.from_scalars(s)
    enc = g.to_bytes_folded().reshape(-1, 32)
    idx = np.array([3, 7, 99, 0, 7], dtype=np.uint64)
    assert np.array_equal(g.get_subset(idx).to_bytes_folded().reshape(-1, 32), enc[idx.astype(int)])
    r = g.get_range(5, 50, 7)
    assert len(r) == 7
    assert np.array_equal(r.to_bytes_folded().reshape(-1, 32), enc[5:50:7])
    g2 = ft.GeArray.from_scalars(s)
    g2.assign_zero_to_subset(np.array([1, 2], dtype=np.uint64))
    g2.assign_to_range(10, 13, 1, g.get_subset(np.array([50, 51, 52], dtype=np.uint64)))
    e2 = g2.to_bytes_folded().reshape(-1, 32)
    assert bytes(e2[1]) == IDENT and bytes(e2[2]) == IDENT
    assert np.array_equal(e2[10:13], enc[50:53])
    g2.delete_range(0, 100, 2)
    assert np.array_equal(g2.to_bytes_folded().reshape(-1, 32)[6:], enc[13::2])
    g3 = ft.GeArray.from_scalars(s)
    g3.delete_subset(np.array([0, 5, 5, 99], dtype=np.uint64))
    keep = [i for i in range(100) if i not in (0, 5, 99)]
    assert np.array_equal(g3.to_bytes_folded().reshape(-1, 32), enc[keep])
    g4 = ft.GeArray.from_scalars(s)
    g4.set_length(120)
    e4 = g4.to_bytes_folded().reshape(-1, 32)
    assert np.array_equal(e4[:100], enc) and all(bytes(e4[i]) == IDENT for i in range(100, 120))
    g4.set_length(30)
    assert np.array_equal(g4.to_bytes_folded().reshape(-1, 32), enc[:30])


def test_predicates():
    a = ft.GeArray.from_small_scalars(np.array([3, 5, 8, 2], dtype=np.uint64))
    b = ft.GeArray.from_small_scalars(np.arange(1, 8, dtype=np.uint64))
    assert list(a.contains(b)) == [1, 1, 0, 1]          # run_tests.cc:157-170
    f = ft.GeArray.from_small_scalars(np.array([5], dtype=np.uint64))
    f2 = ft.GeArray.from_small_scalars(np.array([5], dtype=np.uint64))
    assert list(f.equal(f2)) == [1]                    # run_tests.cc:140-152
    f2.add(f)
    assert list(f.equal(f2)) == [0] and list(f.not_equal(f2)) == [1]
    z = ft.GeArray.from_small_scalars(np.array([0, 1, 0, 9], dtype=np.uint64))
    assert list(z.index()) == [0, 1, 0, 1]
    t = ft.GeArray.from_small_scalars(np.array([1, 2, 3], dtype=np.uint64))
    e = ft.GeArray.from_small_scalars(np.array([7, 8, 9], dtype=np.uint64))
    t.mux(np.array([1, 0, 5], dtype=np.int64), e)
    assert bytes(t.to_bytes_folded()) == b"".join(O.scalarmult_base(bytes(sb(v))) for v in [1, 8, 3])


def test_contains_device_set():
    """ft_array_contains through the device hash set, against membership of
    the scalars (k B are distinct for distinct k < l): heavy duplicates in the
    set, the identity, negated points (encodings differing in the sign bit
    only), points equal in value but built by another path, empty sets."""
    rng = np.random.default_rng(21)
    set_k = rng.integers(0, 3000, size=50000).astype(np.uint64)
    set_k[:7] = 0
    probe_k = np.arange(6000, dtype=np.uint64)
    a = ft.GeArray.from_small_scalars(probe_k - 1)           # (k - 1) B + B, k >= 1
    a.add(ft.GeArray.from_small_scalars(np.ones(6000, dtype=np.uint64)))
    a.assign_to_subset([0], ft.GeArray.from_small_scalars([0]))
    want = np.isin(probe_k, set_k).astype(np.int64)
    got = a.contains(ft.GeArray.from_small_scalars(set_k))
    assert np.array_equal(got, want)
    n = ft.GeArray.from_small_scalars(probe_k[1:])
    n.neg()
    assert not n.contains(ft.GeArray.from_small_scalars(set_k)).any()
    dup = ft.GeArray.from_small_scalars(np.full(5000, 7, dtype=np.uint64))
    assert list(np.nonzero(ft.GeArray.from_small_scalars(probe_k[:20]).contains(dup))[0]) == [7]
    empty = ft.GeArray.from_small_scalars(np.zeros(0, dtype=np.uint64))
    assert not ft.GeArray.from_small_scalars(probe_k[:100]).contains(empty).any()
    assert len(empty.contains(dup)) == 0


def test_contains_large():
    n = 1 << 20
    rng = np.random.default_rng(22)
    vals = rng.integers(0, 2**63, size=2 * n, dtype=np.uint64)
    vals = np.unique(vals)[: 2 * n - 64]
    set_k, other = vals[:n], vals[n:]
    probe = np.concatenate([set_k[::2], other[: n // 2]])
    rng.shuffle(probe)
    got = ft.GeArray.from_small_scalars(probe).contains(ft.GeArray.from_small_scalars(set_k))
    assert np.array_equal(got, np.isin(probe, set_k).astype(np.int64))


# ------------------------------------------------------------ the small reference workload

def test_config0_workload():
    """Keypair, encrypt two arrays of 65,536 bits, add them, reduce: every
    intermediate compared byte-exact with the oracle."""
    n = 65536
    x = scalars(1, 0xF7C00000, 1)
    pk = ft.GeArray.from_scalars(x)
    H = bytes(pk.to_bytes_folded())
    key = ft.Key(pk)
    rng = np.random.default_rng(0xF7C00000)
    mA = rng.integers(0, 2, size=n, dtype=np.uint64)
    mB = rng.integers(0, 2, size=n, dtype=np.uint64)
    rA, rB = scalars(n, 0xF7C00000, 2), scalars(n, 0xF7C00000, 3)
    a1, a2 = ft.encrypt(key, mA, rA, msg_bits=1)
    b1, b2 = ft.encrypt(key, mB, rB, msg_bits=1)
    oa1, oa2 = O.encrypt(H, mA, rA)
    ob1, ob2 = O.encrypt(H, mB, rB)
    for g, o in ((a1, oa1), (a2, oa2), (b1, ob1), (b2, ob2)):
        same(g, o)
    a1.add(b1)
    a2.add(b2)
    oa1.add(ob1)
    oa2.add(ob2)
    same(a1, oa1)
    same(a2, oa2)
    want1, want2 = bytes(oa1.reduce()), bytes(oa2.reduce())
    a1.reduce()
    a2.reduce()
    assert bytes(a1.to_bytes_folded()) == want1
    assert bytes(a2.to_bytes_folded()) == want2
    # closed form: the sum decrypts to sum(m) (C2 - x C1 = (sum m) B)
    xs = int.from_bytes(bytes(x[0]), "little")
    msum = int(mA.sum() + mB.sum())
    c1 = edref.decode(want1)
    c2 = edref.decode(want2)
    assert edref.add(c2, edref.neg(edref.mul(xs, c1))) == edref.mul(msum, edref.G)
