"""Every aggregation schedule the tuner may pick gives the oracle's result.

The reduce, scan and segmented sum run with each candidate chunk size and
reduce grid set for the size class of the input, and the outputs are compared
byte for byte (canonical encodings) with the oracle; the autotuner's own run
must leave results unchanged.
"""
import itertools
import os

import numpy as np
import pytest

import edref
import libftcrypto_amd as ft
from oracle import oracle as O

L = edref.L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    ft.init()
    saved = ft.tune_table()
    yield
    for c, v in saved.items():
        ft.tune_set(c, *v)


def same(a, o):
    assert np.array_equal(a.to_bytes_folded(), o.to_bytes())


class _devbuf:
    """device copy of a numpy array (freed with the object)"""

    def __init__(self, arr):
        import ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        arr = np.ascontiguousarray(arr)
        self.p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.p), max(1, arr.nbytes)) == 0
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def __del__(self):
        self.hip.hipFree(self.p)


def seg_case(n, seed):
    rng = np.random.default_rng(seed)
    lens = np.clip(np.round(rng.lognormal(np.log(16), 1.5, size=n)), 1, 900).astype(np.int64)
    lens[rng.integers(0, len(lens), size=len(lens) // 9)] = 0
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    off = off[off <= n]
    if off[-1] != n:
        off = np.concatenate([off, [n]])
    return off


@pytest.mark.parametrize("n", [5000, 70000])
def test_every_segsum_schedule(n):
    c = ft.size_class(n)
    base = ft.tune_table()[c]
    s = ft.random_scalars(n, 4100 + n, 0)
    g, o = ft.GeArray.from_scalars(s), O.Array.from_scalars(s)
    off = seg_case(n, n)
    want = O.segmented_sum(off, o).to_bytes()
    rng = np.random.default_rng(n)
    cols = rng.integers(0, n, size=n, dtype=np.int32)
    want_g = O.segmented_sum(off, o, idx=cols).to_bytes()
    d_off = _devbuf(off)
    for ch1, chu in itertools.product((4, 5, 6, 7), (1, 2, 3, 4)):
        ft.tune_set(c, ch1, chu, *base[2:])
        got = ft.segmented_sum(off, g)
        assert np.array_equal(got.to_bytes_folded(), want), (ch1, chu)
        got = ft.gather_sum(off, cols, g)
        assert np.array_equal(got.to_bytes_folded(), want_g), ("gather", ch1, chu)
        out = ft.GeArray.alloc(len(off) - 1)       # the persistent batched walk
        ft.aggregate_batch([ft.agg_job(ft.AGG_SEGSUM, out, g, d_off.p, len(off) - 1, n)])
        assert np.array_equal(out.to_bytes_folded(), want), ("batch", ch1, chu)
    ft.tune_set(c, *base)


def test_segsum_one_segment_spanning_every_chunk():
    """One segment over all n elements between empty segments: its run of
    chunk partials is as long as there are chunks, so every tree pass of every
    radix combines; the gather form with reversed indices as well."""
    n = 70001
    c = ft.size_class(n)
    base = ft.tune_table()[c]
    s = ft.random_scalars(n, 4300, 0)
    g, o = ft.GeArray.from_scalars(s), O.Array.from_scalars(s)
    off = np.array([0, 0, n, n], dtype=np.int64)
    want = O.segmented_sum(off, o).to_bytes()
    cols = np.arange(n - 1, -1, -1, dtype=np.int32)
    want_g = O.segmented_sum(off, o, idx=cols).to_bytes()
    for ch1, lr in itertools.product((4, 5, 6, 7), (1, 2, 3, 4)):
        ft.tune_set(c, ch1, lr, *base[2:])
        assert np.array_equal(ft.segmented_sum(off, g).to_bytes_folded(), want), (ch1, lr)
        assert np.array_equal(ft.gather_sum(off, cols, g).to_bytes_folded(), want_g), ("gather", ch1, lr)
    ft.tune_set(c, *base)


@pytest.mark.parametrize("n", [1, 63, 5000, 65536, 65537, 70000])
def test_every_scan_schedule(n):
    c = ft.size_class(n)
    base = ft.tune_table()[c]
    s = ft.random_scalars(n, 4200 + n, 0)
    o_inc, o_exc = O.Array.from_scalars(s), O.Array.from_scalars(s)
    o_inc.scan()
    o_exc.scan(exclusive=True)
    for ch in (4, 5, 6, 7):
        ft.tune_set(c, base[0], base[1], ch, *base[3:])
        g = ft.GeArray.from_scalars(s)
        g.scan()
        same(g, o_inc)
        g = ft.GeArray.from_scalars(s)
        g.prescan()
        same(g, o_exc)
    ft.tune_set(c, *base)


def test_scan_upper_chunk_levels_every_chunk_size():
    """n = 2^23 + 5 points k_i B with small k_i: the first chunk level holds
    more than 2^16 sums for every first-level chunk size, so a second chunk
    level (16 per lane) runs below the block scan; inclusive and exclusive
    outputs at chunk, wave and level edges and at random positions equal the
    fixed-base multiples of the prefix sums of the k_i."""
    n = (1 << 23) + 5
    c = ft.size_class(n)
    base = ft.tune_table()[c]
    k = (np.arange(n, dtype=np.uint64) * np.uint64(40503) + np.uint64(7)) % np.uint64(1 << 20)
    inc = np.cumsum(k)
    exc = inc - k
    edges = [0, 1, 15, 16, 63, 64, 1023, 1024, 4095, 4096, 65535, 65536, (1 << 18) - 1, 1 << 18,
             (1 << 20) + 1, (1 << 22) - 1, 1 << 22, (1 << 23) - 1, 1 << 23, n - 1]
    pick = np.array(sorted(set(edges) | set(np.random.default_rng(5).integers(0, n, 300).tolist())))
    want_inc = ft.GeArray.from_small_scalars(inc[pick]).to_bytes_folded().reshape(-1, 32)
    want_exc = ft.GeArray.from_small_scalars(exc[pick]).to_bytes_folded().reshape(-1, 32)
    for ch in (4, 5, 6, 7):
        ft.tune_set(c, base[0], base[1], ch, *base[3:])
        for exclusive, want in ((False, want_inc), (True, want_exc)):
            g = ft.GeArray.from_small_scalars(k)
            g.prescan() if exclusive else g.scan()
            got = g.get_subset(pick.astype(np.uint64)).to_bytes_folded().reshape(-1, 32)
            bad = np.nonzero((got != want).any(axis=1))[0]
            assert len(bad) == 0, (ch, exclusive, pick[bad[:5]].tolist())
    ft.tune_set(c, *base)


@pytest.mark.parametrize("n", [1, 63, 5000, 8192, 8193, 65537, 70000])
def test_lookback_scan_small(n):
    """the single-pass look-back scan (8,192-point tiles): one tile, a tile
    boundary, a few tiles; inclusive and exclusive against the oracle"""
    c = ft.size_class(n)
    before = ft.scan_lookback_table()[c]
    s = ft.random_scalars(n, 4250 + n, 0)
    o_inc, o_exc = O.Array.from_scalars(s), O.Array.from_scalars(s)
    o_inc.scan()
    o_exc.scan(exclusive=True)
    try:
        ft.tune_set_scan_lookback(c, 1)
        g = ft.GeArray.from_scalars(s)
        g.scan()
        same(g, o_inc)
        g = ft.GeArray.from_scalars(s)
        g.prescan()
        same(g, o_exc)
    finally:
        ft.tune_set_scan_lookback(c, before)


@pytest.mark.parametrize("n", [(1 << 20) + 3, (1 << 22) + 7])
def test_lookback_scan_many_tiles(n):
    """129 and 513 tiles: look-backs that cross more than one window of 64
    tiles; canonical outputs equal to the chunk-level scan's, and the last
    prefix equal to the closed form (sum k) B"""
    c = ft.size_class(n)
    before = ft.scan_lookback_table()[c]
    k = (np.arange(n, dtype=np.uint64) * np.uint64(40503) + np.uint64(7)) % np.uint64(1 << 30)
    try:
        ft.tune_set_scan_lookback(c, 0)
        a = ft.GeArray.from_small_scalars(k)
        a.scan()
        want = a.to_bytes_folded()
        ft.tune_set_scan_lookback(c, 1)
        for rep in range(2):
            g = ft.GeArray.from_small_scalars(k)
            g.scan()
            assert np.array_equal(g.to_bytes_folded(), want), rep
        last = O.scalarmult_base(int.to_bytes(int(k.sum(dtype=np.uint64)) % L, 32, "little"))
        assert bytes(want[-32:]) == last
        g = ft.GeArray.from_small_scalars(k)
        g.prescan()
        assert np.array_equal(g.to_bytes_folded()[32:], want[:-32])
    finally:
        ft.tune_set_scan_lookback(c, before)


@pytest.mark.parametrize("n", [1, 300, 70000])
def test_every_reduce_schedule(n):
    c = ft.size_class(n)
    base = ft.tune_table()[c]
    s = ft.random_scalars(n, 4300 + n, 0)
    want = bytes(O.Array.from_scalars(s).reduce())
    for blocks in (1, 3, 512, 1024, 2048, 4096):
        ft.tune_set(c, *base[:3], blocks, *base[4:])
        g = ft.GeArray.from_scalars(s)
        g.reduce()
        assert bytes(g.to_bytes_folded()) == want, blocks
    ft.tune_set(c, *base)


@pytest.mark.parametrize("n", [1, 1000, 20000])
def test_every_u32_window(n):
    """ciphertext x u32 with radix-8 and radix-16 windows against the oracle,
    multipliers including 0, 1, 7, 8, 2^31, 2^32 - 1 and random"""
    c = ft.size_class(n)
    base = ft.tune_table()[c]
    s1, s2 = ft.random_scalars(n, 4500 + n, 1), ft.random_scalars(n, 4500 + n, 2)
    rng = np.random.default_rng(n)
    a = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    edge = np.array([0, 1, 7, 8, 2**31, 2**32 - 1, 4, 5, 0x924924924 & 0xFFFFFFFF, 0xB6DB6DB6], dtype=np.uint32)
    a[:min(n, len(edge))] = edge[:min(n, len(edge))]
    o1, o2 = O.Array.from_scalars(s1), O.Array.from_scalars(s2)
    o1.mul_u32(a)
    o2.mul_u32(a)
    for w in (3, 4):
        ft.tune_set(c, *base[:4], w, base[5])
        g1, g2 = ft.GeArray.from_scalars(s1), ft.GeArray.from_scalars(s2)
        g1.mul_u32(a, g2)
        same(g1, o1)
        same(g2, o2)
    ft.tune_set(c, *base)


def varbase_edge_scalars():
    L = 2**252 + 27742317777372353535851937790883648493
    ks = [0, 1, 2, 3, 4, 5, 7, 8, 9, L - 1, L, 2**252, 2**253 - 1, 2**255 - 1, 2**255, 2**256 - 1,
          2**255 + 2**252 * 3, 15 * 2**252, 9 * 2**252 + 5, 8 * 2**252 + 2**251, int("8" * 64, 16),
          int("7" * 64, 16), int("9" * 64, 16), int("4" * 64, 16), int("c" * 64, 16), int("f0" * 32, 16)]
    return np.stack([np.frombuffer(int.to_bytes(k, 32, "little"), dtype=np.uint8) for k in ks])


@pytest.mark.parametrize("n", [30, 3000])
def test_every_varbase_window(n):
    """k P (ft_mul) and one k for all P (ft_scale) in radix 8 and radix 16
    against the oracle, incl. scalars >= 2^255 whose top radix-16 digit
    exceeds 8 (the reference's cmov8 then selects the identity)"""
    c = ft.size_class(n)
    base = ft.tune_table()[c]
    s = ft.random_scalars(n, 4600 + n, 1)
    k = ft.random_scalars(n, 4600 + n, 2)
    e = varbase_edge_scalars()
    k[:len(e)] = e
    o = O.Array.from_scalars(s)
    o.mul(k)
    for sc in (e[13], e[15], e[18]):
        os_ = O.Array.from_scalars(s)
        os_.scale(sc)
        for w in (3, 4):
            ft.tune_set(c, *base[:5], w)
            g = ft.GeArray.from_scalars(s)
            g.scale(sc)
            same(g, os_)
    for w in (3, 4):
        ft.tune_set(c, *base[:5], w)
        g = ft.GeArray.from_scalars(s)
        g.mul(k)
        same(g, o)
    ft.tune_set(c, *base)


def test_autotune_keeps_results(tmp_path):
    path = str(tmp_path / "tune.txt")
    table = ft.autotune(max_class=2, path=path)
    assert os.path.exists(path)
    for c, (a, b, sc, r, w, v) in table.items():
        assert 4 <= a <= 7 and 1 <= b <= 4 and 4 <= sc <= 7 and 1 <= r <= 4096 and w in (3, 4) and v in (3, 4)
    ft.tune_load(path)
    assert ft.tune_table() == table
    n = 20000
    s = ft.random_scalars(n, 4400, 0)
    g, o = ft.GeArray.from_scalars(s), O.Array.from_scalars(s)
    off = seg_case(n, 5)
    same(ft.segmented_sum(off, g), O.segmented_sum(off, o))
    g.scan()
    o.scan()
    same(g, o)


@pytest.mark.parametrize("n", [1000, 300000])
def test_every_ct_lanes(n):
    """Encryption and rerandomisation with one lane per ciphertext, with two
    (C1 and C2 on separate waves) and with four (each comb's rows split
    over two waves, the C2 split moving with the message's digit count:
    8-bit and full 64-bit messages) against the oracle; at 300,000 the
    one-lane form is the 512-lane-block kernel"""
    c = ft.size_class(n)
    before = ft.ct_lanes_table()[c]
    x = ft.random_scalars(1, 4800, 0)
    pk = ft.GeArray.from_scalars(x)
    H = bytes(pk.to_bytes_folded())
    key = ft.Key(pk)
    rng = np.random.default_rng(n)
    m = rng.integers(0, 256, size=n, dtype=np.uint64)
    r, rho = ft.random_scalars(n, 4800 + n, 1), ft.random_scalars(n, 4800 + n, 2)
    o1, o2 = O.encrypt(H, m, r)
    q1, q2 = o1.copy(), o2.copy()
    O.rerandomize(q1, q2, H, rho)
    try:
        for lanes in (1, 2, 4):
            ft.tune_set_ct_lanes(c, lanes)
            g1, g2 = ft.encrypt(key, m, r, msg_bits=8)
            same(g1, o1)
            same(g2, o2)
            ft.rerandomize(g1, g2, key, rho)
            same(g1, q1)
            same(g2, q2)
        m64 = rng.integers(0, 2**64, size=n, dtype=np.uint64)
        w1, w2 = O.encrypt(H, m64, r)
        for lanes in (2, 4):
            ft.tune_set_ct_lanes(c, lanes)
            g1, g2 = ft.encrypt(key, m64, r)
            same(g1, w1)
            same(g2, w2)
    finally:
        ft.tune_set_ct_lanes(c, before)
