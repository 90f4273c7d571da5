"""ft_reduce_isum (res[idx[i]] += summand[i], ftcrypto.cu:1150-1217) with its
keys sorted on the device, and the device-key form ftx_reduce_isum_dev.

Checks: the oracle's reduce_isum (byte-exact encodings) at sizes it finishes
quickly; linearity at larger sizes (summands k_i B, so res[j] = (r_j +
sum_{idx[i]=j} k_i) B, checked against the library's fixed-base product of the
per-key sums, itself oracle-checked elsewhere); the exact sum order (raw
extended coordinates equal to a gather-sum over the host's stable CSR order of
the keys, so equal keys are added in increasing i); keys >= len(res) ignored
by the device form; capture in a graph replayed with eager work between.
"""
import ctypes

import numpy as np
import pytest

import libftcrypto_amd as ft
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    ft.init()


class DevBuf:
    def __init__(self, arr=None, nbytes=None):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.p = ctypes.c_void_p()
        size = nbytes if arr is None else np.ascontiguousarray(arr).nbytes
        assert self.hip.hipMalloc(ctypes.byref(self.p), max(1, size)) == 0
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            assert self.hip.hipMemcpy(self.p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def __del__(self):
        self.hip.hipFree(self.p)


def stable_csr(keys, m):
    order = np.argsort(keys, kind="stable").astype(np.int32)
    off = np.concatenate([[0], np.cumsum(np.bincount(keys.astype(np.int64), minlength=m))]).astype(np.int64)
    return off, order


def key_cases():
    rng = np.random.default_rng(5)
    return {
        # (m, keys): one, two and three radix passes, ragged block tails,
        # a single hot key, descending keys, a key range with gaps
        "m1": (1, np.zeros(3000, dtype=np.uint64)),
        "m256_hot": (256, np.full(5000, 200, dtype=np.uint64)),
        "m300": (300, rng.integers(0, 300, size=2049).astype(np.uint64)),
        "m70000": (70000, rng.integers(0, 70000, size=9000).astype(np.uint64)),
        "desc": (4096, (4095 - np.arange(4096) % 4096).astype(np.uint64)),
        "gaps": (1 << 20, (rng.integers(0, 50, size=6000) * 20011).astype(np.uint64)),
        "n1": (17, np.array([16], dtype=np.uint64)),
    }


@pytest.mark.parametrize("case", list(key_cases()))
def test_isum_oracle(case):
    m, keys = key_cases()[case]
    n = len(keys)
    s = ft.random_scalars(n, 31, 1)
    if m <= 70000:
        r = ft.random_scalars(m, 31, 0)
        g, o = ft.GeArray.from_scalars(r), O.Array.from_scalars(r)
    else:
        r = np.random.default_rng(m).integers(0, 1 << 62, size=m, dtype=np.uint64)
        g, o = ft.GeArray.from_small_scalars(r), O.Array.from_small_scalars(r)
    g.reduce_isum(keys, ft.GeArray.from_scalars(s))
    assert o.reduce_isum(keys, O.Array.from_scalars(s)) == 0
    assert np.array_equal(g.to_bytes_folded(), o.to_bytes()), case


@pytest.mark.parametrize("case", list(key_cases()))
def test_isum_stable_order(case):
    """Raw coordinates equal those of res + gather_sum over the host's
    stable CSR order: the device sort adds equal keys in increasing i."""
    m, keys = key_cases()[case]
    n = len(keys)
    src = ft.GeArray.from_scalars(ft.random_scalars(n, 32, 1))
    if m <= 70000:
        r = ft.random_scalars(m, 32, 0)
        res0 = lambda: ft.GeArray.from_scalars(r)
    else:                                   # large m (three radix passes): cheaper small-scalar res
        r = np.random.default_rng(m + 32).integers(0, 1 << 62, size=m, dtype=np.uint64)
        res0 = lambda: ft.GeArray.from_small_scalars(r)
    a = res0()
    a.reduce_isum(keys, src)
    off, order = stable_csr(keys, m)
    b = res0()
    b.add(ft.gather_sum(off, order, src))
    assert np.array_equal(a.to_bytes(), b.to_bytes()), case
    c = res0()
    c.reduce_isum(keys, src)
    assert np.array_equal(a.to_bytes(), c.to_bytes()), case


@pytest.mark.parametrize("n,m", [(1 << 20, 1 << 20), (1 << 21, 1000), (300000, 1 << 24)])
def test_isum_linearity_large(n, m):
    rng = np.random.default_rng(n + m)
    keys = rng.integers(0, m, size=n).astype(np.uint64)
    keys[:4096] = 7                          # one long run
    k = rng.integers(0, 1 << 30, size=n).astype(np.uint64)
    r0 = rng.integers(0, 1 << 30, size=m).astype(np.uint64)
    res = ft.GeArray.from_small_scalars(r0)
    res.reduce_isum(keys, ft.GeArray.from_small_scalars(k))
    want = r0.copy()
    np.add.at(want, keys.astype(np.int64), k)
    assert np.array_equal(res.to_bytes_folded(), ft.GeArray.from_small_scalars(want).to_bytes_folded())


def test_isum_dev_ignores_out_of_range_keys_and_replays_in_graph():
    Lb = ft.lib()
    m, n = 5000, 40000
    rng = np.random.default_rng(9)
    keys = rng.integers(0, m, size=n).astype(np.uint64)
    bad = rng.random(n) < 0.1
    keys[bad] = m + rng.integers(0, 1 << 40, size=int(bad.sum())).astype(np.uint64)
    keys[5] = np.uint64(2**64 - 1)
    k = rng.integers(0, 1 << 30, size=n).astype(np.uint64)
    r0 = rng.integers(0, 1 << 30, size=m).astype(np.uint64)
    summand = ft.GeArray.from_small_scalars(k)
    valid = keys < m
    add = np.zeros(m, dtype=np.uint64)
    np.add.at(add, keys[valid].astype(np.int64), k[valid])
    d_keys = DevBuf(keys)
    # library-owned scratch (allocates, synchronises)
    res = ft.GeArray.from_small_scalars(r0)
    ft.check(Lb.ftx_reduce_isum_dev(res.h, d_keys.p, summand.h, None, None, None))
    assert np.array_equal(res.to_bytes_folded(), ft.GeArray.from_small_scalars(r0 + add).to_bytes_folded())
    # caller scratch on a stream, eager then captured
    scratch = ft.GeArray.alloc(Lb.ftx_reduce_isum_scratch_points(n, m))
    work = DevBuf(nbytes=Lb.ftx_reduce_isum_work_bytes(n, m))
    st = ft.Stream()
    res = ft.GeArray.from_small_scalars(r0)
    ft.check(Lb.ftx_reduce_isum_dev(res.h, d_keys.p, summand.h, scratch.h, work.p, st.h))
    st.synchronize()
    assert np.array_equal(res.to_bytes_folded(), ft.GeArray.from_small_scalars(r0 + add).to_bytes_folded())
    with ft.Graph(st) as g:
        ft.check(Lb.ftx_reduce_isum_dev(res.h, d_keys.p, summand.h, scratch.h, work.p, st.h))
    for rep in range(3):
        junk = ft.GeArray.from_small_scalars(np.arange(1000, dtype=np.uint64))   # eager work between replays
        junk.reduce()
        g.launch()
        st.synchronize()
        want = ft.GeArray.from_small_scalars(r0 + (rep + 2) * add).to_bytes_folded()
        assert np.array_equal(res.to_bytes_folded(), want), rep
    # too small a caller scratch is an error, not a fault
    small = ft.GeArray.alloc(10)
    with pytest.raises(ft.FtError):
        ft.check(Lb.ftx_reduce_isum_dev(res.h, d_keys.p, summand.h, small.h, work.p, st.h))


def test_isum_out_of_range_key_leaves_res_unchanged():
    """ft_reduce_isum checks its keys on the device after the copy and before
    adding anything: one key >= len(res) (the last, or 2^64 - 1) is
    FT_ERR_OUT_OF_BOUNDS and res keeps its bytes; an empty res rejects any
    key."""
    rng = np.random.default_rng(3)
    m, n = 1000, 50000
    res = ft.GeArray.from_scalars(ft.random_scalars(m, 41, 0))
    before = res.to_bytes()
    src = ft.GeArray.from_small_scalars(rng.integers(0, 1 << 30, size=n, dtype=np.uint64))
    for bad in (m, 2**64 - 1):
        keys = rng.integers(0, m, size=n).astype(np.uint64)
        keys[-1] = bad
        with pytest.raises(ft.FtError) as e:
            res.reduce_isum(keys, src)
        assert e.value.code == ft.FT_ERR_OUT_OF_BOUNDS
        assert np.array_equal(res.to_bytes(), before)
    empty = ft.GeArray.alloc(0)
    with pytest.raises(ft.FtError) as e:
        empty.reduce_isum(np.zeros(n, dtype=np.uint64), src)
    assert e.value.code == ft.FT_ERR_OUT_OF_BOUNDS
