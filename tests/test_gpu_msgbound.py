"""Message bounds of ElGamal encryption.

msg_bits is a public bound on the bit length of every message (it fixes the
number of message windows, so an encryption's time depends on it and not on
the messages).  0 or 64 means any uint64; a bound 0 < msg_bits < 64 is
checked: on the host by ftx_encrypt and the serving loop (FT_ERR_OUT_OF_BOUNDS,
nothing allocated or queued), on the device by ftx_encrypt_dev and
ftx_encrypt_batch_dev (a fail word; the encryption then writes nothing).
Every launch form of the encryption kernels is covered: one and two lanes per
ciphertext, 256- and 512-lane blocks, and the batched kernels."""
import ctypes

import numpy as np
import pytest

import libftcrypto_amd as ft
from libftcrypto_amd import serving as SV
from oracle import oracle as O

pytestmark = pytest.mark.gpu


class Dev:
    """A device copy of a numpy array."""

    def __init__(self, arr):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        arr = np.ascontiguousarray(arr)
        self.p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.p), max(1, arr.nbytes)) == 0
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def set(self, arr):
        arr = np.ascontiguousarray(arr)
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def __del__(self):
        self.hip.hipFree(self.p)


@pytest.fixture(scope="module")
def key():
    ft.init()
    x = ft.random_scalars(1, 9191, 0)
    pk = ft.GeArray.from_scalars(x)
    return ft.Key(pk), bytes(pk.to_bytes_folded())


def sample_idx(n, seed):
    rng = np.random.default_rng(seed)
    return np.unique(np.concatenate([np.arange(min(n, 48)), np.arange(max(0, n - 48), n),
                                     rng.integers(0, n, size=min(n, 160))]))


def check_sample(c1, c2, H, m, r, seed):
    """c1, c2 equal the oracle's encryption of (m, r) at sampled indices"""
    idx = sample_idx(len(m), seed)
    o1, o2 = O.encrypt(H, m[idx], r[idx])
    e1 = c1.to_bytes_folded().reshape(-1, 32)[idx].reshape(-1)
    e2 = c2.to_bytes_folded().reshape(-1, 32)[idx].reshape(-1)
    assert np.array_equal(e1, o1.to_bytes())
    assert np.array_equal(e2, o2.to_bytes())


@pytest.fixture(params=[1, 2, 4])
def lanes(request):
    """lanes per ciphertext of every size class, restored afterwards"""
    L = ft.lib()
    old = []
    for c in range(4):
        v = ctypes.c_int()
        ft.check(L.ftx_tune_get_ct_lanes(c, ctypes.byref(v)))
        old.append(v.value)
        ft.check(L.ftx_tune_set_ct_lanes(c, request.param))
    yield request.param
    for c, v in enumerate(old):
        ft.check(L.ftx_tune_set_ct_lanes(c, v))


def test_host_unbounded_messages(key):
    """msg_bits 0 and 64: any uint64, including 12 (above one digit) and 2^64 - 1"""
    k, H = key
    m = np.array([12, 0, 1, 9, 15, 16, 2**40 + 5, 2**63, 2**64 - 1, 255, 0xDEADBEEFCAFEF00D], dtype=np.uint64)
    r = ft.random_scalars(len(m), 9191, 1)
    o1, o2 = O.encrypt(H, m, r)
    for bits in (0, 64):
        c1, c2 = ft.encrypt(k, m, r, msg_bits=bits)
        assert np.array_equal(c1.to_bytes_folded(), o1.to_bytes())
        assert np.array_equal(c2.to_bytes_folded(), o2.to_bytes())


def test_host_bound_refused(key):
    k, _ = key
    r = ft.random_scalars(3, 9191, 2)
    with pytest.raises(ft.FtError) as e:
        ft.encrypt(k, np.array([1, 2**9, 3], dtype=np.uint64), r, msg_bits=8)
    assert e.value.code == ft.FT_ERR_OUT_OF_BOUNDS
    with pytest.raises(ft.FtError) as e:
        ft.encrypt(k, np.array([1, 2, 3], dtype=np.uint64), r, msg_bits=65)
    assert e.value.code == ft.FT_ERR_BAD_SIZE
    c1, c2 = ft.encrypt(k, np.array([255, 0, 128], dtype=np.uint64), r, msg_bits=8)   # at the bound
    assert len(c1) == 3


@pytest.mark.parametrize("n", [37, 70000, 262200])
def test_device_bound(key, lanes, n):
    """a message at or above 2^msg_bits: fail = 1 + the largest such index,
    c1 and c2 untouched; in range: the oracle's encryption, fail = 0"""
    k, H = key
    L = ft.lib()
    rng = np.random.default_rng(n + lanes)
    m = rng.integers(0, 256, size=n, dtype=np.uint64)
    r = ft.random_scalars(n, 9191, 3 + n)
    bad = m.copy()
    bad[n // 3] = 2**9
    bad[n - 2] = 300
    dm, dr = Dev(bad), Dev(r)
    fail = ft.DeviceWord()
    c1 = ft.GeArray.from_small_scalars(np.arange(1, n + 1, dtype=np.uint64))
    c2 = ft.GeArray.from_small_scalars(np.arange(2, n + 2, dtype=np.uint64))
    b1, b2 = c1.to_bytes_folded(), c2.to_bytes_folded()
    ft.check(L.ftx_encrypt_dev(c1.h, c2.h, k.h, dm.p, dr.p, 8, fail.p, None))
    ft.check(L.ftx_synchronize(None))
    assert fail.read() == n - 1
    assert np.array_equal(c1.to_bytes_folded(), b1)
    assert np.array_equal(c2.to_bytes_folded(), b2)
    # a bound without a fail word is refused before anything is enqueued
    assert L.ftx_encrypt_dev(c1.h, c2.h, k.h, dm.p, dr.p, 8, None, None) == ft.FT_ERR_BAD_SIZE
    dm.set(m)
    ft.check(L.ftx_encrypt_dev(c1.h, c2.h, k.h, dm.p, dr.p, 8, fail.p, None))
    ft.check(L.ftx_synchronize(None))
    assert fail.read() == 0
    check_sample(c1, c2, H, m, r, n)


def test_device_unbounded(key, lanes):
    """msg_bits 0 on the device: full 64-bit messages; a given fail word is cleared"""
    k, H = key
    L = ft.lib()
    n = 300
    rng = np.random.default_rng(5)
    m = rng.integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    m[:4] = [12, 2**64 - 1, 0, 9]
    r = ft.random_scalars(n, 9191, 4)
    dm, dr = Dev(m), Dev(r)
    fail = ft.DeviceWord()
    fail.write(77)
    c1, c2 = ft.GeArray.alloc(n), ft.GeArray.alloc(n)
    ft.check(L.ftx_encrypt_dev(c1.h, c2.h, k.h, dm.p, dr.p, 0, fail.p, None))
    ft.check(L.ftx_synchronize(None))
    assert fail.read() == 0
    o1, o2 = O.encrypt(H, m, r)
    assert np.array_equal(c1.to_bytes_folded(), o1.to_bytes())
    assert np.array_equal(c2.to_bytes_folded(), o2.to_bytes())
    d1, d2 = ft.GeArray.alloc(n), ft.GeArray.alloc(n)
    ft.check(L.ftx_encrypt_dev(d1.h, d2.h, k.h, dm.p, dr.p, 64, None, None))
    ft.check(L.ftx_synchronize(None))
    assert np.array_equal(d2.to_bytes_folded(), o2.to_bytes())


@pytest.mark.parametrize("sizes", [[10, 33, 5], [70000, 1000, 2000]])
def test_batch_bound(key, sizes):
    """one job with a message above the bound: its bit in the fail word, no
    job written; then in range, every job equal to the oracle"""
    k, H = key
    rng = np.random.default_rng(sum(sizes))
    js = []
    for q, n in enumerate(sizes):
        m = rng.integers(0, 16, size=n, dtype=np.uint64)
        r = ft.random_scalars(n, 9191, 100 + q)
        c1 = ft.GeArray.from_small_scalars(np.arange(1, n + 1, dtype=np.uint64))
        c2 = ft.GeArray.from_small_scalars(np.arange(1, n + 1, dtype=np.uint64))
        js.append([m, r, c1, c2, Dev(m), Dev(r)])
    before = [(j[2].to_bytes_folded(), j[3].to_bytes_folded()) for j in js]
    bad = js[1][0].copy()
    bad[-1] = 16
    js[1][4].set(bad)
    jobs = [ft.enc_job(c1, c2, dm.p, dr.p) for _, _, c1, c2, dm, dr in js]
    with pytest.raises(ft.FtError) as e:
        ft.encrypt_batch(k, jobs, 4)
    assert e.value.code == ft.FT_ERR_OUT_OF_BOUNDS
    for j, (b1, b2) in zip(js, before):
        assert np.array_equal(j[2].to_bytes_folded(), b1)
        assert np.array_equal(j[3].to_bytes_folded(), b2)
    js[1][4].set(js[1][0])
    w = ft.encrypt_batch(k, jobs, 4)
    assert w.read() == 0
    for q, (m, r, c1, c2, _, _) in enumerate(js):
        check_sample(c1, c2, H, m, r, q)


def test_batch_fail_bits(key):
    """the fail word of a batch names every offending job by its bit"""
    k, _ = key
    L = ft.lib()
    js = []
    for q in range(5):
        m = np.full(40, 3, dtype=np.uint64)
        if q in (1, 4):
            m[q] = 1 << 20
        r = ft.random_scalars(40, 9191, 200 + q)
        js.append((ft.GeArray.alloc(40), ft.GeArray.alloc(40), Dev(m), Dev(r)))
    arr = (ft.EncJob * 5)(*[ft.enc_job(c1, c2, dm.p, dr.p) for c1, c2, dm, dr in js])
    fail = ft.DeviceWord()
    ft.check(L.ftx_encrypt_batch_dev(arr, 5, k.h, 8, fail.p, None))
    ft.check(L.ftx_synchronize(None))
    assert fail.read() == (1 << 1) | (1 << 4)


def test_served_bound(key):
    """the serving loop refuses a message above its request's bound at
    submission (nothing queued), and serves an odd count of ciphertexts (the
    scalars staged after the messages stay 16-byte aligned)"""
    k, H = key
    srv = SV.Server(k, budget_bytes=64 << 20, n_streams=2)
    r = ft.random_scalars(7, 9191, 300)
    with pytest.raises(ft.FtError) as e:
        srv.encrypt(np.array([1, 2, 1000, 4, 5, 6, 7], dtype=np.uint64), r, 8)
    assert e.value.code == ft.FT_ERR_OUT_OF_BOUNDS
    for n in (7, 1, 33):
        m = np.arange(n, dtype=np.uint64) * np.uint64(37) % np.uint64(256)
        rr = ft.random_scalars(n, 9191, 310 + n)
        req, h = srv.encrypt(m, rr, 8)
        srv.wait(req)
        a, b = srv.read(h, n)
        o1, o2 = O.encrypt(H, m, rr)
        assert np.array_equal(a, o1.to_bytes())
        assert np.array_equal(b, o2.to_bytes())
        req, h2 = srv.encrypt(m | np.uint64(1 << 40), rr, 0)      # unbounded
        srv.wait(req)
        a, b = srv.read(h2, n)
        o1, o2 = O.encrypt(H, m | np.uint64(1 << 40), rr)
        assert np.array_equal(b, o2.to_bytes())
    srv.close()
