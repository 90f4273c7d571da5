"""Batched encryption (ftx_encrypt_batch_dev): several ciphertext arrays
encrypted under one key in one launch, each equal byte for byte to the
oracle's encryption of the same messages and scalars; both kernels (256- and
512-lane blocks), jobs whose starts are padded to a wave, a captured graph
replayed, and the host-side refusals."""
import ctypes

import numpy as np
import pytest

import libftcrypto_amd as ft
from oracle import oracle as O

pytestmark = pytest.mark.gpu


class DevBuf:
    def __init__(self, arr):
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


@pytest.fixture(scope="module")
def key():
    ft.init()
    x = ft.random_scalars(1, 6060, 0)
    pk = ft.GeArray.from_scalars(x)
    return ft.Key(pk), bytes(pk.to_bytes_folded())


def make_jobs(sizes, seed, bits):
    rng = np.random.default_rng(seed)
    out = []
    for q, n in enumerate(sizes):
        m = rng.integers(0, 1 << bits, size=n, dtype=np.uint64)
        r = ft.random_scalars(n, seed, q)
        c1, c2 = ft.GeArray.alloc(n), ft.GeArray.alloc(n)
        dm, dr = DevBuf(m), DevBuf(r)
        out.append((m, r, c1, c2, dm, dr))
    return out


@pytest.mark.parametrize("sizes", [[10, 33, 5, 1], [1000, 70000, 1, 4097]])
def test_batch_matches_oracle(key, sizes):
    k, H = key
    js = make_jobs(sizes, 61 + len(sizes) + sum(sizes), 8)
    ft.encrypt_batch(k, [ft.enc_job(c1, c2, dm.p, dr.p) for _, _, c1, c2, dm, dr in js], 8)
    for m, r, c1, c2, _, _ in js:
        o1, o2 = O.encrypt(H, m, r)
        assert np.array_equal(c1.to_bytes_folded(), o1.to_bytes())
        assert np.array_equal(c2.to_bytes_folded(), o2.to_bytes())


def test_batch_graph_replay(key):
    k, H = key
    js = make_jobs([3000, 200], 77, 1)
    st = ft.Stream()
    jobs = [ft.enc_job(c1, c2, dm.p, dr.p) for _, _, c1, c2, dm, dr in js]
    with ft.Graph(st) as g:
        ft.encrypt_batch(k, jobs, 1, st)
    for _ in range(2):
        for _, _, c1, c2, _, _ in js:             # overwrite, then replay
            ft.check(ft.lib().ftx_add_dev(c1.h, c2.h, st.h))
        g.launch()
        st.synchronize()
        for m, r, c1, c2, _, _ in js:
            o1, o2 = O.encrypt(H, m, r)
            assert np.array_equal(c1.to_bytes_folded(), o1.to_bytes())
            assert np.array_equal(c2.to_bytes_folded(), o2.to_bytes())


def test_batch_refusals(key):
    k, _ = key
    (m, r, c1, c2, dm, dr), = make_jobs([8], 5, 1)
    short = ft.GeArray.alloc(4)
    Lb = ft.lib()

    fail = ft.DeviceWord()

    def rc(jobs, bits=1):
        arr = (ft.EncJob * max(1, len(jobs)))(*jobs)
        out = Lb.ftx_encrypt_batch_dev(arr, len(jobs), k.h, bits, fail.p, None)
        Lb.ftx_synchronize(None)
        return out

    j = ft.enc_job(c1, c2, dm.p, dr.p)
    BAD_SIZE, SIZE_MISMATCH = 3, 2
    assert rc([j] * (ft.ENC_MAX_JOBS + 1)) == BAD_SIZE
    assert rc([j, j]) == BAD_SIZE                                   # outputs shared
    assert rc([ft.enc_job(c1, c1, dm.p, dr.p)]) == BAD_SIZE         # c1 is c2
    assert rc([ft.enc_job(c1, short, dm.p, dr.p)]) == SIZE_MISMATCH
    assert rc([j], bits=65) == BAD_SIZE
    assert rc([]) == 0
    # a bound below 64 bits needs the fail word
    arr = (ft.EncJob * 1)(j)
    assert Lb.ftx_encrypt_batch_dev(arr, 1, k.h, 8, None, None) == BAD_SIZE
    assert Lb.ftx_encrypt_batch_dev(arr, 1, k.h, 0, None, None) == 0
    Lb.ftx_synchronize(None)
