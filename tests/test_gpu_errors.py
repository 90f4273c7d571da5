"""Error paths of the C ABI that a caller can force from the host
(ftcrypto.h:51-63 codes): undersized output buffers, size mismatches, bad
byte lengths, out-of-range positions and device memory exhaustion -- each
returns its code, leaves its outputs untouched, and the library keeps
working afterwards."""
import ctypes

import numpy as np
import pytest

import libftcrypto_amd as ft

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    ft.init()


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_undersized_output_buffers():
    L = ft.lib()
    g = ft.GeArray.from_small_scalars(np.arange(1, 11, dtype=np.uint64))
    for fn, per in ((L.ft_array_to_bytes_folded, 32), (L.ft_array_to_bytes_affine, 64), (L.ft_array_to_bytes, 160)):
        buf = np.full(per * 10 - 1, 0xAB, dtype=np.uint8)
        assert fn(_p(buf), len(buf), g.h) == ft.FT_ERR_INSUFFICIENT_SPACE
        assert (buf == 0xAB).all()                      # nothing written
        ok = np.zeros(per * 10 + 7, dtype=np.uint8)     # larger than needed is fine
        assert fn(_p(ok), len(ok), g.h) == 0
    out = ft.GeArray.alloc(3)
    off = np.array([0, 2, 5, 7, 10], dtype=np.int64)
    d_off = ctypes.c_void_p()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    assert hip.hipMalloc(ctypes.byref(d_off), off.nbytes) == 0
    assert hip.hipMemcpy(d_off, _p(off), off.nbytes, 1) == 0
    assert L.ftx_segmented_sum_dev(out.h, d_off, 4, 10, None, g.h, None, None) == ft.FT_ERR_INSUFFICIENT_SPACE
    before = g.to_bytes_folded()
    assert L.ftx_segmented_sum_dev(g.h, d_off, 4, 10, None, g.h, None, None) == ft.FT_ERR_BAD_SIZE   # in place
    assert np.array_equal(g.to_bytes_folded(), before)
    hip.hipFree(d_off)


def test_size_and_range_errors():
    L = ft.lib()
    a = ft.GeArray.from_small_scalars(np.arange(1, 11, dtype=np.uint64))
    b = ft.GeArray.from_small_scalars(np.arange(1, 10, dtype=np.uint64))
    before = a.to_bytes_folded()
    assert L.ft_add(a.h, b.h) == ft.FT_ERR_SIZE_MISMATCH
    assert L.ft_sub(a.h, b.h) == ft.FT_ERR_SIZE_MISMATCH
    res = np.zeros(10, dtype=np.int64)
    assert L.ft_equal(_p(res), a.h, b.h) == ft.FT_ERR_SIZE_MISMATCH
    cond = np.zeros(9, dtype=np.int64)
    assert L.ft_array_mux(a.h, _p(cond), b.h) == ft.FT_ERR_SIZE_MISMATCH
    sc = np.zeros(9 * 32, dtype=np.uint8)
    assert L.ft_mul(a.h, _p(sc), 9) == ft.FT_ERR_SIZE_MISMATCH
    assert np.array_equal(a.to_bytes_folded(), before)  # every refused call left a alone
    out = ctypes.c_void_p()
    assert L.ft_array_get_range(ctypes.byref(out), 0, 5, 0, a.h) == ft.FT_ERR_BAD_SIZE
    assert L.ft_array_get_range(ctypes.byref(out), 0, 11, 1, a.h) == ft.FT_ERR_OUT_OF_BOUNDS
    assert L.ft_array_delete_range(a.h, 3, 2, 1) == ft.FT_ERR_OUT_OF_BOUNDS
    idx = np.array([0, 10], dtype=np.uint64)
    assert L.ft_array_get_subset(ctypes.byref(out), _p(idx), 2, a.h) == ft.FT_ERR_OUT_OF_BOUNDS
    assert L.ft_array_assign_to_subset(a.h, _p(idx), 2, b.h) in (ft.FT_ERR_OUT_OF_BOUNDS, ft.FT_ERR_SIZE_MISMATCH)
    aff = np.zeros(65, dtype=np.uint8)
    assert L.ft_array_from_bytes_affine(ctypes.byref(out), _p(aff), 65) == ft.FT_ERR_BAD_SIZE
    raw = np.zeros(161, dtype=np.uint8)
    assert L.ft_array_from_bytes(ctypes.byref(out), _p(raw), 161) == ft.FT_ERR_BAD_SIZE
    assert np.array_equal(a.to_bytes_folded(), before)


def test_device_memory_exhaustion_is_reported_and_recovered():
    L = ft.lib()
    out = ctypes.c_void_p()
    # 2^40 points: 128 TiB, far beyond the device; the allocation fails cleanly
    assert L.ft_array_init_scalar(ctypes.byref(out), 1 << 40, None) == ft.FT_ERR_OUT_OF_MEMORY
    assert L.ftx_array_alloc(ctypes.byref(out), 1 << 40) == ft.FT_ERR_OUT_OF_MEMORY
    # and the next calls work
    g = ft.GeArray.from_small_scalars(np.arange(1, 5, dtype=np.uint64))
    g.reduce()
    assert bytes(g.to_bytes_folded()) == bytes(ft.GeArray.from_small_scalars(np.array([10], dtype=np.uint64))
                                               .to_bytes_folded())


def test_ciphertext_halves_must_differ():
    """encrypt / rerandomise / x u32 write both halves at once: the same
    array as both is refused and left unchanged"""
    L = ft.lib()
    x = ft.random_scalars(1, 44, 0)
    key = ft.Key(ft.GeArray.from_scalars(x))
    g = ft.GeArray.from_small_scalars(np.arange(1, 9, dtype=np.uint64))
    before = g.to_bytes_folded()
    k = np.arange(8, dtype=np.uint32)
    r = ft.random_scalars(8, 44, 1)
    m = np.zeros(8, dtype=np.uint64)
    assert L.ftx_mul_u32(g.h, g.h, _p(k), 8) == ft.FT_ERR_BAD_SIZE
    assert L.ftx_rerandomize(g.h, g.h, key.h, _p(r)) == ft.FT_ERR_BAD_SIZE
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    dm, dr = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dm), m.nbytes) == 0 and hip.hipMalloc(ctypes.byref(dr), r.nbytes) == 0
    assert hip.hipMemcpy(dm, _p(m), m.nbytes, 1) == 0 and hip.hipMemcpy(dr, _p(r), r.nbytes, 1) == 0
    assert L.ftx_encrypt_dev(g.h, g.h, key.h, dm, dr, 1, None, None) == ft.FT_ERR_BAD_SIZE
    L.ftx_synchronize(None)
    hip.hipFree(dm)
    hip.hipFree(dr)
    assert np.array_equal(g.to_bytes_folded(), before)
