"""ftx_reduce_noalloc_dev: the reduce that allocates nothing (its block sums
in the library's static per-GPU array), called through the C ABI and
compared with ftx_reduce_dev (caller scratch) on the same inputs -- two
calls in a row, the second on a smaller input, then an empty and a one-point
input, raw coordinates and canonical encodings both equal."""
import ctypes

import numpy as np
import pytest

import libftcrypto_amd as ft

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    ft.init()


def reduce_with_scratch(src, stream):
    L = ft.lib()
    out = ft.GeArray.alloc(1)
    scr = ft.GeArray.alloc(L.ftx_reduce_scratch_points(len(src)))
    ft.check(L.ftx_reduce_dev(out.h, src.h, scr.h, stream.h))
    stream.synchronize()
    return out


def test_noalloc_reduce_matches_scratch_reduce_twice():
    L = ft.lib()
    st = ft.Stream()
    out = ft.GeArray.alloc(1)
    for n, seed in ((1 << 20) + 37, 1), (5000, 2):          # the second call smaller than the first
        src = ft.GeArray.from_scalars(ft.random_scalars(n, 8800, seed))
        ft.check(L.ftx_reduce_noalloc_dev(out.h, src.h, st.h))
        st.synchronize()
        want = reduce_with_scratch(src, st)
        assert np.array_equal(out.to_bytes(), want.to_bytes()), n
        assert np.array_equal(out.to_bytes_folded(), want.to_bytes_folded()), n


def test_noalloc_reduce_small_inputs_and_refusal():
    L = ft.lib()
    st = ft.Stream()
    out = ft.GeArray.alloc(1)
    for n in (1, 0, 255, 257):
        src = ft.GeArray.from_small_scalars(np.arange(1, n + 1, dtype=np.uint64)) if n else ft.GeArray.alloc(0)
        ft.check(L.ftx_reduce_noalloc_dev(out.h, src.h, st.h))
        st.synchronize()
        want = reduce_with_scratch(src, st)
        assert np.array_equal(out.to_bytes_folded(), want.to_bytes_folded()), n
    empty_out = ft.GeArray.alloc(0)
    assert L.ftx_reduce_noalloc_dev(empty_out.h, out.h, st.h) == ft.FT_ERR_INSUFFICIENT_SPACE
