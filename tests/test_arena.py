"""CPU checks of the serving loop's HBM arena (libftcrypto_amd/csrc/arena.h,
compiled for the host): admission control is its alloc, so it must place
first-fit in 256-byte granules, never overlap live blocks, refuse what does
not fit (back-pressure), keep `used` exact, and coalesce freed neighbours so
that a drained arena is one block again.  A Python model of the same policy
is the reference for every step of random alloc/release sequences."""
import ctypes
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ALIGN = 256


@pytest.fixture(scope="module")
def A(tmp_path_factory):
    out = tmp_path_factory.mktemp("arena") / "arena_host.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(out),
                    os.path.join(HERE, "native", "arena_host.cc")], check=True)
    L = ctypes.CDLL(str(out))
    sz = ctypes.c_size_t
    L.arena_new.restype = ctypes.c_void_p
    L.arena_new.argtypes = [sz]
    L.arena_delete.argtypes = [ctypes.c_void_p]
    L.arena_alloc.restype = ctypes.c_int
    L.arena_alloc.argtypes = [ctypes.c_void_p, sz, ctypes.POINTER(sz)]
    L.arena_release.argtypes = [ctypes.c_void_p, sz, sz]
    L.arena_used.restype = sz
    L.arena_used.argtypes = [ctypes.c_void_p]
    L.arena_nfree.restype = sz
    L.arena_nfree.argtypes = [ctypes.c_void_p]
    L.arena_free_block.argtypes = [ctypes.c_void_p, sz, ctypes.POINTER(sz), ctypes.POINTER(sz)]
    return L


class Model:
    """First fit over free blocks sorted by offset, neighbours merged on free."""

    def __init__(self, size):
        self.free = [(0, size)]
        self.used = 0

    def alloc(self, n):
        n = -(-n // ALIGN) * ALIGN
        for k, (o, ln) in enumerate(self.free):
            if ln >= n:
                self.free[k:k + 1] = [(o + n, ln - n)] if ln > n else []
                self.used += n
                return o
        return None

    def release(self, o, n):
        n = -(-n // ALIGN) * ALIGN
        self.used -= n
        blocks = sorted(self.free + [(o, n)])
        merged = []
        for b in blocks:
            if merged and merged[-1][0] + merged[-1][1] == b[0]:
                merged[-1] = (merged[-1][0], merged[-1][1] + b[1])
            else:
                merged.append(b)
        self.free = merged


def free_list(L, a):
    out = []
    for i in range(L.arena_nfree(a)):
        o, ln = ctypes.c_size_t(), ctypes.c_size_t()
        L.arena_free_block(a, i, ctypes.byref(o), ctypes.byref(ln))
        out.append((o.value, ln.value))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_random_sequences_match_model(A, seed):
    rnd = random.Random(seed)
    size = rnd.choice([1 << 16, 1 << 20, 6 << 20])
    a = A.arena_new(size)
    m = Model(size)
    live = []
    for _ in range(3000):
        if live and rnd.random() < 0.45:
            o, n = live.pop(rnd.randrange(len(live)))
            A.arena_release(a, o, n)
            m.release(o, n)
        else:
            n = rnd.choice([1, 255, 256, 257, 4096, rnd.randrange(1, size // 8)])
            off = ctypes.c_size_t()
            ok = A.arena_alloc(a, n, ctypes.byref(off))
            want = m.alloc(n)
            assert bool(ok) == (want is not None)
            if ok:
                assert off.value == want and off.value % ALIGN == 0
                rn = -(-n // ALIGN) * ALIGN
                assert off.value + rn <= size
                for o2, n2 in live:                 # no overlap with a live block
                    r2 = -(-n2 // ALIGN) * ALIGN
                    assert off.value + rn <= o2 or o2 + r2 <= off.value
                live.append((off.value, n))
        assert A.arena_used(a) == m.used
        assert free_list(A, a) == m.free
    for o, n in live:
        A.arena_release(a, o, n)
    assert A.arena_used(a) == 0 and free_list(A, a) == [(0, size)]
    A.arena_delete(a)


def test_refuses_what_does_not_fit_then_admits_after_release(A):
    a = A.arena_new(4 * ALIGN)
    off = ctypes.c_size_t()
    assert A.arena_alloc(a, 3 * ALIGN, ctypes.byref(off)) and off.value == 0
    assert not A.arena_alloc(a, 2 * ALIGN, ctypes.byref(off))        # back-pressure
    assert A.arena_alloc(a, 1, ctypes.byref(off)) and off.value == 3 * ALIGN
    A.arena_release(a, 0, 3 * ALIGN)
    assert A.arena_alloc(a, 2 * ALIGN, ctypes.byref(off)) and off.value == 0
    assert not A.arena_alloc(a, 5 * ALIGN, ctypes.byref(off))        # larger than the budget
    A.arena_delete(a)
