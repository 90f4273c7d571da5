"""ctypes binding of the CPU oracle (oracle/ftc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.

Parity status of each oracle function (how it is pinned) is listed in
tests/test_oracle.py; every function there is checked against at least one of
RFC 8032 known answers, libsodium's public API (the reference CPU path's
arithmetic dependency, ftcrypto.c:12-14), closed forms in Python big
integers, or the expectations printed by the reference's run_tests.cc.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libftc_oracle.so")
_lib = None

u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


FLAGS_GENERIC = "gcc -O3 -fPIC -fopenmp -std=c11"
FLAGS_NATIVE = FLAGS_GENERIC + " -march=native -mtune=native"
_flags = FLAGS_GENERIC


def use_native():
    """Rebind the oracle to a build for this host's CPU (-march=native, made
    here with make's `native` target) when that builds and loads; returns the
    compile flags of the oracle now in use.  For timing the CPU baseline; the
    generic build (the default) is what the tests check against."""
    global _lib, _flags
    import tempfile
    so = os.path.join(tempfile.mkdtemp(prefix="ftc_oracle_native_"), "libftc_oracle_native.so")
    try:
        subprocess.run(["make", "-s", "-C", _HERE, "native", "NATIVE_OUT=" + so], check=True, capture_output=True)
        _lib = None
        lib(so)
        _flags = FLAGS_NATIVE
    except (subprocess.CalledProcessError, OSError, RuntimeError):
        _lib = None
        lib()
        _flags = FLAGS_GENERIC
    return _flags


def flags():
    return _flags


def lib(path=None):
    global _lib
    if _lib is not None:
        return _lib
    path = path or _SO
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    vp, sz, i32, i64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int64
    sig = {
        "ftco_init": (i32, []),
        "ftco_array_new": (vp, [sz]),
        "ftco_array_free": (None, [vp]),
        "ftco_array_len": (sz, [vp]),
        "ftco_array_copy": (vp, [vp]),
        "ftco_array_from_bytes": (i64, [vp, vp, i32]),
        "ftco_array_to_bytes": (None, [vp, vp, i32]),
        "ftco_from_scalars": (None, [vp, vp, i32]),
        "ftco_from_scalars_varbase": (None, [vp, vp, i32]),
        "ftco_from_small_scalars": (None, [vp, vp, i32]),
        "ftco_add": (None, [vp, vp, i32]),
        "ftco_sub": (None, [vp, vp, i32]),
        "ftco_neg": (None, [vp, i32]),
        "ftco_mul": (None, [vp, vp, i32]),
        "ftco_scale": (None, [vp, vp, i32]),
        "ftco_mul_u32": (None, [vp, vp, i32]),
        "ftco_reduce": (None, [vp, vp, i32]),
        "ftco_scan": (None, [vp, i32, i32]),
        "ftco_segmented_sum": (i32, [vp, vp, sz, vp, vp, i32]),
        "ftco_reduce_isum": (i32, [vp, vp, sz, vp]),
        "ftco_encrypt": (i32, [vp, vp, vp, vp, vp, i32]),
        "ftco_rerandomize": (i32, [vp, vp, vp, vp, i32]),
        "ftco_scalarmult_base_bytes": (i32, [vp, vp]),
        "ftco_scalarmult_bytes": (i32, [vp, vp, vp]),
        "ftco_add_bytes": (i32, [vp, vp, vp]),
        "ftco_decode_roundtrip": (i32, [vp, vp]),
        "ftco_scalar_reduce64": (None, [vp, vp, sz, i32]),
        "ftco_fe_mul_bytes": (None, [vp, vp, vp]),
        "ftco_fe_add_bytes": (None, [vp, vp, vp]),
        "ftco_fe_sub_bytes": (None, [vp, vp, vp]),
        "ftco_fe_invert_bytes": (None, [vp, vp]),
        "ftco_fe_pow22523_bytes": (None, [vp, vp]),
        "ftco_fe_tobytes_roundtrip": (None, [vp, vp]),
        "ftco_constants": (None, [vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.ftco_init() != 0:
        raise RuntimeError("oracle initialisation failed")
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def threads():
    return min(16, os.cpu_count() or 1)


class Array:
    """An oracle-side point array (extended coordinates in host RAM)."""

    def __init__(self, n=0, handle=None):
        self.L = lib()
        self.h = handle if handle is not None else self.L.ftco_array_new(n)
        if not self.h:
            raise MemoryError("oracle array allocation failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ftco_array_free(self.h)
            self.h = None

    def __len__(self):
        return self.L.ftco_array_len(self.h)

    def copy(self):
        return Array(handle=self.L.ftco_array_copy(self.h))

    @classmethod
    def from_bytes(cls, mem, nt=None):
        mem = np.ascontiguousarray(np.frombuffer(bytes(mem), dtype=np.uint8) if not isinstance(mem, np.ndarray) else mem, dtype=np.uint8)
        a = cls(len(mem) // 32)
        bad = a.L.ftco_array_from_bytes(a.h, _ptr(mem), nt or threads())
        return a, bad

    def to_bytes(self, nt=None):
        out = np.zeros(len(self) * 32, dtype=np.uint8)
        self.L.ftco_array_to_bytes(_ptr(out), self.h, nt or threads())
        return out

    @classmethod
    def from_scalars(cls, scalars, nt=None, varbase=False):
        scalars = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1)
        a = cls(len(scalars) // 32)
        f = a.L.ftco_from_scalars_varbase if varbase else a.L.ftco_from_scalars
        f(a.h, _ptr(scalars), nt or threads())
        return a

    @classmethod
    def from_small_scalars(cls, small, nt=None):
        small = np.ascontiguousarray(small, dtype=np.uint64)
        a = cls(len(small))
        a.L.ftco_from_small_scalars(a.h, _ptr(small), nt or threads())
        return a

    def add(self, other, nt=None):
        self.L.ftco_add(self.h, other.h, nt or threads())

    def sub(self, other, nt=None):
        self.L.ftco_sub(self.h, other.h, nt or threads())

    def neg(self, nt=None):
        self.L.ftco_neg(self.h, nt or threads())

    def mul(self, scalars, nt=None):
        scalars = np.ascontiguousarray(scalars, dtype=np.uint8)
        self.L.ftco_mul(self.h, _ptr(scalars), nt or threads())

    def scale(self, scalar, nt=None):
        scalar = np.ascontiguousarray(np.frombuffer(bytes(scalar), dtype=np.uint8))
        self.L.ftco_scale(self.h, _ptr(scalar), nt or threads())

    def mul_u32(self, k, nt=None):
        k = np.ascontiguousarray(k, dtype=np.uint32)
        self.L.ftco_mul_u32(self.h, _ptr(k), nt or threads())

    def reduce(self, nt=None):
        out = np.zeros(32, dtype=np.uint8)
        self.L.ftco_reduce(_ptr(out), self.h, nt or threads())
        return out

    def scan(self, exclusive=False, nt=None):
        self.L.ftco_scan(self.h, int(exclusive), nt or threads())

    def reduce_isum(self, idxs, src):
        idxs = np.ascontiguousarray(idxs, dtype=np.uint64)
        return self.L.ftco_reduce_isum(self.h, _ptr(idxs), len(idxs), src.h)


def segmented_sum(offsets, src, idx=None, nt=None):
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    nseg = len(offsets) - 1
    out = Array(nseg)
    idx_arr = None if idx is None else np.ascontiguousarray(idx, dtype=np.int32)
    rc = lib().ftco_segmented_sum(out.h, _ptr(offsets), nseg, _ptr(idx_arr), src.h, nt or threads())
    if rc != 0:
        raise IndexError("segment offsets or gather indices out of bounds")
    return out


def encrypt(H_enc, m, r, nt=None):
    m = np.ascontiguousarray(m, dtype=np.uint64)
    r = np.ascontiguousarray(r, dtype=np.uint8)
    H = np.ascontiguousarray(np.frombuffer(bytes(H_enc), dtype=np.uint8))
    c1, c2 = Array(len(m)), Array(len(m))
    if lib().ftco_encrypt(c1.h, c2.h, _ptr(H), _ptr(m), _ptr(r), nt or threads()) != 0:
        raise ValueError("invalid public key encoding")
    return c1, c2


def rerandomize(c1, c2, H_enc, r, nt=None):
    r = np.ascontiguousarray(r, dtype=np.uint8)
    H = np.ascontiguousarray(np.frombuffer(bytes(H_enc), dtype=np.uint8))
    if lib().ftco_rerandomize(c1.h, c2.h, _ptr(H), _ptr(r), nt or threads()) != 0:
        raise ValueError("invalid public key encoding")


def scalar_reduce64(wide, nt=None):
    """Reduce 64-byte little-endian integers modulo the group order l."""
    wide = np.ascontiguousarray(wide, dtype=np.uint8).reshape(-1)
    n = len(wide) // 64
    out = np.zeros(n * 32, dtype=np.uint8)
    lib().ftco_scalar_reduce64(_ptr(out), _ptr(wide), n, nt or threads())
    return out.reshape(n, 32)


def _b32(x):
    return np.frombuffer(bytes(x), dtype=np.uint8).copy() if not isinstance(x, np.ndarray) else np.ascontiguousarray(x, dtype=np.uint8)


def scalarmult_base(k):
    out = np.zeros(32, dtype=np.uint8)
    lib().ftco_scalarmult_base_bytes(_ptr(out), _ptr(_b32(k)))
    return bytes(out)


def scalarmult(k, P):
    out = np.zeros(32, dtype=np.uint8)
    if lib().ftco_scalarmult_bytes(_ptr(out), _ptr(_b32(k)), _ptr(_b32(P))) != 0:
        raise ValueError("invalid point")
    return bytes(out)


def point_add(P, Q):
    out = np.zeros(32, dtype=np.uint8)
    if lib().ftco_add_bytes(_ptr(out), _ptr(_b32(P)), _ptr(_b32(Q))) != 0:
        raise ValueError("invalid point")
    return bytes(out)


def decode_roundtrip(P):
    out = np.zeros(32, dtype=np.uint8)
    rc = lib().ftco_decode_roundtrip(_ptr(out), _ptr(_b32(P)))
    return rc, bytes(out)


def fe_op(name, *args):
    out = np.zeros(32, dtype=np.uint8)
    getattr(lib(), "ftco_fe_" + name)(_ptr(out), *[_ptr(_b32(a)) for a in args])
    return bytes(out)


def constants():
    bufs = [np.zeros(32, dtype=np.uint8) for _ in range(4)]
    lib().ftco_constants(*[_ptr(b) for b in bufs])
    return [bytes(b) for b in bufs]
