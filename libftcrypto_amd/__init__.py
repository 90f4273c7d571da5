"""Python binding of libftcrypto.so, the gfx950 batched Ed25519 ElGamal library.

The library's product is its C ABI (include/ftcrypto.h, include/ftcrypto_ext.h);
this module binds it with ctypes for tests, the benchmark and Python callers.
`GeArray` mirrors the reference's ft_ge25519_array API (ftcrypto.h:78-435):
same operations, same argument meaning, and FtError raised with the
reference's error code where the C call returns one.

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback in the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libftcrypto.so")

FT_ERR = {
    0: "FT_ERR_NO_ERROR", 1: "FT_ERR_OUT_OF_MEMORY", 2: "FT_ERR_SIZE_MISMATCH", 3: "FT_ERR_BAD_SIZE",
    4: "FT_ERR_OUT_OF_BOUNDS", 5: "FT_ERR_INVALID_POINT", 6: "FT_ERR_SODIUM_INIT",
    7: "FT_ERR_INSUFFICIENT_SPACE", 8: "FT_ERR_BAD_CIPHERTEXT", 9: "FT_ERR_INTERNAL_LIBRARY_ERROR",
    10: "FT_ERR_INSUFFICIENT_DEVICE_OR_DRIVER",
}
for _k, _v in FT_ERR.items():
    globals()[_v] = _k


class FtError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__("%s (%d)%s" % (FT_ERR.get(code, "?"), code, (": " + what) if what else ""))


_lib = None
_initialised = False

vp, sz, i32, i64, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
_SIGS = {
    # ftcrypto.h
    "ft_error_str": (ctypes.c_char_p, [i32]),
    "ft_crypto_init": (i32, []),
    "ft_crypto_device_memory": (i32, [vp, vp]),
    "ft_array_init_scalar": (i32, [vp, sz, vp]),
    "ft_array_init_point": (i32, [vp, sz, vp, sz]),
    "ft_array_get_subset": (i32, [vp, vp, sz, vp]),
    "ft_array_get_range": (i32, [vp, sz, sz, sz, vp]),
    "ft_array_assign_to_subset": (i32, [vp, vp, sz, vp]),
    "ft_array_assign_zero_to_subset": (i32, [vp, vp, sz]),
    "ft_array_assign_to_range": (i32, [vp, sz, sz, sz, vp]),
    "ft_array_delete_subset": (i32, [vp, vp, sz]),
    "ft_array_delete_range": (i32, [vp, sz, sz, sz]),
    "ft_array_to_bytes_folded": (i32, [vp, sz, vp]),
    "ft_array_to_bytes_affine": (i32, [vp, sz, vp]),
    "ft_array_to_bytes": (i32, [vp, sz, vp]),
    "ft_array_from_bytes_folded": (i32, [vp, vp, vp, sz]),
    "ft_array_from_bytes_affine": (i32, [vp, vp, sz]),
    "ft_array_from_bytes": (i32, [vp, vp, sz]),
    "ft_array_validate": (i32, [vp, vp]),
    "ft_array_from_scalars": (i32, [vp, vp, sz]),
    "ft_array_from_small_scalars": (i32, [vp, vp, sz]),
    "ft_array_free": (None, [vp]),
    "ft_array_get_length": (sz, [vp]),
    "ft_array_set_length": (i32, [vp, sz]),
    "ft_array_mux": (i32, [vp, vp, vp]),
    "ft_array_contains": (i32, [vp, vp, vp]),
    "ft_add": (i32, [vp, vp]),
    "ft_reduce_isum": (i32, [vp, vp, sz, vp]),
    "ft_sub": (i32, [vp, vp]),
    "ft_scale": (None, [vp, vp]),
    "ft_neg": (None, [vp]),
    "ft_mul": (i32, [vp, vp, sz]),
    "ft_index": (None, [vp, vp]),
    "ft_equal": (i32, [vp, vp, vp]),
    "ft_not_equal": (i32, [vp, vp, vp]),
    "ft_array_prescan": (i32, [vp]),
    "ft_array_scan": (i32, [vp]),
    "ft_array_reduce": (i32, [vp]),
    # ftcrypto_ext.h
    "ftx_array_alloc": (i32, [vp, sz]),
    "ftx_array_device_view": (i32, [vp, vp, vp]),
    "ftx_default_stream": (vp, []),
    "ftx_synchronize": (i32, [vp]),
    "ftx_key_create": (i32, [vp, vp, sz]),
    "ftx_key_free": (None, [vp]),
    "ftx_encrypt": (i32, [vp, vp, vp, vp, vp, sz, i32]),
    "ftx_rerandomize": (i32, [vp, vp, vp, vp]),
    "ftx_mul_u32": (i32, [vp, vp, vp, sz]),
    "ftx_segmented_sum": (i32, [vp, vp, sz, vp]),
    "ftx_gather_sum": (i32, [vp, vp, vp, sz, vp]),
    "ftx_from_scalars_dev": (i32, [vp, vp, vp]),
    "ftx_mul_dev": (i32, [vp, vp, vp]),
    "ftx_mul_u32_dev": (i32, [vp, vp, vp, vp]),
    "ftx_add_dev": (i32, [vp, vp, vp]),
    "ftx_encrypt_dev": (i32, [vp, vp, vp, vp, vp, i32, vp, vp]),
    "ftx_rerandomize_dev": (i32, [vp, vp, vp, vp, vp]),
    "ftx_segsum_scratch_points": (sz, [sz]),
    "ftx_segmented_sum_dev": (i32, [vp, vp, sz, sz, vp, vp, vp, vp]),
    "ftx_encrypt_batch_dev": (i32, [vp, sz, vp, i32, vp, vp]),
    "ftx_aggregate_batch_scratch_points": (sz, [vp, sz]),
    "ftx_aggregate_batch_dev": (i32, [vp, sz, vp, vp]),
    "ftx_reduce_isum_scratch_points": (sz, [sz, sz]),
    "ftx_reduce_isum_work_bytes": (sz, [sz, sz]),
    "ftx_reduce_isum_dev": (i32, [vp, vp, vp, vp, vp, vp]),
    "ftx_reduce_scratch_points": (sz, [sz]),
    "ftx_reduce_dev": (i32, [vp, vp, vp, vp]),
    "ftx_reduce_noalloc_dev": (i32, [vp, vp, vp]),
    "ftx_scan_scratch_points": (sz, [sz]),
    "ftx_scan_dev": (i32, [vp, i32, vp, vp]),
    "ftx_gather_scratch_points": (sz, [sz, sz]),
    "ftx_encode_dev": (i32, [vp, vp, vp]),
    "ftx_decode_dev": (i32, [vp, vp, vp, vp]),
    "ftx_stream_create": (i32, [vp]),
    "ftx_stream_destroy": (None, [vp]),
    "ftx_capture_begin": (i32, [vp]),
    "ftx_capture_end": (i32, [vp, vp]),
    "ftx_graph_launch": (i32, [vp, vp]),
    "ftx_graph_free": (None, [vp]),
    "ftx_autotune": (i32, [i32, ctypes.c_char_p]),
    "ftx_tune_save": (i32, [ctypes.c_char_p]),
    "ftx_tune_load": (i32, [ctypes.c_char_p]),
    "ftx_tune_get": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "ftx_tune_set": (i32, [i32, i32, i32, i32, i32, i32, i32]),
    "ftx_tune_size_class": (i32, [sz]),
    "ftx_tune_get_ct_lanes": (i32, [i32, vp]),
    "ftx_tune_set_ct_lanes": (i32, [i32, i32]),
    "ftx_tune_get_scan_lookback": (i32, [i32, vp]),
    "ftx_tune_set_scan_lookback": (i32, [i32, i32]),
    "ftx_random_scalars": (None, [vp, sz, u64, u64]),
    "ftx_scalar_reduce64": (None, [vp, vp, sz]),
}


def lib():
    """Load libftcrypto.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libftcrypto.so is not built; run `python -m libftcrypto_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        raise FtError(rc, what)


def init():
    global _initialised
    if not _initialised:
        check(lib().ft_crypto_init(), "ft_crypto_init")
        _initialised = True


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _u8(x):
    if isinstance(x, (bytes, bytearray)):
        return np.frombuffer(bytes(x), dtype=np.uint8).copy()
    return np.ascontiguousarray(x, dtype=np.uint8).reshape(-1)


class GeArray:
    """Handle of a device-resident array of Ed25519 points (ft_ge25519_array)."""

    def __init__(self, handle):
        self.h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def free(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().ft_array_free(self.h)
            self.h = None

    @staticmethod
    def _make(fn, *args, what=""):
        out = ctypes.c_void_p()
        check(fn(ctypes.byref(out), *args), what)
        return GeArray(out)

    # constructors ---------------------------------------------------
    @classmethod
    def alloc(cls, n):
        init()
        return cls._make(lib().ftx_array_alloc, n, what="ftx_array_alloc")

    @classmethod
    def init_scalar(cls, n, scalar=None):
        init()
        s = None if scalar is None else _u8(scalar)
        return cls._make(lib().ft_array_init_scalar, n, _p(s), what="ft_array_init_scalar")

    @classmethod
    def init_point(cls, n, src, idx):
        init()
        return cls._make(lib().ft_array_init_point, n, src.h, idx, what="ft_array_init_point")

    @classmethod
    def from_scalars(cls, scalars):
        init()
        s = _u8(scalars)
        return cls._make(lib().ft_array_from_scalars, _p(s), len(s) // 32, what="ft_array_from_scalars")

    @classmethod
    def from_small_scalars(cls, small):
        init()
        s = np.ascontiguousarray(small, dtype=np.uint64)
        return cls._make(lib().ft_array_from_small_scalars, _p(s), len(s), what="ft_array_from_small_scalars")

    @classmethod
    def from_bytes_folded(cls, mem):
        init()
        m = _u8(mem)
        res = ctypes.c_int64(0)
        out = ctypes.c_void_p()
        check(lib().ft_array_from_bytes_folded(ctypes.byref(res), ctypes.byref(out), _p(m), len(m)),
              "ft_array_from_bytes_folded")
        return GeArray(out), res.value

    @classmethod
    def from_bytes_affine(cls, mem):
        init()
        m = _u8(mem)
        return cls._make(lib().ft_array_from_bytes_affine, _p(m), len(m), what="ft_array_from_bytes_affine")

    @classmethod
    def from_bytes(cls, mem):
        init()
        m = _u8(mem)
        return cls._make(lib().ft_array_from_bytes, _p(m), len(m), what="ft_array_from_bytes")

    def get_subset(self, idxs):
        i = np.ascontiguousarray(idxs, dtype=np.uint64)
        return GeArray._make(lib().ft_array_get_subset, _p(i), len(i), self.h, what="ft_array_get_subset")

    def get_range(self, start, stop, step):
        return GeArray._make(lib().ft_array_get_range, start, stop, step, self.h, what="ft_array_get_range")

    # properties -----------------------------------------------------
    def __len__(self):
        return lib().ft_array_get_length(self.h)

    def set_length(self, n):
        check(lib().ft_array_set_length(self.h, n), "ft_array_set_length")

    def device_view(self):
        planes, cap = ctypes.c_void_p(), ctypes.c_size_t()
        check(lib().ftx_array_device_view(self.h, ctypes.byref(planes), ctypes.byref(cap)))
        return planes.value, cap.value

    # serialisation --------------------------------------------------
    def to_bytes_folded(self):
        out = np.zeros(len(self) * 32, dtype=np.uint8)
        check(lib().ft_array_to_bytes_folded(_p(out), len(out), self.h), "ft_array_to_bytes_folded")
        return out

    def to_bytes_affine(self):
        out = np.zeros(len(self) * 64, dtype=np.uint8)
        check(lib().ft_array_to_bytes_affine(_p(out), len(out), self.h), "ft_array_to_bytes_affine")
        return out

    def to_bytes(self):
        out = np.zeros(len(self) * 160, dtype=np.uint8)
        check(lib().ft_array_to_bytes(_p(out), len(out), self.h), "ft_array_to_bytes")
        return out

    def validate(self):
        r = ctypes.c_int64(0)
        check(lib().ft_array_validate(ctypes.byref(r), self.h), "ft_array_validate")
        return r.value

    # operations -----------------------------------------------------
    def add(self, other):
        check(lib().ft_add(self.h, other.h), "ft_add")

    def sub(self, other):
        check(lib().ft_sub(self.h, other.h), "ft_sub")

    def neg(self):
        lib().ft_neg(self.h)

    def mul(self, scalars):
        s = _u8(scalars)
        check(lib().ft_mul(self.h, _p(s), len(s) // 32), "ft_mul")

    def scale(self, scalar):
        s = _u8(scalar)
        lib().ft_scale(self.h, _p(s))

    def reduce(self):
        check(lib().ft_array_reduce(self.h), "ft_array_reduce")

    def scan(self):
        check(lib().ft_array_scan(self.h), "ft_array_scan")

    def prescan(self):
        check(lib().ft_array_prescan(self.h), "ft_array_prescan")

    def reduce_isum(self, idxs, summand):
        i = np.ascontiguousarray(idxs, dtype=np.uint64)
        check(lib().ft_reduce_isum(self.h, _p(i), len(i), summand.h), "ft_reduce_isum")

    def assign_to_subset(self, idxs, src):
        i = np.ascontiguousarray(idxs, dtype=np.uint64)
        check(lib().ft_array_assign_to_subset(self.h, _p(i), len(i), src.h), "ft_array_assign_to_subset")

    def assign_zero_to_subset(self, idxs):
        i = np.ascontiguousarray(idxs, dtype=np.uint64)
        check(lib().ft_array_assign_zero_to_subset(self.h, _p(i), len(i)), "ft_array_assign_zero_to_subset")

    def assign_to_range(self, start, stop, step, src):
        check(lib().ft_array_assign_to_range(self.h, start, stop, step, src.h), "ft_array_assign_to_range")

    def delete_subset(self, idxs):
        i = np.ascontiguousarray(idxs, dtype=np.uint64)
        check(lib().ft_array_delete_subset(self.h, _p(i), len(i)), "ft_array_delete_subset")

    def delete_range(self, start, stop, step):
        check(lib().ft_array_delete_range(self.h, start, stop, step), "ft_array_delete_range")

    def mux(self, cond, iffalse):
        c = np.ascontiguousarray(cond, dtype=np.int64)
        check(lib().ft_array_mux(self.h, _p(c), iffalse.h), "ft_array_mux")

    def contains(self, other):
        r = np.zeros(len(self), dtype=np.int64)
        check(lib().ft_array_contains(_p(r), self.h, other.h), "ft_array_contains")
        return r

    def index(self):
        r = np.zeros(len(self), dtype=np.int64)
        lib().ft_index(_p(r), self.h)
        return r

    def equal(self, other):
        r = np.zeros(len(self), dtype=np.int64)
        check(lib().ft_equal(_p(r), self.h, other.h), "ft_equal")
        return r

    def not_equal(self, other):
        r = np.zeros(len(self), dtype=np.int64)
        check(lib().ft_not_equal(_p(r), self.h, other.h), "ft_not_equal")
        return r

    def mul_u32(self, k, other=None):
        kk = np.ascontiguousarray(k, dtype=np.uint32)
        check(lib().ftx_mul_u32(self.h, None if other is None else other.h, _p(kk), len(kk)), "ftx_mul_u32")


class Key:
    """Comb table of a public key H (ftx_key)."""

    def __init__(self, pk, idx=0):
        init()
        k = ctypes.c_void_p()
        check(lib().ftx_key_create(ctypes.byref(k), pk.h, idx), "ftx_key_create")
        self.h = k

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().ftx_key_free(self.h)
            self.h = None


def encrypt(key, m, r, msg_bits=0):
    m = np.ascontiguousarray(m, dtype=np.uint64)
    r = _u8(r)
    c1, c2 = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().ftx_encrypt(ctypes.byref(c1), ctypes.byref(c2), key.h, _p(m), _p(r), len(m), msg_bits),
          "ftx_encrypt")
    return GeArray(c1), GeArray(c2)


def rerandomize(c1, c2, key, r):
    r = _u8(r)
    check(lib().ftx_rerandomize(c1.h, c2.h, key.h, _p(r)), "ftx_rerandomize")


def segmented_sum(offsets, src):
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    out = ctypes.c_void_p()
    check(lib().ftx_segmented_sum(ctypes.byref(out), _p(off), len(off) - 1, src.h), "ftx_segmented_sum")
    return GeArray(out)


ENC_MAX_JOBS = 16


class EncJob(ctypes.Structure):
    """ftx_enc_job: one encryption of ftx_encrypt_batch_dev."""
    _fields_ = [("c1", ctypes.c_void_p), ("c2", ctypes.c_void_p), ("d_m", ctypes.c_void_p),
                ("d_r", ctypes.c_void_p)]


def enc_job(c1, c2, d_m, d_r):
    """One job; it holds references to its arrays (the structure keeps only
    the raw handles)."""
    j = EncJob(c1.h, c2.h, d_m, d_r)
    j._arrays = (c1, c2)
    return j


class DeviceWord:
    """One device uint64: the fail word of the bounded device encryptions
    (ftx_encrypt_dev, ftx_encrypt_batch_dev)."""

    def __init__(self):
        self._hip = ctypes.CDLL("libamdhip64.so")
        self._hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self._hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self._hip.hipFree.argtypes = [ctypes.c_void_p]
        self.p = ctypes.c_void_p()
        if self._hip.hipMalloc(ctypes.byref(self.p), 8) != 0:
            raise FtError(FT_ERR_OUT_OF_MEMORY, "hipMalloc of a device word")

    def write(self, value):
        v = ctypes.c_uint64(value)
        if self._hip.hipMemcpy(self.p, ctypes.byref(v), 8, 1) != 0:
            raise FtError(FT_ERR_INTERNAL_LIBRARY_ERROR, "hipMemcpy of a device word")

    def read(self):
        """The word's value (the stream that wrote it must be synchronised)."""
        v = ctypes.c_uint64()
        if self._hip.hipMemcpy(ctypes.byref(v), self.p, 8, 2) != 0:
            raise FtError(FT_ERR_INTERNAL_LIBRARY_ERROR, "hipMemcpy of a device word")
        return v.value

    def __del__(self):
        if getattr(self, "p", None) is not None and self.p.value:
            self._hip.hipFree(self.p)
            self.p = None


def encrypt_batch(key, jobs, msg_bits, stream=None, fail=None):
    """All jobs (EncJob list) in one launch on `stream` (default: the library
    stream, then synchronised).  A bound 0 < msg_bits < 64 is checked on the
    device into `fail` (a DeviceWord; one is made when none is given): bit j
    set when a message of job j is out of range, and then nothing is written.
    Without a stream, a nonzero fail word raises FT_ERR_OUT_OF_BOUNDS.
    Returns the fail word."""
    arr = (EncJob * max(1, len(jobs)))(*jobs)
    st = stream.h if stream is not None else None
    if fail is None:
        fail = DeviceWord()
    check(lib().ftx_encrypt_batch_dev(arr, len(jobs), key.h, msg_bits, fail.p, st), "ftx_encrypt_batch_dev")
    if stream is None:
        check(lib().ftx_synchronize(None), "ftx_synchronize")
        if fail.read():
            raise FtError(FT_ERR_OUT_OF_BOUNDS, "ftx_encrypt_batch_dev: message above the bound")
    return fail


AGG_SEGSUM, AGG_REDUCE, AGG_SCAN, AGG_XSCAN = 0, 1, 2, 3
AGG_MAX_JOBS = 32


class AggJob(ctypes.Structure):
    """ftx_agg_job: one job of ftx_aggregate_batch_dev."""
    _fields_ = [("kind", ctypes.c_int), ("out", ctypes.c_void_p), ("src", ctypes.c_void_p),
                ("d_offsets", ctypes.c_void_p), ("n_seg", ctypes.c_size_t), ("n_elems", ctypes.c_size_t),
                ("seg_len", ctypes.c_size_t)]


def agg_job(kind, out, src=None, d_offsets=None, n_seg=0, n_elems=0, seg_len=0):
    """One job; it holds references to its arrays, so an array made inline
    (e.g. a get_range) lives as long as the job does, not just until this
    call returns (the structure itself keeps only the raw handles)."""
    j = AggJob(kind, out.h, src.h if src is not None else None, d_offsets, n_seg, n_elems, seg_len)
    j._arrays = (out, src)
    return j


def aggregate_batch(jobs, scratch=None, stream=None):
    """Run the jobs (AggJob list) in one persistent launch on `stream`
    (default: the library stream, then synchronised)."""
    arr = (AggJob * max(1, len(jobs)))(*jobs)
    st = stream.h if stream is not None else None
    check(lib().ftx_aggregate_batch_dev(arr, len(jobs), scratch.h if scratch is not None else None, st),
          "ftx_aggregate_batch_dev")
    if stream is None:
        check(lib().ftx_synchronize(None), "ftx_synchronize")


def aggregate_batch_scratch_points(jobs):
    arr = (AggJob * max(1, len(jobs)))(*jobs)
    return lib().ftx_aggregate_batch_scratch_points(arr, len(jobs))


def gather_sum(row_off, cols, src):
    off = np.ascontiguousarray(row_off, dtype=np.int64)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    out = ctypes.c_void_p()
    check(lib().ftx_gather_sum(ctypes.byref(out), _p(off), _p(c), len(off) - 1, src.h), "ftx_gather_sum")
    return GeArray(out)


def random_scalars(n, seed, stream=0):
    """n scalars uniform mod l from the library's counter-based SplitMix64 stream."""
    out = np.zeros(n * 32, dtype=np.uint8)
    lib().ftx_random_scalars(_p(out), n, seed, stream)
    return out.reshape(n, 32)


def scalar_reduce64(wide):
    wide = np.ascontiguousarray(wide, dtype=np.uint8).reshape(-1)
    n = len(wide) // 64
    out = np.zeros(n * 32, dtype=np.uint8)
    lib().ftx_scalar_reduce64(_p(out), _p(wide), n)
    return out.reshape(n, 32)


def autotune(max_class=3, path=None):
    """Time every reduce / scan / segmented-sum schedule and both window
    widths of the variable-base kernels on the device for size classes
    0..max_class, keep the fastest (and save them to path)."""
    init()
    check(lib().ftx_autotune(max_class, path.encode() if path else None), "ftx_autotune")
    return tune_table()


def tune_table():
    """{class: (seg1, segu, scan, reduce_blocks, u32_window, varbase_window)}
    of the schedules in use."""
    out = {}
    for c in range(4):
        v = [ctypes.c_int() for _ in range(6)]
        check(lib().ftx_tune_get(c, *[ctypes.byref(x) for x in v]), "ftx_tune_get")
        out[c] = tuple(x.value for x in v)
    return out


def tune_set(size_class, seg1, segu, scan, reduce_blocks, u32_window=3, varbase_window=3):
    check(lib().ftx_tune_set(size_class, seg1, segu, scan, reduce_blocks, u32_window, varbase_window),
          "ftx_tune_set")


def ct_lanes_table():
    """{class: lanes per ciphertext of an encryption (1 or 2)}"""
    out = {}
    for c in range(4):
        v = ctypes.c_int()
        check(lib().ftx_tune_get_ct_lanes(c, ctypes.byref(v)), "ftx_tune_get_ct_lanes")
        out[c] = v.value
    return out


def tune_set_ct_lanes(size_class, lanes):
    check(lib().ftx_tune_set_ct_lanes(size_class, lanes), "ftx_tune_set_ct_lanes")


def scan_lookback_table():
    """{class: scan form (0 chunk levels, 1 single-pass look-back)}"""
    out = {}
    for c in range(4):
        v = ctypes.c_int()
        check(lib().ftx_tune_get_scan_lookback(c, ctypes.byref(v)), "ftx_tune_get_scan_lookback")
        out[c] = v.value
    return out


def tune_set_scan_lookback(size_class, lookback):
    check(lib().ftx_tune_set_scan_lookback(size_class, lookback), "ftx_tune_set_scan_lookback")


def tune_load(path):
    check(lib().ftx_tune_load(path.encode()), "ftx_tune_load")


def tune_save(path):
    check(lib().ftx_tune_save(path.encode()), "ftx_tune_save")


def size_class(n):
    return lib().ftx_tune_size_class(n)


def error_str(code):
    return lib().ft_error_str(code).decode()


class Stream:
    """A HIP stream owned by the library (default priority)."""

    def __init__(self):
        init()
        s = ctypes.c_void_p()
        check(lib().ftx_stream_create(ctypes.byref(s)), "ftx_stream_create")
        self.h = s

    def synchronize(self):
        check(lib().ftx_synchronize(self.h), "ftx_synchronize")

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().ftx_stream_destroy(self.h)
            self.h = None


class Graph:
    """Capture of library _dev calls issued on `stream` (context manager)."""

    def __init__(self, stream):
        self.stream = stream
        self.h = None

    def __enter__(self):
        check(lib().ftx_capture_begin(self.stream.h), "ftx_capture_begin")
        return self

    def __exit__(self, et, ev, tb):
        g = ctypes.c_void_p()
        rc = lib().ftx_capture_end(self.stream.h, ctypes.byref(g))
        if et is None:
            check(rc, "ftx_capture_end")
            self.h = g
        return False

    def launch(self, stream=None):
        check(lib().ftx_graph_launch(self.h, (stream or self.stream).h), "ftx_graph_launch")

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().ftx_graph_free(self.h)
            self.h = None
