"""Mixed-request serving workload over the library's request loop (ftx_server).

Trace (seeded): requests arrive as a Poisson process; kinds are 40 % encrypt,
30 % ciphertext add, 15 % multiply by a 32-bit constant, 15 % segmented sum.
Sizes are lognormal(ln 2^14, 2) clipped to [2^8, 2^22] ciphertexts and rounded
to a power of two, so that an add can pair two resident vectors of the same
length (an add or multiply with no suitable resident vector becomes an
encrypt).  A segmented sum splits a vector into n/16 lognormal segments.  The
client keeps at most `max_live` vectors alive, releasing the oldest first.

Three ways to run the same trace:
  run_oneshot  every request as its own plain kernel launch, back to back on
               one stream, inputs already in HBM (the throughput reference)
  run_served   through the server, all requests submitted at once (saturation)
               or at Poisson arrival times (latency at a fixed load)
"""
import ctypes
import threading
import time

import numpy as np

from . import FT_ERR_OUT_OF_BOUNDS, DeviceWord, FtError, GeArray, check, init, lib, random_scalars
from .workloads import lognormal_segments

ENC, ADD, MUL, SEG, FREE = range(5)
# the library's default batch cap (FTX_SERVER_DEFAULT_BATCH): longer encryptions
# and multiplies run as slices of this many ciphertexts
DEFAULT_BATCH = 1 << 18


class Server:
    def __init__(self, key, budget_bytes=64 << 30, n_streams=4, max_batch=0, borrow=False):
        init()
        L = lib()
        for name, res, args in [
            ("ftx_server_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_int]),
            ("ftx_server_destroy", None, [ctypes.c_void_p]),
            ("ftx_server_encrypt", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
            ("ftx_server_add", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]),
            ("ftx_server_mul_u32", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
            ("ftx_server_segsum", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]),
            ("ftx_server_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]),
            ("ftx_server_wait_ex", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]),
            ("ftx_server_release", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
            ("ftx_server_read", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]),
            ("ftx_server_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L
        self.h = ctypes.c_void_p()
        check(L.ftx_server_create(ctypes.byref(self.h), key.h, budget_bytes, n_streams, max_batch, 1 if borrow else 0),
              "ftx_server_create")

    def close(self):
        if self.h and self.h.value:
            self.L.ftx_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def encrypt(self, m, r, msg_bits):
        req, hnd = ctypes.c_uint64(), ctypes.c_uint64()
        check(self.L.ftx_server_encrypt(self.h, m.ctypes.data, r.ctypes.data, len(m), msg_bits,
                                        ctypes.byref(req), ctypes.byref(hnd)), "ftx_server_encrypt")
        return req.value, hnd.value

    def add(self, dst, src):
        req = ctypes.c_uint64()
        check(self.L.ftx_server_add(self.h, dst, src, ctypes.byref(req)), "ftx_server_add")
        return req.value

    def mul_u32(self, h, k):
        req = ctypes.c_uint64()
        check(self.L.ftx_server_mul_u32(self.h, h, k.ctypes.data, len(k), ctypes.byref(req)), "ftx_server_mul_u32")
        return req.value

    def segsum(self, h, off):
        req, hnd = ctypes.c_uint64(), ctypes.c_uint64()
        check(self.L.ftx_server_segsum(self.h, h, off.ctypes.data, len(off) - 1, ctypes.byref(req),
                                       ctypes.byref(hnd)), "ftx_server_segsum")
        return req.value, hnd.value

    def wait(self, req):
        lat = ctypes.c_double()
        check(self.L.ftx_server_wait(self.h, req, ctypes.byref(lat)), "ftx_server_wait")
        return lat.value

    def wait_ex(self, req):
        """(latency, time queued before the first launch), microseconds"""
        lat, q = ctypes.c_double(), ctypes.c_double()
        check(self.L.ftx_server_wait_ex(self.h, req, ctypes.byref(lat), ctypes.byref(q)), "ftx_server_wait_ex")
        return lat.value, q.value

    def release(self, h):
        check(self.L.ftx_server_release(self.h, h), "ftx_server_release")

    def read(self, h, n):
        a, b = np.zeros(n * 32, dtype=np.uint8), np.zeros(n * 32, dtype=np.uint8)
        check(self.L.ftx_server_read(self.h, h, a.ctypes.data, b.ctypes.data), "ftx_server_read")
        return a, b

    def stats(self):
        u, b, d = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(self.L.ftx_server_stats(self.h, ctypes.byref(u), ctypes.byref(b), ctypes.byref(d)))
        return {"bytes_in_use": u.value, "batches": b.value, "completed": d.value}


def make_trace(n_req, seed, mu_log2=14, sigma=2.0, lo_log2=8, hi_log2=22, max_live=256, msg_bits=8):
    """List of requests (kind, ...) referring to vectors by trace index."""
    rng = np.random.default_rng(seed)
    ops = []
    live = []            # trace vector ids, oldest first
    size_of = {}
    next_vec = 0
    stream = 0
    for _ in range(n_req):
        u = rng.random()
        kind = ENC if u < 0.40 else ADD if u < 0.70 else MUL if u < 0.85 else SEG
        lg = int(np.clip(np.round(rng.normal(mu_log2, sigma / np.log(2))), lo_log2, hi_log2))
        n = 1 << lg
        if kind == ADD:
            same = [v for v in live if size_of[v] == n]
            if len(same) < 2:
                sizes = {}
                for v in live:
                    sizes.setdefault(size_of[v], []).append(v)
                pairs = [vs for vs in sizes.values() if len(vs) >= 2]
                if pairs:
                    same = pairs[rng.integers(len(pairs))]
            if len(same) >= 2:
                i, j = rng.choice(len(same), 2, replace=False)
                ops.append((ADD, same[i], same[j]))
                continue
            kind = ENC
        if kind in (MUL, SEG) and live:
            v = live[rng.integers(len(live))]
            n = size_of[v]
            if kind == MUL:
                k = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
                ops.append((MUL, v, k))
            else:
                nseg = max(1, n // 16)
                off = lognormal_segments(nseg, n, int(rng.integers(2**31)))
                ops.append((SEG, v, off, next_vec))
                size_of[next_vec] = nseg
                live.append(next_vec)
                next_vec += 1
        else:
            m = rng.integers(0, 2**msg_bits, size=n, dtype=np.uint64)
            stream += 1
            r = random_scalars(n, seed, stream)
            ops.append((ENC, m, r, next_vec, msg_bits))
            size_of[next_vec] = n
            live.append(next_vec)
            next_vec += 1
        while len(live) > max_live:
            ops.append((FREE, live.pop(0)))
    return ops, size_of


def trace_elements(ops):
    tot = 0
    for op in ops:
        if op[0] == ENC:
            tot += len(op[1])
        elif op[0] == MUL:
            tot += len(op[2])
        elif op[0] == SEG:
            tot += int(op[2][-1])
    return tot


# ------------------------------------------------------------------ served

def run_served(server, ops, arrivals=None, info=None):
    """Submit the trace (at arrival times in seconds, or all at once) and wait
    for every request.  Returns (wall seconds, per-request latencies in us,
    trace vector -> server handle).  `info` (a dict) receives the largest delay
    of a submission behind its arrival time (ms), the time of the last
    submission (s) and each request's time queued before its first launch
    (us)."""
    hmap = {}
    reqs = []
    t0 = time.perf_counter()
    pending_free = []
    lag = 0.0
    for idx, op in enumerate(ops):
        if arrivals is not None and op[0] != FREE:
            target = t0 + arrivals[idx]
            while True:
                dt = target - time.perf_counter()
                if dt <= 0:
                    break
                time.sleep(dt if dt > 0.002 else 0)
        kind = op[0]
        if kind == ENC:
            req, h = server.encrypt(op[1], op[2], op[4])
            hmap[op[3]] = h
        elif kind == ADD:
            req = server.add(hmap[op[1]], hmap[op[2]])
        elif kind == MUL:
            req = server.mul_u32(hmap[op[1]], op[2])
        elif kind == SEG:
            req, h = server.segsum(hmap[op[1]], op[2])
            hmap[op[3]] = h
        else:
            pending_free.append(op[1])
            server.release(hmap[op[1]])
            continue
        reqs.append(req)
        if arrivals is not None:
            lag = max(lag, time.perf_counter() - t0 - arrivals[idx])
    t_sub = time.perf_counter() - t0
    done = [server.wait_ex(q) for q in reqs]
    wall = time.perf_counter() - t0
    lats = [d[0] for d in done]
    if info is not None:
        info.update(submit_lag_ms=lag * 1e3, last_submit_s=t_sub, queued_us=np.array([d[1] for d in done]))
    return wall, np.array(lats), hmap


def poisson_arrivals(ops, rate, seed):
    """Arrival times (s) of a Poisson process of `rate` requests per second;
    a release takes no time of its own (it follows the request before it)."""
    rng = np.random.default_rng(seed)
    is_req = np.array([op[0] != FREE for op in ops], dtype=np.float64)
    gaps = rng.exponential(1.0 / rate, size=len(ops)) * is_req
    return np.cumsum(gaps)


# ------------------------------------------------------------------ one-shot

class _Hip:
    def __init__(self):
        L = ctypes.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.L = L

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        if self.L.hipMalloc(ctypes.byref(p), max(1, a.nbytes)) != 0:
            raise MemoryError
        self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1)
        return p


def run_oneshot(key, ops, size_of):
    """Execute each request as one plain kernel launch, back to back on the
    library stream, with every input already in HBM.  Returns (GPU seconds,
    trace vector -> (c1, c2) arrays)."""
    L = lib()
    hip = _Hip()
    st = L.ftx_default_stream()
    vec = {}
    dev = []
    prepared = []
    for op in ops:
        kind = op[0]
        if kind == ENC:
            n = len(op[1])
            c1, c2 = GeArray.alloc(n), GeArray.alloc(n)
            vec[op[3]] = (c1, c2)
            dm, dr = hip.upload(op[1]), hip.upload(op[2])
            dev += [dm, dr]
            prepared.append((ENC, c1, c2, dm, dr, op[4]))
        elif kind == ADD:
            prepared.append((ADD, op[1], op[2]))
        elif kind == MUL:
            dk = hip.upload(op[2])
            dev.append(dk)
            prepared.append((MUL, op[1], dk))
        elif kind == SEG:
            nseg = len(op[2]) - 1
            o1, o2 = GeArray.alloc(nseg), GeArray.alloc(nseg)
            vec[op[3]] = (o1, o2)
            doff = hip.upload(op[2])
            dev.append(doff)
            scr = GeArray.alloc(L.ftx_segsum_scratch_points(int(op[2][-1])))
            prepared.append((SEG, op[1], doff, nseg, int(op[2][-1]), o1, o2, scr))
    fail = DeviceWord()
    check(L.ftx_synchronize(None))
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.L.hipEventCreate(ctypes.byref(e0))
    hip.L.hipEventCreate(ctypes.byref(e1))
    hip.L.hipEventRecord(e0, st)
    for p in prepared:
        if p[0] == ENC:
            check(L.ftx_encrypt_dev(p[1].h, p[2].h, key.h, p[3], p[4], p[5], fail.p, st))
        elif p[0] == ADD:
            d, s = vec[p[1]], vec[p[2]]
            check(L.ftx_add_dev(d[0].h, s[0].h, st))
            check(L.ftx_add_dev(d[1].h, s[1].h, st))
        elif p[0] == MUL:
            a = vec[p[1]]
            check(L.ftx_mul_u32_dev(a[0].h, a[1].h, p[2], st))
        else:
            a = vec[p[1]]
            check(L.ftx_segmented_sum_dev(p[5].h, p[2], p[3], p[4], None, a[0].h, p[7].h, st))
            check(L.ftx_segmented_sum_dev(p[6].h, p[2], p[3], p[4], None, a[1].h, p[7].h, st))
    hip.L.hipEventRecord(e1, st)
    hip.L.hipEventSynchronize(e1)
    if fail.read():
        raise FtError(FT_ERR_OUT_OF_BOUNDS, "one-shot encryption: message above its bound")
    ms = ctypes.c_float()
    hip.L.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    for p in dev:
        hip.L.hipFree(p)
    return ms.value * 1e-3, vec


class PinnedPool:
    """One hipHostMalloc allocation carved into numpy views (page-locked
    inputs for the borrowed-input server mode)."""

    def __init__(self, nbytes):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        self.L.hipHostFree.argtypes = [ctypes.c_void_p]
        self.p = ctypes.c_void_p()
        if self.L.hipHostMalloc(ctypes.byref(self.p), max(1, nbytes), 0) != 0:
            raise MemoryError("hipHostMalloc(%d)" % nbytes)
        self.buf = (ctypes.c_uint8 * max(1, nbytes)).from_address(self.p.value)
        self.off = 0

    def copy(self, a):
        a = np.ascontiguousarray(a)
        n = (a.nbytes + 255) // 256 * 256
        view = np.frombuffer(self.buf, dtype=a.dtype, count=a.size, offset=self.off).reshape(a.shape)
        view[...] = a
        self.off += n
        return view

    def __del__(self):
        if self.p and self.p.value:
            self.L.hipHostFree(self.p)
            self.p = None


def pin_inputs(ops):
    """Copy every request input of a trace into one pinned pool."""
    total = 0
    for op in ops:
        for a in op[1:]:
            if isinstance(a, np.ndarray):
                total += (a.nbytes + 255) // 256 * 256
    pool = PinnedPool(total)
    out = []
    for op in ops:
        out.append(tuple(pool.copy(a) if isinstance(a, np.ndarray) else a for a in op))
    return pool, out
