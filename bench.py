#!/usr/bin/env python3
"""Benchmark of the gfx950 batched Ed25519 ElGamal library on one MI355X.

Workload (one "step"): the scalar-multiplication configuration with
N = 2^22 of each kind, every input resident in HBM before timing starts:
  fixed-base      k_i * B                       (N scalar-mults)
  variable-base   k_i * P_i, P_i = s_i * B      (N scalar-mults)
  ct x constant   (a_i C1_i, a_i C2_i), a_i u32 (2N scalar-mults)
`value` = scalar-mults per second over the whole step.  The other
configurations are measured at full size under "detail.configs": config 0
(encrypt 2 x 65,536 bits, add, reduce), config 2 (reduce 2^26 ciphertexts,
inclusive scan 2^24, CSR segmented sum 2^24 in 2^19 lognormal segments) and
config 3 (sparse gather-sum over a 2^22-node / 2^25-edge power-law graph plus
rerandomisation, replayed as one hipGraph).

Timing uses HIP events recorded on the stream the kernels run on.  The
roofline object is for the kernel with the largest share of the step.  The
CPU baseline is the oracle (a C port of the reference CPU path's algorithm)
on a bounded sample, all host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n LOG2N]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X: 256 CUs, 2.4 GHz peak engine clock, 64 lanes issued per CU per
# clock for full-rate VALU ops; v_mad_u64_u32 (one 32x32->64 product) issues
# at that full rate on gfx950 (measured, see profiles/).
PEAK_MUL32 = 256 * 64 * 2.4e9
HBM_PEAK = 8.0e12

# algorithmic cost model: 32x32->64 products per field mul / square (8x32-bit
# schoolbook + fold), group operations counted from the reference formulas.
M, S = 72, 44
PRODUCTS = {
    "fixed_base": 461 * M + 16 * S,
    "variable_base": 1316 * M + 1024 * S,
    "mul_u32": 216 * M + 144 * S,
    "point_add": 9 * M,
}
POINT_BYTES = 128

# HBM bytes per launch at N = 2^22 (rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, one pass each, scaled by the same counters' reading for a plain
# copy kernel of known size) and VALU busy, read from the committed counter
# summary below: its table has one row per kernel.
PMC_PROFILE = os.path.join(ROOT, "profiles", "scalarmult_n22_pmc_counters.md")
PMC_KERNEL = {"fixed_base": "k_fixed_base32<0>", "variable_base": {3: "k_varbase_r8", 4: "k_varbase"},
              "mul_u32": {3: "k_mul_u32_r8", 4: "k_mul_u32_r16"}}


def pmc_row(kernel, path=PMC_PROFILE):
    """{'read', 'write' (bytes per launch), 'valu_busy'} of `kernel` from the
    profile's table, or None when the profile has no row for it."""
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    head = None
    for ln in lines:
        if not ln.startswith("|"):
            continue
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if head is None and "kernel" in cells[0]:
            head = cells
            continue
        if head and cells[0].strip("`") == kernel:
            row = dict(zip(head, cells))
            try:
                return {"read": float(row["HBM read (MB)"]) * 1e6, "write": float(row["HBM write (MB)"]) * 1e6,
                        "valu_busy": float(row["VALU busy"])}
            except (KeyError, ValueError):
                return None
    return None


ALG_BYTES = {"fixed_base": 32 + 128, "variable_base": 32 + 2 * 128, "mul_u32": 4 + 4 * 128}


class Hip:
    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        L = self.L
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        L.hipStreamWaitEvent.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        if self.L.hipMalloc(ctypes.byref(p), nbytes) != 0:
            raise MemoryError("hipMalloc(%d)" % nbytes)
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.malloc(arr.nbytes)
        if self.L.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy")
        return p

    def event(self):
        e = ctypes.c_void_p()
        self.L.hipEventCreate(ctypes.byref(e))
        return e

    def elapsed_ms(self, a, b):
        self.L.hipEventSynchronize(b)
        ms = ctypes.c_float()
        self.L.hipEventElapsedTime(ctypes.byref(ms), a, b)
        return ms.value


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run(args):
    import libftcrypto_amd as ft
    hip = Hip()
    ft.init()
    L = ft.lib()
    stream = L.ftx_default_stream()
    n = 1 << args.n
    seed = 0xF7C00001
    schedules = None
    if not args.no_autotune:
        t0 = time.time()
        tab = ft.autotune(3)
        lanes = ft.ct_lanes_table()
        table = {c: tab[c] + (lanes[c],) for c in tab}
        schedules = {"picked_by": "ftx_autotune on this GPU", "seconds": round(time.time() - t0, 2),
                     "classes": {"n<2^14": table[0], "n<2^18": table[1], "n<2^22": table[2], "n>=2^22": table[3]},
                     "fields": "(seg1, scan: log2 elements per lane; segu: log2 radix of the segmented-sum run tree; reduce: max blocks; window width of ct x u32 and of variable-base; lanes per ciphertext of an encryption)"}
        log("[bench] autotuned aggregation schedules in %.1fs: %s" % (schedules["seconds"], table))

    log("[bench] generating %d x 4 inputs" % n)
    t0 = time.time()
    k_fixed = ft.random_scalars(n, seed, 1)
    k_var = ft.random_scalars(n, seed, 2)
    s_var = ft.random_scalars(n, seed, 3)
    r_ct = ft.random_scalars(n, seed, 4)
    rng = np.random.default_rng(seed)
    a_u32 = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    m_ct = rng.integers(0, 2, size=n, dtype=np.uint64)
    log("[bench] inputs generated in %.1fs" % (time.time() - t0))

    d_kf, d_kv, d_a = hip.upload(k_fixed), hip.upload(k_var), hip.upload(a_u32)
    out_fixed = ft.GeArray.alloc(n)
    pts = ft.GeArray.from_scalars(s_var)                    # P_i = s_i B (setup)
    x = ft.random_scalars(1, seed, 5)
    key = ft.Key(ft.GeArray.from_scalars(x))
    c1, c2 = ft.encrypt(key, m_ct, r_ct, msg_bits=1)          # setup ciphertexts
    ft.check(L.ftx_synchronize(None))

    ops = {
        "fixed_base": lambda: ft.check(L.ftx_from_scalars_dev(out_fixed.h, d_kf, stream)),
        "variable_base": lambda: ft.check(L.ftx_mul_dev(pts.h, d_kv, stream)),
        "mul_u32": lambda: ft.check(L.ftx_mul_u32_dev(c1.h, c2.h, d_a, stream)),
    }
    mults = {"fixed_base": n, "variable_base": n, "mul_u32": 2 * n}

    ev = [hip.event() for _ in range(2)]

    def timed(fn, reps):
        hip.L.hipEventRecord(ev[0], stream)
        for _ in range(reps):
            fn()
        hip.L.hipEventRecord(ev[1], stream)
        return hip.elapsed_ms(ev[0], ev[1]) / reps

    if args.calibration_copy:
        # a plain copy kernel of known size (2 x N x 128 B moved) for PMC byte-counter calibration
        cp = pts.get_range(0, n, 1)
        ft.check(L.ftx_synchronize(None))
        del cp
    log("[bench] warmup %d" % args.warmup)
    for _ in range(args.warmup):
        for f in ops.values():
            f()
    ft.check(L.ftx_synchronize(None))

    log("[bench] timing %d steps" % args.steps)
    per = {k: 0.0 for k in ops}
    hip.L.hipEventRecord(ev[0], stream)
    t_wall = time.time()
    for _ in range(args.steps):
        for k, f in ops.items():
            f()
    hip.L.hipEventRecord(ev[1], stream)
    step_ms = hip.elapsed_ms(ev[0], ev[1]) / args.steps
    wall_ms = (time.time() - t_wall) * 1e3 / args.steps
    for k, f in ops.items():
        per[k] = timed(f, max(1, args.steps))
        log("[bench] %-14s %.3f ms" % (k, per[k]))
    total_mults = sum(mults.values())
    value = total_mults / (step_ms * 1e-3)

    # the reference-compatible calls take host buffers: the same mults through
    # ft_array_from_scalars / ft_mul, PCIe copy of the scalars and the final
    # synchronisation included (wall clock)
    ft.check(L.ftx_synchronize(None))
    t0 = time.time()
    hb = ft.GeArray.from_scalars(k_fixed)
    t_fixed_host = time.time() - t0
    t0 = time.time()
    hb.mul(k_var)
    t_var_host = time.time() - t0
    del hb
    host_rates = {"fixed_base_mults_per_s": n / t_fixed_host, "variable_base_mults_per_s": n / t_var_host,
                  "note": "ft_array_from_scalars / ft_mul with host scalar buffers, PCIe and sync included"}

    detail = {"host_buffer_calls": host_rates,
              "arithmetic": "exact GF(2^255-19): nine 29-bit limbs in u32, 32x32->64 products into u64 column sums"}
    for k in ops:
        rate = mults[k] / (per[k] * 1e-3)
        ach = rate * PRODUCTS[k]
        detail[k] = {"ms": round(per[k], 4), "scalar_mults_per_s": rate,
                     "mul32_per_s": ach, "valu_frac": ach / PEAK_MUL32}
    dom = max(per, key=lambda k: per[k])
    ach = mults[dom] / (per[dom] * 1e-3) * PRODUCTS[dom]
    roofline = {"kernel": dom, "bound": "valu", "achieved": ach / 1e12, "peak": PEAK_MUL32 / 1e12,
                "unit": "Tmul32/s", "frac": ach / PEAK_MUL32, "traffic": None,
                "algorithmic_bytes": ALG_BYTES[dom] * (mults[dom] // (2 if dom == "mul_u32" else 1))}
    if n == 1 << 22:
        kname = PMC_KERNEL[dom]
        if isinstance(kname, dict):            # the window width the schedule table picks at this size
            got = [ctypes.c_int() for _ in range(6)]
            ft.check(L.ftx_tune_get(L.ftx_tune_size_class(n), *[ctypes.byref(g) for g in got]))
            kname = kname[got[5].value if dom == "variable_base" else got[4].value]
        pm = pmc_row(kname)
        if pm:
            roofline["traffic"] = pm["read"] + pm["write"]
            roofline["valu_busy"] = pm["valu_busy"]
            roofline["counters"] = "profiles/scalarmult_n22_pmc_counters.md (%s)" % kname

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, k_fixed, k_var, s_var, a_u32)
    del out_fixed, pts, c1, c2
    configs = {}
    sections = set(args.sections.split(",")) if not args.headline_only else set()
    if "cfg0" in sections:
        configs["cfg0_encrypt_add_reduce_n65536"] = bench_cfg0(ft, hip, key, seed)
    if "cfg2" in sections:
        configs["cfg2_aggregation"] = bench_cfg2(ft, hip, key, seed, args)
    if "cfg3" in sections:
        configs["cfg3_sparse_step_graph"] = bench_cfg3(ft, hip, key, seed, args)
    if "isum" in sections:
        configs["reduce_isum"] = bench_isum(ft, hip, seed, args)
    if "io" in sections:
        configs["encode_decode"] = bench_io(ft, hip, seed, args)
    if "cfg4" in sections:
        configs["cfg4_serving"] = bench_cfg4(ft, key, seed, args)
    detail["configs"] = configs
    if not args.no_cpu and configs:
        log("[bench] CPU port on samples of configs 0, 2 and 3")
        H = bytes(ft.GeArray.from_scalars(x).to_bytes_folded())
        detail["cpu_port_by_config"] = cpu_config_rates(H, seed, configs)
    detail["aggregation_schedules"] = schedules
    return {
        "metric": "scalar-mults/s (fixed-base, variable-base, ciphertext x u32) on 1 MI355X",
        "value": value, "unit": "scalar-mults/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": step_ms, "wall_ms_per_step": wall_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded SplitMix64 scalars mod l)",
        "config": {"workload": "scalar-mult arrays N=2^%d each: k*B, k*P, (a*C1, a*C2) with a u32" % args.n,
                   "n": n, "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "detail": detail,
    }


class Timer:
    def __init__(self, hip, stream):
        self.hip, self.stream = hip, stream
        self.ev = [hip.event(), hip.event()]

    def __call__(self, fn, reps=3):
        fn()
        self.hip.L.hipStreamSynchronize(self.stream)
        self.hip.L.hipEventRecord(self.ev[0], self.stream)
        for _ in range(reps):
            fn()
        self.hip.L.hipEventRecord(self.ev[1], self.stream)
        return self.hip.elapsed_ms(self.ev[0], self.ev[1]) / reps


def adds_entry(ms, adds, bytes_per_add=None):
    rate = adds / (ms * 1e-3)
    d = {"ms": round(ms, 4), "point_adds": adds, "point_adds_per_s": rate,
         "valu_frac": rate * PRODUCTS["point_add"] / PEAK_MUL32}
    if bytes_per_add:
        d["hbm_GBps"] = rate * bytes_per_add / 1e9
        d["hbm_frac"] = rate * bytes_per_add / HBM_PEAK
    return d


def bench_cfg0(ft, hip, key, seed):
    """Encrypt 2 x 65,536 bits, add the two arrays, reduce both halves of the
    sum: on one stream, with the independent halves on two streams, and
    batched (both encryptions in one launch, both reduces in one batch)."""
    L = ft.lib()
    stream = L.ftx_default_stream()
    T = Timer(hip, stream)
    n = 65536
    rng = np.random.default_rng(seed)
    m = hip.upload(rng.integers(0, 2, size=n, dtype=np.uint64))
    r = hip.upload(ft.random_scalars(n, seed, 10))
    a1, a2, b1, b2 = (ft.GeArray.alloc(n) for _ in range(4))
    out1, out2 = ft.GeArray.alloc(1), ft.GeArray.alloc(1)
    rscr1 = ft.GeArray.alloc(L.ftx_reduce_scratch_points(n))
    rscr2 = ft.GeArray.alloc(L.ftx_reduce_scratch_points(n))
    # message-bound fail words of the bounded encryptions (one per stream)
    fa, fb = ft.DeviceWord(), ft.DeviceWord()
    enc_ms = T(lambda: (ft.check(L.ftx_encrypt_dev(a1.h, a2.h, key.h, m, r, 1, fa.p, stream)),
                        ft.check(L.ftx_encrypt_dev(b1.h, b2.h, key.h, m, r, 1, fa.p, stream))))
    add_ms = T(lambda: (ft.check(L.ftx_add_dev(a1.h, b1.h, stream)), ft.check(L.ftx_add_dev(a2.h, b2.h, stream))))
    red_ms = T(lambda: (ft.check(L.ftx_reduce_dev(out1.h, a1.h, rscr1.h, stream)),
                        ft.check(L.ftx_reduce_dev(out2.h, a2.h, rscr2.h, stream))))
    # compression of the outputs (A, B, A+B: 6 arrays of n points), timed apart from the ops
    enc_out = hip.malloc(n * 32)
    cmp_ms = T(lambda: [ft.check(L.ftx_encode_dev(enc_out, x.h, stream)) for x in (a1, a2, b1, b2, a1, a2)])
    hip.L.hipFree(enc_out)
    # the same step with A's and B's work (then the two halves' reduces) on two streams
    p = ctypes.c_void_p()
    ft.check(L.ftx_stream_create(ctypes.byref(p)))
    s2 = p
    ev = [hip.event() for _ in range(2)]

    def fork():
        hip.L.hipEventRecord(ev[0], stream)
        hip.L.hipStreamWaitEvent(s2, ev[0], 0)

    def join():
        hip.L.hipEventRecord(ev[1], s2)
        hip.L.hipStreamWaitEvent(stream, ev[1], 0)

    def step2():
        fork()
        ft.check(L.ftx_encrypt_dev(a1.h, a2.h, key.h, m, r, 1, fa.p, stream))
        ft.check(L.ftx_encrypt_dev(b1.h, b2.h, key.h, m, r, 1, fb.p, s2))
        join()
        fork()
        ft.check(L.ftx_add_dev(a1.h, b1.h, stream))
        ft.check(L.ftx_add_dev(a2.h, b2.h, s2))
        ft.check(L.ftx_reduce_dev(out1.h, a1.h, rscr1.h, stream))
        ft.check(L.ftx_reduce_dev(out2.h, a2.h, rscr2.h, s2))
        join()

    two_ms = T(step2)
    ft.check(L.ftx_synchronize(None))
    L.ftx_stream_destroy(s2)
    # the same step batched: A's and B's encryptions in one launch
    # (ftx_encrypt_batch_dev), the adds, both halves' reduces as one
    # aggregation batch (ftx_aggregate_batch_dev)
    ejobs = (ft.EncJob * 2)(ft.enc_job(a1, a2, m, r), ft.enc_job(b1, b2, m, r))
    rjobs = (ft.AggJob * 2)(ft.agg_job(ft.AGG_REDUCE, out1, a1), ft.agg_job(ft.AGG_REDUCE, out2, a2))
    bscr = ft.GeArray.alloc(L.ftx_aggregate_batch_scratch_points(rjobs, 2))

    def step_batched():
        ft.check(L.ftx_encrypt_batch_dev(ejobs, 2, key.h, 1, fa.p, stream))
        ft.check(L.ftx_add_dev(a1.h, b1.h, stream))
        ft.check(L.ftx_add_dev(a2.h, b2.h, stream))
        ft.check(L.ftx_aggregate_batch_dev(rjobs, 2, bscr.h, stream))

    batched_ms = T(step_batched)
    enc_batched_ms = T(lambda: ft.check(L.ftx_encrypt_batch_dev(ejobs, 2, key.h, 1, fa.p, stream)))
    ft.check(L.ftx_synchronize(None))
    assert fa.read() == 0 and fb.read() == 0, "config 0: messages above their 1-bit bound"
    hip.L.hipFree(m)
    hip.L.hipFree(r)
    # elementwise ciphertext add at 2^22 (bandwidth-sized: 2 x 384 B per ct)
    nb = 1 << 22
    big = [ft.GeArray.from_small_scalars(np.arange(1, nb + 1, dtype=np.uint64)) for _ in range(4)]
    big_ms = T(lambda: (ft.check(L.ftx_add_dev(big[0].h, big[2].h, stream)),
                        ft.check(L.ftx_add_dev(big[1].h, big[3].h, stream))))
    del big
    return {"compress": {"ms": round(cmp_ms, 4), "points": 6 * n, "points_per_s": 6 * n / (cmp_ms * 1e-3),
                         "note": "canonical 32-byte encodings, batch inversion; not part of step_ms"},
            "encrypt": {"ms": round(enc_ms, 4), "ciphertexts_per_s": 2 * n / (enc_ms * 1e-3),
                        "valu_frac": 2 * n * (2 * PRODUCTS["fixed_base"]) / (enc_ms * 1e-3) / PEAK_MUL32},
            "ct_add": adds_entry(add_ms, 2 * n, 3 * POINT_BYTES),
            "ct_add_n2^22": adds_entry(big_ms, 2 * nb, 3 * POINT_BYTES),
            "reduce": adds_entry(red_ms, 2 * (n - 1), POINT_BYTES),
            "step_ms_one_stream": round(enc_ms + add_ms + red_ms, 4),
            "step_ms_two_streams": round(two_ms, 4),
            "ciphertexts_per_s_two_streams": 2 * n / (two_ms * 1e-3),
            "step_ms_batched": round(batched_ms, 4),
            "encrypt_batched": {"ms": round(enc_batched_ms, 4), "ciphertexts_per_s": 2 * n / (enc_batched_ms * 1e-3)},
            "ciphertexts_per_s_batched": 2 * n / (batched_ms * 1e-3)}


def bench_cfg2(ft, hip, key, seed, args):
    """reduce 2^26 ct, inclusive scan 2^24 ct, segmented sum 2^24 ct / 2^19
    segments: each job alone on one stream, then all three at once on three
    streams (the inputs of every job resident together)."""
    from libftcrypto_amd import workloads as W
    L = ft.lib()
    stream = L.ftx_default_stream()
    T = Timer(hip, stream)
    rng = np.random.default_rng(seed + 2)
    nr, ns, nseg = 1 << args.cfg2_reduce_log2, 1 << args.cfg2_log2, 1 << (args.cfg2_log2 - 5)
    res = {}
    log("[bench] cfg2: encrypting 2^%d + 2 x 2^%d ciphertexts" % (args.cfg2_reduce_log2, args.cfg2_log2))

    def make(n, sid):
        c1, c2 = ft.GeArray.alloc(n), ft.GeArray.alloc(n)
        m = hip.upload(rng.integers(0, 256, size=n, dtype=np.uint64))
        r = hip.upload(ft.random_scalars(n, seed, sid))
        fail = ft.DeviceWord()
        ft.check(L.ftx_encrypt_dev(c1.h, c2.h, key.h, m, r, 8, fail.p, stream))
        ft.check(L.ftx_synchronize(stream))
        assert fail.read() == 0
        hip.L.hipFree(m)
        hip.L.hipFree(r)
        return c1, c2

    r1, r2 = make(nr, 20)
    g1, g2 = make(ns, 21)          # segmented-sum input
    q1, q2 = make(ns, 22)          # scanned in place
    out = ft.GeArray.alloc(1)
    rscr = ft.GeArray.alloc(L.ftx_reduce_scratch_points(nr))
    off = W.lognormal_segments(nseg, ns, seed)
    d_off = hip.upload(off)
    s1, s2 = ft.GeArray.alloc(nseg), ft.GeArray.alloc(nseg)
    gscr = ft.GeArray.alloc(L.ftx_segsum_scratch_points(ns))
    sscr = ft.GeArray.alloc(L.ftx_scan_scratch_points(ns))

    def job_reduce(st):
        ft.check(L.ftx_reduce_dev(out.h, r1.h, rscr.h, st))
        ft.check(L.ftx_reduce_dev(out.h, r2.h, rscr.h, st))

    def job_segsum(st):
        ft.check(L.ftx_segmented_sum_dev(s1.h, d_off, nseg, ns, None, g1.h, gscr.h, st))
        ft.check(L.ftx_segmented_sum_dev(s2.h, d_off, nseg, ns, None, g2.h, gscr.h, st))

    def job_scan(st):
        ft.check(L.ftx_scan_dev(q1.h, 0, sscr.h, st))
        ft.check(L.ftx_scan_dev(q2.h, 0, sscr.h, st))

    log("[bench] cfg2: each job alone")
    ms_r = T(lambda: job_reduce(stream))
    res["reduce"] = adds_entry(ms_r, 2 * (nr - 1), POINT_BYTES)
    ms_g = T(lambda: job_segsum(stream))
    seg_bytes = 2 * ns * POINT_BYTES + (nseg + 1) * 8 + 2 * nseg * POINT_BYTES
    res["segmented_sum"] = adds_entry(ms_g, 2 * (ns - nseg), seg_bytes / (2 * (ns - nseg)))
    res["segmented_sum"]["segments"] = nseg
    ms_s = T(lambda: job_scan(stream))
    res["scan"] = adds_entry(ms_s, 2 * (ns - 1), 2 * POINT_BYTES)
    # both halves of the segmented sum as two jobs of one persistent launch
    gjobs = (ft.AggJob * 2)(ft.agg_job(ft.AGG_SEGSUM, s1, g1, d_off, nseg, ns),
                            ft.agg_job(ft.AGG_SEGSUM, s2, g2, d_off, nseg, ns))
    gbscr = ft.GeArray.alloc(L.ftx_aggregate_batch_scratch_points(gjobs, 2))
    ms_gb = T(lambda: ft.check(L.ftx_aggregate_batch_dev(gjobs, 2, gbscr.h, stream)))
    res["segmented_sum_one_launch"] = adds_entry(ms_gb, 2 * (ns - nseg), seg_bytes / (2 * (ns - nseg)))
    res["segmented_sum_one_launch"]["note"] = "ftx_aggregate_batch_dev with the two halves as two jobs"
    del gbscr

    if args.cfg2_alone_only:
        ft.check(L.ftx_synchronize(None))
        hip.L.hipFree(d_off)
        return res
    log("[bench] cfg2: all three at once on three streams")
    sts = []
    for _ in range(2):
        p = ctypes.c_void_p()
        ft.check(L.ftx_stream_create(ctypes.byref(p)))
        sts.append(p)
    ev = [hip.event() for _ in range(4)]

    def together():
        hip.L.hipEventRecord(ev[2], stream)
        for st in sts:
            hip.L.hipStreamWaitEvent(st, ev[2], 0)
        job_reduce(stream)
        job_segsum(sts[0])
        job_scan(sts[1])
        for st, e in zip(sts, ev[:2]):
            hip.L.hipEventRecord(e, st)
            hip.L.hipStreamWaitEvent(stream, e, 0)

    ms_all = T(together)
    adds = 2 * (nr - 1) + 2 * (ns - nseg) + 2 * (ns - 1)
    res["all_three_concurrent"] = adds_entry(ms_all, adds)
    res["all_three_concurrent"]["sum_of_alone_ms"] = round(ms_r + ms_g + ms_s, 4)

    log("[bench] cfg2: all three in one persistent launch")
    outb = [ft.GeArray.alloc(1) for _ in range(2)]
    jobs = [ft.agg_job(ft.AGG_REDUCE, outb[0], r1), ft.agg_job(ft.AGG_REDUCE, outb[1], r2),
            ft.agg_job(ft.AGG_SEGSUM, s1, g1, d_off, nseg, ns), ft.agg_job(ft.AGG_SEGSUM, s2, g2, d_off, nseg, ns),
            ft.agg_job(ft.AGG_SCAN, q1), ft.agg_job(ft.AGG_SCAN, q2)]
    arr = (ft.AggJob * len(jobs))(*jobs)
    bscr = ft.GeArray.alloc(L.ftx_aggregate_batch_scratch_points(arr, len(jobs)))
    ms_b = T(lambda: ft.check(L.ftx_aggregate_batch_dev(arr, len(jobs), bscr.h, stream)))
    res["one_persistent_launch"] = adds_entry(ms_b, adds)
    res["one_persistent_launch"]["note"] = (
        "ftx_aggregate_batch_dev: both halves of the reduce, segmented sum and scan as six jobs; their "
        "element walks (the scan's first level included) share one persistent launch whose waves take "
        "tiles from a device counter, then run-tree passes, the reduces' partial sums and the scans' "
        "upper levels and second walk")
    ft.check(L.ftx_synchronize(None))
    for st in sts:
        L.ftx_stream_destroy(st)
    hip.L.hipFree(d_off)
    return res


def bench_isum(ft, hip, seed, args):
    """ft_reduce_isum, res[key[i]] += src[i], for 2^22 summands over 2^19
    result elements with uniform random keys: the device-key form (keys
    sorted on the device, caller scratch) and the reference-signature form
    (host keys, checked and copied per call).  Each summand is one point add."""
    L = ft.lib()
    stream = L.ftx_default_stream()
    T = Timer(hip, stream)
    rng = np.random.default_rng(seed + 5)
    n, m = 1 << args.isum_log2, 1 << (args.isum_log2 - 3)
    keys = rng.integers(0, m, size=n, dtype=np.uint64)
    src = ft.GeArray.from_small_scalars(rng.integers(0, 1 << 40, size=n, dtype=np.uint64))
    res = ft.GeArray.from_small_scalars(np.zeros(m, dtype=np.uint64))
    d_keys = hip.upload(keys)
    scr = ft.GeArray.alloc(L.ftx_reduce_isum_scratch_points(n, m))
    work = hip.malloc(L.ftx_reduce_isum_work_bytes(n, m))
    log("[bench] isum: 2^%d summands over 2^%d keys" % (args.isum_log2, args.isum_log2 - 3))
    ms_dev = T(lambda: ft.check(L.ftx_reduce_isum_dev(res.h, d_keys, src.h, scr.h, work, stream)), reps=5)
    out = {"n": n, "m": m, "device_keys": adds_entry(ms_dev, n)}
    t0 = time.perf_counter()
    for _ in range(3):
        res.reduce_isum(keys, src)
    out["host_keys_ms"] = round((time.perf_counter() - t0) / 3 * 1e3, 4)
    ft.check(L.ftx_synchronize(None))
    hip.L.hipFree(d_keys)
    hip.L.hipFree(work)
    return out


def bench_io(ft, hip, seed, args):
    """Compressed encode (canonical 32-byte form, batch inversion) and decode
    (square root per point, failure index) of 2^n device-resident points,
    device to device, and validation (on the curve, X(l P) = 0) of the
    decoded points."""
    L = ft.lib()
    stream = L.ftx_default_stream()
    T = Timer(hip, stream)
    n = 1 << args.n
    rng = np.random.default_rng(seed + 6)
    g = ft.GeArray.from_small_scalars(rng.integers(0, 1 << 62, size=n, dtype=np.uint64))
    d_enc, d_fail = hip.malloc(32 * n), hip.malloc(8)
    out = ft.GeArray.alloc(n)
    log("[bench] io: encode / decode of 2^%d points" % args.n)
    ms_e = T(lambda: ft.check(L.ftx_encode_dev(d_enc, g.h, stream)))
    ms_d = T(lambda: ft.check(L.ftx_decode_dev(out.h, d_enc, d_fail, stream)))
    bad = [None]

    def validate():
        bad[0] = out.validate()
    ms_v = T(validate)
    fail = np.zeros(1, dtype=np.uint64)
    hip.L.hipMemcpy(fail.ctypes.data_as(ctypes.c_void_p), d_fail, 8, 2)
    ft.check(L.ftx_synchronize(None))
    hip.L.hipFree(d_enc)
    hip.L.hipFree(d_fail)
    if fail[0] != 0 or bad[0] != -1:
        raise RuntimeError("decode or validation of valid points reported a failure")
    ent = lambda ms, b: {"ms": round(ms, 4), "points_per_s": n / (ms * 1e-3),
                         "hbm_GBps": n * b / (ms * 1e-3) / 1e9}
    return {"n": n, "encode": ent(ms_e, 128 + 32), "decode": ent(ms_d, 32 + 128),
            "validate": ent(ms_v, 128)}


def bench_cfg3(ft, hip, key, seed, args):
    """Gather-sum of in-neighbours + rerandomisation, one hipGraph per step."""
    from libftcrypto_amd import workloads as W
    L = ft.lib()
    nn, ne = 1 << args.cfg3_nodes_log2, 1 << args.cfg3_edges_log2
    log("[bench] cfg3: graph with 2^%d nodes, 2^%d edges" % (args.cfg3_nodes_log2, args.cfg3_edges_log2))
    row_off, cols, deg = W.power_law_graph(nn, ne, seed + 3)
    rng = np.random.default_rng(seed + 3)
    st = ft.Stream()
    T = Timer(hip, st.h.value)
    m = hip.upload((rng.random(nn) < 0.01).astype(np.uint64))
    r = hip.upload(ft.random_scalars(nn, seed, 30))
    rho = hip.upload(ft.random_scalars(nn, seed, 31))
    d_off, d_cols = hip.upload(row_off), hip.upload(cols)
    c1, c2, o1, o2 = (ft.GeArray.alloc(nn) for _ in range(4))
    fail = ft.DeviceWord()
    ft.check(L.ftx_encrypt_dev(c1.h, c2.h, key.h, m, r, 1, fail.p, st.h))
    st.synchronize()
    assert fail.read() == 0
    # scratch with room for the point-contiguous copy of the source
    scr = ft.GeArray.alloc(L.ftx_gather_scratch_points(ne, nn))
    scr_planes = ft.GeArray.alloc(L.ftx_segsum_scratch_points(ne))

    def step(sc=scr):
        ft.check(L.ftx_segmented_sum_dev(o1.h, d_off, nn, ne, d_cols, c1.h, sc.h, st.h))
        ft.check(L.ftx_segmented_sum_dev(o2.h, d_off, nn, ne, d_cols, c2.h, sc.h, st.h))
        ft.check(L.ftx_rerandomize_dev(o1.h, o2.h, key.h, rho, st.h))
    st.synchronize()
    with ft.Graph(st) as g:
        step()
    ms = T(lambda: g.launch(), reps=3)
    eager_ms = T(step, reps=3)
    planes_ms = T(lambda: step(scr_planes), reps=3)
    gather_ms = T(lambda: ft.check(L.ftx_segmented_sum_dev(o1.h, d_off, nn, ne, d_cols, c1.h, scr.h, st.h)), reps=3)
    gather_planes_ms = T(lambda: ft.check(L.ftx_segmented_sum_dev(o1.h, d_off, nn, ne, d_cols, c1.h,
                                                                   scr_planes.h, st.h)), reps=3)
    nonempty = int((np.diff(row_off) > 0).sum())
    adds = 2 * (ne - nonempty) + 2 * nn
    mults = 2 * nn
    return {"graph_ms_per_step": round(ms, 4), "eager_ms_per_step": round(eager_ms, 4),
            "eager_ms_per_step_gather_from_planes": round(planes_ms, 4),
            "gather_sum_ms": round(gather_ms, 4), "gather_sum_from_planes_ms": round(gather_planes_ms, 4),
            "point_adds": adds, "fixed_base_mults": mults,
            "valu_frac": (adds * PRODUCTS["point_add"] + mults * PRODUCTS["fixed_base"]) / (ms * 1e-3) / PEAK_MUL32}


def bench_cfg4(ft, key, seed, args):
    """Mixed request stream through the serving loop vs plain one-shot kernels."""
    from libftcrypto_amd import serving as SV
    log("[bench] cfg4: building a %d-request trace" % args.cfg4_requests)
    ops, size_of = SV.make_trace(args.cfg4_requests, seed + 4, max_live=64)
    pool, ops = SV.pin_inputs(ops)
    n_req = sum(1 for op in ops if op[0] != SV.FREE)
    elems = SV.trace_elements(ops)
    t1, vec = SV.run_oneshot(key, ops, size_of)
    log("[bench] cfg4: one-shot %.3f s; served burst" % t1)
    srv = SV.Server(key, budget_bytes=args.cfg4_budget_gib << 30, n_streams=args.cfg4_streams, borrow=True)
    t2, lat_b, hmap = SV.run_served(srv, ops)
    st = srv.stats()
    # byte-exact check of the served vectors left at the end against the
    # one-shot results: every surviving vector of at most 2^18 ciphertexts
    # (up to 48 of them) and the largest one
    freed = {op[1] for op in ops if op[0] == SV.FREE}
    live = sorted((size_of[v], v) for v in hmap if v not in freed)
    sample = [v for n, v in live if n <= 1 << 18][:48] + ([live[-1][1]] if live else [])
    same = 0
    for v in dict.fromkeys(sample):
        a, b = srv.read(hmap[v], size_of[v])
        same += int(np.array_equal(a, vec[v][0].to_bytes_folded()) and np.array_equal(b, vec[v][1].to_bytes_folded()))
    checked = len(dict.fromkeys(sample))
    srv.close()
    del vec
    if same != checked:
        raise SystemExit("cfg4: %d of %d served vectors differ from the one-shot results" % (checked - same, checked))
    cap = n_req / t1
    rate = args.cfg4_load * cap
    # request sizes in submission order, for the latency of short requests
    size = np.array([len(op[1]) if op[0] == SV.ENC else size_of[op[1]] if op[0] == SV.ADD else
                     len(op[2]) if op[0] == SV.MUL else int(op[2][-1]) for op in ops if op[0] != SV.FREE])
    short = size <= SV.DEFAULT_BATCH
    kinds = [("encrypt", "add", "mul_u32", "segsum")[op[0]] for op in ops if op[0] != SV.FREE]
    runs = []
    for j in range(2):                      # two arrival sequences: is the tail stable?
        log("[bench] cfg4: served burst %.3f s; Poisson at %.0f req/s (run %d)" % (t2, rate, j + 1))
        arr = SV.poisson_arrivals(ops, rate, seed + 44 + j)
        srv = SV.Server(key, budget_bytes=args.cfg4_budget_gib << 30, n_streams=args.cfg4_streams, borrow=True)
        info = {}
        t3, lat, _ = SV.run_served(srv, ops, arrivals=arr, info=info)
        srv.close()
        runs.append({"wall_s": round(t3, 4), "achieved_req_per_s": n_req / t3,
                     "last_arrival_s": round(float(arr[-1]), 4), "last_submit_s": round(info["last_submit_s"], 4),
                     "max_submit_lag_ms": round(info["submit_lag_ms"], 3),
                     "p50_ms": float(np.percentile(lat, 50)) / 1e3, "p99_ms": float(np.percentile(lat, 99)) / 1e3,
                     "max_ms": float(lat.max()) / 1e3,
                     "short_p99_ms": float(np.percentile(lat[short], 99)) / 1e3,
                     "short_queued_p99_ms": float(np.percentile(info["queued_us"][short], 99)) / 1e3,
                     "slowest": [[kinds[i], int(size[i]), round(float(lat[i]) / 1e3, 2),
                                  round(float(info["queued_us"][i]) / 1e3, 2)]
                                 for i in np.argsort(lat)[::-1][:12]]})
    p99 = [r["p99_ms"] for r in runs]
    return {"requests": n_req, "ciphertexts": elems, "mix": "40% encrypt, 30% add, 15% x u32, 15% segsum",
            "sizes": "lognormal(ln 2^14, 2) in [2^8, 2^22], power-of-two rounded",
            "oneshot_s": round(t1, 4), "oneshot_req_per_s": cap,
            "served_saturation_s": round(t2, 4), "served_vs_oneshot_throughput": t1 / t2,
            "served_vs_oneshot_target": 0.8, "served_vs_oneshot_met": bool(t1 / t2 >= 0.8),
            "batches_at_saturation": st["batches"],
            "byte_exact_vs_oneshot": {"vectors_checked": checked, "equal": same},
            "fixed_load": {"offered_req_per_s": rate, "load_vs_oneshot": args.cfg4_load,
                           "short": "requests of at most %d ciphertexts (%d of %d)" % (SV.DEFAULT_BATCH, short.sum(), n_req),
                           "runs": runs, "p99_spread": max(p99) / min(p99) - 1},
            "budget_gib": args.cfg4_budget_gib, "streams": args.cfg4_streams,
            "batch_cap": SV.DEFAULT_BATCH}


def cpu_baseline(n, k_fixed, k_var, s_var, a_u32):
    """The oracle on a bounded sample of the same step, all host cores."""
    from oracle import oracle as O
    flags = O.use_native()                 # built for this host's CPU when gcc can
    nt = O.threads()
    samp = max(256, min(n, 1 << 19))
    t = {}
    t0 = time.time()
    O.Array.from_scalars(k_fixed[:samp], nt=nt)
    t["fixed_base"] = time.time() - t0
    P = O.Array.from_scalars(s_var[:samp], nt=nt)
    t0 = time.time()
    P.mul(k_var[:samp], nt=nt)
    t["variable_base"] = time.time() - t0
    t0 = time.time()
    P.mul_u32(a_u32[:samp], nt=nt)
    t["mul_u32"] = (time.time() - t0) * 2          # two points per ciphertext
    total = sum(t.values())
    rate = 4 * samp / total
    # one thread, as the reference's serial CPU loops run
    s1 = max(64, samp >> 6)
    P = O.Array.from_scalars(s_var[:s1], nt=nt)
    t0 = time.time()
    O.Array.from_scalars(k_fixed[:s1], nt=1)
    t1 = time.time()
    P.mul(k_var[:s1], nt=1)
    t2 = time.time()
    P.mul_u32(a_u32[:s1], nt=1)
    t3 = time.time()
    one = (t1 - t0) + (t2 - t1) + 2 * (t3 - t2)
    return {"value": rate, "unit": "scalar-mults/s", "cores": nt, "kind": "port", "flags": flags,
            "sample": "%d of each of the N=2^k fixed-base, variable-base and u32 mults (ct = 2 point mults)" % samp,
            "seconds": round(total, 2),
            "one_thread": {"value": 4 * s1 / one, "sample": "%d of each kind" % s1, "seconds": round(one, 2)},
            "libsodium_one_thread": libsodium_rates(k_fixed, k_var)}


def cpu_config_rates(H, seed, configs):
    """The oracle (a C port of the reference CPU path's algorithm, all host
    cores) on bounded samples of configs 0, 2 and 3, in the units of the GPU
    lines, with the GPU / CPU ratio where the GPU line was measured."""
    from oracle import oracle as O
    from libftcrypto_amd import workloads as W
    import libftcrypto_amd as ft
    nt = O.threads()
    rng = np.random.default_rng(seed + 9)

    def timed(fn):
        t0 = time.time()
        fn()
        return time.time() - t0

    res = {"cores": nt, "kind": "port", "flags": O.flags()}
    # config 0: encrypt, ciphertext add, reduce
    n0 = 1 << 16
    m = rng.integers(0, 2, size=n0, dtype=np.uint64)
    r = ft.random_scalars(n0, seed, 40)
    box = {}
    t_enc = timed(lambda: box.update(a=O.encrypt(H, m, r, nt=nt)))
    a1, a2 = box["a"]
    b1, b2 = O.encrypt(H, m, r, nt=nt)
    t_add = timed(lambda: (a1.add(b1, nt=nt), a2.add(b2, nt=nt)))
    t_red = timed(lambda: (a1.reduce(nt=nt), a2.reduce(nt=nt)))
    c0 = {"sample": "%d ciphertexts" % n0, "encrypt_ciphertexts_per_s": n0 / t_enc,
          "ct_add_point_adds_per_s": 2 * n0 / t_add, "reduce_point_adds_per_s": 2 * (n0 - 1) / t_red}
    g0 = configs.get("cfg0_encrypt_add_reduce_n65536")
    if g0:
        c0["gpu_over_cpu"] = {"encrypt": g0["encrypt"]["ciphertexts_per_s"] / c0["encrypt_ciphertexts_per_s"],
                              "ct_add": g0["ct_add"]["point_adds_per_s"] / c0["ct_add_point_adds_per_s"],
                              "reduce": g0["reduce"]["point_adds_per_s"] / c0["reduce_point_adds_per_s"]}
    res["cfg0"] = c0
    # config 2: reduce, inclusive scan, segmented sum (lognormal, median 16)
    n2 = 1 << 21
    P = O.Array.from_scalars(ft.random_scalars(n2, seed, 41), nt=nt)
    off = W.lognormal_segments(n2 >> 5, n2, seed)
    t_r = timed(lambda: P.reduce(nt=nt))
    t_g = timed(lambda: O.segmented_sum(off, P, nt=nt))
    Q = P.copy()
    t_s = timed(lambda: Q.scan(nt=nt))
    c2 = {"sample": "2^21 points (one ciphertext half); segmented sum in 2^16 lognormal segments",
          "reduce_point_adds_per_s": (n2 - 1) / t_r, "segmented_sum_point_adds_per_s": (n2 - (n2 >> 5)) / t_g,
          "scan_point_adds_per_s": (n2 - 1) / t_s}
    g2 = configs.get("cfg2_aggregation")
    if g2:
        c2["gpu_over_cpu"] = {"reduce": g2["reduce"]["point_adds_per_s"] / c2["reduce_point_adds_per_s"],
                              "segmented_sum": g2["segmented_sum"]["point_adds_per_s"]
                              / c2["segmented_sum_point_adds_per_s"],
                              "scan": g2["scan"]["point_adds_per_s"] / c2["scan_point_adds_per_s"]}
    res["cfg2"] = c2
    # config 3: gather-sum of in-neighbours (both halves) and rerandomisation,
    # on a graph of the same shape (power law 2.1, 8 edges per node)
    nn, ne = 1 << 17, 1 << 20
    row_off, cols, _ = W.power_law_graph(nn, ne, seed + 3)
    S = O.Array.from_scalars(ft.random_scalars(nn, seed, 42), nt=nt)
    t_gs = 2 * timed(lambda: O.segmented_sum(row_off, S, idx=cols, nt=nt))
    e1, e2 = O.encrypt(H, np.zeros(nn, dtype=np.uint64), ft.random_scalars(nn, seed, 43), nt=nt)
    t_rr = timed(lambda: O.rerandomize(e1, e2, H, ft.random_scalars(nn, seed, 44), nt=nt))
    c3 = {"sample": "2^17 nodes, 2^20 edges", "nodes_per_s": nn / (t_gs + t_rr),
          "gather_s": round(t_gs, 3), "rerandomize_s": round(t_rr, 3)}
    g3 = configs.get("cfg3_sparse_step_graph")
    if g3:
        c3["gpu_over_cpu"] = ((1 << 22) / (g3["graph_ms_per_step"] * 1e-3)) / c3["nodes_per_s"]
    res["cfg3"] = c3
    return res


def libsodium_rates(k_fixed, k_var, count=2048):
    """libsodium 1.0.18's public API on one host thread (the reference CPU
    path's arithmetic library); every call decodes and encodes 32-byte points,
    so these are per-operation rates, not array throughput."""
    lib = None
    for path in ("/opt/conda/lib/libsodium.so.23", "/opt/conda/lib/libsodium.so"):
        try:
            lib = ctypes.CDLL(path)
            break
        except OSError:
            continue
    if lib is None or lib.sodium_init() < 0:
        return None
    out = ctypes.create_string_buffer(32)
    ks = [bytes(k_fixed[i]) for i in range(count)]
    kv = [bytes(k_var[i]) for i in range(count)]
    t0 = time.time()
    for k in ks:
        lib.crypto_scalarmult_ed25519_base_noclamp(out, k)
    t1 = time.time()
    pt = bytes(out.raw)
    for k in kv:
        lib.crypto_scalarmult_ed25519_noclamp(out, k, pt)
    t2 = time.time()
    return {"fixed_base_per_s": count / (t1 - t0), "variable_base_per_s": count / (t2 - t1),
            "sample": "%d of each" % count, "note": "per call: decode and encode of 32-byte points included"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=22, help="log2 of the array length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default aggregation schedules")
    ap.add_argument("--headline-only", action="store_true", help="skip the configurations other than the N=2^22 scalar mults")
    ap.add_argument("--sections", default="cfg0,cfg2,cfg3,isum,io,cfg4",
                    help="configurations measured after the headline step (comma list of cfg0,cfg2,cfg3,cfg4)")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="first run one plain copy kernel over the N points (N x 128 B read and written)")
    ap.add_argument("--cfg2-reduce-log2", type=int, default=26)
    ap.add_argument("--cfg2-log2", type=int, default=24)
    ap.add_argument("--cfg2-alone-only", action="store_true",
                    help="skip the concurrent run of the three aggregation jobs (per-kernel counters stay separate)")
    ap.add_argument("--isum-log2", type=int, default=22)
    ap.add_argument("--cfg3-nodes-log2", type=int, default=22)
    ap.add_argument("--cfg3-edges-log2", type=int, default=25)
    ap.add_argument("--cfg4-requests", type=int, default=1000)
    ap.add_argument("--cfg4-load", type=float, default=0.7)
    ap.add_argument("--cfg4-streams", type=int, default=8)
    ap.add_argument("--cfg4-budget-gib", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this library's hot path runs on a single device "
              "(sharding across GPUs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)
    res = run(args)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
