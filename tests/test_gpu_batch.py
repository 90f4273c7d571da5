"""Batched aggregation (ftx_aggregate_batch_dev): reduce, scan and segmented-sum
jobs whose element walks share one persistent, load-balanced launch.

Every job's output is compared byte-exact, as canonical encodings, with the
oracle (small sizes) or with the same job run alone through the
reference-signature calls (ft_array_reduce, ft_array_scan, ft_array_prescan,
ftx_segmented_sum) at sizes where the oracle would take minutes.  Covered:
CSR segments with empty, single-element and chunk-spanning segments, uniform
segments (the form a reduce or a scan's first level takes), short scans that
skip the persistent phase, all 32 job slots, graph capture, and the host-side
refusals (too many jobs, aliased outputs, wrong uniform segment count,
undersized scratch or output).
"""
import ctypes

import numpy as np
import pytest

import edref
import libftcrypto_amd as ft
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    ft.init()


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


def points(n, seed):
    s = ft.random_scalars(n, 5100 + seed, seed)
    return s, ft.GeArray.from_scalars(s)


def ragged_offsets(rng, n_seg, extra_long=()):
    lens = rng.poisson(5, size=n_seg)
    lens[::7] = 0                                   # empty segments
    lens[1::11] = 1
    for i, ln in extra_long:
        lens[i] = ln
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)


def enc(arr, n=None):
    out = np.asarray(arr.to_bytes_folded())
    return out if n is None else out[: 32 * n]


def test_batch_mixed_jobs_match_oracle():
    """Two CSR segmented sums (one with a 3000-element segment spanning many
    chunks), a uniform one, two reduces, an inclusive and an exclusive scan
    above the block-scan limit (their first levels run in the persistent
    launch) and a short exclusive scan, in one call."""
    rng = np.random.default_rng(11)
    sA, A = points(9000, 1)
    sB, B = points(70001, 2)
    sC, C = points(5000, 3)
    oA, oB, oC = O.Array.from_scalars(sA), O.Array.from_scalars(sB), O.Array.from_scalars(sC)
    offA = ragged_offsets(rng, 900, extra_long=[(5, 3000)])
    assert offA[-1] <= 9000
    offC = ragged_offsets(rng, 400)
    assert offC[-1] <= 5000
    dA, dC = DevBuf(offA), DevBuf(offC)
    segA, segC = ft.GeArray.alloc(900), ft.GeArray.alloc(400)
    uni = ft.GeArray.alloc((5000 + 36) // 37)
    redA, redB = ft.GeArray.alloc(1), ft.GeArray.alloc(1)
    scanB = B.get_range(0, len(B), 1)
    xscanB = B.get_range(0, 66000, 1)
    xscanC = C.get_range(0, 3000, 1)
    jobs = [
        ft.agg_job(ft.AGG_SEGSUM, segA, A, dA.p, 900, int(offA[-1])),
        ft.agg_job(ft.AGG_REDUCE, redA, A),
        ft.agg_job(ft.AGG_SEGSUM, segC, C, dC.p, 400, int(offC[-1])),
        ft.agg_job(ft.AGG_SCAN, scanB),
        ft.agg_job(ft.AGG_SEGSUM, uni, C, None, (5000 + 36) // 37, 5000, 37),
        ft.agg_job(ft.AGG_REDUCE, redB, B),
        ft.agg_job(ft.AGG_XSCAN, xscanC),
        ft.agg_job(ft.AGG_XSCAN, xscanB),
    ]
    ft.aggregate_batch(jobs)
    assert np.array_equal(enc(segA), O.segmented_sum(offA, oA).to_bytes())
    assert np.array_equal(enc(segC), O.segmented_sum(offC, oC).to_bytes())
    uoff = np.minimum(np.arange(0, 5000 + 37, 37), 5000).astype(np.int64)
    assert np.array_equal(enc(uni), O.segmented_sum(uoff, oC).to_bytes())
    assert np.array_equal(enc(redA), oA.reduce())
    assert np.array_equal(enc(redB), oB.reduce())
    o = oB.copy()
    o.scan()
    assert np.array_equal(enc(scanB), o.to_bytes())
    o = O.Array.from_scalars(sC[:3000])
    o.scan(exclusive=True)
    assert np.array_equal(enc(xscanC), o.to_bytes())
    o = O.Array.from_scalars(sB[:66000])
    o.scan(exclusive=True)
    assert np.array_equal(enc(xscanB), o.to_bytes())


# element steps per lane r of a batched reduce's tiles: the smallest r <= 8
# with n / 2^(6 + r) <= 4096 tiles; up to 2^18 elements no tiles (summed
# straight from the source after the launch)
REDUCE_DEPTH_SIZES = [(1 << 18) + 1, 1 << 19, (1 << 20) + 3, 1 << 21, (1 << 22) - 1, (1 << 23) + 5, 1 << 24,
                      (1 << 25) - 64, 1 << 26]


def reduce_depth(n):
    if n <= 64 * 4096:
        return None
    r = 0
    while r < 8 and (n >> (6 + r)) > 4096:
        r += 1
    return r


def test_reduce_depth_sizes_cover_every_tile_depth():
    assert sorted({reduce_depth(n) for n in REDUCE_DEPTH_SIZES}) == list(range(9))


@pytest.mark.parametrize("n", REDUCE_DEPTH_SIZES)
def test_batched_reduce_every_tile_depth(n):
    """A batched reduce at every tile depth (r = 0 .. 8, ragged last tiles
    included) against its closed form (sum k) B and the single-call reduce"""
    # k < 2^32, so the sum of at most 2^26 of them is exact in uint64
    k = (np.arange(n, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(n)) % np.uint64(1 << 32)
    X = ft.GeArray.from_small_scalars(k)
    out = ft.GeArray.alloc(1)
    ft.aggregate_batch([ft.agg_job(ft.AGG_REDUCE, out, X)])
    want = O.scalarmult_base(int.to_bytes(int(k.sum(dtype=np.uint64)) % edref.L, 32, "little"))
    assert bytes(out.to_bytes_folded()) == want, (n, reduce_depth(n))
    X.reduce()
    assert bytes(X.to_bytes_folded()) == want, n


def test_batched_reduces_of_mixed_depths_in_one_launch():
    """reduces of several tile depths and two direct ones share a launch"""
    sizes = [100, (1 << 18) - 3, (1 << 19) + 11, (1 << 21) + 7, (1 << 23) + 5]
    arrs, outs, wants = [], [], []
    for q, n in enumerate(sizes):
        k = (np.arange(n, dtype=np.uint64) * np.uint64(40503 + q) + np.uint64(q)) % np.uint64(1 << 32)
        arrs.append(ft.GeArray.from_small_scalars(k))
        outs.append(ft.GeArray.alloc(1))
        wants.append(O.scalarmult_base(int.to_bytes(int(k.sum(dtype=np.uint64)) % edref.L, 32, "little")))
    ft.aggregate_batch([ft.agg_job(ft.AGG_REDUCE, o, a) for o, a in zip(outs, arrs)])
    for o, w, n in zip(outs, wants, sizes):
        assert bytes(o.to_bytes_folded()) == w, n


def test_batch_full_size_matches_single_jobs():
    """2^22-element reduce, 2^22-element inclusive scan and a segmented sum
    of 2^22 elements in 2^17 heavy-tailed segments, batched, equal to the
    same jobs run alone."""
    n = 1 << 22
    rng = np.random.default_rng(22)
    _, X = points(n, 4)
    lens = np.clip(np.round(rng.lognormal(np.log(16), 1.5, size=1 << 17)), 0, 65536).astype(np.int64)
    lens = np.minimum(lens, np.maximum(0, n - np.concatenate([[0], np.cumsum(lens)[:-1]])))
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    assert off[-1] <= n
    d_off = DevBuf(off)
    seg = ft.GeArray.alloc(1 << 17)
    red = ft.GeArray.alloc(1)
    scan = X.get_range(0, n, 1)
    ft.aggregate_batch([ft.agg_job(ft.AGG_REDUCE, red, X), ft.agg_job(ft.AGG_SCAN, scan),
                        ft.agg_job(ft.AGG_SEGSUM, seg, X, d_off.p, 1 << 17, int(off[-1]))])
    want_seg = ft.segmented_sum(off, X)
    assert np.array_equal(enc(seg), enc(want_seg))
    r = X.get_range(0, n, 1)
    r.reduce()
    assert np.array_equal(enc(red), enc(r))
    s = X.get_range(0, n, 1)
    s.scan()
    assert np.array_equal(enc(scan), enc(s))


def test_batch_all_job_slots():
    """32 segmented sums of sizes 1 .. 2^15 (CSR and uniform alternating) over
    one source: every slot of the job table, each equal to the job alone."""
    rng = np.random.default_rng(32)
    _, X = points(1 << 15, 5)
    jobs, outs, wants, keep = [], [], [], []
    for j in range(ft.AGG_MAX_JOBS):
        ne = int(rng.integers(1, 1 << 15)) if j else 1
        if j % 2:
            off = ragged_offsets(rng, max(1, ne // 6))
            off = np.minimum(off, ne)
            d = DevBuf(off)
            keep.append(d)
            out = ft.GeArray.alloc(len(off) - 1)
            jobs.append(ft.agg_job(ft.AGG_SEGSUM, out, X, d.p, len(off) - 1, int(off[-1])))
        else:
            ln = int(rng.integers(1, 300))
            nseg = (ne + ln - 1) // ln
            off = np.minimum(np.arange(nseg + 1, dtype=np.int64) * ln, ne)
            out = ft.GeArray.alloc(nseg)
            jobs.append(ft.agg_job(ft.AGG_SEGSUM, out, X, None, nseg, ne, ln))
        outs.append(out)
        wants.append(enc(ft.segmented_sum(off, X)))
    ft.aggregate_batch(jobs)
    for j, (o, w) in enumerate(zip(outs, wants)):
        assert np.array_equal(enc(o), w), j


def test_batch_graph_replays():
    """A batch captured in a hipGraph with caller scratch (the scanned array
    regenerated from device scalars first, so every replay starts from the
    same input), replayed three times with eager work in between: every
    replay equals the eager call, whose scans ran on forked streams."""
    rng = np.random.default_rng(3)
    _, X = points(100000, 6)
    off = ragged_offsets(rng, 15000, extra_long=[(10, 20000)])
    assert off[-1] <= 100000
    d_off = DevBuf(off)
    ks = ft.random_scalars(80000, 61, 6)
    d_ks = DevBuf(ks)
    seg, red, scan = ft.GeArray.alloc(15000), ft.GeArray.alloc(1), ft.GeArray.alloc(80000)
    jobs = [ft.agg_job(ft.AGG_SEGSUM, seg, X, d_off.p, 15000, int(off[-1])), ft.agg_job(ft.AGG_REDUCE, red, X),
            ft.agg_job(ft.AGG_SCAN, scan)]
    scratch = ft.GeArray.alloc(ft.aggregate_batch_scratch_points(jobs))
    st = ft.Stream()
    Lb = ft.lib()

    def step():
        ft.check(Lb.ftx_from_scalars_dev(scan.h, d_ks.p, st.h))
        ft.aggregate_batch(jobs, scratch, st)

    step()
    st.synchronize()
    want_seg, want_red, want_scan = enc(seg), enc(red), enc(scan)
    assert np.array_equal(want_seg, enc(ft.segmented_sum(off, X)))
    o = O.Array.from_scalars(ks)
    o.scan()
    assert np.array_equal(want_scan, o.to_bytes())
    with ft.Graph(st) as g:
        step()
    junk = ft.GeArray.from_scalars(ft.random_scalars(15000, 9, 9))
    for _ in range(3):
        # eager work that overwrites both outputs before each replay
        ft.aggregate_batch([ft.agg_job(ft.AGG_SEGSUM, seg, junk, None, 15000, 15000, 1),
                            ft.agg_job(ft.AGG_REDUCE, red, junk)])
        assert not np.array_equal(enc(red), want_red)
        g.launch()
        st.synchronize()
        assert np.array_equal(enc(seg), want_seg)
        assert np.array_equal(enc(red), want_red)
        assert np.array_equal(enc(scan), want_scan)


def test_batch_refusals():
    _, X = points(1000, 7)
    off = np.array([0, 10, 20], dtype=np.int64)
    d = DevBuf(off)
    out = ft.GeArray.alloc(2)
    Lb = ft.lib()

    def rc(jobs, scratch=None):
        arr = (ft.AggJob * max(1, len(jobs)))(*jobs)
        r = Lb.ftx_aggregate_batch_dev(arr, len(jobs), scratch.h if scratch else None, None)
        Lb.ftx_synchronize(None)
        return r

    seg = ft.agg_job(ft.AGG_SEGSUM, out, X, d.p, 2, 20)
    BAD_SIZE, OUT_OF_BOUNDS, INSUFFICIENT_SPACE = 3, 4, 7
    before = enc(out)
    assert rc([seg] * (ft.AGG_MAX_JOBS + 1)) == BAD_SIZE
    assert rc([seg, ft.agg_job(ft.AGG_REDUCE, out, X)]) == BAD_SIZE          # same output twice
    assert rc([ft.agg_job(ft.AGG_REDUCE, X, X), seg]) == BAD_SIZE            # output is another job's input
    out3 = ft.GeArray.alloc(3)
    assert rc([ft.agg_job(ft.AGG_SEGSUM, out3, X, None, 3, 20, 10)]) == BAD_SIZE  # ceil(20/10) != 3
    assert rc([ft.agg_job(7, out, X)]) == BAD_SIZE                           # unknown kind
    assert rc([ft.agg_job(ft.AGG_SEGSUM, X, X, d.p, 2, 20)]) == BAD_SIZE     # summed in place
    assert rc([ft.agg_job(ft.AGG_SEGSUM, out, X, d.p, 5, 20)]) == INSUFFICIENT_SPACE   # output too short
    assert rc([ft.agg_job(ft.AGG_SEGSUM, out, X, d.p, 2, 2000)]) == OUT_OF_BOUNDS      # beyond len(src)
    assert rc([seg], ft.GeArray.alloc(1)) == INSUFFICIENT_SPACE              # scratch too small
    assert np.array_equal(enc(out), before)                                  # nothing written
    assert rc([]) == 0


def test_batch_from_several_threads():
    """Four host threads, each on its own stream and arrays, call batches
    whose scans fork onto the library's auxiliary streams: every result
    equals the same batch run alone."""
    import threading
    rng = np.random.default_rng(44)
    cases = []
    for t in range(4):
        ks = ft.random_scalars(70000 + 1000 * t, 71, t)
        _, X = points(20000, 20 + t)
        off = ragged_offsets(rng, 3000)
        assert off[-1] <= 20000
        cases.append((ks, X, off, DevBuf(ks), DevBuf(off)))
    Lb = ft.lib()

    def run(case, st):
        ks, X, off, d_ks, d_off = case
        scan, seg = ft.GeArray.alloc(len(ks)), ft.GeArray.alloc(len(off) - 1)
        short = X.get_range(0, 100, 1)
        ft.check(Lb.ftx_from_scalars_dev(scan.h, d_ks.p, st.h))
        ft.aggregate_batch([ft.agg_job(ft.AGG_SCAN, scan), ft.agg_job(ft.AGG_XSCAN, short),
                            ft.agg_job(ft.AGG_SEGSUM, seg, X, d_off.p, len(off) - 1, int(off[-1]))], None, st)
        st.synchronize()
        return enc(scan), enc(seg)

    st0 = ft.Stream()
    want = [run(c, st0) for c in cases]
    got = [None] * 4
    errs = []

    def worker(t):
        try:
            st = ft.Stream()
            for _ in range(3):
                got[t] = run(cases[t], st)
                assert np.array_equal(got[t][0], want[t][0]) and np.array_equal(got[t][1], want[t][1])
        except Exception as e:          # reported below
            errs.append((t, e))

    th = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errs, errs


@pytest.mark.parametrize("lens", [(1,), (3,), (1, 2), (5, 1, 1, 64)])
def test_segment_ends_at_every_lane(lens):
    """Segments of lengths cycling through `lens` over 70,000 elements k_i B
    (small k_i): many lanes of a wave end a segment at the same element
    (all 64 for length 1), more than a wave's flush queue holds, so queued
    and directly stored sums mix.  CSR through ftx_segmented_sum and as a
    batched job, and uniform segments as a batched job, equal the fixed-base
    multiples of the segment sums of the k_i."""
    n = 70000
    k = (np.arange(n, dtype=np.uint64) * np.uint64(7919) + np.uint64(3)) % np.uint64(1 << 24)
    X = ft.GeArray.from_small_scalars(k)
    reps = -(-n // sum(lens)) * len(lens)
    ln = np.resize(np.array(lens, dtype=np.int64), reps)
    off = np.concatenate([[0], np.cumsum(ln)])
    off = np.concatenate([off[off < n], [n]]).astype(np.int64)
    want = ft.GeArray.from_small_scalars(np.add.reduceat(k, off[:-1])).to_bytes_folded()
    got = ft.segmented_sum(off, X)
    assert np.array_equal(got.to_bytes_folded(), want)
    d = DevBuf(off)
    out = ft.GeArray.alloc(len(off) - 1)
    jobs = [ft.agg_job(ft.AGG_SEGSUM, out, X, d.p, len(off) - 1, n)]
    if len(lens) == 1:
        u = lens[0]
        outu = ft.GeArray.alloc(-(-n // u))
        jobs.append(ft.agg_job(ft.AGG_SEGSUM, outu, X, None, -(-n // u), n, u))
    ft.aggregate_batch(jobs)
    assert np.array_equal(out.to_bytes_folded(), want)
    if len(lens) == 1:
        wu = ft.GeArray.from_small_scalars(np.add.reduceat(k, np.arange(0, n, lens[0]))).to_bytes_folded()
        assert np.array_equal(outu.to_bytes_folded(), wu)
