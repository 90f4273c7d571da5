"""Full-size runs of the benchmark configurations on the GPU, checked through
size-independent properties (exact closed forms of homomorphic sums) plus
byte-exact oracle comparisons on sampled elements.

Closed forms used (x = secret key, H = x B):
  sum_i Enc(m_i; r_i)           = Enc(sum m; sum r)
  prefix j of a scan            = Enc(sum_{i<=j} m; sum_{i<=j} r)
  segment s                     = Enc(sum_s m; sum_s r)
  sparse step, node v           = Enc(sum_{u->v} m_u; sum_{u->v} r_u + rho_v)
  sum_i k_i B                   = (sum k) B
  sum_i k_i (s_i B)             = (sum k s) B
"""
import ctypes
import os

import numpy as np
import pytest

import edref
import libftcrypto_amd as ft
from libftcrypto_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
L = edref.L


@pytest.fixture(scope="module", autouse=True)
def _init():
    ft.init()


def enc_k(k):
    return O.scalarmult_base(int.to_bytes(k % L, 32, "little"))


def reduce_enc(arr):
    """GPU reduce of a copy of arr -> encoding."""
    c = arr.get_range(0, len(arr), 1)
    c.reduce()
    return bytes(c.to_bytes_folded())


def keypair(seed):
    x = ft.random_scalars(1, seed, 99)
    pk = ft.GeArray.from_scalars(x)
    return int.from_bytes(bytes(x[0]), "little"), bytes(pk.to_bytes_folded()), ft.Key(pk)


def enc_expected(H, m_sum, r_sum):
    c1, c2 = O.encrypt(H, np.array([m_sum % 2**64], dtype=np.uint64),
                       np.frombuffer(int.to_bytes(r_sum % L, 32, "little"), dtype=np.uint8))
    # m_sum may exceed 2^64: add the remainder as a scalar multiple of B
    e1, e2 = bytes(c1.to_bytes()), bytes(c2.to_bytes())
    if m_sum >= 2**64:
        e2 = O.point_add(e2, enc_k(m_sum - m_sum % 2**64))
    return e1, e2


def test_config1_scalar_mults_full_size():
    n = 1 << 22
    seed = W.SEED_BASE + 1
    k = ft.random_scalars(n, seed, 1)
    g = ft.GeArray.from_scalars(k)
    assert reduce_enc(g) == enc_k(W.scalar_sum(k))
    idx = np.random.default_rng(1).choice(n, 512, replace=False)
    ge = g.to_bytes_folded().reshape(-1, 32)
    oe = O.Array.from_scalars(k[idx]).to_bytes().reshape(-1, 32)
    assert np.array_equal(ge[idx], oe)
    # variable base: k_i (s_i B)
    s = ft.random_scalars(n, seed, 2)
    p = ft.GeArray.from_scalars(s)
    p.mul(k)
    ks = W.weighted_sum(k, s)
    assert reduce_enc(p) == enc_k(ks)
    pe = p.to_bytes_folded().reshape(-1, 32)
    po = O.Array.from_scalars(s[idx])
    po.mul(k[idx])
    assert np.array_equal(pe[idx], po.to_bytes().reshape(-1, 32))


def test_config1_ct_times_u32_full_size():
    n = 1 << 22
    seed = W.SEED_BASE + 11
    x, H, key = keypair(seed)
    rng = np.random.default_rng(seed)
    m = rng.integers(0, 2, size=n, dtype=np.uint64)
    r = ft.random_scalars(n, seed, 1)
    a = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    c1, c2 = ft.encrypt(key, m, r, msg_bits=1)
    c1.mul_u32(a, c2)
    ar = W.weighted_sum(a, r)
    am = int((a.astype(np.uint64) * m).sum())
    assert reduce_enc(c1) == enc_k(ar)
    assert reduce_enc(c2) == enc_k(am + ar * x)


def test_config2_aggregation_full_size():
    seed = W.SEED_BASE + 2
    x, H, key = keypair(seed)
    rng = np.random.default_rng(seed)
    # reduce over 2^26 ciphertexts, messages in [0, 255]
    n = 1 << 26
    m = rng.integers(0, 256, size=n, dtype=np.uint64)
    r = ft.random_scalars(n, seed, 1)
    c1, c2 = ft.encrypt(key, m, r, msg_bits=8)
    del r
    w1, w2 = enc_expected(H, int(m.sum()), W.scalar_sum(ft.random_scalars(n, seed, 1)))
    c1.reduce()
    c2.reduce()
    assert bytes(c1.to_bytes_folded()) == w1
    assert bytes(c2.to_bytes_folded()) == w2
    del c1, c2, m
    # inclusive scan over 2^24 ciphertexts, sampled prefixes
    n = 1 << 24
    m = rng.integers(0, 256, size=n, dtype=np.uint64)
    r = ft.random_scalars(n, seed, 2)
    c1, c2 = ft.encrypt(key, m, r, msg_bits=8)
    c1.scan()
    c2.scan()
    cm = np.cumsum(m)
    cr = np.cumsum(W.limb_columns(r), axis=0, dtype=np.uint64)
    e1, e2 = c1.to_bytes_folded().reshape(-1, 32), c2.to_bytes_folded().reshape(-1, 32)
    for j in [0, 1, 15, 16, 4095, 4096, 65535, 1 << 20, n // 2 + 7, n - 1]:
        w1, w2 = enc_expected(H, int(cm[j]), W.limbs_to_int(cr[j]))
        assert bytes(e1[j]) == w1 and bytes(e2[j]) == w2, j
    del c1, c2, e1, e2, cr
    # segmented sum: 2^24 ciphertexts in 2^19 lognormal segments
    off = W.lognormal_segments(1 << 19, n, seed)
    c1, c2 = ft.encrypt(key, m, r, msg_bits=8)
    s1, s2 = ft.segmented_sum(off, c1), ft.segmented_sum(off, c2)
    assert reduce_enc(s1) == reduce_enc(c1) and reduce_enc(s2) == reduce_enc(c2)
    segm = np.add.reduceat(m, off[:-1])
    segr = np.add.reduceat(W.limb_columns(r), off[:-1], axis=0)
    lens = np.diff(off)
    pick = list(np.argsort(-lens)[:3]) + list(np.random.default_rng(3).choice(len(lens), 200, replace=False))
    e1, e2 = s1.to_bytes_folded().reshape(-1, 32), s2.to_bytes_folded().reshape(-1, 32)
    for s in pick:
        w1, w2 = enc_expected(H, int(segm[s]), W.limbs_to_int(segr[s]))
        assert bytes(e1[s]) == w1 and bytes(e2[s]) == w2, s


def test_config2_one_persistent_launch_full_size():
    """Config 2 as the benchmark runs it: both halves of the 2^26-ciphertext
    reduce, the 2^24 inclusive scan and the 2^24 segmented sum over 2^19
    lognormal CSR segments, six jobs in ONE ftx_aggregate_batch_dev call (the
    reduces at eight element steps per lane and tile).  Each output is
    checked against its closed form: the full sums, sampled scan prefixes
    (n - 1 included), the longest and sampled segments."""
    seed = W.SEED_BASE + 22
    x, H, key = keypair(seed)
    rng = np.random.default_rng(seed)
    nr, ns = 1 << 26, 1 << 24
    mr = rng.integers(0, 256, size=nr, dtype=np.uint64)
    r1, r2 = ft.encrypt(key, mr, ft.random_scalars(nr, seed, 1), msg_bits=8)
    ms = rng.integers(0, 256, size=ns, dtype=np.uint64)
    rs = ft.random_scalars(ns, seed, 2)
    s1, s2 = ft.encrypt(key, ms, rs, msg_bits=8)
    mg = rng.integers(0, 256, size=ns, dtype=np.uint64)
    rg = ft.random_scalars(ns, seed, 3)
    g1, g2 = ft.encrypt(key, mg, rg, msg_bits=8)
    off = W.lognormal_segments(1 << 19, ns, seed)
    d_off = DevBuf(off)
    nseg = len(off) - 1
    red1, red2 = ft.GeArray.alloc(1), ft.GeArray.alloc(1)
    seg1, seg2 = ft.GeArray.alloc(nseg), ft.GeArray.alloc(nseg)
    jobs = [ft.agg_job(ft.AGG_REDUCE, red1, r1), ft.agg_job(ft.AGG_REDUCE, red2, r2),
            ft.agg_job(ft.AGG_SCAN, s1), ft.agg_job(ft.AGG_SCAN, s2),
            ft.agg_job(ft.AGG_SEGSUM, seg1, g1, d_off.p, nseg, ns),
            ft.agg_job(ft.AGG_SEGSUM, seg2, g2, d_off.p, nseg, ns)]
    scratch = ft.GeArray.alloc(ft.aggregate_batch_scratch_points(jobs))
    ft.aggregate_batch(jobs, scratch)
    # reduces
    w1, w2 = enc_expected(H, int(mr.sum()), W.scalar_sum(ft.random_scalars(nr, seed, 1)))
    assert bytes(red1.to_bytes_folded()) == w1 and bytes(red2.to_bytes_folded()) == w2
    del r1, r2, mr, jobs
    # scans
    cm = np.cumsum(ms)
    cr = np.cumsum(W.limb_columns(rs), axis=0, dtype=np.uint64)
    e1, e2 = s1.to_bytes_folded().reshape(-1, 32), s2.to_bytes_folded().reshape(-1, 32)
    for j in [0, 1, 63, 64, 4095, 4096, 65535, 65536, 1 << 20, (1 << 20) + 17, ns // 2 + 7, ns - 2, ns - 1]:
        w1, w2 = enc_expected(H, int(cm[j]), W.limbs_to_int(cr[j]))
        assert bytes(e1[j]) == w1 and bytes(e2[j]) == w2, j
    del cr, e1, e2
    # segmented sums
    segm = np.add.reduceat(mg, off[:-1])
    segr = np.add.reduceat(W.limb_columns(rg), off[:-1], axis=0)
    lens = np.diff(off)
    pick = list(np.argsort(-lens)[:4]) + list(np.random.default_rng(5).choice(nseg, 300, replace=False))
    pick += [0, nseg - 1]
    e1, e2 = seg1.to_bytes_folded().reshape(-1, 32), seg2.to_bytes_folded().reshape(-1, 32)
    for sg in pick:
        w1, w2 = enc_expected(H, int(segm[sg]), W.limbs_to_int(segr[sg]))
        assert bytes(e1[sg]) == w1 and bytes(e2[sg]) == w2, sg
    assert reduce_enc(seg1) == reduce_enc(g1) and reduce_enc(seg2) == reduce_enc(g2)


def sparse_step(stream, out1, out2, c1, c2, d_off, n_nodes, n_edges, d_cols, key, d_rho, scratch):
    Lb = ft.lib()
    ft.check(Lb.ftx_segmented_sum_dev(out1.h, d_off, n_nodes, n_edges, d_cols, c1.h, scratch.h, stream.h))
    ft.check(Lb.ftx_segmented_sum_dev(out2.h, d_off, n_nodes, n_edges, d_cols, c2.h, scratch.h, stream.h))
    ft.check(Lb.ftx_rerandomize_dev(out1.h, out2.h, key.h, d_rho, stream.h))


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


@pytest.mark.parametrize("n_nodes,n_edges,contiguous", [(1 << 12, 1 << 15, False), (1 << 12, 1 << 15, True),
                                                        (1 << 22, 1 << 25, True)])
def test_config3_sparse_step_graph(n_nodes, n_edges, contiguous):
    """Gather-sum of in-neighbours then rerandomisation, captured in a hipGraph
    and replayed three times with eager work in between; every replay must
    equal the eager result, which must satisfy the closed form.  With
    `contiguous` the scratch also holds the point-contiguous source copy the
    gather then reads (one line per point); without, it gathers from the
    four coordinate planes."""
    seed = W.SEED_BASE + 3
    x, H, key = keypair(seed)
    rng = np.random.default_rng(seed)
    row_off, cols, deg = W.power_law_graph(n_nodes, n_edges, seed)
    m = (rng.random(n_nodes) < 0.01).astype(np.uint64)
    r = ft.random_scalars(n_nodes, seed, 1)
    rho = ft.random_scalars(n_nodes, seed, 2)
    c1, c2 = ft.encrypt(key, m, r, msg_bits=1)
    out1, out2 = ft.GeArray.alloc(n_nodes), ft.GeArray.alloc(n_nodes)
    Lb = ft.lib()
    scratch = ft.GeArray.alloc(Lb.ftx_gather_scratch_points(n_edges, n_nodes) if contiguous
                               else Lb.ftx_segsum_scratch_points(n_edges))
    d_off, d_cols, d_rho = DevBuf(row_off), DevBuf(cols), DevBuf(rho)
    st = ft.Stream()
    sparse_step(st, out1, out2, c1, c2, d_off.p, n_nodes, n_edges, d_cols.p, key, d_rho.p, scratch)
    st.synchronize()
    eager1, eager2 = out1.to_bytes_folded(), out2.to_bytes_folded()
    with ft.Graph(st) as g:
        sparse_step(st, out1, out2, c1, c2, d_off.p, n_nodes, n_edges, d_cols.p, key, d_rho.p, scratch)
    for rep in range(3):
        junk = ft.GeArray.from_small_scalars(np.arange(1000, dtype=np.uint64))   # eager work between replays
        junk.reduce()
        g.launch()
        st.synchronize()
        assert np.array_equal(out1.to_bytes_folded(), eager1), rep
        assert np.array_equal(out2.to_bytes_folded(), eager2), rep
    # closed form on sampled nodes, and on the total
    gm = np.add.reduceat(np.concatenate([m[cols], [0]]), row_off[:-1]) if n_edges else None
    rc = W.limb_columns(r)[cols]
    gr = np.add.reduceat(np.concatenate([rc, np.zeros((1, 8), dtype=np.uint64)]), row_off[:-1], axis=0)
    empty = np.diff(row_off) == 0
    gm[empty] = 0
    gr[empty] = 0
    rhoc = W.limb_columns(rho)
    e1, e2 = eager1.reshape(-1, 32), eager2.reshape(-1, 32)
    pick = list(np.nonzero(empty)[0][:3]) + list(np.random.default_rng(4).choice(n_nodes, 150, replace=False))
    pick.append(int(np.argmax(np.diff(row_off))))
    for v in pick:
        w1, w2 = enc_expected(H, int(gm[v]), W.limbs_to_int(gr[v]) + W.limbs_to_int(rhoc[v]))
        assert bytes(e1[v]) == w1 and bytes(e2[v]) == w2, v
    tm = int((m * deg.astype(np.uint64)).sum())
    tr = W.limbs_to_int((W.limb_columns(r) * deg.astype(np.uint64)[:, None]).sum(axis=0, dtype=np.uint64)) \
        + W.scalar_sum(rho)
    w1, w2 = enc_expected(H, tm, tr)
    assert reduce_enc(out1) == w1 and reduce_enc(out2) == w2


@pytest.mark.parametrize("n_src", [4097, 4098, 4099])
def test_gather_odd_sizes_unaligned_indices(n_src):
    """Gather sums through the _dev entry with a source length that is not a
    multiple of 4 and an index array starting 4 bytes past a 16-byte boundary
    (the kernel then loads indices one at a time), from the planes and from
    the point-contiguous copy; against the oracle."""
    rng = np.random.default_rng(n_src)
    n_rows = 700
    deg = rng.poisson(6, size=n_rows)
    deg[3] = 500
    deg[4] = 0
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    ne = int(off[-1])
    cols = rng.integers(0, n_src, size=ne, dtype=np.int32)
    s = ft.random_scalars(n_src, 4700 + n_src, 0)
    g, o = ft.GeArray.from_scalars(s), O.Array.from_scalars(s)
    want = O.segmented_sum(off, o, idx=cols).to_bytes()
    Lb = ft.lib()
    d_off = DevBuf(off)
    d_cols = DevBuf(np.concatenate([np.zeros(1, dtype=np.int32), cols]))     # cols start at byte 4
    cols_ptr = ctypes.c_void_p(d_cols.p.value + 4)
    st = ft.Stream()
    for sc_points in (Lb.ftx_segsum_scratch_points(ne), Lb.ftx_gather_scratch_points(ne, n_src)):
        scratch = ft.GeArray.alloc(sc_points)
        out = ft.GeArray.alloc(n_rows)
        ft.check(Lb.ftx_segmented_sum_dev(out.h, d_off.p, n_rows, ne, cols_ptr, g.h, scratch.h, st.h))
        st.synchronize()
        assert np.array_equal(out.to_bytes_folded(), want), sc_points
