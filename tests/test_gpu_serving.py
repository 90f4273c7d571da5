"""The serving loop (ftx_server): every result is the same computed in the
mixed, batched, multi-stream stream as computed alone by plain one-shot
kernels, and matches the oracle; admission under a tight HBM budget still
completes every request; bad requests are refused with the reference's error
codes."""
import numpy as np
import pytest

import libftcrypto_amd as ft
from libftcrypto_amd import serving as SV
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def key():
    ft.init()
    x = ft.random_scalars(1, 4242, 0)
    pk = ft.GeArray.from_scalars(x)
    return ft.Key(pk), bytes(pk.to_bytes_folded())


def compare_all(server, hmap, vec, size_of, freed):
    checked = 0
    for v, h in hmap.items():
        if v in freed:
            continue
        n = size_of[v]
        a, b = server.read(h, n)
        assert np.array_equal(a, vec[v][0].to_bytes_folded()), v
        assert np.array_equal(b, vec[v][1].to_bytes_folded()), v
        checked += 1
    return checked


@pytest.mark.parametrize("streams,max_batch,budget,max_live",
                         [(4, 1 << 22, 1 << 30, 10**9), (2, 3000, 1 << 30, 10**9), (3, 1 << 22, 6 << 20, 6)])
def test_mixed_stream_equals_alone(key, streams, max_batch, budget, max_live):
    k, H = key
    ops, size_of = SV.make_trace(120, seed=streams * 7 + max_batch % 97, mu_log2=9, sigma=1.0, lo_log2=8,
                                 hi_log2=12, max_live=max_live)
    freed = {op[1] for op in ops if op[0] == SV.FREE}
    srv = SV.Server(k, budget_bytes=budget, n_streams=streams, max_batch=max_batch)
    wall, lats, hmap = SV.run_served(srv, ops)
    assert len(lats) == sum(1 for op in ops if op[0] != SV.FREE)
    assert (lats > 0).all()
    _, vec = SV.run_oneshot(k, ops, size_of)
    assert compare_all(srv, hmap, vec, size_of, freed) == len(hmap) - len(freed)
    st = srv.stats()
    assert st["batches"] <= st["completed"]
    srv.close()


def unmodified_encryptions(ops, freed):
    modified = {o[1] for o in ops if o[0] in (SV.ADD, SV.MUL)}
    return [op for op in ops if op[0] == SV.ENC and op[3] not in freed and op[3] not in modified]


def slices(ops, cap):
    """batches the sliced requests of a trace take at a batch cap of `cap`"""
    tot = 0
    for op in ops:
        n = len(op[1]) if op[0] == SV.ENC else len(op[2]) if op[0] == SV.MUL else 0
        if n > cap:
            tot += -(-n // cap)
    return tot


@pytest.mark.parametrize("streams", [1, 3])
def test_sliced_requests_equal_alone(key, streams):
    """A batch cap of 1,000 ciphertexts: encryptions and multiplies of 1,024 to
    8,192 ciphertexts run as slices of 1,000 (the last one short), one at a
    time on one stream or two in flight on three, interleaved with the other
    requests and followed by adds, multiplies and segmented sums of the same
    vectors.  Every result equals the one-shot kernels', and sliced
    encryptions equal the oracle's."""
    k, H = key
    ops, size_of = SV.make_trace(100, seed=60 + streams, mu_log2=11, sigma=1.0, lo_log2=8, hi_log2=13,
                                 max_live=40)
    freed = {op[1] for op in ops if op[0] == SV.FREE}
    n_slices = slices(ops, 1000)
    assert n_slices > 150
    srv = SV.Server(k, budget_bytes=1 << 30, n_streams=streams, max_batch=1000)
    wall, lats, hmap = SV.run_served(srv, ops)
    assert len(lats) == sum(1 for op in ops if op[0] != SV.FREE)
    st = srv.stats()
    assert st["completed"] == len(lats)
    assert st["batches"] >= n_slices
    _, vec = SV.run_oneshot(k, ops, size_of)
    assert compare_all(srv, hmap, vec, size_of, freed) == len(hmap) - len(freed)
    enc = [op for op in unmodified_encryptions(ops, freed) if len(op[1]) > 1000]
    assert len(enc) >= 2
    for op in enc[:2]:
        a, b = srv.read(hmap[op[3]], len(op[1]))
        o1, o2 = O.encrypt(H, op[1], op[2])
        assert np.array_equal(a, o1.to_bytes()) and np.array_equal(b, o2.to_bytes())
    srv.close()


def test_config4_mix_full_size(key):
    """The serving workload at its stated sizes: 200 requests of the 40 % encrypt
    / 30 % add / 15 % x u32 / 15 % segmented-sum mix, lognormal sizes in
    [2^8, 2^22] ciphertexts (this trace holds two 2^22-ciphertext encryptions
    and a 2^22 multiply, so slicing runs at full length, and a 2^20 add), a
    64 GiB budget, four streams, borrowed pinned inputs, the
    default batch cap.  Every vector left at the end equals the one-shot
    kernels' result byte for byte, and three served encryptions (one of them
    sliced) equal the oracle's."""
    k, H = key
    ops, size_of = SV.make_trace(200, seed=5, max_live=64)
    pool, ops = SV.pin_inputs(ops)
    freed = {op[1] for op in ops if op[0] == SV.FREE}
    assert max(size_of.values()) == 1 << 22
    assert slices(ops, 1 << 18) >= 48
    srv = SV.Server(k, budget_bytes=64 << 30, n_streams=4, borrow=True)
    wall, lats, hmap = SV.run_served(srv, ops)
    assert len(lats) == sum(1 for op in ops if op[0] != SV.FREE)
    _, vec = SV.run_oneshot(k, ops, size_of)
    live = len(hmap) - len(freed)
    assert live > 0
    assert compare_all(srv, hmap, vec, size_of, freed) == live
    del vec
    enc = sorted(unmodified_encryptions(ops, freed), key=lambda op: len(op[1]))
    picks = enc[:2] + [op for op in enc if len(op[1]) > 1 << 18][:1]
    assert len(picks) == 3 and len(picks[2][1]) > 1 << 18
    for op in picks:
        a, b = srv.read(hmap[op[3]], len(op[1]))
        o1, o2 = O.encrypt(H, op[1], op[2])
        assert np.array_equal(a, o1.to_bytes()) and np.array_equal(b, o2.to_bytes()), len(op[1])
    srv.close()
    del pool


def test_releases_and_oracle(key):
    k, H = key
    ops, size_of = SV.make_trace(150, seed=77, mu_log2=9, sigma=1.0, lo_log2=8, hi_log2=11, max_live=12)
    freed = {op[1] for op in ops if op[0] == SV.FREE}
    assert freed
    srv = SV.Server(k, budget_bytes=64 << 20, n_streams=4)
    wall, lats, hmap = SV.run_served(srv, ops)
    _, vec = SV.run_oneshot(k, ops, size_of)
    assert compare_all(srv, hmap, vec, size_of, freed) > 0
    # an encrypt served in the mix equals the oracle's encryption
    modified = {o[1] for o in ops if o[0] in (SV.ADD, SV.MUL)}
    enc = [op for op in ops if op[0] == SV.ENC and op[3] not in freed and op[3] not in modified]
    for op in enc[:3]:
        a, b = srv.read(hmap[op[3]], len(op[1]))
        o1, o2 = O.encrypt(H, op[1], op[2])
        assert np.array_equal(a, o1.to_bytes()) and np.array_equal(b, o2.to_bytes())
    srv.close()


def test_bad_requests(key):
    k, _ = key
    srv = SV.Server(k, budget_bytes=1 << 24, n_streams=2)
    m = np.zeros(300, dtype=np.uint64)
    r = ft.random_scalars(300, 5, 5)
    q1, h1 = srv.encrypt(m, r, 1)
    q2, h2 = srv.encrypt(m[:200].copy(), r[:200].copy(), 1)
    srv.wait(q1)
    srv.wait(q2)
    with pytest.raises(ft.FtError) as e:
        srv.add(h1, h2)
    assert e.value.code == ft.FT_ERR_SIZE_MISMATCH
    with pytest.raises(ft.FtError) as e:
        srv.add(h1, 999999)
    assert e.value.code == ft.FT_ERR_OUT_OF_BOUNDS
    # a vector larger than the whole budget is refused, not queued forever
    big = 1 << 17
    q3, h3 = srv.encrypt(np.zeros(big, dtype=np.uint64), ft.random_scalars(big, 6, 6), 1)
    with pytest.raises(ft.FtError) as e:
        srv.wait(q3)
    assert e.value.code == ft.FT_ERR_OUT_OF_MEMORY
    srv.close()


def test_segsum_requests_batched(key):
    """24 segmented-sum requests over 24 distinct resident vectors, queued
    together: the scheduler runs them as jobs of shared persistent launches
    (fewer batches than requests), and each result equals the one-shot
    ftx_segmented_sum of the same ciphertexts."""
    k, H = key
    rng = np.random.default_rng(2024)
    # a batch cap above 2^20, so that the two large encryptions below run
    # whole (not sliced) and hold both streams
    srv = SV.Server(k, budget_bytes=1 << 30, n_streams=2, max_batch=1 << 21)
    srcs = []
    for i in range(24):
        n = int(rng.integers(300, 6000))
        m = rng.integers(0, 256, size=n, dtype=np.uint64)
        r = ft.random_scalars(n, 900 + i, i)
        q, h = srv.encrypt(m, r, 8)
        srv.wait(q)
        off = SV.lognormal_segments(max(1, n // 16), n, 31 + i)
        srcs.append((h, off, ft.encrypt(k, m, r, msg_bits=8)))
    # two large encryptions occupy both streams, so the segmented sums
    # submitted next are all queued when a stream frees up
    big = 1 << 20
    big_r = [ft.random_scalars(big, 77, j) for j in range(2)]
    busy_q = [srv.encrypt(np.zeros(big, dtype=np.uint64), big_r[j], 1) for j in range(2)]
    busy = [q for q, _ in busy_q]
    before = srv.stats()["batches"]
    reqs = [srv.segsum(h, off) for h, off, _ in srcs]
    for q in busy:
        srv.wait(q)
    for q, _ in reqs:
        srv.wait(q)
    batches = srv.stats()["batches"] - before
    assert batches <= 4, batches
    for (q, h), (_, off, (c1, c2)) in zip(reqs, srcs):
        a, b = srv.read(h, len(off) - 1)
        assert np.array_equal(a, ft.segmented_sum(off, c1).to_bytes_folded())
        assert np.array_equal(b, ft.segmented_sum(off, c2).to_bytes_folded())
    # a large served encryption (two lanes per ciphertext, 512-lane blocks
    # under the default schedule) equals the oracle's
    a, b = srv.read(busy_q[0][1], big)
    o1, o2 = O.encrypt(H, np.zeros(big, dtype=np.uint64), big_r[0])
    assert np.array_equal(a, o1.to_bytes()) and np.array_equal(b, o2.to_bytes())
    srv.close()


@pytest.mark.parametrize("lanes", [1, 2])
def test_served_encrypt_both_lane_schedules(key, lanes):
    """Served encryptions of 1,000 and 300,000 ciphertexts (one batch each)
    with one lane per ciphertext (256- and 512-lane-block batch kernels) and
    with two, against the oracle"""
    k, H = key
    before = ft.ct_lanes_table()
    try:
        for c in range(4):
            ft.tune_set_ct_lanes(c, lanes)
        srv = SV.Server(k, budget_bytes=1 << 30, n_streams=2)
        rng = np.random.default_rng(lanes)
        for n in (1000, 300000):
            m = rng.integers(0, 256, size=n, dtype=np.uint64)
            r = ft.random_scalars(n, 930 + lanes, n)
            q, h = srv.encrypt(m, r, 8)
            srv.wait(q)
            a, b = srv.read(h, n)
            o1, o2 = O.encrypt(H, m, r)
            assert np.array_equal(a, o1.to_bytes()) and np.array_equal(b, o2.to_bytes()), n
        srv.close()
    finally:
        for c, v in before.items():
            ft.tune_set_ct_lanes(c, v)


def test_destroy_with_unadmittable_request(key):
    """a create request that fits the budget but not the free arena while the
    client keeps its vectors: destroy fails it instead of waiting forever"""
    import threading
    k, _ = key
    srv = SV.Server(k, budget_bytes=1 << 20, n_streams=2)             # 4096 ciphertexts
    q1, h1 = srv.encrypt(np.zeros(3000, dtype=np.uint64), ft.random_scalars(3000, 8, 1), 1)
    srv.wait(q1)
    srv.encrypt(np.zeros(2000, dtype=np.uint64), ft.random_scalars(2000, 8, 2), 1)   # cannot be admitted
    done = threading.Event()

    def close():
        srv.close()
        done.set()

    t = threading.Thread(target=close, daemon=True)
    t.start()
    assert done.wait(30.0), "ftx_server_destroy did not return"
    t.join()


def test_read_ordered_against_writer(key):
    """reads of a vector while another thread multiplies it by 2 request
    after request: each read returns one whole state 2^j C (both halves the
    same j, j never decreasing), never a mix of states"""
    import threading
    k, H = key
    n, steps = 2048, 12
    m = np.arange(n, dtype=np.uint64) % np.uint64(256)
    r = ft.random_scalars(n, 8, 3)
    srv = SV.Server(k, budget_bytes=1 << 26, n_streams=4)
    q, h = srv.encrypt(m, r, 8)
    srv.wait(q)
    o1, o2 = O.encrypt(H, m, r)
    states = []
    for j in range(steps + 1):
        kk = np.full(n, 1 << j, dtype=np.uint32)
        a, b = o1.copy(), o2.copy()
        a.mul_u32(kk)
        b.mul_u32(kk)
        states.append((bytes(a.to_bytes()), bytes(b.to_bytes())))
    index = {s: j for j, s in enumerate(states)}
    two = np.full(n, 2, dtype=np.uint32)
    errors = []

    def writer():
        try:
            for _ in range(steps):
                srv.wait(srv.mul_u32(h, two))
        except Exception as e:                       # reported by the main thread
            errors.append(e)

    t = threading.Thread(target=writer)
    t.start()
    seen = []
    while t.is_alive() or not seen:
        a, b = srv.read(h, n)
        key_ab = (bytes(a), bytes(b))
        assert key_ab in index, "a read returned a state that is no whole 2^j multiple"
        seen.append(index[key_ab])
    t.join()
    assert not errors, errors
    assert seen == sorted(seen)
    a, b = srv.read(h, n)
    assert index[(bytes(a), bytes(b))] == steps
    srv.close()
