"""Reference-signature calls from several host threads at once (ctypes
releases the GIL during each call): calls that stage host input in the
library's shared staging buffer (which grows, and is reallocated, as sizes
increase) or use its flag word must give the same results as the same calls
made one at a time."""
import threading

import numpy as np
import pytest

import libftcrypto_amd as ft

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    ft.init()


def job(seed):
    rng = np.random.default_rng(seed)
    out = []
    for step in range(6):
        n = int(rng.integers(1, 3000)) << step        # growing: the staging buffer is reallocated
        s = ft.random_scalars(n, seed, step)
        g = ft.GeArray.from_scalars(s)
        g.mul(ft.random_scalars(n, seed, step + 100))
        g.scale(ft.random_scalars(1, seed, step + 200)[0])
        small = ft.GeArray.from_small_scalars(rng.integers(0, 2**63, size=n, dtype=np.uint64))
        g.add(small)
        enc = g.to_bytes_folded()
        back, bad = ft.GeArray.from_bytes_folded(enc)
        out.append((enc.tobytes(), bad, back.validate()))
    return out


def test_concurrent_calls_match_serial():
    seeds = [11, 12, 13, 14, 15, 16]
    serial = {sd: job(sd) for sd in seeds}
    got, errs = {}, []

    def run(sd):
        try:
            got[sd] = job(sd)
        except Exception as e:                        # noqa: BLE001 - reported below
            errs.append(repr(e))

    th = [threading.Thread(target=run, args=(sd,)) for sd in seeds]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for sd in seeds:
        assert got[sd] == serial[sd], sd
