"""Seeded synthetic inputs of the benchmark configurations.

Rules for the two ambiguous shapes (stated once, used by bench.py and the
full-size GPU tests):

* segment lengths: lognormal(mu = ln 16, sigma = 1.5) clipped to
  [1, 65536]; then scaled by total / sum, rounded, clamped to >= 1, and the
  residual added to (or taken from) the longest segments, one element each in
  descending-length order, so the lengths sum to exactly `total`.
* graph out-degrees: discrete power law with exponent 2.1 and minimum 1
  (d = floor(U^(-1/1.1)), clipped to the node count), rescaled to exactly
  `n_edges` by the same rule; destinations uniform; stored as CSR by
  destination (int64 row offsets, int32 source indices).
"""
import numpy as np

SEED_BASE = 0xF7C00000


def fix_total(lens, total, rng=None):
    lens = np.asarray(lens, dtype=np.float64)
    scaled = np.maximum(1, np.round(lens * (total / lens.sum()))).astype(np.int64)
    resid = int(total - scaled.sum())
    order = np.argsort(-scaled, kind="stable")
    i = 0
    while resid != 0:
        j = order[i % len(order)]
        if resid > 0:
            scaled[j] += 1
            resid -= 1
        elif scaled[j] > 1:
            scaled[j] -= 1
            resid += 1
        i += 1
    return scaled


def lognormal_segments(n_seg, total, seed, median=16.0, sigma=1.5, lo=1, hi=65536):
    rng = np.random.default_rng(seed)
    raw = np.clip(np.round(rng.lognormal(np.log(median), sigma, size=n_seg)), lo, hi)
    lens = fix_total(raw, total)
    off = np.zeros(n_seg + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    return off


def power_law_graph(n_nodes, n_edges, seed, alpha=2.1):
    rng = np.random.default_rng(seed)
    u = rng.random(n_nodes)
    deg = np.minimum(np.floor(u ** (-1.0 / (alpha - 1.0))), n_nodes)
    deg = fix_total(deg, n_edges)
    src = np.repeat(np.arange(n_nodes, dtype=np.int32), deg)
    dst = rng.integers(0, n_nodes, size=n_edges, dtype=np.int64)
    order = np.argsort(dst, kind="stable")
    cols = src[order].astype(np.int32)
    counts = np.bincount(dst, minlength=n_nodes)
    row_off = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(counts, out=row_off[1:])
    return row_off, cols, deg


def limb_columns(scalars):
    """(n, 32) little-endian bytes -> (n, 8) uint64 holding the 32-bit words."""
    s = np.ascontiguousarray(scalars, dtype=np.uint8).reshape(-1, 32)
    return s.view("<u4").astype(np.uint64)


def limbs_to_int(cols_sum):
    return sum(int(v) << (32 * j) for j, v in enumerate(cols_sum))


def scalar_sum(scalars):
    """Exact integer sum of many 32-byte scalars (32-bit limb columns)."""
    c = limb_columns(scalars)
    return limbs_to_int(c.sum(axis=0, dtype=np.uint64))


def _limbs16(x):
    """(n, 32)-byte scalars or (n,) unsigned ints -> (n, k) uint64 16-bit limbs."""
    x = np.asarray(x)
    if x.ndim == 2:
        return np.ascontiguousarray(x, dtype=np.uint8).view("<u2").astype(np.uint64)
    x = x.astype(np.uint64)
    return np.stack([(x >> np.uint64(16 * j)) & np.uint64(0xFFFF) for j in range(4)], axis=1)


def weighted_sum(a, b):
    """Exact sum_i a_i * b_i for up to 2^22 pairs (16-bit limb products)."""
    A, B = _limbs16(a), _limbs16(b)
    assert len(A) <= (1 << 32), "limb sums must stay below 2^64"
    total = 0
    for j in range(A.shape[1]):
        col = (A[:, j:j + 1] * B).sum(axis=0, dtype=np.uint64)
        total += sum(int(v) << (16 * (j + l)) for l, v in enumerate(col))
    return total
