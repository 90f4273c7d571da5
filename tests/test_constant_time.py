"""Evidence that the secret-dependent scalar multiplications run in constant
time (the reference's rule: a select over every table entry and a masked
negate, sodium_ge.cu:458-511, 650-731).

CPU part: the kernels that take secret scalars or messages -- fixed-base
comb, both variable-base radices, ciphertext x u32, every encryption and
rerandomisation kernel, and the serving loop's batched forms -- are compiled
to gfx950 assembly with the library's own flags, and from the first to the
last field multiply of each kernel (table builds, digit streams, selects,
window loops) there is no instruction that changes the exec mask and no
branch on it.  A branch on a lane's own data needs exec masking on this ISA (the
compiler emits s_cbranch_scc / s_cbranch_vcc only for conditions it proved
the same on every lane: kernel arguments and loop counters); so no branch in
that region depends on a scalar, digit or message.  The table selects are
v_cndmask merges of every entry.

GPU part: the same kernels timed with HIP events at 2^20 elements on scalar
classes that would expose data-dependent work -- all digits zero, every
radix-16 digit -8, alternating +7/-7 digits, a top digit above 8 (k >= 2^255),
l - 1, uniform random -- interleaved, medians compared."""
import ctypes
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "libftcrypto_amd", "csrc")
sys.path.insert(0, ROOT)

SECRET_KERNELS = {
    "kernels.hip": ["k_fixed_base32", "k_encrypt_w32", "k_mul_u32_r8", "k_mul_u32_r16", "k_encrypt",
                    "k_encrypt_wide", "k_encrypt_batch32", "k_encrypt_w32q"],
    "varbase.hip": ["k_varbase", "k_varbase_r8"],
    "serve.hip": ["kb_encrypt", "kb_encrypt_w32", "kb_encrypt_wide", "kb_mul_u32_r8",
                  "kb_mul_u32_r16"],
}
EXEC = re.compile(r"^\s+(s_cbranch_exec\w+|s_\w+_saveexec_b64|v_cmpx_\w+|s_\w+_b64\s+exec\b)")


def _asm(src):
    from libftcrypto_amd import build as B
    out = os.path.join(B.BUILD, src + ".ct.s")
    os.makedirs(B.BUILD, exist_ok=True)
    cmd = [B.HIPCC] + B.CFLAGS + B.EXTRA.get(src, []) + ["--cuda-device-only", "-S", os.path.join(CSRC, src),
                                                         "-o", out]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def _functions(path):
    cur, out = None, {}
    with open(path) as f:
        for line in f:
            m = re.match(r"^(_Z\S+):", line)
            if m:
                cur = m.group(1)
                out[cur] = []
                continue
            if cur:
                if line.startswith(".Lfunc_end"):
                    cur = None
                    continue
                out[cur].append(line.rstrip("\n"))
    return out


def _short(mangled):
    m = re.match(r"_ZN2ft(\d+)(\w+)", mangled)
    return m.group(2)[:int(m.group(1))] if m else mangled


def _field_region(lines, w=64, dense=32):
    """lines from the first field multiply to the end of the last one: a field
    multiply is a run of at least `dense` v_mad_u64_u32 in `w` instructions
    (isolated multiply-adds are address arithmetic, e.g. a public table search)"""
    mad = np.array(["v_mad_u64_u32" in l for l in lines], dtype=np.int32)
    win = np.convolve(mad, np.ones(w, dtype=np.int32), mode="valid")
    hot = np.nonzero(win >= dense)[0]
    if not len(hot):
        return []
    return lines[hot[0]:hot[-1] + w]


@pytest.fixture(scope="module")
def asm():
    with ThreadPoolExecutor(3) as ex:
        paths = dict(zip(SECRET_KERNELS, ex.map(_asm, SECRET_KERNELS)))
    return {src: _functions(p) for src, p in paths.items()}


def test_secret_kernels_have_no_data_dependent_branches(asm):
    seen = set()
    for src, names in SECRET_KERNELS.items():
        for mangled, lines in asm[src].items():
            name = _short(mangled)
            if name not in names:
                continue
            region = _field_region(lines)
            # at least a loop body of seven field multiplies
            assert sum("v_mad_u64_u32" in l for l in region) > 600, (src, name)
            bad = [l.strip() for l in region if EXEC.search(l)]
            assert not bad, (src, name, bad[:5])
            branches = {l.split()[0] for l in region if re.match(r"^\s+s_cbranch", l)}
            assert branches <= {"s_cbranch_scc0", "s_cbranch_scc1", "s_cbranch_vccz", "s_cbranch_vccnz"}, \
                (name, branches)
            seen.add(name)
    assert seen == {n for v in SECRET_KERNELS.values() for n in v}, seen


def test_checker_flags_a_divergent_branch():
    """the pattern catches exec-masked code"""
    assert EXEC.search("\ts_and_saveexec_b64 s[2:3], vcc")
    assert EXEC.search("\ts_cbranch_execz .LBB1_2")
    assert EXEC.search("\ts_andn2_b64 exec, exec, s[4:5]")
    assert not EXEC.search("\ts_cbranch_scc1 .LBB1_2")
    assert not EXEC.search("\tv_cndmask_b32_e64 v1, v2, v3, s[0:1]")


# ---------------------------------------------------------------- GPU timing

L_ORDER = 2**252 + 27742317777372353535851937790883648493


def _bytes(vals):
    return np.frombuffer(b"".join(int(v).to_bytes(32, "little") for v in vals), dtype=np.uint8).reshape(-1, 32)


def scalar_classes(n, seed):
    """Two families of scalar arrays.  Lane-identical: every lane holds the
    same scalar (all digits zero, every radix-16 digit -8, alternating +7/-7,
    top digit above 8, l - 1, and one random scalar broadcast).  Lane-diverse:
    a different scalar per lane (uniform mod l, below 2^128 so the top 32
    digits are zero, top digit above 8, digits drawn from the extremes -8 and
    +7, and sparse: nine nibbles in ten zero).  A kernel that skipped zero
    digits or branched on a digit's value would separate the classes of a
    family; data-dependent power (identical lanes switch fewer bits than
    diverse ones) separates only the two families."""
    import libftcrypto_amd as ft
    rng = np.random.default_rng(seed)
    minus8 = sum(-8 * 16**i for i in range(63)) + 16**63          # digits -8 .. -8, 1
    alt7 = sum((7 if i % 2 == 0 else -7) * 16**i for i in range(63)) + 16**63
    top = 2**255 + 2**254 + 12345                                   # top radix-16 digit 12
    rnd = int.from_bytes(bytes(ft.random_scalars(1, seed, 7)[0]), "little")
    same = {"zero": 0, "minus8": minus8, "alt7": alt7, "top>8": top, "l-1": L_ORDER - 1, "rand-bcast": rnd}
    identical = {k: np.tile(_bytes([v]), (n, 1)) for k, v in same.items()}
    m = 4096                                                        # distinct values, tiled over n lanes
    half = _bytes(int.from_bytes(rng.bytes(16), "little") for _ in range(m))
    topd = _bytes((int.from_bytes(rng.bytes(32), "little") % 2**252) + (int(rng.integers(9, 16)) << 252)
                  for _ in range(m))
    ext = _bytes(sum(int(d) * 16**i for i, d in enumerate(rng.choice([-8, 7], size=63))) + 16**63
                 for _ in range(m))
    sparse = _bytes(sum(int(d) << (4 * i) for i, d in enumerate(np.where(rng.random(63) < 0.1,
                                                                         rng.integers(1, 16, 63), 0)))
                    for _ in range(m))
    diverse = {"random": ft.random_scalars(n, seed, 1), "k<2^128": np.tile(half, (n // m, 1)),
               "top>8": np.tile(topd, (n // m, 1)), "extreme": np.tile(ext, (n // m, 1)),
               "sparse": np.tile(sparse, (n // m, 1))}
    return identical, diverse


class _Timer:
    def __init__(self, stream):
        self.hip = ctypes.CDLL("libamdhip64.so")
        pv = ctypes.POINTER(ctypes.c_void_p)
        for f, a in [("hipEventCreate", [pv]), ("hipEventRecord", [ctypes.c_void_p, ctypes.c_void_p]),
                     ("hipEventSynchronize", [ctypes.c_void_p]),
                     ("hipEventElapsedTime", [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]),
                     ("hipMalloc", [pv, ctypes.c_size_t]),
                     ("hipMemcpy", [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]),
                     ("hipFree", [ctypes.c_void_p])]:
            getattr(self.hip, f).argtypes = a
        self.st = stream
        self.e = [ctypes.c_void_p(), ctypes.c_void_p()]
        for e in self.e:
            self.hip.hipEventCreate(ctypes.byref(e))

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return p

    def ms(self, fn):
        self.hip.hipEventRecord(self.e[0], self.st)
        fn()
        self.hip.hipEventRecord(self.e[1], self.st)
        self.hip.hipEventSynchronize(self.e[1])
        t = ctypes.c_float()
        self.hip.hipEventElapsedTime(ctypes.byref(t), self.e[0], self.e[1])
        return t.value


def _spread(T, fns, reps=5):
    """median time per class, classes interleaved so clock drift spreads evenly"""
    times = {k: [] for k in fns}
    for k, f in fns.items():                   # warm-up
        T.ms(f)
    for _ in range(reps):
        for k, f in fns.items():
            times[k].append(T.ms(f))
    med = {k: float(np.median(v)) for k, v in times.items()}
    lo, hi = min(med.values()), max(med.values())
    return med, hi / lo - 1.0


# Timing on this GPU also follows data-dependent power: identical lanes and
# low-entropy digits switch fewer bits, and the clock then runs a little
# higher (measured up to ~3 % between classes whose instruction counts per
# dispatch are identical under rocprofv3 --pmc).  A branch on a digit or a
# skipped zero window would show 20 % and more.
TOL = 0.05


def _families(T, make, same, diverse, reps=3):
    """(lane-identical medians and spread, lane-diverse medians and spread)"""
    a = _spread(T, {k: make(d) for k, d in same.items()}, reps)
    b = _spread(T, {k: make(d) for k, d in diverse.items()}, reps)
    return a, b


def measure(n=1 << 20, reps=3):
    """Time every secret-input kernel over the scalar, multiplier and message
    classes; returns {kernel: ((identical medians, spread), (diverse medians,
    spread))}.  Each family runs one warm-up pass, then `reps` passes over
    its classes in dictionary order."""
    import libftcrypto_amd as ft
    ft.init()
    L = ft.lib()
    st = ft.Stream()
    T = _Timer(st.h)
    same_h, diverse_h = scalar_classes(n, 7070)
    same = {k: T.upload(v) for k, v in same_h.items()}
    diverse = {k: T.upload(v) for k, v in diverse_h.items()}
    out = ft.GeArray.alloc(n)
    pts = ft.GeArray.from_scalars(diverse_h["random"])
    report = {}
    report["fixed-base"] = _families(T, lambda d: (lambda: ft.check(L.ftx_from_scalars_dev(out.h, d, st.h))),
                                     same, diverse, reps=reps + 2)
    cls = L.ftx_tune_size_class(n)
    got = [ctypes.c_int() for _ in range(6)]
    ft.check(L.ftx_tune_get(cls, *[ctypes.byref(g) for g in got]))
    base = [g.value for g in got]
    rng = np.random.default_rng(3)
    du = []
    try:
        for w in (3, 4):
            ft.check(L.ftx_tune_set(cls, base[0], base[1], base[2], base[3], base[4], w))
            report["variable-base w%d" % w] = _families(
                T, lambda d: (lambda: ft.check(L.ftx_mul_dev(pts.h, d, st.h))), same, diverse, reps=reps)
        u_same = {"zero": np.zeros(n, np.uint32), "ones": np.full(n, 0xFFFFFFFF, np.uint32),
                  "eights": np.full(n, 0x88888888, np.uint32), "alt": np.full(n, 0x79797979, np.uint32),
                  "rand-bcast": np.full(n, 0x5D2C91E7, np.uint32)}
        u_div = {"random": rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32),
                 "below 2^16": rng.integers(0, 2**16, size=n, dtype=np.uint64).astype(np.uint32),
                 "extreme": np.tile(np.array([sum(int(d) * 16**i for i, d in enumerate(rng.choice([8, 15], 8)))
                                              for _ in range(4096)], dtype=np.uint64).astype(np.uint32), n // 4096)}
        us = {k: T.upload(v) for k, v in u_same.items()}
        ud = {k: T.upload(v) for k, v in u_div.items()}
        du = list(us.values()) + list(ud.values())
        c2 = ft.GeArray.from_scalars(diverse_h["random"])
        for w in (3, 4):
            ft.check(L.ftx_tune_set(cls, base[0], base[1], base[2], base[3], w, base[5]))
            report["ct x u32 w%d" % w] = _families(
                T, lambda d: (lambda: ft.check(L.ftx_mul_u32_dev(pts.h, c2.h, d, st.h))), us, ud, reps=reps + 2)
    finally:
        ft.check(L.ftx_tune_set(cls, *base))
    # encryption: secret r (the classes above) and secret messages (msg_bits 8)
    key = ft.Key(ft.GeArray.from_scalars(ft.random_scalars(1, 7070, 9)))
    e1, e2 = ft.GeArray.alloc(n), ft.GeArray.alloc(n)
    fail = ft.DeviceWord()
    dm = {"m=0": T.upload(np.zeros(n, np.uint64)), "m=255": T.upload(np.full(n, 255, np.uint64)),
          "m=136": T.upload(np.full(n, 0x88, np.uint64))}
    dmr = {"m random": T.upload(rng.integers(0, 256, size=n, dtype=np.uint64)),
           "m random 0/255": T.upload(rng.choice([0, 255], size=n).astype(np.uint64))}
    rand_r = diverse["random"]
    report["encrypt (r)"] = _families(
        T, lambda d: (lambda: ft.check(L.ftx_encrypt_dev(e1.h, e2.h, key.h, dmr["m random"], d, 8, fail.p, st.h))),
        same, diverse, reps=reps)
    report["encrypt (m)"] = _families(
        T, lambda d: (lambda: ft.check(L.ftx_encrypt_dev(e1.h, e2.h, key.h, d, rand_r, 8, fail.p, st.h))), dm, dmr,
        reps=reps)
    report["rerandomize"] = _families(
        T, lambda d: (lambda: ft.check(L.ftx_rerandomize_dev(e1.h, e2.h, key.h, d, st.h))), same, diverse,
        reps=reps)
    st.synchronize()
    assert fail.read() == 0
    for p in list(same.values()) + list(diverse.values()) + du + list(dm.values()) + list(dmr.values()):
        T.hip.hipFree(p)
    return report


def print_report(report):
    for k, ((m1, s1), (m2, s2)) in report.items():
        gap = np.median(list(m2.values())) / np.median(list(m1.values())) - 1.0
        print("%-18s identical lanes: spread %.4f | diverse lanes: spread %.4f | diverse/identical %+.4f" %
              (k, s1, s2, gap))
        print("    " + " ".join("%s=%.3f" % (c, t) for c, t in list(m1.items()) + list(m2.items())))


@pytest.mark.gpu
def test_scalar_mult_time_independent_of_scalars():
    report = measure()
    print_report(report)
    worst = {k: max(s1, s2) for k, ((_, s1), (_, s2)) in report.items()}
    assert max(worst.values()) <= TOL, worst


if __name__ == "__main__":
    # one warm-up and one timed pass per class (for a counter run: every
    # kernel's dispatches follow the class order printed here)
    rep = measure(reps=1)
    print_report(rep)
