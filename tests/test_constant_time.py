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
    "kernels.hip": ["k_fixed_base", "k_mul_u32_r8", "k_mul_u32_r16", "k_encrypt", "k_encrypt_split",
                    "k_encrypt_split_wide", "k_encrypt_wide", "k_encrypt_batch", "k_encrypt_batch_wide"],
    "varbase.hip": ["k_varbase", "k_varbase_r8"],
    "serve.hip": ["kb_encrypt", "kb_encrypt_split", "kb_encrypt_split_wide", "kb_encrypt_wide", "kb_mul_u32_r8",
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
            assert sum("v_mad_u64_u32" in l for l in region) > 1000, (src, name)
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


def scalar_classes(n, seed):
    rng = np.random.default_rng(seed)
    minus8 = sum(-8 * 16**i for i in range(63)) + 16**63          # digits -8 .. -8, 1
    alt7 = sum((7 if i % 2 == 0 else -7) * 16**i for i in range(63)) + 16**63
    top = 2**255 + 2**254 + 12345                                   # top radix-16 digit 12
    fixed = {"zero": 0, "minus8": minus8, "alt7": alt7, "top>8": top, "l-1": L_ORDER - 1}
    out = {}
    for name, v in fixed.items():
        out[name] = np.tile(np.frombuffer(v.to_bytes(32, "little"), dtype=np.uint8), (n, 1))
    import libftcrypto_amd as ft
    out["random"] = ft.random_scalars(n, seed, 1)
    return out


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


TOL = 0.03


@pytest.mark.gpu
def test_scalar_mult_time_independent_of_scalars():
    import libftcrypto_amd as ft
    ft.init()
    L = ft.lib()
    n = 1 << 20
    st = ft.Stream()
    T = _Timer(st.h)
    classes = scalar_classes(n, 7070)
    dev = {k: T.upload(v) for k, v in classes.items()}
    out = ft.GeArray.alloc(n)
    pts = ft.GeArray.from_scalars(classes["random"])
    report = {}
    med, spread = _spread(T, {k: (lambda d=d: ft.check(L.ftx_from_scalars_dev(out.h, d, st.h)))
                              for k, d in dev.items()})
    report["fixed-base"] = (med, spread)
    cls = L.ftx_tune_size_class(n)
    got = [ctypes.c_int() for _ in range(6)]
    ft.check(L.ftx_tune_get(cls, *[ctypes.byref(g) for g in got]))
    base = [g.value for g in got]
    try:
        for w in (3, 4):
            ft.check(L.ftx_tune_set(cls, base[0], base[1], base[2], base[3], base[4], w))
            med, spread = _spread(T, {k: (lambda d=d: ft.check(L.ftx_mul_dev(pts.h, d, st.h)))
                                      for k, d in dev.items()}, reps=3)
            report["variable-base w%d" % w] = (med, spread)
        rng = np.random.default_rng(3)
        u32 = {"zero": np.zeros(n, np.uint32), "ones": np.full(n, 0xFFFFFFFF, np.uint32),
               "eights": np.full(n, 0x88888888, np.uint32), "alt": np.full(n, 0x79797979, np.uint32),
               "random": rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)}
        du = {k: T.upload(v) for k, v in u32.items()}
        c2 = ft.GeArray.from_scalars(classes["random"])
        for w in (3, 4):
            ft.check(L.ftx_tune_set(cls, base[0], base[1], base[2], base[3], w, base[5]))
            med, spread = _spread(T, {k: (lambda d=d: ft.check(L.ftx_mul_u32_dev(pts.h, c2.h, d, st.h)))
                                      for k, d in du.items()})
            report["ct x u32 w%d" % w] = (med, spread)
    finally:
        ft.check(L.ftx_tune_set(cls, *base))
    # encryption: secret r (classes above) and secret messages (msg_bits 8)
    key = ft.Key(ft.GeArray.from_scalars(ft.random_scalars(1, 7070, 9)))
    e1, e2 = ft.GeArray.alloc(n), ft.GeArray.alloc(n)
    fail = ft.DeviceWord()
    msgs = {"m=0": np.zeros(n, np.uint64), "m=255": np.full(n, 255, np.uint64),
            "m=136": np.full(n, 0x88, np.uint64), "m random": rng.integers(0, 256, size=n, dtype=np.uint64)}
    dm = {k: T.upload(v) for k, v in msgs.items()}
    fns = {("r " + k): (lambda d=d: ft.check(L.ftx_encrypt_dev(e1.h, e2.h, key.h, dm["m random"], d, 8, fail.p,
                                                                st.h))) for k, d in dev.items()}
    fns.update({k: (lambda d=d: ft.check(L.ftx_encrypt_dev(e1.h, e2.h, key.h, d, dev["random"], 8, fail.p, st.h)))
                for k, d in dm.items()})
    report["encrypt"] = _spread(T, fns, reps=3)
    fns = {k: (lambda d=d: ft.check(L.ftx_rerandomize_dev(e1.h, e2.h, key.h, d, st.h))) for k, d in dev.items()}
    report["rerandomize"] = _spread(T, fns, reps=3)
    st.synchronize()
    assert fail.read() == 0
    for k, (med, spread) in report.items():
        print("%-18s spread %.4f  %s" % (k, spread, " ".join("%s=%.3f" % (c, t) for c, t in med.items())))
    worst = max(s for _, s in report.values())
    assert worst <= TOL, {k: s for k, (_, s) in report.items()}
    for p in list(dev.values()) + list(du.values()) + list(dm.values()):
        T.hip.hipFree(p)
