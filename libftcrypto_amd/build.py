"""Build libftcrypto.so (gfx950) in-tree with hipcc.

Every .hip source under csrc/ is compiled to an object in _build/ and linked
into libftcrypto_amd/libftcrypto.so.  `python -m libftcrypto_amd.build`
rebuilds; objects are recompiled when a source or header is newer.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libftcrypto.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("FT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# per-source extra flags: the variable-base kernels and the aggregation walks
# (two waves per SIMD, long dependent carry chains) are scheduled for
# instruction-level parallelism; the fixed-base comb loses 13 % with it
ILP = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
EXTRA = {"varbase.hip": ILP, "aggregate.hip": ILP}


def _deps():
    hdrs = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(HERE, "..", "include", "*.h"))
    return max((os.path.getmtime(h) for h in hdrs), default=0)


def _compile(src, force):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(src), _deps()):
        return obj
    cmd = [HIPCC] + CFLAGS + EXTRA.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-4000:]))
    return obj


def build(force=True, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


REF_CALLER_SRC = os.path.join(HERE, "..", "tests", "native", "ref_caller.cc")
REF_CALLER = os.path.join(HERE, "..", "tests", "native", "ref_caller")


def build_ref_caller(verbose=False):
    """Compile the reference-style C++ caller (tests/native/ref_caller.cc) with
    g++ against include/ftcrypto.h, linked to the in-tree libftcrypto.so by a
    path relative to the binary, so it runs wherever the tree is copied."""
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-I" + os.path.join(HERE, "..", "include"), REF_CALLER_SRC,
           "-o", REF_CALLER, "-L" + HERE, "-l:libftcrypto.so", "-Wl,-rpath,$ORIGIN/../../libftcrypto_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for ref_caller.cc:\n" + r.stderr[-4000:])
    if verbose:
        print("built", os.path.normpath(REF_CALLER))
    return REF_CALLER


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
