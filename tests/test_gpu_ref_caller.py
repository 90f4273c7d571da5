"""The drop-in boundary from compiled code: tests/native/ref_caller, a C++
program written like the reference's run_tests.cc (header in extern "C",
only ft_ calls), built with g++ against include/ftcrypto.h and linked to the
in-tree libftcrypto.so, runs the driver's flow and asserts its checks and the
closed forms the driver only prints."""
import os
import subprocess

import pytest

from libftcrypto_amd import build as B

pytestmark = pytest.mark.gpu


def test_reference_style_caller():
    exe = B.REF_CALLER
    if not os.path.exists(exe):
        B.build_ref_caller()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
