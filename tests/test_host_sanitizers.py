"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
MatrixMarket reader (the library's one parser of untrusted input) on valid
and malformed files, and the scalar Zolotarev math, built with g++
(`make -C zolo-svd_amd sanitize`) and run without a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "zolo-svd_amd")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_sanitizers(tmp_path):
    b = subprocess.run(["make", "-s", "-C", PKG, "sanitize"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([os.path.join(PKG, "_build", "host_checks"), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    print(r.stderr[-4000:])
    assert r.returncode == 0
    assert "host checks passed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
