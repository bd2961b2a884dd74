"""The reference's C++17 wrapper (its include/notorious_fft.hpp, unchanged)
over this repository's notorious_fft.h and libnotorious_fft_hip_f64.so:
oracle/_ref/cpp_wrapper_check (built by `make -C oracle wrapper` where the
reference tree is present) checks the wrapper's vector, plan and 2-D calls
against long-double direct sums and known round-trip scales."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "cpp_wrapper_check")


def test_wrapper_check_links_engine():
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/cpp_wrapper_check not built (reference tree absent)")
    out = subprocess.run(["ldd", BIN], stdout=subprocess.PIPE, text=True).stdout
    assert "not found" not in out, out
    assert "libnotorious_fft_hip_f64.so" in out


@pytest.mark.gpu
def test_wrapper_check_runs(dev):
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/cpp_wrapper_check not built (reference tree absent)")
    r = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:]
    assert " 0 failed" in r.stdout
