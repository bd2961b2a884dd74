"""The C-ABI libraries load and export every function their headers declare
(no compute calls: CPU only), and the product does not reach the oracle."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:notorious_fft|nfft)_\w+)\s*\(", text)))


def test_headers_declare_expected_counts():
    assert len(declared("notorious_fft.h")) == 27
    assert "nfft_execute" in declared("notorious_fft_hip.h")


def test_core_exports_everything(pkg):
    lib = ctypes.CDLL(pkg.CORE_SO)
    missing = [f for f in declared("notorious_fft_hip.h") if not hasattr(lib, f)]
    assert not missing, missing


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dropin_exports_reference_api(pkg, prec):
    lib = ctypes.CDLL(pkg.DROPIN_SO[prec])
    missing = [f for f in declared("notorious_fft.h") if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(pkg.DROPIN_FUNCS) == declared("notorious_fft.h")


def test_product_does_not_use_the_oracle():
    for d in ("notorious-fft_amd", os.path.join("notorious-fft_amd", "csrc")):
        for name in os.listdir(os.path.join(ROOT, d)):
            path = os.path.join(ROOT, d, name)
            if os.path.isfile(path) and name.endswith((".py", ".cpp", ".hip", ".h", "Makefile")):
                text = open(path).read()
                assert "oracle" not in text.lower(), path
