"""GPU parity of the N-D transforms (row-major, last index fastest) against
the oracle's tree recursion (src/07_legacy.h:109-140 and the DCT/DST N-D
loops), including non-power-of-two axes, unit axes and batches.

FP64 cases with a non-power-of-two axis are judged against pocketfft
(numpy.fft) at the FP64 tolerance: the oracle restates the reference's
Bluestein, whose chirp phase pi k^2 / N is not reduced mod 2N
(src/06_plan.h:91-112) and carries an error growing with N; the oracle
comparison stays as a parity cross-check at the width of that error."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import pocket, rounded, run_plan, tol_for

pytestmark = pytest.mark.gpu

DFT_DIMS = [[4, 4], [8, 16], [32, 32], [3, 5], [6, 10, 4], [4, 8, 8], [16, 1], [1, 16], [1, 1],
            [64, 64, 64], [512, 256], [12, 1024], [2048, 8], [5, 20000]]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dims", DFT_DIMS)
def test_dft_nd(pkg, dev, prec, dims):
    p = int(np.prod(dims))
    batch = 2
    x = rounded(ol.uniform(31, 2 * p * batch), prec)
    for op in ("dft", "invdft"):
        got = run_plan(pkg, dev, op, dims, x, batch, prec)
        ref = ol.run_oracle(op, dims, x, batch)
        err = ol.rel_l2(got, ref)
        blue = any(n & (n - 1) for n in dims)
        print(f"{prec} {op} {dims}: rel-L2 {err:.3e}")
        if prec == "f64" and blue:
            errp = ol.rel_l2(got, pocket(op, dims, x, batch))
            print(f"    vs pocketfft {errp:.3e}")
            assert errp < tol_for(prec, p), (op, errp)
            assert err < tol_for(prec, p) * 1e4, (op, err)
        else:
            assert err < tol_for(prec, p), (op, err)


REAL_DIMS = [[4, 8], [8, 16], [3, 8], [4, 2, 8], [64, 128], [16, 16, 16], [1, 8], [8, 1], [4096, 64]]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dims", REAL_DIMS)
def test_realdft_nd(pkg, dev, prec, dims):
    p = int(np.prod(dims))
    batch = 2
    x = rounded(ol.uniform(32, p * batch), prec)
    got = run_plan(pkg, dev, "realdft", dims, x, batch, prec)
    ref = ol.run_oracle("realdft", dims, x, batch)
    err = ol.rel_l2(got, ref)
    h = p // dims[-1] * (dims[-1] // 2 + 1)
    z = rounded(ol.uniform(33, 2 * h * batch), prec)
    goti = run_plan(pkg, dev, "invrealdft", dims, z, batch, prec)
    refi = ol.run_oracle("invrealdft", dims, z, batch)
    erri = ol.rel_l2(goti, refi)
    print(f"{prec} realdft {dims}: fwd {err:.3e} inv {erri:.3e}")
    blue = any(n & (n - 1) for n in dims)
    if prec == "f64" and blue:
        errp = ol.rel_l2(got, pocket("realdft", dims, x, batch))
        errpi = ol.rel_l2(goti, pocket("invrealdft", dims, z, batch))
        print(f"    vs pocketfft: fwd {errp:.3e} inv {errpi:.3e}")
        assert errp < tol_for(prec, p) and errpi < tol_for(prec, p)
        assert err < tol_for(prec, p) * 1e4 and erri < tol_for(prec, p) * 1e4
    else:
        assert err < tol_for(prec, p)
        assert erri < tol_for(prec, p)


TRIG_DIMS = [[4, 4], [8, 16], [4, 8, 8], [16, 1], [1, 16], [1, 1], [64, 256], [2, 1, 4]]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dims", TRIG_DIMS)
@pytest.mark.parametrize("op", ["dct2", "dst2", "dct3", "dst3", "dct4", "dst4"])
def test_trig_nd(pkg, dev, prec, dims, op):
    p = int(np.prod(dims))
    batch = 2
    x = rounded(ol.uniform(34, p * batch), prec)
    got = run_plan(pkg, dev, op, dims, x, batch, prec)
    ref = ol.run_oracle(op, dims, x, batch)
    err = ol.rel_l2(got, ref)
    print(f"{prec} {op} {dims}: rel-L2 {err:.3e}")
    assert err < tol_for(prec, p)


def test_dropin_nd_host(pkg, dev):
    lib = pkg.dropin("f64")
    a = lib.notorious_fft_mkaux_dft_3d(4, 8, 8)
    x = ol.uniform(35, 2 * 256)
    y = np.zeros_like(x)
    lib.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
    assert ol.rel_l2(y, ol.run_oracle("dft", [4, 8, 8], x)) < tol_for("f64", 256)
    lib.notorious_fft_free_aux(a)
    a = lib.notorious_fft_mkaux_t2t3_2d(8, 16)
    x = ol.uniform(36, 128)
    y = np.zeros_like(x)
    lib.notorious_fft_dct2(x.ctypes.data, y.ctypes.data, a)
    assert ol.rel_l2(y, ol.run_oracle("dct2", [8, 16], x)) < tol_for("f64", 128)
    lib.notorious_fft_free_aux(a)
    a = lib.notorious_fft_mkaux_realdft_2d(8, 16)
    x = ol.uniform(37, 128)
    z = np.zeros(2 * 8 * 9)
    lib.notorious_fft_realdft(x.ctypes.data, z.ctypes.data, a)
    assert ol.rel_l2(z, ol.run_oracle("realdft", [8, 16], x)) < tol_for("f64", 128)
    lib.notorious_fft_free_aux(a)
