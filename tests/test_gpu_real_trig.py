"""GPU parity of realdft / invrealdft and the DCT/DST family (types 2, 3, 4)
against the oracle, single-pass and multi-pass lengths, FP32 and FP64.
Tolerance: rel-L2 <= 1e-6 log2 N (FP32), 1e-14 log2 N (FP64)."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import rounded, run_plan, tol_for

pytestmark = pytest.mark.gpu

R32 = [1, 2, 4, 8, 16, 64, 256, 1024, 4096, 16384, 32768, 1 << 16, 1 << 17, 1 << 20, 1 << 21]
R64 = [1, 2, 4, 16, 1024, 8192, 16384, 1 << 15, 1 << 17, 1 << 20]


@pytest.mark.parametrize("prec,n", [("f32", n) for n in R32] + [("f64", n) for n in R64])
def test_realdft(pkg, dev, prec, n):
    """Single-pass lines, and the four-step with the unpack fused into its
    row pass (row pairs p, n1 - p) from N = 2^16 (FP32) / 2^15 (FP64)."""
    batch = 3 if n <= (1 << 17) else 2
    x = rounded(ol.uniform(21, n * batch), prec)
    got = run_plan(pkg, dev, "realdft", [n], x, batch, prec)
    ref = ol.run_oracle("realdft", [n], x, batch)
    err = ol.rel_l2(got, ref)
    # inverse on an arbitrary half spectrum: the imaginary parts of bins 0 and
    # N/2 are ignored, as the reference does (src/07_legacy.h:257-258)
    h = n // 2 + 1
    z = rounded(ol.uniform(22, 2 * h * batch), prec)
    goti = run_plan(pkg, dev, "invrealdft", [n], z, batch, prec)
    refi = ol.run_oracle("invrealdft", [n], z, batch)
    erri = ol.rel_l2(goti, refi)
    print(f"{prec} realdft N={n}: fwd {err:.3e} inv {erri:.3e}")
    assert err < tol_for(prec, n)
    assert erri < tol_for(prec, n)


@pytest.mark.parametrize("prec,n,exact", [("f32", 16384, True), ("f64", 16384, True), ("f32", 32768, True),
                                          ("f32", 65536, False), ("f64", 32768, False)])
def test_realdft_line_halves_bitwise(pkg, dev, monkeypatch, prec, n, exact):
    """The half-LDS realdft / invrealdft kernels (NFFT_LINE_H=2) against the
    LDS-resident ones (0): the same arithmetic per point, so bit-identical
    outputs, for both transforms and both vector-Jacobian products (the
    adjoint weightings of bins 0 < k < N/2).  At twice the LDS-resident
    length the alternative is the four-step (agreement within the
    tolerance)."""
    batch = 300 if exact else 40
    h = n // 2 + 1
    x = rounded(ol.uniform(31, n * batch), prec)
    z = rounded(ol.uniform(32, 2 * h * batch), prec)
    outs = {}
    for mode in ("2", "0"):
        monkeypatch.setenv("NFFT_LINE_H", mode)
        plan = pkg.Plan("realdft", prec, [n], batch)
        assert ("two LDS halves" in plan.describe()) == (mode == "2"), plan.describe()
        plan.close()
        outs[mode] = [run_plan(pkg, dev, "realdft", [n], x, batch, prec),
                      run_plan(pkg, dev, "invrealdft", [n], z, batch, prec),
                      run_plan(pkg, dev, "realdft", [n], z, batch, prec, vjp=True),
                      run_plan(pkg, dev, "invrealdft", [n], x, batch, prec, vjp=True)]
    for i, what in enumerate(("realdft", "invrealdft", "realdft vjp", "invrealdft vjp")):
        if exact:
            assert np.array_equal(outs["2"][i], outs["0"][i]), what
        else:
            assert ol.rel_l2(outs["2"][i], outs["0"][i]) < tol_for(prec, n), what
    ref = ol.run_oracle("realdft", [n], x, batch)
    assert ol.rel_l2(outs["2"][0], ref) < tol_for(prec, n)


TRIG = ["dct2", "dst2", "dct3", "dst3", "dct4", "dst4"]
T32 = [1, 2, 4, 8, 16, 128, 1024, 2048, 4096, 16384, 32768, 1 << 16]
T64 = [1, 2, 4, 32, 1024, 2048, 16384, 1 << 15]


@pytest.mark.parametrize("op", TRIG)
@pytest.mark.parametrize("prec,n", [("f32", n) for n in T32] + [("f64", n) for n in T64])
def test_trig(pkg, dev, op, prec, n):
    batch = 3
    x = rounded(ol.uniform(23, n * batch), prec)
    got = run_plan(pkg, dev, op, [n], x, batch, prec)
    ref = ol.run_oracle(op, [n], x, batch)
    err = ol.rel_l2(got, ref)
    print(f"{prec} {op} N={n}: rel-L2 {err:.3e}")
    assert err < tol_for(prec, n)


def test_dropin_real_and_trig_host(pkg, dev):
    lib = pkg.dropin("f64")
    for n in (1, 2, 4, 16, 1024):
        a = lib.notorious_fft_mkaux_realdft_1d(n)
        x = ol.uniform(24, n)
        z = np.zeros(2 * (n // 2 + 1))
        lib.notorious_fft_realdft(x.ctypes.data, z.ctypes.data, a)
        assert ol.rel_l2(z, ol.run_oracle("realdft", [n], x)) < tol_for("f64", n)
        y = np.zeros(n)
        lib.notorious_fft_invrealdft(z.ctypes.data, y.ctypes.data, a)
        assert ol.rel_l2(y, n * x) < tol_for("f64", n)
        lib.notorious_fft_free_aux(a)
        b = lib.notorious_fft_mkaux_t2t3_1d(n)
        c = lib.notorious_fft_mkaux_t4_1d(n)
        for op, aux in (("dct2", b), ("dst2", b), ("dct3", b), ("dst3", b), ("dct4", c), ("dst4", c)):
            y = np.zeros(n)
            getattr(lib, "notorious_fft_" + op)(x.ctypes.data, y.ctypes.data, aux)
            assert ol.rel_l2(y, ol.run_oracle(op, [n], x)) < tol_for("f64", n), (op, n)
            w = x.copy()  # in place
            getattr(lib, "notorious_fft_" + op)(w.ctypes.data, w.ctypes.data, aux)
            assert np.array_equal(w, y)
        lib.notorious_fft_free_aux(b)
        lib.notorious_fft_free_aux(c)
