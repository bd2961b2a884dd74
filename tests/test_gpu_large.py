"""Lengths beyond the four-step and the three-pass Bluestein: the six-step
(N = N1 x N2 over two engines) for powers of two above 2^23 (FP32) / 2^22
(FP64), and Bluestein through an M-point engine for M above 4096 x 4096
(FP32) / 4096 x 2048 (FP64).  The reference accepts every N >= 1 in its
constructors (src/07_legacy.h:1405-1420, src/06_plan.h:28-33); so does the
drop-in here up to 2^30.

Checks: the oracle (FP64 split radix) where it finishes in seconds,
pocketfft for Bluestein lengths (the reference's own chirp phase is not
reduced mod 2N), and at the largest sizes the round trip, Parseval and
direct-sum bins.  Tolerances: rel-L2 <= 1e-6 log2 N (FP32), 1e-14 log2 N
(FP64)."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import rounded, tol_for

pytestmark = pytest.mark.gpu


def cx(a):
    return a[0::2] + 1j * a[1::2]


def run(pkg, dev, prec, n, batch, op, xd, norm="reference", inplace=False):
    import torch
    plan = pkg.Plan("dft", prec, [n], batch, norm)
    desc = plan.describe()
    out = xd if inplace else torch.empty_like(xd)
    torch.cuda.synchronize()
    plan.execute(op, xd.data_ptr(), out.data_ptr())
    plan.check()
    plan.close()
    return out, desc


@pytest.mark.parametrize("prec,lg,batch", [("f32", 24, 2), ("f32", 25, 1), ("f64", 23, 2), ("f64", 24, 1)])
def test_six_step_against_oracle(pkg, dev, prec, lg, batch):
    import torch
    n = 1 << lg
    dt = torch.float32 if prec == "f32" else torch.float64
    xd = torch.empty(2 * n * batch, dtype=dt, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 21, 0, prec)
    x = xd.cpu().numpy().astype(np.float64)
    tol = tol_for(prec, n)
    for op in ("dft", "invdft"):
        y, desc = run(pkg, dev, prec, n, batch, op, xd)
        assert desc.startswith("six-step"), desc
        got = y.cpu().numpy().astype(np.float64)
        for t in range(batch):
            sl = slice(2 * n * t, 2 * n * (t + 1))
            err = ol.rel_l2(got[sl], ol.run_oracle(op, [n], x[sl]))
            print(f"{prec} {op} N=2^{lg} [{desc}] transform {t}: rel-L2 {err:.3e} (tol {tol:.1e})")
            assert err < tol
    # in place, with the 1/N normalisation
    y2, _ = run(pkg, dev, prec, n, batch, "invdft", xd.clone(), norm="inverse", inplace=True)
    ref = ol.run_oracle("invdft", [n], x[:2 * n]) / n
    assert ol.rel_l2(y2[:2 * n].cpu().numpy().astype(np.float64), ref) < tol


@pytest.mark.parametrize("prec,lg", [("f32", 27), ("f64", 26)])
def test_six_step_largest_properties(pkg, dev, prec, lg):
    """1 GiB (FP32 2^27) / 1 GiB (FP64 2^26) transforms: round trip,
    Parseval and three bins by direct sums in FP64."""
    import torch
    n = 1 << lg
    dt = torch.float32 if prec == "f32" else torch.float64
    xd = torch.empty(2 * n, dtype=dt, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 22, 0, prec)
    y, desc = run(pkg, dev, prec, n, 1, "dft", xd)
    z, _ = run(pkg, dev, prec, n, 1, "invdft", y)
    tol = tol_for(prec, n)
    x64 = xd.double()
    rt = float((z.double() - n * x64).norm() / (n * x64).norm())
    pe = abs(float((y.double() ** 2).sum()) / (n * float((x64 ** 2).sum())) - 1)
    print(f"{prec} N=2^{lg} [{desc}]: round trip {rt:.3e}, Parseval {pe:.3e} (tol {tol:.1e})")
    assert rt < 2 * tol and pe < tol
    xc = cx(xd.cpu().numpy().astype(np.float64))
    yc = cx(y.cpu().numpy().astype(np.float64))
    nrm = np.sqrt(np.sum(np.abs(xc) ** 2))
    idx = np.arange(n, dtype=np.int64)
    for k in (0, 12345, n - 7):
        d = np.sum(xc * np.exp(-2j * np.pi * ((k * idx) % n) / n))
        err = abs(yc[k] - d) / nrm
        print(f"  bin {k}: |got - direct| / |x| = {err:.3e}")
        assert err < 10 * tol


@pytest.mark.parametrize("prec,n", [("f64", 4194319), ("f32", 8388617), ("f64", 5000011)])
def test_bluestein_beyond_three_pass(pkg, dev, prec, n):
    import torch
    dt = torch.float32 if prec == "f32" else torch.float64
    xd = torch.empty(2 * n, dtype=dt, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 23, 0, prec)
    x = xd.cpu().numpy().astype(np.float64)
    y, desc = run(pkg, dev, prec, n, 1, "dft", xd)
    assert "bluestein" in desc and "[" in desc, desc
    ref = np.fft.fft(cx(x))
    tol = tol_for(prec, n)
    err = ol.rel_l2(cx(y.cpu().numpy().astype(np.float64)), ref)
    # the reference's invdft scales non-power-of-two lengths by 1/N: x back
    z, _ = run(pkg, dev, prec, n, 1, "invdft", y)
    rt = ol.rel_l2(z.cpu().numpy().astype(np.float64), x)
    print(f"{prec} N={n} [{desc}]: vs pocketfft {err:.3e}, round trip {rt:.3e} (tol {tol:.1e})")
    assert err < tol and rt < 2 * tol


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_dropin_large_lengths(pkg, dev, prec):
    """The drop-in constructors return working auxes beyond the four-step and
    the three-pass Bluestein, as the reference's do; transforms run on device
    pointers and match the oracle (power of two) / pocketfft (prime)."""
    import torch
    lib = pkg.dropin(prec)
    dt = torch.float32 if prec == "f32" else torch.float64
    for n in (1 << 24, 4194319):
        a = lib.notorious_fft_mkaux_dft_1d(n)
        assert a, n
        xd = torch.empty(2 * n, dtype=dt, device=dev)
        pkg.fill_uniform(xd.data_ptr(), xd.numel(), 24, 0, prec)
        yd = torch.empty_like(xd)
        torch.cuda.synchronize()
        lib.notorious_fft_dft(xd.data_ptr(), yd.data_ptr(), a)
        lib.notorious_fft_free_aux(a)
        x = xd.cpu().numpy().astype(np.float64)
        got = yd.cpu().numpy().astype(np.float64)
        if n & (n - 1) == 0:
            err = ol.rel_l2(got, ol.run_oracle("dft", [n], x))
        else:
            err = ol.rel_l2(cx(got), np.fft.fft(cx(x)))
        print(f"dropin {prec} N={n}: rel-L2 {err:.3e}")
        assert err < tol_for(prec, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_realdft_beyond_fused_four_step(pkg, dev, prec):
    """realdft whose half-length complex engine is a six-step (unfused
    pack / transform / unpack), against numpy.fft.rfft."""
    import torch
    n = 1 << (26 if prec == "f32" else 25)
    dt = torch.float32 if prec == "f32" else torch.float64
    xd = torch.empty(n, dtype=dt, device=dev)
    pkg.fill_uniform(xd.data_ptr(), n, 25, 0, prec)
    zd = torch.empty(n + 2, dtype=dt, device=dev)
    plan = pkg.Plan("realdft", prec, [n], 1)
    desc = plan.describe()
    torch.cuda.synchronize()
    plan.execute("realdft", xd.data_ptr(), zd.data_ptr())
    plan.check()
    plan.close()
    x = xd.cpu().numpy().astype(np.float64)
    err = ol.rel_l2(cx(zd.cpu().numpy().astype(np.float64)), np.fft.rfft(x))
    print(f"realdft {prec} N=2^{int(np.log2(n))} [{desc}]: vs numpy {err:.3e}")
    assert err < tol_for(prec, n)
