"""Allocation failures: every device allocation the library makes while
building a plan, a ragged batch, a drop-in aux and its host staging buffers,
or the weighted-power partial sums is made to fail in turn
(NFFT_FAIL_ALLOC=k:tag fails the k-th allocation counted from the first one
that sees the value; a fresh tag re-arms it).  Each failure
must come back as the documented error (NFFT_E_ALLOC, a NULL aux, a no-op
drop-in transform) with nothing half-built left behind, and the next attempt
must succeed and compute the right result (against pocketfft / the oracle)."""
import ctypes
import itertools

import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import tol_for

pytestmark = pytest.mark.gpu

KMAX = 60
_tags = itertools.count()


def arm(monkeypatch, k):
    """Fail the k-th device allocation from here on (a fresh tag, so the
    library restarts its count even when it never saw the variable unset)."""
    monkeypatch.setenv("NFFT_FAIL_ALLOC", f"{k}:{next(_tags)}")


def walk(monkeypatch, attempt):
    """attempt() -> None on failure (after checking its status), else the
    built object; returns (object, number of injected failures)."""
    for k in range(1, KMAX + 1):
        arm(monkeypatch, k)
        obj = attempt()
        print(f"  NFFT_FAIL_ALLOC={k}: {'built' if obj is not None else 'failed as expected'}")
        if obj is not None:
            monkeypatch.delenv("NFFT_FAIL_ALLOC")
            return obj, k - 1
    raise AssertionError("still failing after %d injected failures" % KMAX)


def plan_attempt(pkg, kind, prec, dims, batch):
    lib = pkg.core()

    def attempt():
        h = ctypes.c_void_p()
        arr = (ctypes.c_int * len(dims))(*dims)
        st = lib.nfft_plan_create(ctypes.byref(h), pkg.KIND[kind], 0 if prec == "f32" else 1, len(dims), arr,
                                  batch, 0)
        if st != pkg.OK:
            assert st == pkg.E_ALLOC, (st, pkg.last_error())
            assert not h.value
            return None
        return h
    return attempt


@pytest.mark.parametrize("kind,prec,dims,batch", [
    ("dft", "f32", [1 << 20], 16),     # four-step tables, workspace, XCD-affine counter block
    ("dft", "f64", [1009], 2),         # Bluestein filter built on the device
    ("dft", "f64", [6, 10], 2),        # N-D with a full-array copy
    ("dft", "f32", [1 << 24], 1),      # six-step (two sub-engines)
    ("realdft", "f32", [1 << 16], 4),
    ("t2t3", "f64", [1024], 3),
])
def test_plan_create_allocation_failures(pkg, dev, monkeypatch, kind, prec, dims, batch):
    import torch
    h, fails = walk(monkeypatch, plan_attempt(pkg, kind, prec, dims, batch))
    print(f"{kind} {prec} {dims} x {batch}: {fails} injected failures before success")
    assert fails >= 1
    lib = pkg.core()
    if kind == "dft":
        p = int(np.prod(dims))
        x = ol.uniform(31, 2 * p * batch)
        xd = torch.from_numpy(x.astype(np.float32 if prec == "f32" else np.float64)).to(dev)
        yd = torch.empty_like(xd)
        torch.cuda.synchronize()
        assert lib.nfft_execute(h, pkg.OP["dft"], xd.data_ptr(), yd.data_ptr(), None) == pkg.OK
        assert lib.nfft_plan_check(h, None) == pkg.OK
        xr = xd.cpu().numpy().astype(np.float64)
        ref = np.fft.fftn((xr[0::2] + 1j * xr[1::2]).reshape([batch] + dims), axes=tuple(range(1, len(dims) + 1)))
        got = yd.cpu().numpy().astype(np.float64)
        assert ol.rel_l2(got[0::2] + 1j * got[1::2], ref.reshape(-1)) < tol_for(prec, p)
    lib.nfft_plan_destroy(h)


def test_ragged_create_allocation_failures(pkg, dev, monkeypatch):
    import torch
    lib = pkg.core()
    sizes = [16, 100, 1009, 65521, 8192, 30000]
    arr = (ctypes.c_int * len(sizes))(*sizes)

    def attempt():
        h = ctypes.c_void_p()
        st = lib.nfft_ragged_create(ctypes.byref(h), 1, len(sizes), arr, 0)
        if st != pkg.OK:
            assert st == pkg.E_ALLOC, (st, pkg.last_error())
            return None
        return h
    h, fails = walk(monkeypatch, attempt)
    print(f"ragged: {fails} injected failures before success")
    assert fails >= 1
    total = sum(sizes)
    x = ol.uniform(32, 2 * total)
    xd = torch.from_numpy(x).to(dev)
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    assert lib.nfft_ragged_execute(h, 0, xd.data_ptr(), yd.data_ptr(), None) == pkg.OK
    assert lib.nfft_ragged_check(h, None) == pkg.OK
    got = yd.cpu().numpy()
    off = 0
    for n in sizes:
        want = np.fft.fft(x[2 * off:2 * (off + n):2] + 1j * x[2 * off + 1:2 * (off + n):2])
        assert ol.rel_l2(got[2 * off:2 * (off + n):2] + 1j * got[2 * off + 1:2 * (off + n):2], want) < 1e-13
        off += n
    lib.nfft_ragged_destroy(h)


def test_dropin_allocation_failures(pkg, dev, monkeypatch):
    lib = pkg.dropin("f64")

    def attempt():
        # 2^20 points: a four-step with a workspace of its own
        a = lib.notorious_fft_mkaux_dft_1d(1 << 20)
        return a if a else None
    a, fails = walk(monkeypatch, attempt)
    print(f"dropin mkaux: {fails} injected failures before success")
    assert fails >= 1
    lib.notorious_fft_free_aux(a)
    n = 1 << 15
    a = lib.notorious_fft_mkaux_dft_1d(n)
    assert a
    x = ol.uniform(33, 2 * n)
    y = np.full_like(x, 7.0)
    # the first host-pointer call of 512 KiB stages through device buffers
    # made on first use: a failed staging allocation leaves y untouched (the
    # reference's transforms are void and have no error channel)
    arm(monkeypatch, 1)
    lib.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
    assert np.all(y == 7.0)
    arm(monkeypatch, 2)
    lib.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
    assert np.all(y == 7.0)
    monkeypatch.delenv("NFFT_FAIL_ALLOC")
    lib.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
    assert ol.rel_l2(y, ol.run_oracle("dft", [n], x)) < tol_for("f64", n)
    lib.notorious_fft_free_aux(a)


def test_weighted_power_partials_allocation_failure(pkg, dev, monkeypatch):
    import torch
    dims, batch = [64, 32], 2
    p = 64 * 32
    plan = pkg.Plan("dft", "f64", dims, batch)
    xd = torch.from_numpy(ol.uniform(34, 2 * p * batch)).to(dev)
    wd = torch.ones(p, dtype=torch.float64, device=dev)
    gd = torch.empty_like(xd)
    ld = torch.empty(batch, dtype=torch.float64, device=dev)
    arm(monkeypatch, 1)
    st = pkg.core().nfft_dft_weighted_power(plan.handle, xd.data_ptr(), wd.data_ptr(), gd.data_ptr(),
                                            ld.data_ptr(), None)
    assert st == pkg.E_ALLOC
    monkeypatch.delenv("NFFT_FAIL_ALLOC")
    plan.weighted_power(xd.data_ptr(), wd.data_ptr(), gd.data_ptr(), ld.data_ptr())
    plan.check()
    x = xd.cpu().numpy()
    y = np.fft.fft2((x[0::2] + 1j * x[1::2]).reshape(batch, 64, 32))
    assert np.allclose(ld.cpu().numpy(), np.sum(np.abs(y) ** 2, axis=(1, 2)), rtol=1e-13)
    plan.close()
