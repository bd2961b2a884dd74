"""Weighted spectral power and its gradient (nfft_dft_weighted_power):
    loss[t] = sum_k |W_k Y_k|^2,  Y = DFT(x_t),   grad = 2 DFT^H (W^2 Y)
with the Gaussian-bump mask W[k] = exp(-sum_a d(k_a)^2 / (2 sigma_a^2)),
d(k) = min(k, N - k), sigma_a = N_a / 16.

References: pocketfft in FP64 on the same rounded inputs (DFT^H = P ifftn),
the oracle's own dft and invdft (the reference's inverse,
src/07_legacy.h:146-152: grad = 2 invdft(W^2 dft(x)) for power-of-two axes),
FP64 central finite differences of the loss, and the fused pass against the
unfused one (forward, elementwise weighting, adjoint; NFFT_FUSE=0).
Tolerances: rel error <= 1e-6 log2 P (FP32), 1e-14 log2 P (FP64)."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import tol_for

pytestmark = pytest.mark.gpu


def gauss_mask(dims):
    w = np.ones(dims)
    for a, n in enumerate(dims):
        k = np.arange(n)
        d = np.minimum(k, n - k).astype(np.float64)
        sh = [1] * len(dims)
        sh[a] = n
        w = w * np.exp(-(d ** 2) / (2.0 * (n / 16.0) ** 2)).reshape(sh)
    return w


def np_ref(x, w, dims, batch):
    """FP64 loss per transform and gradient 2 P ifftn(W^2 fftn(x)), interleaved."""
    axes = tuple(range(1, len(dims) + 1))
    z = (x[0::2] + 1j * x[1::2]).reshape([batch] + list(dims))
    y = np.fft.fftn(z, axes=axes)
    w2 = w ** 2
    loss = np.sum(w2 * np.abs(y) ** 2, axis=axes)
    g = 2 * np.prod(dims) * np.fft.ifftn(w2 * y, axes=axes)
    out = np.empty(2 * g.size)
    out[0::2], out[1::2] = g.real.reshape(-1), g.imag.reshape(-1)
    return loss, out


def run_wpow(pkg, dev, prec, dims, batch, x, w, monkeypatch, fuse="1"):
    import torch
    monkeypatch.setenv("NFFT_FUSE", fuse)
    npdt = np.float32 if prec == "f32" else np.float64
    xd = torch.from_numpy(np.ascontiguousarray(x.astype(npdt))).to(dev)
    wd = torch.from_numpy(np.ascontiguousarray(w.reshape(-1).astype(npdt))).to(dev)
    gd = torch.empty_like(xd)
    ld = torch.empty(batch, dtype=torch.float64, device=dev)
    plan = pkg.Plan("dft", prec, dims, batch)
    torch.cuda.synchronize()
    plan.weighted_power(xd.data_ptr(), wd.data_ptr(), gd.data_ptr(), ld.data_ptr())
    plan.check()
    plan.close()
    return ld.cpu().numpy(), gd.cpu().numpy().astype(np.float64)


def rounded(a, prec):
    return a.astype(np.float32).astype(np.float64) if prec == "f32" else a


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dims,batch", [([64, 32], 3), ([8, 16, 12], 2), ([1000], 4), ([6, 10], 3),
                                        ([4096, 4], 2), ([256, 256], 2)])
def test_weighted_power_against_pocketfft(pkg, dev, prec, dims, batch, monkeypatch):
    p = int(np.prod(dims))
    x = rounded(ol.uniform(11, 2 * p * batch), prec)
    w = rounded(gauss_mask(dims), prec)
    lref, gref = np_ref(x, w, dims, batch)
    tol = tol_for(prec, p)
    for fuse in ("1", "0"):
        loss, grad = run_wpow(pkg, dev, prec, dims, batch, x, w, monkeypatch, fuse)
        el = np.max(np.abs(loss - lref) / lref)
        eg = ol.rel_l2(grad, gref)
        print(f"{prec} {dims} x {batch} fuse={fuse}: loss rel {el:.3e}, grad rel-L2 {eg:.3e} (tol {tol:.1e})")
        assert el < tol and eg < tol


def test_c5_full_size_fp32(pkg, dev, monkeypatch):
    """2-D C2C FP32 2048 x 2048 x 16 with the Gaussian-bump mask: the fused
    path against pocketfft (entries 0 and 15), the oracle's dft / invdft and
    the unfused path."""
    import torch
    dims, batch = [2048, 2048], 16
    p = 2048 * 2048
    tol = tol_for("f32", p)
    xd = torch.empty(2 * p * batch, dtype=torch.float32, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 0, 0, "f32")
    w = gauss_mask(dims).astype(np.float32)
    wd = torch.from_numpy(w.reshape(-1)).to(dev)
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("NFFT_FUSE", fuse)
        gd = torch.empty_like(xd)
        ld = torch.empty(batch, dtype=torch.float64, device=dev)
        plan = pkg.Plan("dft", "f32", dims, batch)
        torch.cuda.synchronize()
        plan.weighted_power(xd.data_ptr(), wd.data_ptr(), gd.data_ptr(), ld.data_ptr())
        plan.check()
        plan.close()
        out[fuse] = (ld.cpu().numpy(), gd)
    loss, gd = out["1"]
    e_fu = float((gd.double() - out["0"][1].double()).norm() / out["0"][1].double().norm())
    print(f"C5 fused vs unfused: grad rel-L2 {e_fu:.3e}, loss max rel {np.max(np.abs(loss / out['0'][0] - 1)):.3e}")
    assert e_fu < tol and np.max(np.abs(loss / out["0"][0] - 1)) < tol
    w64 = w.astype(np.float64)
    for t in (0, batch - 1):
        x = xd[2 * p * t:2 * p * (t + 1)].cpu().numpy().astype(np.float64)
        g = gd[2 * p * t:2 * p * (t + 1)].cpu().numpy().astype(np.float64)
        lref, gref = np_ref(x, w64, dims, 1)
        el = abs(loss[t] - lref[0]) / lref[0]
        eg = ol.rel_l2(g, gref)
        # the reference's own transforms: grad = 2 invdft(W^2 dft(x)) (unnormalised for 2^k axes)
        y = ol.run_oracle("dft", dims, x)
        w2 = np.repeat(w64.reshape(-1) ** 2, 2)
        go = 2 * ol.run_oracle("invdft", dims, w2 * y)
        eo = ol.rel_l2(g, go)
        print(f"C5 entry {t}: loss rel {el:.3e}, grad vs pocketfft {eg:.3e}, vs oracle invdft(W^2 dft) {eo:.3e}")
        assert el < tol and eg < tol and eo < tol


def test_weighted_power_finite_differences(pkg, dev, monkeypatch):
    """FP64: (L(x + h v) - L(x - h v)) / 2h against <grad, v> (real inner
    product of the interleaved arrays), for the fused and the unfused path."""
    dims, batch = [64, 48], 2
    p = 64 * 48
    x = ol.uniform(12, 2 * p * batch)
    v = ol.uniform(13, 2 * p * batch)
    w = gauss_mask(dims)
    h = 1e-4
    for fuse in ("1", "0"):
        l0, g = run_wpow(pkg, dev, "f64", dims, batch, x, w, monkeypatch, fuse)
        lp, _ = run_wpow(pkg, dev, "f64", dims, batch, x + h * v, w, monkeypatch, fuse)
        lm, _ = run_wpow(pkg, dev, "f64", dims, batch, x - h * v, w, monkeypatch, fuse)
        for t in range(batch):
            sl = slice(2 * p * t, 2 * p * (t + 1))
            fd = (lp[t] - lm[t]) / (2 * h)
            an = float(np.dot(g[sl], v[sl]))
            print(f"fuse={fuse} t={t}: FD {fd:.10e} vs <grad, v> {an:.10e} (rel {abs(fd - an) / abs(an):.2e})")
            # the loss is quadratic: the central difference is exact up to rounding
            assert abs(fd - an) < 1e-8 * abs(an)
