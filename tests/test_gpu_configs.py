"""The benchmark workloads themselves, at full size, against FP64 references:
  - 3-D C2C FP32 512^3: every bin against pocketfft (scipy.fft.fftn in FP64),
    the unnormalised round trip (= 512^3 x), Parseval, and direct sums of a
    few bins computed separably in FP64;
  - 2-D realdft FP32 4096 x 4096 of Gaussian N(0, 1) inputs (the device
    Box-Muller generator, checked against its numpy restatement) against the
    oracle's multi-D realdft and numpy.fft.rfft2;
  - 2-D C2C FP32 2048 x 2048 x 16 forward + vjp: the adjoint identity, the
    vjp against the oracle's invdft (the reference's unnormalised inverse,
    src/07_legacy.h:146-152) on the first and last batch entries, and FP64
    central finite differences of the FP64 forward map;
  - DCT-2 vjp FP32 N = 2^14 x 1024 against the oracle's dct3 with the DC
    input doubled (REDFT10^T = REDFT01 diag(2, 1, ..., 1),
    src/07_legacy.h:951-1066);
  - FP64 batches of 2^28 complex points (4 GiB in, 4 GiB out): the first and
    the last transform against the oracle, so that element offsets above
    4 GiB are compared too.
Tolerances: rel-L2 <= 1e-6 log2 N (FP32), 1e-14 log2 N (FP64)."""
import os

import numpy as np
import pytest
import scipy.fft as sfft

import oracle_lib as ol
from gpu_util import tol_for

pytestmark = pytest.mark.gpu

WORKERS = min(16, os.cpu_count() or 1)


def cx(a):
    return a[0::2] + 1j * a[1::2]


def dev_f64_dot(a, b):
    return float((a.double() * b.double()).sum().item())


def dev_norm(a):
    return float(a.double().norm().item())


def test_fill_gaussian_matches_numpy(pkg, dev):
    import torch
    n = 1 << 20
    for prec, dt in (("f64", torch.float64), ("f32", torch.float32)):
        for start in (0, 12345):
            g = torch.empty(n, dtype=dt, device=dev)
            pkg.fill_gaussian(g.data_ptr(), n, 3, start, prec)
            torch.cuda.synchronize()
            want = ol.gaussian_np(3, n, start)
            got = g.cpu().numpy().astype(np.float64)
            assert np.max(np.abs(got - want.astype(np.float32 if prec == "f32" else np.float64))) < \
                (4e-6 if prec == "f32" else 1e-13)
    assert abs(got.mean()) < 5e-3 and abs(got.std() - 1) < 5e-3


def test_3d_512_cube_fp32(pkg, dev):
    import torch
    dims = [512, 512, 512]
    p = 512 ** 3
    xd = torch.empty(2 * p, dtype=torch.float32, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 0, 0, "f32")
    yd = torch.empty_like(xd)
    zd = torch.empty_like(xd)
    plan = pkg.Plan("dft", "f32", dims, 1)
    torch.cuda.synchronize()
    plan.execute("dft", xd.data_ptr(), yd.data_ptr())
    plan.execute("invdft", yd.data_ptr(), zd.data_ptr())
    plan.check()
    plan.close()
    tol = tol_for("f32", p)
    # round trip and Parseval on the device, accumulated in FP64
    rt = dev_norm(zd.double() - p * xd.double()) / dev_norm(p * xd.double())
    ex, ey = dev_f64_dot(xd, xd), dev_f64_dot(yd, yd)
    print(f"512^3: round trip {rt:.3e}, Parseval {abs(ey / (p * ex) - 1):.3e}")
    assert rt < 2 * tol
    assert abs(ey / (p * ex) - 1) < tol
    x = xd.cpu().numpy()
    y = yd.cpu().numpy()
    del xd, yd, zd
    xc = cx(x.astype(np.float64)).reshape(dims)
    # a few bins by direct sums, separably in FP64
    n = np.arange(512)
    for k in [(0, 0, 0), (1, 2, 3), (511, 256, 7), (100, 400, 511)]:
        w = [np.exp(-2j * np.pi * k[a] * n / 512) for a in range(3)]
        want = np.einsum("ijk,k->ij", xc, w[2]) @ w[1] @ w[0]
        got = y[2 * ((k[0] * 512 + k[1]) * 512 + k[2]):][:2]
        err = abs(got[0] + 1j * got[1] - want) / np.sqrt(p)
        assert err < 1e-5, (k, err)
    ref = sfft.fftn(xc, workers=WORKERS)
    err = ol.rel_l2(cx(y.astype(np.float64)), ref.reshape(-1))
    print(f"512^3 FP32 vs pocketfft: rel-L2 {err:.3e} (tol {tol:.1e})")
    assert err < tol


def test_2d_realdft_4096_gaussian(pkg, dev):
    import torch
    dims = [4096, 4096]
    p = 4096 * 4096
    xd = torch.empty(p, dtype=torch.float32, device=dev)
    pkg.fill_gaussian(xd.data_ptr(), p, 0, 0, "f32")
    zd = torch.empty(2 * 4096 * 2049, dtype=torch.float32, device=dev)
    plan = pkg.Plan("realdft", "f32", dims, 1)
    torch.cuda.synchronize()
    plan.execute("realdft", xd.data_ptr(), zd.data_ptr())
    plan.check()
    plan.close()
    x = xd.cpu().numpy().astype(np.float64)
    z = zd.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(x - ol.gaussian_np(0, p).astype(np.float32))) < 4e-6
    tol = tol_for("f32", p)
    err = ol.rel_l2(z, ol.run_oracle("realdft", dims, x))
    errp = ol.rel_l2(cx(z), np.fft.rfft2(x.reshape(dims)).reshape(-1))
    print(f"4096^2 realdft Gaussian: vs oracle {err:.3e}, vs numpy {errp:.3e} (tol {tol:.1e})")
    assert err < tol and errp < tol


def test_c5_forward_vjp_2048sq_x16(pkg, dev):
    import torch
    dims, batch = [2048, 2048], 16
    p = 2048 * 2048
    tol = tol_for("f32", p)
    xd = torch.empty(2 * p * batch, dtype=torch.float32, device=dev)
    gd = torch.empty_like(xd)
    vd = torch.empty_like(xd)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 0, 0, "f32")
    pkg.fill_uniform(gd.data_ptr(), gd.numel(), 1, 0, "f32")
    pkg.fill_uniform(vd.data_ptr(), vd.numel(), 2, 0, "f32")
    fx = torch.empty_like(xd)
    vg = torch.empty_like(xd)
    plan = pkg.Plan("dft", "f32", dims, batch)
    torch.cuda.synchronize()
    plan.execute("dft", xd.data_ptr(), fx.data_ptr())
    plan.vjp("dft", gd.data_ptr(), vg.data_ptr())
    plan.check()
    plan.close()
    # adjoint identity <vjp(g), x> = <g, f(x)>
    lhs, rhs = dev_f64_dot(vg, xd), dev_f64_dot(gd, fx)
    scale = dev_norm(vg) * dev_norm(xd) + dev_norm(gd) * dev_norm(fx)
    print(f"C5 adjoint identity: {abs(lhs - rhs) / scale:.3e} of scale")
    assert abs(lhs - rhs) < tol * scale
    # vjp and forward against the oracle on the first and last batch entries
    g = gd.cpu().numpy().astype(np.float64)
    x = xd.cpu().numpy().astype(np.float64)
    vgh = vg.cpu().numpy().astype(np.float64)
    fxh = fx.cpu().numpy().astype(np.float64)
    for b in (0, batch - 1):
        sl = slice(2 * p * b, 2 * p * (b + 1))
        e1 = ol.rel_l2(vgh[sl], ol.run_oracle("invdft", dims, g[sl]))
        e2 = ol.rel_l2(fxh[sl], ol.run_oracle("dft", dims, x[sl]))
        print(f"C5 entry {b}: vjp vs oracle invdft {e1:.3e}, forward vs oracle dft {e2:.3e}")
        assert e1 < tol and e2 < tol
    # FP64 central finite differences of the forward map along v, against the
    # FP32 vjp: <g, (f(x + h v) - f(x - h v)) / 2h> = <vjp(g), v>
    h = 1e-3
    plan64 = pkg.Plan("dft", "f64", dims, batch)
    x64 = xd.double()
    v64 = vd.double()
    fp = torch.empty_like(x64)
    fm = torch.empty_like(x64)
    xp = x64 + h * v64
    xm = x64 - h * v64
    torch.cuda.synchronize()
    plan64.execute("dft", xp.data_ptr(), fp.data_ptr())
    plan64.execute("dft", xm.data_ptr(), fm.data_ptr())
    plan64.check()
    plan64.close()
    fd = float((gd.double() * (fp - fm)).sum().item()) / (2 * h)
    an = dev_f64_dot(vg, vd)
    sc = dev_norm(vg) * dev_norm(vd)
    print(f"C5 FD: {fd:.6e} vs vjp {an:.6e} ({abs(fd - an) / sc:.3e} of scale)")
    assert abs(fd - an) < tol * sc


def test_dct2_vjp_16384_x1024_fp32(pkg, dev):
    import torch
    n, batch = 1 << 14, 1024
    gd = torch.empty(n * batch, dtype=torch.float32, device=dev)
    pkg.fill_uniform(gd.data_ptr(), gd.numel(), 4, 0, "f32")
    out = torch.empty_like(gd)
    plan = pkg.Plan("t2t3", "f32", [n], batch)
    torch.cuda.synchronize()
    plan.vjp("dct2", gd.data_ptr(), out.data_ptr())
    plan.check()
    plan.close()
    g = gd.cpu().numpy().astype(np.float64).reshape(batch, n)
    g2 = g.copy()
    g2[:, 0] *= 2
    want = ol.run_oracle("dct3", [n], g2.reshape(-1), batch)
    err = ol.rel_l2(out.cpu().numpy().astype(np.float64), want)
    print(f"DCT-2 vjp 2^14 x 1024: rel-L2 {err:.3e}")
    assert err < tol_for("f32", n)


@pytest.mark.parametrize("lg", [20, 10])
def test_fp64_full_batch_first_and_last(pkg, dev, lg):
    import torch
    n = 1 << lg
    batch = 1 << (28 - lg)
    xd = torch.empty(2 * n * batch, dtype=torch.float64, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 0, 0, "f64")
    yd = torch.empty_like(xd)
    plan = pkg.Plan("dft", "f64", [n], batch)
    torch.cuda.synchronize()
    plan.execute("dft", xd.data_ptr(), yd.data_ptr())
    plan.check()
    plan.close()
    for b in (0, batch // 2, batch - 1):
        x = ol.uniform(0, 2 * n, start=2 * n * b)
        got = yd[2 * n * b:2 * n * (b + 1)].cpu().numpy()
        err = ol.rel_l2(got, ol.run_oracle("dft", [n], x))
        print(f"FP64 N=2^{lg} x {batch}: transform {b} (byte offset {16 * n * b:.3e}) rel-L2 {err:.3e}")
        assert err < tol_for("f64", n), (b, err)
