"""Adjoint (vector-Jacobian product) kernels through nfft_vjp, FP64:
  - the adjoint identity <vjp(g), x> = <g, f(x)> (real inner products) for
    random x and g, for every op, 1-D and N-D;
  - FP64 central finite differences of the GPU forward transform along a
    random direction v: <g, (f(x + h v) - f(x - h v)) / 2h> = <vjp(g), v>;
  - the reference's own inverse transforms: vjp(dft) equals the
    unnormalised invdft, vjp(dct2) equals dct3 with the DC term doubled.
Also a torch.autograd.Function wrapper (test harness only) checked with
torch.autograd.gradcheck on a tiny FP64 case.
"""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import run_plan

pytestmark = pytest.mark.gpu

CASES = [("dft", [64]), ("dft", [100]), ("dft", [1 << 16]), ("dft", [8, 16]), ("dft", [3, 8]),
         ("invdft", [64]), ("invdft", [100]), ("invdft", [4, 6]),
         ("realdft", [64]), ("realdft", [1 << 17]), ("realdft", [8, 16]), ("realdft", [3, 8]),
         ("invrealdft", [64]), ("invrealdft", [1 << 17]), ("invrealdft", [8, 16]),
         ("dct2", [64]), ("dst2", [64]), ("dct3", [64]), ("dst3", [64]), ("dct4", [64]), ("dst4", [64]),
         ("dct2", [1 << 16]), ("dct3", [1 << 16]), ("dct2", [8, 16]), ("dst3", [4, 8]), ("dct4", [8, 8]),
         ("dct2", [1]), ("dct3", [2]), ("dst2", [2])]


def sizes(op, dims):
    return ol.out_shape(op, dims)


@pytest.mark.parametrize("op,dims", CASES)
def test_adjoint_identity_and_fd(pkg, dev, op, dims):
    nin, nout = sizes(op, dims)
    x = ol.uniform(41, nin)
    g = ol.uniform(42, nout)
    fx = run_plan(pkg, dev, op, dims, x, 1, "f64")
    vg = run_plan(pkg, dev, op, dims, g, 1, "f64", vjp=True)
    lhs, rhs = np.dot(vg, x), np.dot(g, fx)
    scale = np.linalg.norm(vg) * np.linalg.norm(x) + np.linalg.norm(g) * np.linalg.norm(fx)
    assert abs(lhs - rhs) < 1e-13 * scale, (lhs, rhs)
    v = ol.uniform(43, nin)
    h = 1e-3
    fp = run_plan(pkg, dev, op, dims, x + h * v, 1, "f64")
    fm = run_plan(pkg, dev, op, dims, x - h * v, 1, "f64")
    fd = np.dot(g, (fp - fm) / (2 * h))
    assert abs(fd - np.dot(vg, v)) < 1e-9 * scale, (fd, np.dot(vg, v))


def test_vjp_matches_reference_inverses(pkg, dev):
    n = 256
    g = ol.uniform(44, 2 * n)
    vg = run_plan(pkg, dev, "dft", [n], g, 1, "f64", vjp=True)
    assert ol.rel_l2(vg, ol.run_oracle("invdft", [n], g)) < 1e-14 * 8
    r = ol.uniform(45, n)
    vr = run_plan(pkg, dev, "dct2", [n], r, 1, "f64", vjp=True)
    r2 = r.copy()
    r2[0] *= 2
    assert ol.rel_l2(vr, ol.run_oracle("dct3", [n], r2)) < 1e-14 * 8
    vr = run_plan(pkg, dev, "dct4", [n], r, 1, "f64", vjp=True)
    assert ol.rel_l2(vr, ol.run_oracle("dct4", [n], r)) < 1e-14 * 8


def test_autograd_wrapper_gradcheck(pkg, dev):
    import torch

    class DFT(torch.autograd.Function):
        """Harness-only wrapper: forward and backward both call the C ABI."""

        @staticmethod
        def forward(ctx, x):
            plan = pkg.Plan("dft", "f64", [x.numel() // 2])
            y = torch.empty_like(x)
            torch.cuda.synchronize()
            plan.execute("dft", x.contiguous().data_ptr(), y.data_ptr())
            torch.cuda.synchronize()
            ctx.n = x.numel() // 2
            return y

        @staticmethod
        def backward(ctx, gy):
            plan = pkg.Plan("dft", "f64", [ctx.n])
            gx = torch.empty_like(gy)
            gy = gy.contiguous()
            torch.cuda.synchronize()
            plan.vjp("dft", gy.data_ptr(), gx.data_ptr())
            torch.cuda.synchronize()
            return gx

    x = torch.randn(2 * 16, dtype=torch.float64, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(DFT.apply, (x,), eps=1e-6, atol=1e-8)
