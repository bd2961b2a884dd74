"""nfft_execute_noalloc (static device scratch, no allocation) against
nfft_execute on the same plans: outputs must be bit-identical (the same
kernels run; only the chunking of the batch over the scratch differs), on
two or more calls in a row with the later inputs smaller, so that a second
call finds the first call's data in the shared array.  Also: one output
against the FP64 oracle, and the size limit reported as NFFT_E_SIZE."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import tol_for

pytestmark = pytest.mark.gpu


def _both(pkg, dev, prec, dims, batch, op="dft", seed=7):
    import torch
    dt = torch.float32 if prec == "f32" else torch.float64
    n = int(np.prod(dims)) * batch
    x = torch.empty(2 * n, dtype=dt, device=dev)
    pkg.fill_uniform(x.data_ptr(), x.numel(), seed, 0, prec)
    a = torch.empty_like(x)
    b = torch.empty_like(x)
    plan = pkg.Plan("dft", prec, dims, batch)
    torch.cuda.synchronize()
    plan.execute(op, x.data_ptr(), a.data_ptr())
    plan.execute_noalloc(op, x.data_ptr(), b.data_ptr())
    plan.check()
    desc = plan.describe()
    plan.close()
    same = torch.equal(a, b)
    print(f"{prec} {dims} x {batch} {op} [{desc}]: bit-identical {same}")
    return same, x, b


@pytest.mark.parametrize("calls", [
    # FP32 four-step (chunks of 8 transforms in the static array), then smaller
    [("f32", [1 << 20], 24), ("f32", [1 << 16], 3)],
    # FP64: Bluestein (three-pass), then a four-step of 2 transforms of 32 MiB each, then a line length
    [("f64", [100003], 9), ("f64", [1 << 21], 2), ("f64", [4096], 5)],
    # N-D: an axis transformed through the full-array copy, then a smaller one
    [("f64", [3, 10000], 4), ("f32", [512, 512], 4), ("f32", [6, 10, 4], 2)],
])
def test_noalloc_matches_execute(pkg, dev, calls):
    for prec, dims, batch in calls:
        for op in ("dft", "invdft"):
            same, _, _ = _both(pkg, dev, prec, dims, batch, op)
            assert same, (prec, dims, batch, op)


def test_noalloc_against_oracle(pkg, dev):
    n, batch = 1 << 18, 5
    _, x, y = _both(pkg, dev, "f32", [n], batch, "dft", seed=8)
    xh = x.cpu().numpy().astype(np.float64)
    yh = y.cpu().numpy().astype(np.float64)
    for t in (0, batch - 1):
        sl = slice(2 * n * t, 2 * n * (t + 1))
        err = ol.rel_l2(yh[sl], ol.run_oracle("dft", [n], xh[sl]))
        print(f"noalloc N=2^18 transform {t}: rel-L2 {err:.3e}")
        assert err < tol_for("f32", n)


def test_noalloc_rejects_plans_beyond_its_scratch(pkg, dev):
    import torch
    # axis 0 (length 5, mixed radix) goes through a full-array copy of
    # 5 * 2^22 complex64 = 160 MiB > 64 MiB
    dims = [5, 1 << 22]
    plan = pkg.Plan("dft", "f32", dims, 1)
    x = torch.zeros(2 * 5 * (1 << 22), dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    st = pkg.core().nfft_execute_noalloc(plan.handle, pkg.OP["dft"], x.data_ptr(), y.data_ptr(), None)
    msg = pkg.last_error()
    plan.close()
    print(f"status {st}: {msg}")
    assert st == pkg.E_SIZE
    assert "bytes of scratch" in msg
    # other plan kinds are not covered by this entry point
    rp = pkg.Plan("realdft", "f32", [1024], 1)
    st = pkg.core().nfft_execute_noalloc(rp.handle, pkg.OP["realdft"], x.data_ptr(), y.data_ptr(), None)
    rp.close()
    assert st == pkg.E_INVALID
