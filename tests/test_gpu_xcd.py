"""The XCD-affine single-launch four-step (FP32, N = N1 x 1024, batches of
at least 16 transforms) against the two-launch four-step and the oracle.

Both paths run the same arithmetic (column FFT with the stage twiddles
rounded once to FP32, the inter-pass twiddle and row FFT in the row tiles),
so their outputs must agree bit for bit; sampled transforms (first, middle,
last) are compared with the FP64 oracle at rel-L2 <= 1e-6 log2 N.  In-place
execution, repeated launches (the counter block is reset by each launch),
odd batch sizes and the inverse with a scale are covered."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import tol_for

pytestmark = pytest.mark.gpu


def _run(pkg, dev, n, batch, op, xd, xcd, monkeypatch, norm="reference", inplace=False, reps=1):
    import torch
    # 2: the XCD-affine launch for every length that has one (default: 2^20 only)
    monkeypatch.setenv("NFFT_XCD", "2" if xcd else "0")
    plan = pkg.Plan("dft", "f32", [n], batch, norm)
    desc = plan.describe()
    assert ("XCD-affine" in desc) == bool(xcd), desc
    src = xd.clone()
    out = src if inplace else torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        if inplace:
            src.copy_(xd)
        plan.execute(op, src.data_ptr(), out.data_ptr())
    plan.check()
    plan.close()
    return out


@pytest.mark.parametrize("n,batch,row_twiddle", [(1 << 20, 16, "1"), (1 << 20, 37, "1"), (1 << 19, 40, "1"),
                                                (1 << 21, 17, "1"), (1 << 20, 19, "0"), (1 << 21, 16, "0")])
def test_xcd_matches_two_launch_and_oracle(pkg, dev, n, batch, row_twiddle, monkeypatch):
    """row_twiddle selects where both paths apply the inter-pass twiddle (row
    tiles from an FP32 table, or column tiles from the FP64 table)."""
    import torch
    monkeypatch.setenv("NFFT_ROW_TWIDDLE", row_twiddle)
    xd = torch.empty(2 * n * batch, dtype=torch.float32, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 5, 0, "f32")
    for op, norm in (("dft", "reference"), ("invdft", "inverse")):
        a = _run(pkg, dev, n, batch, op, xd, True, monkeypatch, norm, reps=2)
        b = _run(pkg, dev, n, batch, op, xd, False, monkeypatch, norm)
        same = torch.equal(a, b)
        print(f"N={n} x {batch} {op}: XCD-affine bit-identical to two launches: {same}")
        assert same
        x = xd.cpu().numpy().astype(np.float64)
        got = a.cpu().numpy().astype(np.float64)
        for t in (0, batch // 2, batch - 1):
            sl = slice(2 * n * t, 2 * n * (t + 1))
            ref = ol.run_oracle(op, [n], x[sl])
            if norm == "inverse":
                ref = ref / n
            err = ol.rel_l2(got[sl], ref)
            print(f"  transform {t}: rel-L2 {err:.3e}")
            assert err < tol_for("f32", n)


@pytest.mark.parametrize("n,batch", [(1 << 16, 1024), (1 << 17, 512)])
def test_xcd_slot_reuse_small_columns(pkg, dev, n, batch, monkeypatch):
    """N1 = 64 / 128: a transform is only 8 / 16 tiles, so each intermediate
    slot is reused within microseconds and a CU's L1 can still hold lines of
    the slot's previous transform when a row tile reads it.  The row tiles
    read the slot past the L1 (agent-scope loads); plain loads after a
    workgroup-scope invalidate returned whole stale rows here, different ones
    on each run.  Three launches must each match the two-launch four-step
    bit for bit."""
    import torch
    xd = torch.empty(2 * n * batch, dtype=torch.float32, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 9, 0, "f32")
    ref = _run(pkg, dev, n, batch, "dft", xd, False, monkeypatch)
    for r in range(3):
        got = _run(pkg, dev, n, batch, "dft", xd, True, monkeypatch)
        bad = (got.view(batch, n, 2) != ref.view(batch, n, 2)).any(dim=2).any(dim=1)
        print(f"N={n} x {batch} run {r}: transforms differing from two launches: {int(bad.sum())}")
        assert not bool(bad.any())


def test_xcd_in_place(pkg, dev, monkeypatch):
    import torch
    n, batch = 1 << 20, 24
    xd = torch.empty(2 * n * batch, dtype=torch.float32, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 6, 0, "f32")
    a = _run(pkg, dev, n, batch, "dft", xd, True, monkeypatch, inplace=True, reps=3)
    b = _run(pkg, dev, n, batch, "dft", xd, False, monkeypatch)
    assert torch.equal(a, b)


def test_xcd_full_headline_batch(pkg, dev, monkeypatch):
    """N = 2^20 x 256 (the benchmark batch): bit-identical to the two-launch
    path, and transforms 0 and 255 against the oracle."""
    import torch
    n, batch = 1 << 20, 256
    xd = torch.empty(2 * n * batch, dtype=torch.float32, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 0, 0, "f32")
    a = _run(pkg, dev, n, batch, "dft", xd, True, monkeypatch, reps=3)
    b = _run(pkg, dev, n, batch, "dft", xd, False, monkeypatch)
    assert torch.equal(a, b)
    for t in (0, batch - 1):
        x = ol.uniform(0, 2 * n, start=2 * n * t).astype(np.float32).astype(np.float64)
        got = a[2 * n * t:2 * n * (t + 1)].cpu().numpy().astype(np.float64)
        err = ol.rel_l2(got, ol.run_oracle("dft", [n], x))
        print(f"2^20 x 256: transform {t} rel-L2 {err:.3e}")
        assert err < tol_for("f32", n)
