"""GPU parity of the complex transforms (single-pass lines, four-step,
Bluestein) against the oracle, through the C ABI.

Tolerances (rel-L2 against the FP64 oracle applied to the same FP32-rounded
inputs): FP32 1e-6 * log2 N, FP64 1e-14 * log2 N.  Bluestein FP64 is checked
against numpy.fft (pocketfft), because the reference's own chirp phase is not
reduced mod 2N and carries an error growing with N; the distance to the
reference's algorithm is bounded by its own test tolerance (1e-9,
tests/test_notoriousfft.c:659) instead.
"""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import rounded, run_plan, tol_for

pytestmark = pytest.mark.gpu

LINE32 = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384]
LINE64 = [1, 2, 4, 8, 16, 64, 256, 1024, 4096, 8192]


def cx(a):
    return a[0::2] + 1j * a[1::2]


def il(z):
    z = np.asarray(z).reshape(-1)
    out = np.empty(2 * z.size)
    out[0::2], out[1::2] = z.real, z.imag
    return out


@pytest.mark.parametrize("prec,n", [("f32", n) for n in LINE32] + [("f64", n) for n in LINE64])
def test_pow2_single_pass(pkg, dev, prec, n):
    batch = 5
    x = rounded(ol.uniform(1, 2 * n * batch), prec)
    for op in ("dft", "invdft"):
        got = run_plan(pkg, dev, op, [n], x, batch, prec)
        ref = ol.run_oracle(op, [n], x, batch)
        err = ol.rel_l2(got, ref)
        print(f"{prec} {op} N={n} rel-L2 {err:.3e}")
        assert err < tol_for(prec, n), (op, err)


@pytest.mark.parametrize("prec,n,batch", [("f32", 1 << 15, 3), ("f32", 1 << 16, 2), ("f32", 1 << 18, 2),
                                          ("f32", 1 << 20, 2), ("f32", 1 << 21, 1), ("f64", 1 << 14, 3),
                                          ("f64", 1 << 17, 2), ("f64", 1 << 20, 1)])
@pytest.mark.parametrize("row_twiddle", ["1", "0"])
def test_pow2_four_step(pkg, dev, prec, n, batch, row_twiddle, monkeypatch):
    """FP32 four-steps with 1024-point rows apply the inter-pass twiddle in the
    row pass (NFFT_ROW_TWIDDLE=1, the default); 0 keeps it in the column pass
    as every other four-step does.  Both against the oracle.  FP32 2^15 and
    FP64 2^14 default to one line per workgroup (test_pow2_line_halves); the
    four-step is forced here."""
    monkeypatch.setenv("NFFT_ROW_TWIDDLE", row_twiddle)
    monkeypatch.setenv("NFFT_LINE_H", "0")
    x = rounded(ol.uniform(2, 2 * n * batch), prec)
    for op in ("dft", "invdft"):
        got = run_plan(pkg, dev, op, [n], x, batch, prec)
        ref = ol.run_oracle(op, [n], x, batch)
        err = ol.rel_l2(got, ref)
        print(f"{prec} {op} N={n} rel-L2 {err:.3e}")
        assert err < tol_for(prec, n), (op, err)


@pytest.mark.parametrize("prec,n", [("f32", 1 << 15), ("f64", 1 << 14)])
def test_pow2_line_halves(pkg, dev, prec, n):
    """Twice the LDS-resident line length: one line per workgroup, its
    exchanges made through half a line of LDS in two rounds (k_line_h).  A
    batch of many workgroups, both directions, in place and out of place,
    against the oracle."""
    batch = 67
    plan = pkg.Plan("dft", prec, [n], batch, "reference")
    desc = plan.describe()
    plan.close()
    assert "two LDS halves" in desc, desc
    x = rounded(ol.uniform(11, 2 * n * batch), prec)
    for op in ("dft", "invdft"):
        ref = ol.run_oracle(op, [n], x, batch)
        for inplace in (False, True):
            got = run_plan(pkg, dev, op, [n], x, batch, prec, inplace=inplace)
            err = ol.rel_l2(got, ref)
            print(f"{prec} {op} N={n} x{batch} in-place={inplace}: rel-L2 {err:.3e} [{desc}]")
            assert err < tol_for(prec, n), (op, inplace, err)


@pytest.mark.parametrize("prec,n,default_halves", [("f32", 8192, True), ("f64", 8192, True), ("f32", 16384, True),
                                                    ("f32", 4096, False)])
def test_line_halves_bitwise(pkg, dev, monkeypatch, prec, n, default_halves):
    """The half-LDS line and the LDS-resident line run the same arithmetic in
    the same order per point: bit-identical outputs (NFFT_LINE_H=0 turns the
    half form off).  The default picks it wherever it exists (from 8192
    points; 4096 has no such form)."""
    import torch
    batch = 300
    dt = torch.float32 if prec == "f32" else torch.float64
    x = torch.empty(2 * n * batch, dtype=dt, device=dev)
    pkg.fill_uniform(x.data_ptr(), x.numel(), 12, 0, prec)
    plan = pkg.Plan("dft", prec, [n], batch)
    assert ("two LDS halves" in plan.describe()) == default_halves, plan.describe()
    plan.close()
    outs = {}
    for mode in ("2", "0"):
        monkeypatch.setenv("NFFT_LINE_H", mode)
        plan = pkg.Plan("dft", prec, [n], batch)
        assert ("two LDS halves" in plan.describe()) == (mode == "2" and default_halves), plan.describe()
        for op in ("dft", "invdft"):
            y = torch.empty_like(x)
            plan.execute(op, x.data_ptr(), y.data_ptr())
            outs[(mode, op)] = y
        plan.check()
        plan.close()
    for op in ("dft", "invdft"):
        assert torch.equal(outs[("2", op)], outs[("0", op)]), op


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_largest_pow2_properties(pkg, dev, prec):
    # at the largest supported length: round trip and Parseval (size-free checks)
    n = 1 << 22 if prec == "f32" else 1 << 21
    x = rounded(ol.uniform(3, 2 * n), prec)
    y = run_plan(pkg, dev, "dft", [n], x, 1, prec)
    back = run_plan(pkg, dev, "invdft", [n], y, 1, prec)
    tol = tol_for(prec, n)
    assert ol.rel_l2(back / n, x) < 2 * tol
    assert abs(np.dot(y, y) / (n * np.dot(x, x)) - 1) < 2 * tol
    # spot-check a handful of bins against a direct sum
    z = cx(x)
    k = np.array([0, 1, 7, n // 3, n - 1])
    direct = np.array([np.sum(z * np.exp(-2j * np.pi * ((kk * np.arange(n)) % n) / n)) for kk in k])
    assert np.max(np.abs(cx(y)[k] - direct)) / np.sqrt(np.sum(np.abs(z) ** 2)) < 10 * tol


BLUE = [3, 5, 6, 7, 10, 12, 15, 17, 20, 100, 127, 1000, 1009, 4097, 10007, 65521]


@pytest.mark.parametrize("prec,n", [("f32", n) for n in BLUE] + [("f64", n) for n in BLUE])
def test_bluestein(pkg, dev, prec, n):
    batch = 3
    x = rounded(ol.uniform(4, 2 * n * batch), prec)
    hi = il(np.fft.fft(cx(x).reshape(batch, n), axis=1))
    got = run_plan(pkg, dev, "dft", [n], x, batch, prec)
    err = ol.rel_l2(got, hi)
    ref = ol.run_oracle("dft", [n], x, batch)
    dref = ol.rel_l2(got, ref)
    print(f"{prec} bluestein N={n}: rel-L2 vs pocketfft {err:.3e}, vs reference algorithm {dref:.3e}, "
          f"reference vs pocketfft {ol.rel_l2(ref, hi):.3e}")
    assert err < tol_for(prec, n)
    assert dref < (tol_for(prec, n) if prec == "f32" else 1e-9)
    # inverse: normalised by 1/N for non-power-of-two lengths (reference behaviour)
    goti = run_plan(pkg, dev, "invdft", [n], x, batch, prec)
    hii = il(np.fft.ifft(cx(x).reshape(batch, n), axis=1))
    assert ol.rel_l2(goti, hii) < tol_for(prec, n)


@pytest.mark.parametrize("norm,expect", [("none", 1.0), ("inverse", None), ("reference", None)])
def test_norm_flags(pkg, dev, norm, expect):
    for n in (64, 100):
        x = ol.uniform(5, 2 * n)
        got = run_plan(pkg, dev, "invdft", [n], x, 1, "f64", norm=norm)
        raw = il(np.fft.ifft(cx(x))) * n
        if norm == "none":
            want = raw
        elif norm == "inverse":
            want = raw / n
        else:
            want = raw if n == 64 else raw / n
        assert ol.rel_l2(got, want) < 1e-13, (norm, n)


def test_inplace_and_odd_batches(pkg, dev):
    for n, batch in ((1024, 7), (1 << 16, 3), (17, 9), (1 << 20, 1)):
        x = rounded(ol.uniform(6, 2 * n * batch), "f32")
        got = run_plan(pkg, dev, "dft", [n], x, batch, "f32", inplace=True)
        ref = ol.run_oracle("dft", [n], x, batch)
        assert ol.rel_l2(got, ref) < tol_for("f32", n), (n, batch)


def test_dropin_f64_host_pointers(pkg, dev):
    """The reference API through libnotorious_fft_hip_f64.so on host memory:
    the C1 configuration (N = 4096 dft then invdft, seed 0) and edge sizes."""
    lib = pkg.dropin("f64")
    for n in (1, 2, 8, 16, 4096, 100, 1009, 1 << 16):
        a = lib.notorious_fft_mkaux_dft_1d(n)
        assert a
        x = ol.uniform(0, 2 * n)
        x0 = x.copy()
        y = np.zeros_like(x)
        z = np.zeros_like(x)
        lib.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
        lib.notorious_fft_invdft(y.ctypes.data, z.ctypes.data, a)
        assert np.array_equal(x, x0)  # input untouched
        ref = ol.run_oracle("dft", [n], x)
        e1 = ol.rel_l2(y, ref)
        scale = n if (n & (n - 1)) == 0 else 1
        e2 = ol.rel_l2(z, scale * x)
        print(f"dropin f64 N={n}: dft rel-L2 {e1:.3e}, round trip {e2:.3e}")
        blue = (n & (n - 1)) != 0
        assert e1 < (1e-9 if blue else tol_for("f64", n))
        assert e2 < tol_for("f64", n) * (4 if blue else 1)
        # x == y in place
        w = x.copy()
        lib.notorious_fft_dft(w.ctypes.data, w.ctypes.data, a)
        assert np.array_equal(w, y)
        lib.notorious_fft_free_aux(a)
    assert not lib.notorious_fft_mkaux_dft_1d(0)
    assert not lib.notorious_fft_mkaux_realdft_1d(12)
    lib.notorious_fft_free_aux(None)


def test_dropin_f64_device_pointers(pkg, dev):
    import torch
    lib = pkg.dropin("f64")
    n = 2048
    x = ol.uniform(0, 2 * n)
    xd = torch.from_numpy(x).to(dev)
    yd = torch.empty_like(xd)
    a = lib.notorious_fft_mkaux_dft_1d(n)
    torch.cuda.synchronize()
    lib.notorious_fft_dft(xd.data_ptr(), yd.data_ptr(), a)
    lib.notorious_fft_free_aux(a)
    assert ol.rel_l2(yd.cpu().numpy(), ol.run_oracle("dft", [n], x)) < tol_for("f64", n)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dropin_host_paths_bitwise(pkg, dev, prec, monkeypatch):
    """Host buffers reach the same kernels three ways -- 1-D transforms up
    to 256 KiB run on mapped staging buffers, larger ones and N-D ones go
    through DMA transfers (NFFT_DROPIN_ZC_MAX=0 forces those) -- and every
    way gives the device-pointer result bit for bit, in place too."""
    import torch
    lib = pkg.dropin(prec)
    dt = np.float64 if prec == "f64" else np.float32
    cases = [("dft_1d", (4096,), "dft", 8192, 8192), ("dft_1d", (4096,), "invdft", 8192, 8192),
             ("dft_1d", (1009,), "dft", 2018, 2018), ("realdft_1d", (8192,), "realdft", 8192, 8194),
             ("realdft_1d", (8192,), "invrealdft", 8194, 8192), ("t2t3_1d", (2048,), "dct3", 2048, 2048),
             ("t4_1d", (4096,), "dst4", 4096, 4096), ("dft_2d", (64, 32), "dft", 4096, 4096),
             ("dft_1d", (1 << 15,), "dft", 1 << 16, 1 << 16)]
    for i, (mk, dims, op, ni, no) in enumerate(cases):
        a = getattr(lib, "notorious_fft_mkaux_" + mk)(*dims)
        assert a
        fn = getattr(lib, "notorious_fft_" + op)
        x = ol.uniform(40 + i, ni).astype(dt)
        xd = torch.from_numpy(x).to(dev)
        yd = torch.zeros(no, dtype=xd.dtype, device=dev)
        torch.cuda.synchronize()
        fn(xd.data_ptr(), yd.data_ptr(), a)
        want = yd.cpu().numpy()
        for zc in ("0", "262144"):
            monkeypatch.setenv("NFFT_DROPIN_ZC_MAX", zc)
            y = np.full(no, 3.0, dt)
            fn(x.ctypes.data, y.ctypes.data, a)
            assert np.array_equal(y, want), (mk, dims, op, zc)
            if ni == no:
                w = x.copy()
                fn(w.ctypes.data, w.ctypes.data, a)
                assert np.array_equal(w, want), (mk, dims, op, zc, "in place")
            # host input, device output
            yd2 = torch.zeros_like(yd)
            torch.cuda.synchronize()
            fn(x.ctypes.data, yd2.data_ptr(), a)
            assert torch.equal(yd2, yd), (mk, dims, op, zc, "device output")
        monkeypatch.delenv("NFFT_DROPIN_ZC_MAX")
        lib.notorious_fft_free_aux(a)


@pytest.mark.parametrize("side_stream", [False, True])
def test_dropin_orders_after_queued_writes(pkg, dev, side_stream):
    """The caller queues the write of x (a copy behind a long chain of its own
    kernels, on torch's current stream or on a stream of its own) and calls
    the drop-in with no synchronisation: the synchronous reference contract
    says the transform sees the written x."""
    import torch
    lib = pkg.dropin("f64")
    n = 4096
    x = ol.uniform(5, 2 * n)
    src = torch.from_numpy(x).to(dev)
    xd = torch.zeros_like(src)
    yd = torch.zeros_like(src)
    a = lib.notorious_fft_mkaux_dft_1d(n)
    busy = torch.ones(1 << 24, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    st = torch.cuda.Stream() if side_stream else torch.cuda.current_stream()
    with torch.cuda.stream(st):
        for _ in range(40):
            busy.mul_(1.0000001)
        xd.copy_(src)
    lib.notorious_fft_dft(xd.data_ptr(), yd.data_ptr(), a)
    got = yd.cpu().numpy()
    torch.cuda.synchronize()
    lib.notorious_fft_free_aux(a)
    assert ol.rel_l2(got, ol.run_oracle("dft", [n], x)) < tol_for("f64", n)


def test_dropin_f32(pkg, dev):
    lib = pkg.dropin("f32")
    for n in (16, 1024, 100):
        a = lib.notorious_fft_mkaux_dft_1d(n)
        x = ol.uniform(0, 2 * n).astype(np.float32)
        y = np.zeros_like(x)
        lib.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
        lib.notorious_fft_free_aux(a)
        ref = ol.run_oracle("dft", [n], x.astype(np.float64))
        assert ol.rel_l2(y.astype(np.float64), ref) < tol_for("f32", n)


@pytest.mark.parametrize("prec,n,batch", [("f32", 1 << 20, 5), ("f64", 1 << 18, 7), ("f32", 1 << 15, 9)])
@pytest.mark.parametrize("persist", ["1", "0"])
def test_four_step_small_scratch(pkg, dev, prec, n, batch, persist, monkeypatch):
    """A 1 MiB scratch: the persistent four-step then runs with a single
    rotating slot (every pass-1 tile waits for the previous transform's pass
    2), the two-launch path with many chunks; in place too."""
    monkeypatch.setenv("NFFT_CHUNK_MB", "1")
    monkeypatch.setenv("NFFT_LINE_H", "0")
    monkeypatch.setenv("NFFT_PERSIST_MB", "1")
    monkeypatch.setenv("NFFT_PERSIST", persist)
    x = rounded(ol.uniform(7, 2 * n * batch), prec)
    ref = ol.run_oracle("dft", [n], x, batch)
    for inplace in (False, True):
        got = run_plan(pkg, dev, "dft", [n], x, batch, prec, inplace=inplace)
        assert ol.rel_l2(got, ref) < tol_for(prec, n), (inplace, ol.rel_l2(got, ref))
    refi = ol.run_oracle("invdft", [n], x, batch)
    got = run_plan(pkg, dev, "invdft", [n], x, batch, prec)
    assert ol.rel_l2(got, refi) < tol_for(prec, n)


def persist_variants(prec, n):
    """Shapes built for length n (variant 0 is the default)."""
    return {("f32", 1 << 20): [0, 1, 2, 3, 4, 5], ("f64", 1 << 20): [0, 1]}.get((prec, n), [0])


@pytest.mark.parametrize("prec,n,batch", [("f32", 1 << 20, 3), ("f64", 1 << 20, 2)])
def test_persist_variants(pkg, dev, prec, n, batch, monkeypatch):
    """Every alternative tile shape of the persistent four-step."""
    monkeypatch.setenv("NFFT_PERSIST", "1")
    x = rounded(ol.uniform(8, 2 * n * batch), prec)
    ref = ol.run_oracle("dft", [n], x, batch)
    for v in persist_variants(prec, n):
        monkeypatch.setenv("NFFT_PERSIST_VAR", str(v))
        got = run_plan(pkg, dev, "dft", [n], x, batch, prec)
        err = ol.rel_l2(got, ref)
        print(f"{prec} N={n} variant {v}: rel-L2 {err:.3e}")
        assert err < tol_for(prec, n), (v, err)


def test_persist_many_transforms(pkg, dev, monkeypatch):
    """More transforms than rotating slots and a batch that is not a multiple
    of anything: N = 2^15 x 300 with a 2 MiB scratch (8 slots)."""
    n, batch = 1 << 15, 300
    x = rounded(ol.uniform(9, 2 * n * batch), "f32")
    ref = ol.run_oracle("dft", [n], x, batch)
    monkeypatch.setenv("NFFT_PERSIST_MB", "2")
    monkeypatch.setenv("NFFT_LINE_H", "0")
    monkeypatch.setenv("NFFT_PERSIST", "1")
    got = run_plan(pkg, dev, "dft", [n], x, batch, "f32")
    assert ol.rel_l2(got, ref) < tol_for("f32", n)


@pytest.mark.parametrize("grab", [2, 5])
def test_persist_grab(pkg, dev, grab, monkeypatch):
    """Several queue items per access (runs that straddle the two passes and
    the end of the queue)."""
    monkeypatch.setenv("NFFT_PERSIST", "1")
    monkeypatch.setenv("NFFT_PERSIST_GRAB", str(grab))
    for prec, n, batch in (("f32", 1 << 16, 7), ("f64", 1 << 17, 3)):
        x = rounded(ol.uniform(10, 2 * n * batch), prec)
        for op in ("dft", "invdft"):
            got = run_plan(pkg, dev, op, [n], x, batch, prec)
            assert ol.rel_l2(got, ol.run_oracle(op, [n], x, batch)) < tol_for(prec, n), (prec, op)


SMOOTH = [6, 9, 12, 15, 25, 27, 45, 49, 96, 100, 210, 243, 343, 625, 1000, 1715, 2187, 3000, 4725, 6561, 8000]


@pytest.mark.parametrize("prec,n", [("f32", n) for n in SMOOTH] + [("f64", n) for n in SMOOTH if n <= 4096])
def test_mixed_radix(pkg, dev, prec, n, monkeypatch):
    """7-smooth lengths run as mixed-radix lines (several lines per
    workgroup, a ragged last workgroup): vs pocketfft, and the plan says so."""
    monkeypatch.setenv("NFFT_SMOOTH", "2")  # mixed radix wherever possible
    batch = 7
    x = rounded(ol.uniform(11, 2 * n * batch), prec)
    hi = il(np.fft.fft(cx(x).reshape(batch, n), axis=1))
    got = run_plan(pkg, dev, "dft", [n], x, batch, prec)
    err = ol.rel_l2(got, hi)
    print(f"{prec} mixed-radix N={n}: rel-L2 vs pocketfft {err:.3e}")
    assert err < tol_for(prec, n)
    goti = run_plan(pkg, dev, "invdft", [n], x, batch, prec)
    assert ol.rel_l2(goti, il(np.fft.ifft(cx(x).reshape(batch, n), axis=1))) < tol_for(prec, n)
    p = pkg.Plan("dft", prec, [n], batch)
    assert p.describe().startswith("mixed-radix"), p.describe()
    p.close()


SMOOTH4 = [10000, 15625, 40000, 65610, 196608, 1000000, 531441]


@pytest.mark.parametrize("prec,n", [("f32", n) for n in SMOOTH4] + [("f64", n) for n in SMOOTH4[:5]])
def test_mixed_radix_four_step(pkg, dev, prec, n, monkeypatch):
    """7-smooth lengths above one workgroup: mixed-radix column pass with the
    inter-pass twiddle, then row pass with a transposed store."""
    monkeypatch.setenv("NFFT_SMOOTH", "2")
    batch = 3 if n <= 200000 else 1
    x = rounded(ol.uniform(12, 2 * n * batch), prec)
    hi = il(np.fft.fft(cx(x).reshape(batch, n), axis=1))
    got = run_plan(pkg, dev, "dft", [n], x, batch, prec)
    err = ol.rel_l2(got, hi)
    print(f"{prec} mixed-radix four-step N={n}: rel-L2 vs pocketfft {err:.3e}")
    assert err < tol_for(prec, n)
    goti = run_plan(pkg, dev, "invdft", [n], x, batch, prec, inplace=True)
    assert ol.rel_l2(goti, il(np.fft.ifft(cx(x).reshape(batch, n), axis=1))) < tol_for(prec, n)
    p = pkg.Plan("dft", prec, [n], batch)
    assert p.describe().startswith("mixed-radix four-step"), p.describe()
    p.close()


@pytest.mark.parametrize("prec,lg", [("f32", 16), ("f64", 18)])
def test_four_step_autotuner(pkg, dev, tmp_path, monkeypatch, prec, lg):
    """NFFT_TUNE=1 times every row length (256..2048 points) with both row
    passes at plan creation and records the fastest as
    "c2c-four p<prec> n<N> <log2 rows> <prefetching grid>"; later plans
    follow the record (also one written by hand), and every choice matches
    the oracle."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("NFFT_TUNE_CACHE", str(cache))
    monkeypatch.setenv("NFFT_TUNE", "1")
    n, batch = 1 << lg, 4
    x = rounded(ol.uniform(17, 2 * n * batch), prec)
    want = ol.run_oracle("dft", [n], x[:2 * n])
    got = run_plan(pkg, dev, "dft", [n], x, batch, prec)
    assert ol.rel_l2(got[:2 * n], want) < tol_for(prec, n)
    lines = [l.split() for l in cache.read_text().splitlines()]
    pk = "p0" if prec == "f32" else "p1"
    rec = [l for l in lines if l[:3] == ["c2c-four", pk, f"n{n}"]]
    assert len(rec) == 1 and 8 <= int(rec[0][3]) <= 11 and rec[0][4] in ("0", "1"), lines
    monkeypatch.setenv("NFFT_TUNE", "0")
    p = pkg.Plan("dft", prec, [n], batch)
    assert f"x{1 << int(rec[0][3])}" in p.describe(), (p.describe(), rec)
    p.close()
    assert len(cache.read_text().splitlines()) == len(lines)  # no second measurement
    # a record written by hand: 256-point rows on the one-workgroup-per-tile pass
    cache.write_text(f"c2c-four {pk} n{n} 8 0\n")
    p = pkg.Plan("dft", prec, [n], batch)
    assert p.describe().startswith(f"four-step {n // 256}x256"), p.describe()
    p.close()
    got = run_plan(pkg, dev, "dft", [n], x, batch, prec)
    assert ol.rel_l2(got[:2 * n], want) < tol_for(prec, n)


def test_smooth_autotuner(pkg, dev, tmp_path, monkeypatch):
    """NFFT_TUNE=1 times mixed radix against Bluestein at plan creation and
    records the winner in NFFT_TUNE_CACHE; later plans follow the record,
    and both choices give the right answer."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("NFFT_TUNE_CACHE", str(cache))
    monkeypatch.setenv("NFFT_TUNE", "1")
    n, batch = 3000, 4
    x = rounded(ol.uniform(13, 2 * n * batch), "f32")
    want = il(np.fft.fft(cx(x).reshape(batch, n), axis=1))
    got = run_plan(pkg, dev, "dft", [n], x, batch, "f32")
    assert ol.rel_l2(got, want) < tol_for("f32", n)
    lines = [l.split() for l in cache.read_text().splitlines()]
    rec = [l for l in lines if l[:3] == ["c2c-smooth", "p0", f"n{n}"]]
    assert len(rec) == 1 and rec[0][3] in ("0", "1"), lines
    monkeypatch.setenv("NFFT_TUNE", "0")
    p = pkg.Plan("dft", "f32", [n], batch)
    assert p.describe().startswith("mixed-radix") == (rec[0][3] == "1"), p.describe()
    p.close()
    assert len(cache.read_text().splitlines()) == len(lines)  # no second measurement


ROWT_PF_CASES = ([("f32", 16, n2, 26) for n2 in (256, 512, 1024, 2048)] +
                 [("f32", lg, 0, 26) for lg in range(15, 23)] +
                 [("f32", 20, n2, 26) for n2 in (512, 2048)] +
                 [("f64", 16, n2, 25) for n2 in (256, 512, 1024)] +
                 [("f64", lg, 0, 25) for lg in range(14, 22)] +
                 [("f64", 19, 512, 25)])


@pytest.mark.parametrize("prec,lg,n2,lgtot", ROWT_PF_CASES)
def test_rowT_prefetch_grid_bitwise(pkg, dev, prec, lg, n2, lgtot, monkeypatch):
    """The resident prefetching row pass against the one-workgroup-per-tile
    row pass: the same arithmetic in the same order, so the outputs agree bit
    for bit.  Batches of 2^lgtot points give every workgroup many tiles; row
    lengths n2 (0: the planner's split) cover every prefetching shape."""
    import torch
    n = 1 << lg
    batch = 1 << (lgtot - lg)
    dt = torch.float32 if prec == "f32" else torch.float64
    x = torch.empty(2 * n * batch, dtype=dt, device=dev)
    pkg.fill_uniform(x.data_ptr(), x.numel(), 11, 0, prec)
    monkeypatch.setenv("NFFT_PERSIST", "0")
    monkeypatch.setenv("NFFT_LINE_H", "0")  # FP32 2^15 / FP64 2^14 would run as one line per workgroup
    if n2:
        monkeypatch.setenv("NFFT_FOUR_N2", str(n2))
    outs = {}
    for pf in ("1", "0"):
        monkeypatch.setenv("NFFT_ROWT_PF", pf)
        plan = pkg.Plan("dft", prec, [n], batch)
        assert "four-step" in plan.describe(), plan.describe()
        if n2:
            assert f"x{n2}" in plan.describe(), plan.describe()
        for op in ("dft", "invdft"):
            y = torch.empty_like(x)
            plan.execute(op, x.data_ptr(), y.data_ptr())
            outs[(pf, op)] = y
        torch.cuda.synchronize()
        plan.check()
        plan.close()
    for op in ("dft", "invdft"):
        assert torch.equal(outs[("1", op)], outs[("0", op)]), (op, (outs[("1", op)] - outs[("0", op)]).abs().max())


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_plans_created_while_others_execute(pkg, dev, prec):
    """Plans built (tables and Bluestein filters on the setup stream) while
    earlier plans' transforms still run on the caller's stream, with no
    synchronisation in between; every transform against pocketfft.  FP64
    stores of 16-byte elements through a scalar offset were once corrupted
    intermittently in exactly this situation."""
    import torch
    from ragged_sizes import ragged_sizes
    sizes = ragged_sizes(200, seed=2)
    total = sum(sizes)
    dt = torch.float32 if prec == "f32" else torch.float64
    xd = torch.empty(2 * total, dtype=dt, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 3, 0, prec)
    x = xd.cpu().numpy().astype(np.float64)
    for rnd in range(3):
        yd = torch.empty_like(xd)
        plans, off = {}, 0
        for n in sizes:
            if n not in plans:
                plans[n] = pkg.Plan("dft", prec, [n], 1)
            plans[n].execute("dft", xd[2 * off:].data_ptr(), yd[2 * off:].data_ptr())
            off += n
        torch.cuda.synchronize()
        y = yd.cpu().numpy().astype(np.float64)
        off = 0
        for n in sizes:
            want = np.fft.fft(x[2 * off:2 * (off + n):2] + 1j * x[2 * off + 1:2 * (off + n):2])
            got = y[2 * off:2 * (off + n):2] + 1j * y[2 * off + 1:2 * (off + n):2]
            err = ol.rel_l2(got, want)
            assert err < tol_for(prec, n), (rnd, n, err, plans[n].describe())
            off += n
        for p in plans.values():
            p.close()
