"""Ragged batches (one persistent launch replayed from a HIP graph): every
transform against pocketfft (numpy.fft, FP64), the replays against eager
per-size plans, with the program's own work between replays."""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import rounded, tol_for
from ragged_sizes import ragged_sizes

pytestmark = pytest.mark.gpu


def make(pkg, prec, sizes, flags=0):
    lib = pkg.core()
    h = ctypes.c_void_p()
    arr = (ctypes.c_int * len(sizes))(*sizes)
    pkg.check(lib.nfft_ragged_create(ctypes.byref(h), 0 if prec == "f32" else 1, len(sizes), arr, flags),
              "ragged_create")
    return h


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ragged_vs_pocketfft(pkg, dev, prec):
    sizes = ragged_sizes(300, seed=1)
    sizes += [1, 2, 3, 16, 4096, 4097, 8192, 65536, 65521, 2 * 3 * 5 * 7]
    check_vs_pocketfft(pkg, dev, prec, sizes)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ragged_two_level_radices(pkg, dev, prec):
    """7-smooth lengths whose mixed-radix schedules use every two-level radix
    (6, 9, 10, 12, 15, 16, 20, 25) and 7, as lines and as four-step tiles,
    against pocketfft in both directions."""
    sizes = [6, 9, 10, 12, 15, 16 * 3, 20, 25, 50, 36, 60, 150, 400, 900, 980, 1568, 2025, 3600, 6125,
             10000, 14400, 20736, 50625, 2 * 3 * 5 * 7 * 16, 7 * 7 * 25 * 16]
    check_vs_pocketfft(pkg, dev, prec, sizes)


@pytest.mark.parametrize("prec,mb,gpt,lag", [("f32", 2, 1, 8), ("f32", 4, 3, 1), ("f64", 4, 2, 8),
                                             ("f64", 2, 64, 0)])
def test_ragged_slot_reuse(pkg, dev, prec, mb, gpt, lag, monkeypatch):
    """A scratch budget of a few slots: most phase-1 tiles wait for their
    slot's previous user; tiles of one group up to whole phases."""
    monkeypatch.setenv("NFFT_RAGGED_MB", str(mb))
    monkeypatch.setenv("NFFT_RAGGED_GPT", str(gpt))
    monkeypatch.setenv("NFFT_RAGGED_LAG", str(lag))
    sizes = [65521, 8192, 40000, 65536, 5003, 16384, 32771, 9000, 30030, 65536, 12289, 4099, 16, 100]
    check_vs_pocketfft(pkg, dev, prec, sizes)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ragged_autotuner(pkg, dev, prec, tmp_path, monkeypatch):
    """flags bit 0: the tile size and pipeline lag are measured over five
    candidates (one "ragged-gl" record, value groups x 100 + lag quarters),
    then the workgroups per CU ("ragged" record); a second create with the
    same sizes reads both records and measures nothing; both batches match
    pocketfft."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("NFFT_TUNE_CACHE", str(cache))
    sizes = ragged_sizes(120, seed=4) + [65521, 65536, 40000, 30030]
    check_vs_pocketfft(pkg, dev, prec, sizes, flags=1)
    lines = [l.split() for l in cache.read_text().splitlines()]
    pk = "p0" if prec == "f32" else "p1"
    gl = [l for l in lines if l[0] == "ragged-gl" and l[1] == pk]
    pc = [l for l in lines if l[0] == "ragged" and l[1] == pk]
    assert len(gl) == 1 and len(pc) == 1, lines
    v = int(gl[0][4])
    assert (v // 100, v % 100) in ((16, 4), (8, 4), (32, 4), (16, 2), (16, 8)), gl
    assert 1 <= int(pc[0][4]) <= 4, pc
    check_vs_pocketfft(pkg, dev, prec, sizes, flags=1)
    assert len(cache.read_text().splitlines()) == len(lines)  # read back, not measured again


def check_vs_pocketfft(pkg, dev, prec, sizes, flags=0):
    import torch
    lib = pkg.core()
    h = make(pkg, prec, sizes, flags)
    total = lib.nfft_ragged_total(h)
    assert total == sum(sizes)
    x = rounded(ol.uniform(51, 2 * total), prec)
    dt = torch.float32 if prec == "f32" else torch.float64
    xd = torch.from_numpy(x.astype(np.float32 if prec == "f32" else np.float64)).to(dev)
    yd = torch.empty_like(xd)
    for inverse in (0, 1):
        torch.cuda.synchronize()
        pkg.check(lib.nfft_ragged_execute(h, inverse, xd.data_ptr(), yd.data_ptr(), None), "ragged_execute")
        pkg.check(lib.nfft_ragged_check(h, None), "ragged_check")
        y = yd.cpu().numpy().astype(np.float64)
        off = 0
        worst = 0.0
        for n in sizes:
            z = x[2 * off:2 * (off + n)]
            zc = z[0::2] + 1j * z[1::2]
            if inverse:
                want = np.fft.ifft(zc) * (n if (n & (n - 1)) == 0 else 1)
            else:
                want = np.fft.fft(zc)
            got = y[2 * off:2 * (off + n)]
            err = ol.rel_l2(got[0::2] + 1j * got[1::2], want)
            worst = max(worst, err / tol_for(prec, n))
            assert err < tol_for(prec, n), (prec, inverse, n, err)
            off += n
        print(f"{prec} inverse={inverse}: worst error / tolerance = {worst:.3f}")
    assert lib.nfft_ragged_destroy(h) == 0
    del dt


def test_ragged_graph_replays_match_eager(pkg, dev):
    import torch
    sizes = ragged_sizes(200, seed=2)
    lib = pkg.core()
    h = make(pkg, "f32", sizes)
    total = sum(sizes)
    xd = torch.empty(2 * total, dtype=torch.float32, device=dev)
    pkg.fill_uniform(xd.data_ptr(), xd.numel(), 3, 0, "f32")
    # eager reference: one batched plan per distinct size, transform by transform
    ref = torch.empty_like(xd)
    off = 0
    plans = {}
    for n in sizes:
        if n not in plans:
            plans[n] = pkg.Plan("dft", "f32", [n], 1)
        plans[n].execute("dft", xd[2 * off:].data_ptr(), ref[2 * off:].data_ptr())
        off += n
    torch.cuda.synchronize()
    for rep in range(4):
        yd = torch.full_like(xd, float("nan"))
        junk = torch.randn(1 << 20, device=dev).sum()  # program's own work between replays
        torch.cuda.synchronize()
        pkg.check(lib.nfft_ragged_execute(h, 0, xd.data_ptr(), yd.data_ptr(), None), "ragged_execute")
        pkg.check(lib.nfft_ragged_check(h, None), "ragged_check")
        d = (yd - ref).abs().max().item()
        if not d <= 1e-4 * ref.abs().max().item():
            off, worst = 0, []
            for k, n in enumerate(sizes):
                dk = (yd[2 * off:2 * (off + n)] - ref[2 * off:2 * (off + n)]).abs().max().item()
                if not dk <= 1e-4 * ref.abs().max().item():
                    worst.append((k, n, round(dk, 3)))
                off += n
            # which side is off: both against pocketfft for the worst transforms
            xs = xd.cpu().numpy().astype(np.float64)
            ys, rs = yd.cpu().numpy().astype(np.float64), ref.cpu().numpy().astype(np.float64)
            starts = np.concatenate([[0], np.cumsum(sizes)])
            side = []
            for k, n, _ in worst[:6]:
                o = int(starts[k])
                want = np.fft.fft(xs[2 * o:2 * (o + n):2] + 1j * xs[2 * o + 1:2 * (o + n):2])
                side.append((k, n, float(ol.rel_l2(ys[2 * o:2 * (o + n):2] + 1j * ys[2 * o + 1:2 * (o + n):2], want)),
                             float(ol.rel_l2(rs[2 * o:2 * (o + n):2] + 1j * rs[2 * o + 1:2 * (o + n):2], want))))
            raise AssertionError(f"replay {rep}: max diff {d}; transforms (index, size, diff): {worst[:12]}; "
                                 f"(index, size, ragged rel-L2, eager rel-L2) vs pocketfft: {side}")
        del junk
    for p in plans.values():
        p.close()
    lib.nfft_ragged_destroy(h)


def test_ragged_rejects_oversize(pkg):
    lib = pkg.core()
    h = ctypes.c_void_p()
    arr = (ctypes.c_int * 2)(16, 131072)
    assert lib.nfft_ragged_create(ctypes.byref(h), 0, 2, arr, 0) == pkg.E_UNSUPPORTED
    arr = (ctypes.c_int * 1)(0)
    assert lib.nfft_ragged_create(ctypes.byref(h), 0, 1, arr, 0) == pkg.E_SIZE
