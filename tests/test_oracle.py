"""Pins of the oracle (oracle/nfft_oracle.c, the FP64 restatement of the
reference algorithms) before anything is compared with it:
  - bit-for-bit against the reference compiled from its own header
    (oracle/_ref, when it was built) and against the committed golden vectors
    that build produced (tests/golden/);
  - against a long-double brute-force DFT (as tests/test_notoriousfft.c:609-621
    checks Bluestein), numpy.fft and scipy.fft (FFTW REDFT/RODFT definitions,
    README.md:93-98 of the reference);
  - closed forms: round trips, Parseval, impulse and single tone.
CPU only.
"""
import json
import os

import numpy as np
import pytest
import scipy.fft as sfft

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden_ref_fp64.npz")
MANIFEST = os.path.join(os.path.dirname(__file__), "golden", "golden_ref_fp64.json")

POW2 = [1, 2, 4, 8, 16, 32, 64, 128, 256, 1024, 4096]
BLUE = [3, 5, 6, 7, 10, 12, 15, 17, 20, 100, 127, 1000, 1009]


def cx(a):
    a = np.asarray(a, dtype=np.float64)
    return a[0::2] + 1j * a[1::2]


def il(z):
    z = np.asarray(z, dtype=np.complex128).reshape(-1)
    out = np.empty(2 * z.size)
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


def test_rng_matches_numpy_restatement():
    a = ol.uniform(0, 1000, start=12345)
    b = ol.uniform_np(0, 1000, start=12345)
    assert np.array_equal(a, b)
    assert a.min() >= -1 and a.max() < 1
    assert abs(a.mean()) < 0.1


def test_rel_l2_sees_both_parts():
    """A complex comparison must see the imaginary part: a conjugation, a
    sign flip of the imaginary half or a swap of the parts is a large
    error, as it is for the interleaved float view of the same data."""
    rng = np.random.default_rng(0)
    z = rng.standard_normal(500) + 1j * rng.standard_normal(500)
    assert ol.rel_l2(np.conj(z), z) > 0.5
    assert ol.rel_l2(1j * np.conj(z), z) > 0.5
    assert ol.rel_l2(z.astype(np.complex64), z) < 1e-6
    assert abs(ol.rel_l2(np.conj(z), z) - ol.rel_l2(il(np.conj(z)), il(z))) < 1e-15
    with pytest.raises(ValueError):
        ol.rel_l2(z, il(z)[:10])


def test_golden_vectors_bitwise():
    man = json.load(open(MANIFEST))
    data = np.load(GOLDEN, allow_pickle=False)
    assert len(man["cases"]) > 100
    for c in man["cases"]:
        x = data[c["key"] + "_in"]
        want = data[c["key"] + "_out"]
        got = ol.run_oracle(c["op"], c["dims"], x)
        assert np.array_equal(got, want), (c["key"], np.max(np.abs(got - want)))


@pytest.mark.skipif(not ol.Ref.available(), reason="reference not built here (oracle/_ref)")
@pytest.mark.parametrize("op", list(ol.OP))
def test_oracle_bitwise_vs_reference(op):
    ref = ol.Ref()
    sizes = [1, 2, 4, 8, 16, 32, 64, 256, 2048, 65536]
    if op in ("dft", "invdft"):
        sizes += [3, 6, 17, 100, 1009, 4097]
    for n in sizes:
        nin, _ = ol.out_shape(op, [n])
        x = ol.uniform(11, 3 * nin)
        a = ol.run_oracle(op, [n], x, batch=3)
        b = ref.run(op, [n], x, batch=3)
        assert np.array_equal(a, b), (op, n)


@pytest.mark.skipif(not ol.Ref.available(), reason="reference not built here (oracle/_ref)")
def test_oracle_bitwise_vs_reference_nd():
    ref = ol.Ref()
    for dims in ([4, 4], [8, 16], [3, 5], [4, 8, 8], [16, 1], [1, 16], [2, 3, 4]):
        for op in ("dft", "invdft"):
            x = ol.uniform(3, 2 * int(np.prod(dims)))
            assert np.array_equal(ol.run_oracle(op, dims, x), ref.run(op, dims, x)), (op, dims)
    for dims in ([4, 4], [8, 16], [4, 8, 8], [8, 1], [1, 8], [1, 1]):
        for op in ("dct2", "dst2", "dct3", "dst3", "dct4", "dst4"):
            x = ol.uniform(4, int(np.prod(dims)))
            assert np.array_equal(ol.run_oracle(op, dims, x), ref.run(op, dims, x)), (op, dims)


def test_reference_rejects_like_oracle():
    # size rules of the constructors (src/07_legacy.h:1406,1442,1529,1579)
    for kind, n, ok in [("dft", 0, False), ("dft", 7, True), ("realdft", 6, False),
                        ("realdft", 8, True), ("t2t3", 12, False), ("t4", 3, False), ("t4", 1, True)]:
        assert ol.oracle_mkaux_ok(kind, [n]) == ok, (kind, n)
        if ol.Ref.available():
            assert ol.Ref().mkaux_ok(kind, [n]) == ok


@pytest.mark.parametrize("n", POW2 + BLUE)
def test_dft_vs_bruteforce_and_numpy(n):
    x = ol.uniform(5, 2 * n)
    y = ol.run_oracle("dft", [n], x)
    bf = ol.bruteforce_dft(x)
    tol = 1e-13 if (n & (n - 1)) == 0 else 1e-11
    assert ol.rel_l2(y, bf) < tol
    assert ol.rel_l2(y, il(np.fft.fft(cx(x)))) < tol
    yi = ol.run_oracle("invdft", [n], x)
    scale = 1.0 if (n & (n - 1)) == 0 else 1.0 / n   # normalised for Bluestein n
    assert ol.rel_l2(yi, scale * ol.bruteforce_dft(x, inverse=True)) < tol


def test_dft_catches_planted_sign_error():
    # a transposed sign (inverse used for forward) must fail the same check
    n = 64
    x = ol.uniform(9, 2 * n)
    wrong = ol.run_oracle("invdft", [n], x)
    assert ol.rel_l2(wrong, ol.bruteforce_dft(x)) > 0.5


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 1024, 4096])
def test_realdft_vs_numpy(n):
    x = ol.uniform(6, n)
    z = ol.run_oracle("realdft", [n], x)
    assert ol.rel_l2(z, il(np.fft.rfft(x))) < 1e-14
    back = ol.run_oracle("invrealdft", [n], z)
    assert ol.rel_l2(back, n * x) < 1e-14


@pytest.mark.parametrize("op,typ,fn", [("dct2", 2, sfft.dct), ("dct3", 3, sfft.dct), ("dct4", 4, sfft.dct),
                                       ("dst2", 2, sfft.dst), ("dst3", 3, sfft.dst), ("dst4", 4, sfft.dst)])
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 128, 4096])
def test_trig_vs_scipy(op, typ, fn, n):
    x = ol.uniform(7, n)
    y = ol.run_oracle(op, [n], x)
    assert ol.rel_l2(y, fn(x, type=typ, norm=None)) < 1e-13


@pytest.mark.parametrize("n", [4, 64, 1024])
def test_trig_round_trips(n):
    x = ol.uniform(8, n)
    for a, b in (("dct2", "dct3"), ("dst2", "dst3"), ("dct4", "dct4"), ("dst4", "dst4")):
        y = ol.run_oracle(b, [n], ol.run_oracle(a, [n], x))
        assert ol.rel_l2(y, 2 * n * x) < 1e-13, (a, b)


@pytest.mark.parametrize("dims", [[4, 4], [8, 16], [3, 5], [4, 8, 8], [6, 10, 4]])
def test_nd_dft_vs_numpy(dims):
    x = ol.uniform(2, 2 * int(np.prod(dims)))
    y = ol.run_oracle("dft", dims, x)
    assert ol.rel_l2(y, il(np.fft.fftn(cx(x).reshape(dims)))) < 1e-12


@pytest.mark.parametrize("dims", [[4, 8], [8, 16], [3, 8], [4, 2, 8]])
def test_nd_realdft_vs_numpy(dims):
    # the reference leaves multi-D realdft unimplemented; the oracle uses the
    # FFTW layout (last axis halved), pinned here by numpy.fft.rfftn/irfftn
    x = ol.uniform(1, int(np.prod(dims)))
    z = ol.run_oracle("realdft", dims, x)
    assert ol.rel_l2(z, il(np.fft.rfftn(x.reshape(dims)))) < 1e-13
    back = ol.run_oracle("invrealdft", dims, z)
    # unnormalised, except that a non-power-of-two axis is scaled by 1/n
    # (the reference's Bluestein inverse, src/05_algorithms.h:315-333)
    scale = np.prod([n if (n & (n - 1)) == 0 else 1 for n in dims])
    assert ol.rel_l2(back, scale * x) < 1e-13


@pytest.mark.parametrize("dims", [[4, 8], [8, 8, 4]])
def test_nd_trig_separable(dims):
    x = ol.uniform(12, int(np.prod(dims))).reshape(dims)
    for op, typ, fn in (("dct2", 2, sfft.dctn), ("dst3", 3, sfft.dstn), ("dct4", 4, sfft.dctn)):
        y = ol.run_oracle(op, dims, x.reshape(-1))
        assert ol.rel_l2(y, fn(x, type=typ, norm=None).reshape(-1)) < 1e-13


def test_closed_forms():
    n = 256
    imp = np.zeros(2 * n)
    imp[0] = 1.0
    assert np.allclose(ol.run_oracle("dft", [n], imp)[0::2], 1.0)
    k0 = 5
    tone = il(np.exp(2j * np.pi * k0 * np.arange(n) / n))
    y = cx(ol.run_oracle("dft", [n], tone))
    assert abs(y[k0] - n) < 1e-10 and np.max(np.abs(np.delete(y, k0))) < 1e-10
    x = ol.uniform(13, 2 * n)
    y = ol.run_oracle("dft", [n], x)
    assert abs(np.sum(y * y) - n * np.sum(x * x)) < 1e-9 * n * np.sum(x * x)
