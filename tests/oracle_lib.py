"""ctypes bindings for the oracle (oracle/liboracle.so) and, where it was
built, the reference compiled from its own header (oracle/_ref/).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline import this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

KIND = {"dft": 0, "realdft": 1, "t2t3": 2, "t4": 3}
OP = {"dft": 0, "invdft": 1, "realdft": 2, "invrealdft": 3, "dct2": 4, "dst2": 5,
      "dct3": 6, "dst3": 7, "dct4": 8, "dst4": 9}
OP_KIND = {"dft": "dft", "invdft": "dft", "realdft": "realdft", "invrealdft": "realdft",
           "dct2": "t2t3", "dst2": "t2t3", "dct3": "t2t3", "dst3": "t2t3",
           "dct4": "t4", "dst4": "t4"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        lib = ctypes.CDLL(ORACLE_SO)
        lib.orc_mkaux.restype = ctypes.c_void_p
        lib.orc_mkaux.argtypes = [ctypes.c_int, ctypes.c_int, _ip]
        lib.orc_free_aux.argtypes = [ctypes.c_void_p]
        lib.orc_exec.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        lib.orc_exec_batch.restype = ctypes.c_int
        lib.orc_exec_batch.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int,
                                       ctypes.c_longlong, _dp, _dp, ctypes.c_int]
        lib.orc_bruteforce_dft.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        lib.orc_fill_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_longlong, _dp]
        _oracle = lib
    return _oracle


def _ints(Ns):
    arr = (ctypes.c_int * len(Ns))(*Ns)
    return arr


def out_shape(op, Ns):
    """Per-transform (in_reals, out_reals) counts for op over dims Ns."""
    p = int(np.prod(Ns))
    half = p // Ns[-1] * (Ns[-1] // 2 + 1)
    if op in ("dft", "invdft"):
        return 2 * p, 2 * p
    if op == "realdft":
        return p, 2 * half
    if op == "invrealdft":
        return 2 * half, p
    return p, p


def run_oracle(op, Ns, x, batch=1, nthreads=0):
    """Run the oracle on `batch` contiguous transforms of dims Ns.  x holds
    interleaved complex (as float64 pairs) or real data."""
    Ns = list(Ns)
    lib = oracle()
    nin, nout = out_shape(op, Ns)
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    assert x.size == nin * batch, (x.size, nin, batch)
    y = np.zeros(nout * batch, dtype=np.float64)
    err = lib.orc_exec_batch(KIND[OP_KIND[op]], len(Ns), _ints(Ns), OP[op], batch,
                             x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), nthreads)
    if err:
        raise ValueError(f"oracle rejected {op} {Ns}")
    return y


def oracle_mkaux_ok(kind, Ns):
    lib = oracle()
    a = lib.orc_mkaux(KIND[kind], len(Ns), _ints(list(Ns)))
    if a:
        lib.orc_free_aux(a)
        return True
    return False


def bruteforce_dft(x, inverse=False):
    lib = oracle()
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    n = x.size // 2
    y = np.zeros_like(x)
    lib.orc_bruteforce_dft(n, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), 1 if inverse else 0)
    return y


def uniform(seed, count, start=0):
    """Counter-based SplitMix64 uniform [-1, 1) draws (same bits as the C and
    HIP generators)."""
    lib = oracle()
    out = np.empty(count, dtype=np.float64)
    lib.orc_fill_uniform(seed, start, count, out.ctypes.data_as(_dp))
    return out


def uniform_np(seed, count, start=0):
    """Pure numpy restatement of the same generator (for cross-checking)."""
    with np.errstate(over="ignore"):
        idx = np.arange(start, start + count, dtype=np.uint64)
        z = ((np.uint64(seed) << np.uint64(40)) | idx) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def gaussian_np(seed, count, start=0):
    """Box-Muller N(0, 1) over the same stream (the device generator's
    definition): value j = start + i uses the pair p = j // 2 of uniforms
    u1 = V(2p), u2 = V(2p + 1) in (0, 1], V(k) = ((z_k >> 11) + 1) 2^-53;
    sqrt(-2 ln u1) cos(2 pi u2) for even j, sin for odd j."""
    j = np.arange(start, start + count, dtype=np.uint64)
    p = j >> np.uint64(1)

    def v(k):
        with np.errstate(over="ignore"):
            z = ((np.uint64(seed) << np.uint64(40)) | k) + np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return ((z >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * (1.0 / 9007199254740992.0)

    u1 = v(np.uint64(2) * p)
    u2 = v(np.uint64(2) * p + np.uint64(1))
    r = np.sqrt(-2.0 * np.log(u1))
    return np.where((j & np.uint64(1)) == 0, r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2))


class Ref:
    """The reference library compiled from its own header (oracle/_ref)."""

    def __init__(self, name="libnfft_ref_exact.so"):
        path = os.path.join(REF_DIR, name)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = ctypes.CDLL(path)
        self.real = ctypes.c_double if lib.ref_real_size() == 8 else ctypes.c_float
        self.np_real = np.float64 if self.real is ctypes.c_double else np.float32
        rp = ctypes.POINTER(self.real)
        lib.ref_mkaux.restype = ctypes.c_void_p
        lib.ref_mkaux.argtypes = [ctypes.c_int, ctypes.c_int, _ip]
        lib.ref_free.argtypes = [ctypes.c_void_p]
        lib.ref_exec.argtypes = [ctypes.c_void_p, ctypes.c_int, rp, rp]
        lib.ref_exec_batch.restype = ctypes.c_int
        lib.ref_exec_batch.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int,
                                       ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                       rp, rp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        lib.ref_exec_list.restype = ctypes.c_int
        lib.ref_exec_list.argtypes = [ctypes.c_int, _ip, ctypes.c_int, rp, rp, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
        self.lib = lib
        self._rp = rp

    @staticmethod
    def available(name="libnfft_ref_exact.so"):
        return os.path.exists(os.path.join(REF_DIR, name))

    def mkaux_ok(self, kind, Ns):
        a = self.lib.ref_mkaux(KIND[kind], len(Ns), _ints(list(Ns)))
        if a:
            self.lib.ref_free(a)
            return True
        return False

    def run(self, op, Ns, x, batch=1, nthreads=0):
        return self.run_timed(op, Ns, x, batch, nthreads)[0]

    def run_list(self, sizes, x, op="dft", nthreads=0):
        """Mixed-size 1-D complex transforms packed back to back: (output,
        seconds in the transform loop, auxes made beforehand)."""
        sizes = list(sizes)
        x = np.ascontiguousarray(np.asarray(x, dtype=self.np_real).reshape(-1)).copy()
        assert x.size == 2 * sum(sizes)
        y = np.zeros_like(x)
        t = ctypes.c_double(0.0)
        err = self.lib.ref_exec_list(len(sizes), _ints(sizes), OP[op], x.ctypes.data_as(self._rp),
                                     y.ctypes.data_as(self._rp), nthreads, ctypes.byref(t))
        if err:
            raise ValueError("reference rejected a size of the list")
        return y, t.value

    def run_timed(self, op, Ns, x, batch=1, nthreads=0):
        """(output, seconds spent in the transform loop alone)."""
        Ns = list(Ns)
        nin, nout = out_shape(op, Ns)
        x = np.ascontiguousarray(np.asarray(x, dtype=self.np_real).reshape(-1)).copy()
        assert x.size == nin * batch
        y = np.zeros(nout * batch, dtype=self.np_real)
        t = ctypes.c_double(0.0)
        err = self.lib.ref_exec_batch(KIND[OP_KIND[op]], len(Ns), _ints(Ns), OP[op], batch,
                                      nin, nout, x.ctypes.data_as(self._rp),
                                      y.ctypes.data_as(self._rp), nthreads, ctypes.byref(t))
        if err:
            raise ValueError(f"reference rejected {op} {Ns}")
        return y, t.value


def as_reals(x):
    """Flatten to float64 reals; complex input becomes interleaved (re, im)
    pairs, so that both parts take part in a comparison."""
    x = np.asarray(x)
    if np.iscomplexobj(x):
        x = np.stack([x.real, x.imag], axis=-1)
    return np.asarray(x, dtype=np.float64).reshape(-1)


def rel_l2(got, ref):
    got = as_reals(got)
    ref = as_reals(ref)
    if got.size != ref.size:
        raise ValueError(f"rel_l2: sizes differ ({got.size} vs {ref.size})")
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(got - ref) / (den if den > 0 else 1.0))
