"""notorious-fft_amd: Python bindings (ctypes) of the MI355X FFT engine.

The product is the C ABI built from csrc/ into this directory:
  libnotorious_fft_hip.so      device-resident plans, batched / N-D / adjoint
                               transforms (include/notorious_fft_hip.h)
  libnotorious_fft_hip_f64.so  the reference's 27-function C API, double
  libnotorious_fft_hip_f32.so  the same API, float (include/notorious_fft.h)

This module only loads those libraries and declares their signatures; it has
no fallback: every entry point raises if the native library is missing.
The directory name is not a Python identifier, so load it with
importlib (see __graft_entry__.py): it registers as ``notorious_fft_amd``.
"""
import ctypes
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
CORE_SO = os.path.join(PKG_DIR, "libnotorious_fft_hip.so")
DROPIN_SO = {"f64": os.path.join(PKG_DIR, "libnotorious_fft_hip_f64.so"),
             "f32": os.path.join(PKG_DIR, "libnotorious_fft_hip_f32.so")}

# constants of include/notorious_fft_hip.h
OK, E_INVALID, E_SIZE, E_ALLOC, E_HIP, E_UNSUPPORTED = range(6)
F32, F64 = 0, 1
KIND = {"dft": 0, "realdft": 1, "t2t3": 2, "t4": 3}
OP = {"dft": 0, "invdft": 1, "realdft": 2, "invrealdft": 3, "dct2": 4, "dst2": 5,
      "dct3": 6, "dst3": 7, "dct4": 8, "dst4": 9}
OP_KIND = {"dft": "dft", "invdft": "dft", "realdft": "realdft", "invrealdft": "realdft",
           "dct2": "t2t3", "dst2": "t2t3", "dct3": "t2t3", "dst3": "t2t3",
           "dct4": "t4", "dst4": "t4"}
NORM = {"reference": 0, "none": 1, "inverse": 2}

DROPIN_FUNCS = [
    "notorious_fft_dft", "notorious_fft_invdft", "notorious_fft_realdft",
    "notorious_fft_invrealdft", "notorious_fft_dct2", "notorious_fft_dst2",
    "notorious_fft_dct3", "notorious_fft_dst3", "notorious_fft_dct4", "notorious_fft_dst4",
    "notorious_fft_mkaux_dft_1d", "notorious_fft_mkaux_dft_2d", "notorious_fft_mkaux_dft_3d",
    "notorious_fft_mkaux_dft", "notorious_fft_mkaux_realdft_1d", "notorious_fft_mkaux_realdft_2d",
    "notorious_fft_mkaux_realdft_3d", "notorious_fft_mkaux_realdft", "notorious_fft_mkaux_t2t3_1d",
    "notorious_fft_mkaux_t2t3_2d", "notorious_fft_mkaux_t2t3_3d", "notorious_fft_mkaux_t2t3",
    "notorious_fft_mkaux_t4_1d", "notorious_fft_mkaux_t4_2d", "notorious_fft_mkaux_t4_3d",
    "notorious_fft_mkaux_t4", "notorious_fft_free_aux"]


class NativeError(RuntimeError):
    pass


def build(jobs=8, verbose=False, force=True):
    """Compile every HIP source for gfx950 into this directory (make; with
    force, every target is rebuilt even where it is up to date)."""
    cmd = ["make", "-C", CSRC, f"-j{jobs}"] + (["-B"] if force else [])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if verbose or r.returncode:
        print(r.stdout[-4000:])
    if r.returncode:
        raise NativeError(f"native build failed ({r.returncode})")


_core = None
_dropins = {}

_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_i = ctypes.c_int


def core():
    """The extended C ABI library (raises if it was not built)."""
    global _core
    if _core is None:
        if not os.path.exists(CORE_SO):
            raise NativeError(f"{CORE_SO} is missing: build the package first")
        lib = ctypes.CDLL(CORE_SO)
        lib.nfft_plan_create.restype = _i
        lib.nfft_plan_create.argtypes = [ctypes.POINTER(_vp), _i, _i, _i,
                                         ctypes.POINTER(_i), _ll, _i]
        lib.nfft_plan_destroy.argtypes = [_vp]
        lib.nfft_execute.restype = _i
        lib.nfft_execute.argtypes = [_vp, _i, _vp, _vp, _vp]
        lib.nfft_vjp.restype = _i
        lib.nfft_vjp.argtypes = [_vp, _i, _vp, _vp, _vp]
        lib.nfft_dft_weighted_power.restype = _i
        lib.nfft_dft_weighted_power.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        lib.nfft_execute_noalloc.restype = _i
        lib.nfft_execute_noalloc.argtypes = [_vp, _i, _vp, _vp, _vp]
        lib.nfft_plan_workspace_bytes.restype = _ll
        lib.nfft_plan_workspace_bytes.argtypes = [_vp]
        lib.nfft_plan_describe.restype = ctypes.c_char_p
        lib.nfft_plan_describe.argtypes = [_vp]
        lib.nfft_plan_check.restype = ctypes.c_int
        lib.nfft_plan_check.argtypes = [_vp, _vp]
        lib.nfft_last_error.restype = ctypes.c_char_p
        lib.nfft_last_error.argtypes = []
        lib.nfft_fill_uniform.restype = _i
        lib.nfft_fill_uniform.argtypes = [_vp, _ll, ctypes.c_ulonglong, ctypes.c_ulonglong, _i, _vp]
        lib.nfft_fill_gaussian.restype = _i
        lib.nfft_fill_gaussian.argtypes = [_vp, _ll, ctypes.c_ulonglong, ctypes.c_ulonglong, _i, _vp]
        lib.nfft_prof_enable.restype = _i
        lib.nfft_prof_enable.argtypes = [_i, _i]
        lib.nfft_prof_count.restype = _i
        lib.nfft_prof_count.argtypes = []
        lib.nfft_prof_get.restype = _i
        lib.nfft_prof_get.argtypes = [_i, ctypes.c_char_p, _i, ctypes.POINTER(_ll),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        if hasattr(lib, "nfft_ragged_create"):
            lib.nfft_ragged_create.restype = _i
            lib.nfft_ragged_create.argtypes = [ctypes.POINTER(_vp), _i, _i, ctypes.POINTER(_i), _i]
            lib.nfft_ragged_execute.restype = _i
            lib.nfft_ragged_execute.argtypes = [_vp, _i, _vp, _vp, _vp]
            lib.nfft_ragged_destroy.restype = _i
            lib.nfft_ragged_destroy.argtypes = [_vp]
            lib.nfft_ragged_total.restype = _ll
            lib.nfft_ragged_total.argtypes = [_vp]
            lib.nfft_ragged_check.restype = ctypes.c_int
            lib.nfft_ragged_check.argtypes = [_vp, _vp]
        _core = lib
    return _core


def dropin(prec="f64"):
    """The drop-in library of the reference API for 'f64' or 'f32'."""
    if prec not in _dropins:
        path = DROPIN_SO[prec]
        if not os.path.exists(path):
            raise NativeError(f"{path} is missing: build the package first")
        lib = ctypes.CDLL(path)
        for name in DROPIN_FUNCS:
            fn = getattr(lib, name)
            if name.startswith("notorious_fft_mkaux"):
                fn.restype = _vp
                if name.endswith("_1d"):
                    fn.argtypes = [_i]
                elif name.endswith("_2d"):
                    fn.argtypes = [_i, _i]
                elif name.endswith("_3d"):
                    fn.argtypes = [_i, _i, _i]
                else:
                    fn.argtypes = [_i, ctypes.POINTER(_i)]
            elif name == "notorious_fft_free_aux":
                fn.restype = None
                fn.argtypes = [_vp]
            else:
                fn.restype = None
                fn.argtypes = [_vp, _vp, _vp]
        _dropins[prec] = lib
    return _dropins[prec]


def last_error():
    return core().nfft_last_error().decode()


def check(status, what="call"):
    if status != OK:
        raise NativeError(f"{what} failed with status {status}: {last_error()}")


class Plan:
    """Device-resident plan of the extended API (one per shape/precision)."""

    def __init__(self, kind, prec, dims, batch=1, norm="reference"):
        lib = core()
        self.kind, self.prec, self.dims, self.batch = kind, prec, list(dims), batch
        arr = (_i * len(self.dims))(*self.dims)
        h = _vp()
        st = lib.nfft_plan_create(ctypes.byref(h), KIND[kind], F32 if prec == "f32" else F64,
                                  len(self.dims), arr, batch, NORM[norm])
        self.status = st
        if st != OK:
            raise NativeError(f"plan_create({kind},{prec},{dims},{batch}) -> {st}: {last_error()}")
        self.handle = h

    def execute(self, op, src_ptr, dst_ptr, stream=0):
        check(core().nfft_execute(self.handle, OP[op], src_ptr, dst_ptr, stream), f"execute {op}")

    def execute_noalloc(self, op, src_ptr, dst_ptr, stream=0):
        """nfft_execute_noalloc: the library's static scratch instead of the plan's."""
        check(core().nfft_execute_noalloc(self.handle, OP[op], src_ptr, dst_ptr, stream), f"execute_noalloc {op}")

    def weighted_power(self, x_ptr, w_ptr, grad_ptr, loss_ptr, stream=0):
        """nfft_dft_weighted_power: loss[t] = sum |W DFT(x_t)|^2 (FP64 device array) and
        grad = 2 DFT^H (W^2 DFT(x))."""
        check(core().nfft_dft_weighted_power(self.handle, x_ptr, w_ptr, grad_ptr, loss_ptr, stream),
              "dft_weighted_power")

    def vjp(self, op, g_ptr, out_ptr, stream=0):
        check(core().nfft_vjp(self.handle, OP[op], g_ptr, out_ptr, stream), f"vjp {op}")

    def describe(self):
        return core().nfft_plan_describe(self.handle).decode()

    def check(self, stream=0):
        """Synchronise `stream`; raise if a device-side wait of this plan timed out."""
        check(core().nfft_plan_check(self.handle, stream), "plan_check")

    @property
    def workspace_bytes(self):
        return core().nfft_plan_workspace_bytes(self.handle)

    def close(self):
        if getattr(self, "handle", None):
            core().nfft_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fill_uniform(ptr, count, seed, start=0, prec="f32", stream=0):
    check(core().nfft_fill_uniform(ptr, count, seed, start, F32 if prec == "f32" else F64, stream),
          "fill_uniform")


def fill_gaussian(ptr, count, seed, start=0, prec="f32", stream=0):
    check(core().nfft_fill_gaussian(ptr, count, seed, start, F32 if prec == "f32" else F64, stream),
          "fill_gaussian")


def prof_enable(on=True, capacity=8192):
    """Start (clear) or stop the library's per-launch kernel timer."""
    check(core().nfft_prof_enable(1 if on else 0, capacity), "prof_enable")


def prof_read():
    """{family: (launches, total_ms, algorithmic_bytes)} of the recorded launches."""
    lib = core()
    n = lib.nfft_prof_count()
    if n >= 1000000:
        raise NativeError("kernel timer capacity exceeded")
    out = {}
    buf = ctypes.create_string_buffer(256)
    for i in range(n):
        cnt, ms, by = _ll(), ctypes.c_double(), ctypes.c_double()
        check(lib.nfft_prof_get(i, buf, 256, ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(by)), "prof_get")
        out[buf.value.decode()] = (cnt.value, ms.value, by.value)
    return out
