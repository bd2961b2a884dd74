"""Helpers for the GPU parity tests: run a plan of the extended C ABI on
torch-allocated device buffers (torch is only the allocator here) and
compare with the oracle."""
import numpy as np

import oracle_lib as ol


def tol_for(prec, n):
    lg = max(1.0, float(np.log2(max(2, n))))
    return (1e-6 if prec == "f32" else 1e-14) * lg


def np_dtype(prec):
    return np.float32 if prec == "f32" else np.float64


def to_dev(x, prec, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x).astype(np_dtype(prec)))).to(dev)


def run_plan(pkg, dev, op, dims, x, batch=1, prec="f32", norm="reference", vjp=False, inplace=False):
    import torch
    nin, nout = ol.out_shape(op, dims)
    if vjp:
        nin, nout = nout, nin
    xin = to_dev(x, prec, dev)
    assert xin.numel() == nin * batch, (xin.numel(), nin, batch)
    out = xin if inplace else torch.empty(nout * batch, dtype=xin.dtype, device=dev)
    plan = pkg.Plan(ol.OP_KIND[op], prec, dims, batch, norm)
    torch.cuda.synchronize()
    if vjp:
        plan.vjp(op, xin.data_ptr(), out.data_ptr())
    else:
        plan.execute(op, xin.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    plan.check()  # no device-side dependency wait timed out
    plan.close()
    return out.cpu().numpy().astype(np.float64)


def rounded(x, prec):
    """The FP64 view of the values the device actually received."""
    return np.asarray(x).astype(np_dtype(prec)).astype(np.float64)


def pocket(op, dims, x, batch=1):
    """numpy.fft (pocketfft, FP64) reference for dft / invdft / realdft /
    invrealdft over dims with the reference's normalisation: invdft is
    unnormalised along power-of-two axes and scaled by 1/N along the others
    (src/05_algorithms.h:315-333, also for the leading axes of invrealdft);
    realdft / invrealdft are unnormalised (src/07_legacy.h:162-293).
    Returns interleaved reals like the device output."""
    dims = list(dims)
    x = np.asarray(x, dtype=np.float64)
    axes = tuple(range(1, len(dims) + 1))
    pow2 = float(np.prod([n for n in dims if n & (n - 1) == 0]))
    if op in ("dft", "invdft", "invrealdft"):
        if op == "invrealdft":
            h = dims[:-1] + [dims[-1] // 2 + 1]
            z = (x[0::2] + 1j * x[1::2]).reshape([batch] + h)
            return (np.fft.irfftn(z, s=dims, axes=axes) * pow2).reshape(-1)
        z = (x[0::2] + 1j * x[1::2]).reshape([batch] + dims)
        out = np.fft.fftn(z, axes=axes) if op == "dft" else np.fft.ifftn(z, axes=axes) * pow2
    elif op == "realdft":
        out = np.fft.rfftn(x.reshape([batch] + dims), axes=axes)
    else:
        raise ValueError(op)
    out = out.reshape(-1)
    r = np.empty(2 * out.size)
    r[0::2] = out.real
    r[1::2] = out.imag
    return r
