"""Benchmark of the batched 1D C2C hot path on one MI355X.

Workload (the headline configuration): batched 1D complex FFT, FP32,
N = 2^20, batch = 256 (2^28 complex elements, 2 GiB in + 2 GiB out),
inputs i.i.d. uniform [-1, 1) from the seed-0 SplitMix64 stream, resident
in HBM before the timed region.  A step is one forward transform of the
whole batch.  value = 5 N log2 N * batch / t (GFLOP/s).

Also printed in the same JSON line:
  roofline      the algorithmic bytes of one step (one read and one write of
                the batch) over the summed average durations of the kernels of
                a step, timed live with HIP events by the library's kernel
                timer on the launch stream (second timed pass); the headline
                step is one launch of the XCD-affine four-step k_four_xcd
  cpu_baseline  the reference library compiled from its own header
                (oracle/_ref, FP64, OpenMP over the batch, one aux per thread)
                on a bounded sample of the same workload on this host
  sweep         the rest of the batched C2C configuration (N = 2^10..2^20,
                FP32 and FP64, 2^28 elements each) with rel-L2 vs the oracle
  sweep_r2c     batched realdft N = 2^10..2^20, FP32 and FP64, 2^28 reals each
  accuracy      rel-L2 of sampled output transforms vs the FP64 oracle

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-sweep] [--no-cpu] [--calib]
                       [--only NAME[,NAME...]]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "batched 1D C2C GFLOP/s (5NlogN) + %HBM BW, FP32 vs FP64 at stated rel-L2 err"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_config(pkg, torch, n, prec, total_elems, steps, warmup, profile=False, check=0):
    """Time `steps` batched forward transforms of length n; returns a dict."""
    dt = torch.float32 if prec == "f32" else torch.float64
    batch = total_elems // n
    dev = torch.device("cuda:0")
    x = torch.empty(2 * n * batch, dtype=dt, device=dev)
    y = torch.empty_like(x)
    pkg.fill_uniform(x.data_ptr(), x.numel(), 0, 0, prec)
    plan = pkg.Plan("dft", prec, [n], batch)
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(warmup):
        plan.execute("dft", x.data_ptr(), y.data_ptr(), stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        plan.execute("dft", x.data_ptr(), y.data_ptr(), stream)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1)
    res = {"n": n, "prec": prec, "batch": batch, "steps": steps, "wall_s": wall,
           "ms_per_step": 1e3 * wall / steps, "event_ms_per_step": ev_ms / steps,
           "plan": plan.describe()}
    flops = 5.0 * n * math.log2(n) * batch
    alg_bytes = 2.0 * n * batch * (8 if prec == "f32" else 16)
    res["gflops"] = flops / (res["ms_per_step"] * 1e-3) / 1e9
    res["alg_gbs"] = alg_bytes / (res["ms_per_step"] * 1e-3) / 1e9
    res["hbm_frac"] = res["alg_gbs"] / HBM_PEAK_GBS
    if profile:
        pkg.prof_enable(True, capacity=max(4096, 64 * steps))
        torch.cuda.synchronize()
        for _ in range(steps):
            plan.execute("dft", x.data_ptr(), y.data_ptr(), stream)
        torch.cuda.synchronize()
        rec = pkg.prof_read()
        pkg.prof_enable(False)
        res["kernels"] = {k: {"launches": v[0], "avg_us": 1e3 * v[1] / v[0], "total_ms": v[1],
                              "bytes_per_launch": v[2] / v[0], "gbs": v[2] / (v[1] * 1e-3) / 1e9}
                          for k, v in rec.items()}
    if check:
        import numpy as np
        import oracle_lib as ol
        idx = [0, batch - 1][:check]
        errs = []
        for b in idx:
            xs = x[2 * n * b:2 * n * (b + 1)].cpu().numpy().astype(np.float64)
            ys = y[2 * n * b:2 * n * (b + 1)].cpu().numpy().astype(np.float64)
            errs.append(ol.rel_l2(ys, ol.run_oracle("dft", [n], xs)))
        res["rel_l2"] = max(errs)
        res["rel_l2_tol"] = (1e-6 if prec == "f32" else 1e-14) * math.log2(n)
    plan.close()
    del x, y
    torch.cuda.empty_cache()
    return res


def run_r2c(pkg, torch, n, prec, total_reals, steps=5, warmup=2):
    """Batched realdft of length n over total_reals reals (FFTW layout out:
    n/2+1 complex per transform); rel-L2 of one transform vs the oracle."""
    import numpy as np
    import oracle_lib as ol
    dt = torch.float32 if prec == "f32" else torch.float64
    eb = 4 if prec == "f32" else 8
    batch = total_reals // n
    dev = torch.device("cuda:0")
    x = torch.empty(n * batch, dtype=dt, device=dev)
    y = torch.empty((n + 2) * batch, dtype=dt, device=dev)
    pkg.fill_uniform(x.data_ptr(), x.numel(), 0, 0, prec)
    plan = pkg.Plan("realdft", prec, [n], batch)
    t = timed(torch, lambda: plan.execute("realdft", x.data_ptr(), y.data_ptr()), steps, warmup)
    xs = x[:n].cpu().numpy().astype(np.float64)
    ys = y[:n + 2].cpu().numpy().astype(np.float64)
    err = ol.rel_l2(ys, ol.run_oracle("realdft", [n], xs))
    out = {"n": n, "prec": prec, "batch": batch, "ms": round(1e3 * t, 4),
           "gflops": round(2.5 * n * math.log2(n) * batch / t / 1e9, 1),
           "alg_gbs": round((2 * n + 2) * batch * eb / t / 1e9, 1),
           "hbm_frac": round((2 * n + 2) * batch * eb / t / 1e9 / HBM_PEAK_GBS, 4),
           "rel_l2": float(f"{err:.3e}"), "plan": plan.describe()}
    plan.close()
    del x, y
    torch.cuda.empty_cache()
    return out


def timed(torch, fn, steps=5, warmup=2, warm_s=0.05):
    """Seconds per call of fn over `steps` calls, after at least `warmup`
    calls and at least warm_s seconds of them: after host-side work (the CPU
    baselines and oracle checks run between configurations) the GPU needs
    tens of milliseconds of load to return to full clocks, and a fixed count
    of sub-millisecond calls measured up to 20 % slow."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t_end = time.perf_counter() + warm_s
    while time.perf_counter() < t_end:
        for _ in range(4):
            fn()
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def ref_lib():
    """(ctypes wrapper of the reference built with its Release flags, name), or
    (None, reason) where no build for this host exists."""
    import oracle_lib as ol
    flags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    name = "libnfft_ref_fast64.so" if ("avx512f" in flags and "avx512dq" in flags) else "libnfft_ref_fast64_avx2.so"
    try:
        return ol.Ref(name), name
    except (FileNotFoundError, OSError):
        return None, "no reference build"


def cpu_config(op, dims, batch_full, sample, flops_per, post=None):
    """The reference (FP64, Release flags, OpenMP over the sampled batch with
    one aux per thread; a single N-D transform runs on one thread, as the
    reference does) on `sample` transforms of a configuration: GFLOP/s at the
    configuration's nominal flops and the estimated time of the full batch.
    post(ref) adds work timed with it (C5: mask + inverse)."""
    import oracle_lib as ol
    ref, name = ref_lib()
    if ref is None:
        return None
    cores = min(16, os.cpu_count() or 1)
    nin, _ = ol.out_shape(op, dims)
    x = ol.uniform(0, nin * sample)
    y, t = ref.run_timed(op, dims, x, sample, cores)
    if post is not None:
        t += post(ref, y, cores)
    return {"value": round(flops_per * sample / t / 1e9, 3), "unit": "GFLOP/s",
            "cores": cores if sample > 1 else 1, "kind": "reference",
            "sample": f"{sample} of {batch_full} transforms, {name}, {t:.2f} s in the transform loop",
            "est_full_ms": round(1e3 * t * batch_full / sample, 2)}


def gauss_mask(dims):
    """W[k] = exp(-sum_a d(k_a)^2 / (2 sigma_a^2)), d(k) = min(k, N - k), sigma_a = N_a / 16."""
    import numpy as np
    w = np.ones(dims)
    for a, n in enumerate(dims):
        k = np.arange(n)
        d = np.minimum(k, n - k).astype(np.float64)
        sh = [1] * len(dims)
        sh[a] = n
        w = w * np.exp(-(d ** 2) / (2.0 * (n / 16.0) ** 2)).reshape(sh)
    return w


def masked_power(pkg, torch, out, wanted, with_cpu):
    """Loss sum |W DFT2(x)|^2 and its gradient 2 DFT2^H(W^2 DFT2(x)), 2-D C2C
    2048^2 x 16 with the Gaussian-bump mask: FP32 fused (axis-0 forward,
    weighting and adjoint in one pass), FP32 unfused, FP64 fused and unfused,
    and the reference on the CPU (dft, weighting, invdft)."""
    import numpy as np
    dims, batch = [2048, 2048], 16
    p = 2048 * 2048
    flops = 2 * 5 * p * 22  # forward + adjoint
    dev = torch.device("cuda:0")
    for name, prec, fuse in (("c5_loss_grad_f32_2048^2_x16_fused", "f32", "1"),
                             ("c5_loss_grad_f32_2048^2_x16_unfused", "f32", "0"),
                             ("c5_loss_grad_f64_2048^2_x16_fused", "f64", "1"),
                             ("c5_loss_grad_f64_2048^2_x16_unfused", "f64", "0")):
        if not wanted(name):
            continue
        os.environ["NFFT_FUSE"] = fuse
        dt = torch.float32 if prec == "f32" else torch.float64
        x = torch.empty(2 * p * batch, dtype=dt, device=dev)
        g = torch.empty_like(x)
        loss = torch.empty(batch, dtype=torch.float64, device=dev)
        pkg.fill_uniform(x.data_ptr(), x.numel(), 0, 0, prec)
        w = torch.from_numpy(gauss_mask(dims).reshape(-1).astype(np.float32 if prec == "f32" else np.float64)).to(dev)
        pl = pkg.Plan("dft", prec, dims, batch)
        t = timed(torch, lambda: pl.weighted_power(x.data_ptr(), w.data_ptr(), g.data_ptr(), loss.data_ptr()),
                  steps=20, warmup=10)
        eb = 8 if prec == "f32" else 16
        alg = 2 * p * batch * eb + p * eb // 2  # x read, grad written, W read once
        out[name] = {"ms": round(1e3 * t, 4), "gflops": round(flops * batch / t / 1e9, 1),
                     "alg_gbs": round(alg / t / 1e9, 1), "hbm_frac": round(alg / t / 1e9 / HBM_PEAK_GBS, 4),
                     "loss_0": float(loss[0].item()), "plan": pl.describe()}
        log(f"  {name}: {out[name]}")
        pl.close()
        del x, g, w
        torch.cuda.empty_cache()
    os.environ.pop("NFFT_FUSE", None)
    name = "c5_loss_grad_f32_2048^2_x16_fused"
    if with_cpu and name in out:
        w2 = np.repeat(gauss_mask(dims).reshape(-1) ** 2, 2)

        def post(ref, y, cores):
            # weighting (numpy, timed) and the reference's invdft of the weighted spectra
            t0 = time.perf_counter()
            v = 2 * w2 * y.reshape(16, -1)
            t1 = time.perf_counter()
            _, ti = ref.run_timed("invdft", dims, v.reshape(-1), 16, cores)
            return (t1 - t0) + ti
        try:
            c = cpu_config("dft", dims, batch, batch, flops, post=post)
        except Exception as e:
            c = {"error": str(e)}
        if c is not None:
            out[name]["cpu_baseline"] = c
            if "value" in c:
                out[name]["gpu_over_cpu"] = round(out[name]["gflops"] / c["value"], 1)
            log(f"  {name} cpu: {c}")


def config_extras(pkg, torch, only=None, with_cpu=True):
    """The other workloads of the configuration list, one timing each (FP32
    unless stated): R2C batches, 3D C2C 512^3, 2D R2C 4096^2, DCT-2/DST-2/DCT-4
    rows, the ragged batch of 10,000 mixed sizes (FP32 and FP64), the adjoint
    pair (forward + vjp) of the 2D C2C and the DCT-2, and the masked-power
    loss + gradient of the 2D C2C (fused FP32, unfused FP32 and the plain FP64
    path).  With with_cpu, each carries the reference's CPU time on a bounded
    sample of the same workload ("cpu_baseline")."""
    import ctypes
    from ragged_sizes import ragged_sizes

    dev = torch.device("cuda:0")
    out = {}

    def wanted(name):
        return not only or any(name.startswith(o) for o in only)

    def with_cpu_line(name, *args, **kw):
        if with_cpu and name in out:
            try:
                c = cpu_config(*args, **kw)
            except Exception as e:  # the baseline is a report, never a reason to fail the bench
                c = {"error": str(e)}
            if c is not None:
                out[name]["cpu_baseline"] = c
                if "value" in c and "gflops" in out[name]:
                    out[name]["gpu_over_cpu"] = round(out[name]["gflops"] / c["value"], 1)
                log(f"  {name} cpu: {c}")

    def plan_run(name, kind, prec, dims, batch, op, nin, nout, flops, vjp=False):
        if not wanted(name):
            return
        dt = torch.float32 if prec == "f32" else torch.float64
        x = torch.empty(nin * batch, dtype=dt, device=dev)
        y = torch.empty(nout * batch, dtype=dt, device=dev)
        pkg.fill_uniform(x.data_ptr(), x.numel(), 0, 0, prec)
        p = pkg.Plan(kind, prec, dims, batch)
        fn = (lambda: p.vjp(op, x.data_ptr(), y.data_ptr())) if vjp else \
            (lambda: p.execute(op, x.data_ptr(), y.data_ptr()))
        # 10 warmup calls and at least 50 ms of them (timed): after host-side
        # work the GPU needs load to return to full clocks
        t = timed(torch, fn, steps=20, warmup=10)
        eb = 4 if prec == "f32" else 8
        out[name] = {"ms": round(1e3 * t, 4), "gflops": round(flops * batch / t / 1e9, 1),
                     "alg_gbs": round((nin + nout) * batch * eb / t / 1e9, 1),
                     "hbm_frac": round((nin + nout) * batch * eb / t / 1e9 / HBM_PEAK_GBS, 4),
                     "plan": p.describe()}
        log(f"  {name}: {out[name]}")
        p.close()
        del x, y
        torch.cuda.empty_cache()

    n = 1 << 16
    fl = 2.5 * n * math.log2(n)
    plan_run("r2c_f32_N2^16_x4096", "realdft", "f32", [n], 4096, "realdft", n, n + 2, fl)
    with_cpu_line("r2c_f32_N2^16_x4096", "realdft", [n], 4096, 4096, fl)
    n = 1 << 20
    fl = 2.5 * n * math.log2(n)
    plan_run("r2c_f32_N2^20_x256", "realdft", "f32", [n], 256, "realdft", n, n + 2, fl)
    with_cpu_line("r2c_f32_N2^20_x256", "realdft", [n], 256, 256, fl)
    p3 = 512 ** 3
    plan_run("c2c3d_f32_512^3", "dft", "f32", [512, 512, 512], 1, "dft", 2 * p3, 2 * p3, 5 * p3 * 27)
    with_cpu_line("c2c3d_f32_512^3", "dft", [512, 512, 512], 1, 1, 5 * p3 * 27)
    p2 = 4096 * 4096
    plan_run("r2c2d_f32_4096^2", "realdft", "f32", [4096, 4096], 1, "realdft", p2, 4096 * 2049 * 2,
             2.5 * p2 * 24)
    if with_cpu and "r2c2d_f32_4096^2" in out:
        # the reference declares but does not implement multi-D realdft: the
        # oracle's restatement (FFTW semantics, one thread) stands in
        import oracle_lib as ol
        xr = ol.gaussian_np(0, p2)
        t0 = time.perf_counter()
        ol.run_oracle("realdft", [4096, 4096], xr)
        t = time.perf_counter() - t0
        out["r2c2d_f32_4096^2"]["cpu_baseline"] = {
            "value": round(2.5 * p2 * 24 / t / 1e9, 3), "unit": "GFLOP/s", "cores": 1, "kind": "port",
            "sample": f"the whole transform through the oracle's multi-D realdft, {t:.2f} s"}
        out["r2c2d_f32_4096^2"]["gpu_over_cpu"] = round(
            out["r2c2d_f32_4096^2"]["gflops"] / out["r2c2d_f32_4096^2"]["cpu_baseline"]["value"], 1)
    n = 1 << 14
    fl = 2.5 * n * math.log2(n)
    for op in ("dct2", "dst2", "dct3", "dct4"):
        plan_run(f"{op}_f32_N2^14_x8192", "t2t3" if op != "dct4" else "t4", "f32", [n], 8192, op, n, n, fl)
        with_cpu_line(f"{op}_f32_N2^14_x8192", op, [n], 8192, 8192, fl)
    p2 = 2048 * 2048
    plan_run("adj_c2c2d_f32_2048^2_x16_fwd", "dft", "f32", [2048, 2048], 16, "dft", 2 * p2, 2 * p2, 5 * p2 * 22)
    with_cpu_line("adj_c2c2d_f32_2048^2_x16_fwd", "dft", [2048, 2048], 16, 16, 5 * p2 * 22)
    plan_run("adj_c2c2d_f32_2048^2_x16_vjp", "dft", "f32", [2048, 2048], 16, "dft", 2 * p2, 2 * p2, 5 * p2 * 22,
             vjp=True)
    with_cpu_line("adj_c2c2d_f32_2048^2_x16_vjp", "invdft", [2048, 2048], 16, 16, 5 * p2 * 22)
    n = 1 << 14
    fl = 2.5 * n * math.log2(n)
    plan_run("adj_dct2_f32_N2^14_x1024_vjp", "t2t3", "f32", [n], 1024, "dct2", n, n, fl, vjp=True)
    # the vjp of dct2 is the reference's dct3 with the DC input doubled
    with_cpu_line("adj_dct2_f32_N2^14_x1024_vjp", "dct3", [n], 1024, 1024, fl)
    masked_power(pkg, torch, out, wanted, with_cpu)
    # the drop-in C API on host buffers (staged through the aux's device
    # buffers, so PCIe is in the loop): configuration C1 (FP64 N = 4096,
    # dft then invdft, seed 0) and one large FP32 transform
    if wanted("dropin"):
        import numpy as np
        import oracle_lib as ol
        lib64 = pkg.dropin("f64")
        n = 4096
        a = lib64.notorious_fft_mkaux_dft_1d(n)
        x = ol.uniform(0, 2 * n)
        y = np.zeros_like(x)
        z = np.zeros_like(x)

        def c1():
            lib64.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
            lib64.notorious_fft_invdft(y.ctypes.data, z.ctypes.data, a)
        for _ in range(5):
            c1()
        t0 = time.perf_counter()
        reps = 200
        for _ in range(reps):
            c1()
        t = (time.perf_counter() - t0) / reps
        rt = ol.rel_l2(z, n * x)
        lib64.notorious_fft_free_aux(a)
        out["dropin_c1_f64_N4096_dft_invdft_host"] = {"us_per_round_trip": round(1e6 * t, 1),
                                                       "round_trip_rel_l2": float(f"{rt:.3e}")}
        ref, rname = ref_lib() if with_cpu else (None, "")
        if ref is not None:
            # the reference itself on one thread: 1000 dft + 1000 invdft of N = 4096
            xs = ol.uniform(0, 2 * n * 1000)
            ys, t_f = ref.run_timed("dft", [n], xs, 1000, 1)
            _, t_i = ref.run_timed("invdft", [n], ys, 1000, 1)
            out["dropin_c1_f64_N4096_dft_invdft_host"]["cpu_baseline"] = {
                "us_per_round_trip": round(1e6 * (t_f + t_i) / 1000, 2), "cores": 1, "kind": "reference",
                "sample": f"1000 dft + 1000 invdft of N = 4096 on one thread, {rname}"}
        log(f"  dropin C1: {out['dropin_c1_f64_N4096_dft_invdft_host']}")
        lib32 = pkg.dropin("f32")
        n = 1 << 22
        a = lib32.notorious_fft_mkaux_dft_1d(n)
        x = np.asarray(ol.uniform(0, 2 * n), dtype=np.float32)
        y = np.zeros_like(x)
        for _ in range(2):
            lib32.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
        t0 = time.perf_counter()
        for _ in range(10):
            lib32.notorious_fft_dft(x.ctypes.data, y.ctypes.data, a)
        t = (time.perf_counter() - t0) / 10
        lib32.notorious_fft_free_aux(a)
        out["dropin_f32_N2^22_host"] = {"ms": round(1e3 * t, 3), "pcie_inclusive_gbs": round(2 * n * 8 / t / 1e9, 2)}
        log(f"  dropin host 2^22: {out['dropin_f32_N2^22_host']}")

    # 7-smooth lengths (mixed radix) beside the power-of-two sweep; the
    # reference runs them through Bluestein
    for prec, n in (("f32", 100), ("f64", 100), ("f64", 10000), ("f32", 10000)):
        batch = (1 << 27) // n
        plan_run(f"mixed_radix_{prec}_N{n}_x{batch}", "dft", prec, [n], batch, "dft", 2 * n, 2 * n,
                 5 * n * math.log2(n))
        with_cpu_line(f"mixed_radix_{prec}_N{n}_x{batch}", "dft", [n], batch, 262144 if n == 100 else 4096,
                      5 * n * math.log2(n))
    if not wanted("ragged"):
        return out
    sizes = ragged_sizes(10000)
    lib = pkg.core()
    for prec in ("f32", "f64"):
        h = ctypes.c_void_p()
        arr = (ctypes.c_int * len(sizes))(*sizes)
        pkg.check(lib.nfft_ragged_create(ctypes.byref(h), 0 if prec == "f32" else 1, len(sizes), arr, 0),
                  "ragged_create")
        total = sum(sizes)
        dt = torch.float32 if prec == "f32" else torch.float64
        x = torch.empty(2 * total, dtype=dt, device=dev)
        y = torch.empty_like(x)
        pkg.fill_uniform(x.data_ptr(), x.numel(), 0, 0, prec)
        t = timed(torch, lambda: pkg.check(lib.nfft_ragged_execute(h, 0, x.data_ptr(), y.data_ptr(), None)),
                  steps=10, warmup=5)
        flops = sum(5.0 * m * math.log2(m) for m in sizes)
        eb = 8 if prec == "f32" else 16
        out[f"ragged_{prec}_10000"] = {"ms": round(1e3 * t, 4), "gflops": round(flops / t / 1e9, 1),
                                       "alg_gbs": round(2 * total * eb / t / 1e9, 1),
                                       "hbm_frac": round(2 * total * eb / t / 1e9 / HBM_PEAK_GBS, 4),
                                       "elements": total}
        log(f"  ragged {prec}: {out[f'ragged_{prec}_10000']}")
        lib.nfft_ragged_destroy(h)
        del x, y
        torch.cuda.empty_cache()
        # the same transforms through the plain path: one batched plan per
        # distinct size, executed back to back on the stream
        groups = {}
        for m in sizes:
            groups[m] = groups.get(m, 0) + 1
        plans = [(pkg.Plan("dft", prec, [m], c), m, c) for m, c in sorted(groups.items())]
        x = torch.empty(2 * total, dtype=dt, device=dev)
        y = torch.empty_like(x)
        pkg.fill_uniform(x.data_ptr(), x.numel(), 0, 0, prec)

        def plain():
            off = 0
            for p, m, c in plans:
                p.execute("dft", x.data_ptr() + off * eb, y.data_ptr() + off * eb)
                off += m * c
        t = timed(torch, plain, steps=3, warmup=1)
        out[f"ragged_plain_{prec}_10000"] = {"ms": round(1e3 * t, 4), "gflops": round(flops / t / 1e9, 1),
                                             "alg_gbs": round(2 * total * eb / t / 1e9, 1),
                                             "plans": len(plans)}
        log(f"  ragged plain {prec}: {out[f'ragged_plain_{prec}_10000']}")
        for p, _, _ in plans:
            p.close()
        del x, y
        torch.cuda.empty_cache()
    out["ragged_cpu_sample"] = ragged_cpu(sizes)
    log(f"  ragged cpu: {out['ragged_cpu_sample']}")
    return out


def ragged_cpu(sizes, budget=2000):
    """The reference (FP64) on a bounded random sample of the ragged list:
    one aux per transform made beforehand, the transforms OpenMP-parallel
    over the list; scaled to the whole list by nominal flops."""
    import random
    import oracle_lib as ol
    cores = min(16, os.cpu_count() or 1)
    flags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    name = "libnfft_ref_fast64.so" if ("avx512f" in flags and "avx512dq" in flags) else "libnfft_ref_fast64_avx2.so"
    rnd = random.Random(0)
    sample = rnd.sample(sizes, min(budget, len(sizes)))
    x = ol.uniform(0, 2 * sum(sample))
    try:
        _, t = ol.Ref(name).run_list(sample, x, nthreads=cores)
        kind = "reference"
    except (FileNotFoundError, OSError):
        kind, name = "port", "oracle"
        t0 = time.perf_counter()
        off = 0
        for m in sample:
            ol.run_oracle("dft", [m], x[off:off + 2 * m])
            off += 2 * m
        t = time.perf_counter() - t0
    fl_s = sum(5.0 * m * math.log2(m) for m in sample)
    fl_all = sum(5.0 * m * math.log2(m) for m in sizes)
    return {"value_gflops": round(fl_s / t / 1e9, 3), "est_full_ms": round(1e3 * t * fl_all / fl_s, 1),
            "cores": cores if kind == "reference" else 1, "kind": kind,
            "sample": f"{len(sample)} of {len(sizes)} transforms (seeded random), {name}, "
                      f"{t:.2f} s in the transform loop (auxes made beforehand)"}


def cpu_baseline(n=1 << 20, budget_s=10.0):
    """The reference compiled from its own header, FP64, OpenMP over the batch
    (one aux per thread), on a bounded sample of the workload."""
    import ctypes
    import numpy as np
    import oracle_lib as ol

    cores = min(16, os.cpu_count() or 1)
    flags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    name = "libnfft_ref_fast64.so" if ("avx512f" in flags and "avx512dq" in flags) else "libnfft_ref_fast64_avx2.so"
    kind = "reference"
    try:
        ref = ol.Ref(name)
        run = lambda x, b: ref.run_timed("dft", [n], x, batch=b, nthreads=cores)[1]  # noqa: E731
    except (FileNotFoundError, OSError):
        kind, name = "port", "oracle"

        def run(x, b):
            t0 = time.perf_counter()
            ol.run_oracle("dft", [n], x, batch=b, nthreads=cores)
            return time.perf_counter() - t0
    x1 = ol.uniform(0, 2 * n * cores)
    per_round = run(x1, cores)
    rounds = max(1, min(8, int(budget_s / max(per_round, 1e-3))))
    batch = rounds * cores
    x = ol.uniform(0, 2 * n * batch)
    t = run(x, batch)
    gflops = 5.0 * n * math.log2(n) * batch / t / 1e9
    cpu_model = ""
    for line in flags.splitlines():
        if line.startswith("model name"):
            cpu_model = line.split(":", 1)[1].strip()
            break
    del ctypes
    return {"value": round(gflops, 3), "unit": "GFLOP/s", "cores": cores, "kind": kind,
            "sample": f"{batch} x N=2^20 FP64 C2C forward ({name}, OpenMP over the batch, one aux "
                      f"per thread, {t:.2f} s in the transform loop); host {cpu_model}; "
                      f"nproc {os.cpu_count()}",
            "alg_gbs": round(2 * n * 16 * batch / t / 1e9, 3)}


def measured_traffic(kernel, n, total):
    """HBM bytes per launch of `kernel` for this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain
    copy), or None when no summary covers it."""
    import glob
    import re
    tag = f"N2^{int(math.log2(n))}_x{total // n}"
    for path in glob.glob(os.path.join(ROOT, "profiles", f"*{tag}*pmc_traffic.md")):
        for line in open(path):
            m = re.match(r"traffic_bytes_per_launch (\S+) (\d+)", line.strip())
            if m and m.group(1) == kernel:
                return int(m.group(2))
    return None


def transform_roofline(head, n, total, steps):
    """Roofline of the whole transform: the algorithmic bytes of one step (one
    read and one write of the batch) over the summed average durations of
    every kernel launched per step (HIP events on the launch stream), against
    the 8 TB/s HBM peak.  Per-kernel figures stay under "kernels"; "traffic"
    is the bytes per step between the L2 caches and the memory side (memory-
    side cache + HBM: an upper bound on HBM traffic) from the committed PMC
    summary when it covers every kernel of the step."""
    kern = head["kernels"]
    alg = 2.0 * total * 8
    step_us = sum(v["avg_us"] * v["launches"] / steps for v in kern.values())
    achieved = alg / (step_us * 1e-6) / 1e9
    traffic = 0
    for k, v in kern.items():
        t = measured_traffic(k, n, total)
        if t is None:
            traffic = None
            break
        traffic += t * v["launches"] / steps
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_over_algorithmic": round(traffic / alg, 3) if traffic else None,
            "kernel": " + ".join(sorted(kern)), "step_kernel_us": round(step_us, 2),
            "bytes_per_step": alg,
            "kernels": {k: {"avg_us": round(v["avg_us"], 2), "launches": v["launches"],
                            "bytes_per_launch": v["bytes_per_launch"], "gbs": round(v["gbs"], 1),
                            "traffic_per_launch": measured_traffic(k, n, total)}
                        for k, v in kern.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--calib", action="store_true",
                    help="first run a plain device copy of 1 GiB (reference kernel for PMC byte counters)")
    ap.add_argument("--only", default="",
                    help="comma-separated configuration names (prefixes) to run after the headline; "
                         "implies --no-sweep --no-cpu")
    args = ap.parse_args()
    only = [o for o in args.only.split(",") if o]
    if only:
        args.no_sweep = args.no_cpu = True
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "batched 1D transforms shard by splitting the batch, with no exchange", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as ge

    pkg = ge.load_package()
    if args.calib:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)  # reads 1 GiB, writes 1 GiB per launch
        torch.cuda.synchronize()
        del a, b
        torch.cuda.empty_cache()
    total = 1 << 28
    log(f"headline: N={args.n} FP32 batch {total // args.n}, warmup {args.warmup}, steps {args.steps}")
    head = run_config(pkg, torch, args.n, "f32", total, args.steps, args.warmup, profile=True, check=2)
    log(f"  {head['ms_per_step']:.3f} ms/step, {head['gflops']:.1f} GFLOP/s, "
        f"{head['alg_gbs']:.1f} GB/s ({100 * head['hbm_frac']:.1f}% HBM), rel-L2 {head['rel_l2']:.2e}")
    roofline = transform_roofline(head, args.n, total, args.steps)

    sweep_r2c = []
    if not args.no_sweep:
        for prec in ("f32", "f64"):
            for lg in range(10, 21):
                r = run_r2c(pkg, torch, 1 << lg, prec, 1 << 28, steps=10, warmup=10)
                sweep_r2c.append(r)
                log(f"  sweep r2c {prec} N=2^{lg}: {r['gflops']:.1f} GFLOP/s {100 * r['hbm_frac']:.1f}% HBM "
                    f"rel-L2 {r['rel_l2']:.2e} [{r['plan']}]")

    sweep = []
    if not args.no_sweep:
        for prec in ("f32", "f64"):
            for lg in range(10, 21):
                n = 1 << lg
                r = run_config(pkg, torch, n, prec, total, 10, 10, profile=True, check=1)
                kern = r["kernels"]
                dom = max(kern, key=lambda k: kern[k]["total_ms"])
                sweep.append({"n": n, "prec": prec, "gflops": round(r["gflops"], 1),
                              "alg_gbs": round(r["alg_gbs"], 1), "hbm_frac": round(r["hbm_frac"], 4),
                              "ms": round(r["ms_per_step"], 4), "event_ms": round(r["event_ms_per_step"], 4),
                              "rel_l2": float(f"{r['rel_l2']:.3e}"), "plan": r["plan"],
                              "kernel": dom, "kernel_gbs": round(kern[dom]["gbs"], 1),
                              "kernel_us": round(kern[dom]["avg_us"], 1)})
                log(f"  sweep {prec} N=2^{lg}: {r['gflops']:.1f} GFLOP/s {100 * r['hbm_frac']:.1f}% HBM "
                    f"(dominant {dom} {kern[dom]['gbs']:.0f} GB/s) rel-L2 {r['rel_l2']:.2e} [{r['plan']}]")

    extras = {}
    if not args.no_sweep or only:
        extras = config_extras(pkg, torch, only, with_cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline()
            log(f"  cpu baseline: {cpu['value']} GFLOP/s on {cpu['cores']} threads ({cpu['kind']})")
        except Exception as exc:  # reported, not fatal for the GPU figure
            cpu = {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "reference",
                   "sample": f"failed: {exc}"}

    out = {
        "metric": METRIC,
        "value": round(head["gflops"], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (SplitMix64 uniform [-1,1), seed 0, generated on device)",
        "config": {"workload": f"batched 1D C2C forward, FP32, N=2^{int(math.log2(args.n))} x "
                               f"{total // args.n} transforms (2^28 complex elements)",
                   "n": args.n, "batch": total // args.n, "plan": head["plan"]},
        "hbm_frac": round(head["hbm_frac"], 4),
        "alg_gbs": round(head["alg_gbs"], 1),
        "accuracy": {"rel_l2": float(f"{head['rel_l2']:.3e}"), "tol": head["rel_l2_tol"],
                     "vs": "FP64 oracle on the same FP32 inputs, 2 sampled transforms"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "sweep": sweep,
        "sweep_r2c": sweep_r2c,
        "configs": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
