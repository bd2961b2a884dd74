"""Writes tests/golden/golden_ref_fp64.npz: inputs and outputs of the
reference library, compiled from its own header into oracle/_ref/ (FP64,
scalar IEEE build), for the cases below.  Run in the container that holds the
reference source:  make -C oracle && python tests/golden/gen_golden.py

Inputs: seed-0 SplitMix64 uniform [-1, 1) (the C1 configuration), and for the
sizes of the reference's own test driver its deterministic pattern
x[i] = sin(1.23456789 i + 0.1 i^2) (tests/test_notoriousfft.c:45-48).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402


def sinfill(n):
    i = np.arange(n, dtype=np.float64)
    return np.sin(i * 1.23456789 + 0.1 * i * i)


def cases():
    out = []
    # C1: N = 4096 dft then invdft, seed 0
    out.append(("dft", [4096], "seed0"))
    out.append(("invdft", [4096], "seed0"))
    for n in [1, 2, 4, 8, 16, 64, 1024]:
        out.append(("dft", [n], "seed0"))
        out.append(("invdft", [n], "seed0"))
    for n in [3, 5, 6, 7, 10, 12, 15, 17, 20, 100, 127, 1009]:
        out.append(("dft", [n], "pattern"))
        out.append(("invdft", [n], "seed0"))
    for n in [1, 2, 4, 8, 16, 1024, 4096]:
        out.append(("realdft", [n], "seed0"))
        if n > 1:
            out.append(("invrealdft", [n], "seed0"))
    for op in ["dct2", "dst2", "dct3", "dst3", "dct4", "dst4"]:
        for n in [1, 2, 4, 8, 16, 1024, 4096]:
            out.append((op, [n], "seed0"))
        out.append((op, [32], "pattern"))
    out.append(("dft", [32, 32], "seed0"))
    out.append(("invdft", [8, 16], "seed0"))
    out.append(("dft", [8, 8, 8], "seed0"))
    out.append(("dft", [6, 10], "seed0"))
    for op in ["dct2", "dst2", "dct3", "dst3", "dct4", "dst4"]:
        out.append((op, [8, 16], "seed0"))
    out.append(("dct2", [4, 8, 8], "seed0"))
    return out


def main():
    ref = ol.Ref("libnfft_ref_exact.so")
    arrays = {}
    manifest = []
    for op, dims, inp in cases():
        nin, nout = ol.out_shape(op, dims)
        if inp == "seed0":
            x = ol.uniform(0, nin)
        else:
            x = sinfill(nin)
        y = ref.run(op, dims, x)
        key = f"{op}_{'x'.join(map(str, dims))}_{inp}"
        arrays[key + "_in"] = x
        arrays[key + "_out"] = y
        manifest.append({"key": key, "op": op, "dims": dims, "input": inp})
    path = os.path.join(HERE, "golden_ref_fp64.npz")
    np.savez(path, **arrays)
    with open(path, "rb") as f:
        digest = hashlib.sha256(f.read()).hexdigest()
    with open(os.path.join(HERE, "golden_ref_fp64.json"), "w") as f:
        json.dump({"source": "reference header compiled by oracle/Makefile (libnfft_ref_exact.so)",
                   "flags": "-O2 -ffp-contract=off -fno-fast-math, no vector ISA flags",
                   "sha256": digest, "cases": manifest}, f, indent=1)
    print(f"wrote {len(manifest)} cases, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
