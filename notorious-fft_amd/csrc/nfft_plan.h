// nfft_plan.h -- plan objects behind the C ABI (device-aux).
#pragma once
#include <string>
#include <vector>

#include "nfft_internal.h"

namespace nfft {

hipStream_t setup_stream();

// realdft / invrealdft along contiguous lines of N reals (N a power of two):
// single pass for N/2 <= Lmax, else complex FFT of N/2 + an elementwise
// unpack (forward) or pack (inverse) pass.
struct R2C {
    int N = 0, prec = 0, L = 0;
    bool single = true;
    int halves = 0;             // single: one line per workgroup, exchanges in two LDS halves
                                // (bit 0: realdft, bit 1: invrealdft)
    const void *tw = nullptr;   // stage twiddles (single)
    const void *rtw = nullptr;  // e^{-2 pi i k/N}, k < N/2
    C2C inner;                  // multi-pass path
    bool fused = false;         // forward: unpack fused into the four-step row pass
    int r2c_lpb = 0;            // rows per workgroup of the fused row pass (0: default)
    long long scratch_per_line = 0;
    bool init(int N, int prec);
    void release();
    // dbl: multiply bins 0 < k < N/2 by 2 (adjoint of invrealdft)
    hipError_t fwd(const void *in, void *out, long long nlines, long long in_dist, long long out_dist, double scale,
                   int dbl, void *scratch, long long scratch_elems, hipStream_t s) const;
    // half: weight 1/2 on bins 0 < k < N/2 (adjoint of realdft)
    hipError_t inv(const void *in, void *out, long long nlines, long long in_dist, long long out_dist, double scale,
                   int half, void *scratch, long long scratch_elems, hipStream_t s) const;
};

// DCT/DST types 2/3 (t4 = 0) or type 4 (t4 = 1) along contiguous lines.
struct Trig {
    int N = 0, prec = 0, t4 = 0, L = 0;
    bool single = true;
    const void *tw = nullptr;
    const void *dw = nullptr, *rw = nullptr;     // types 2/3
    const void *pre = nullptr, *post = nullptr;  // type 4
    R2C real;                                     // types 2/3 multi-pass
    C2C cplx;                                     // type 4 multi-pass
    long long scratch_per_line = 0;               // complex elements
    bool init(int N, int prec, int t4);
    void release();
    hipError_t run(int op, const void *in, void *out, long long nlines, long long in_dist, long long out_dist,
                   double scale, double w0, void *scratch, long long scratch_elems, hipStream_t s) const;
};

}  // namespace nfft

struct nfft_plan {
    int kind = 0, prec = 0, d = 0, norm = 0;
    std::vector<int> dims;
    long long batch = 1;
    long long P = 1;
    int n_active = 1;                 // trig: axes [0, n_active) are transformed
    std::vector<nfft::C2C> ax;        // DFT axes (REALDFT: leading axes)
    nfft::R2C real;                   // REALDFT last axis
    std::vector<nfft::Trig> tax;      // trig axes
    void *work = nullptr;
    size_t work_bytes = 0;
    std::string desc;
    double *wpow_part = nullptr;      // loss partial sums of weighted_power (grown on demand)
    long long wpow_cap = 0;

    bool init(int kind, int prec, int d, const int *dims, long long batch, int norm);
    void release();
    // vjp_mode != 0 runs the vector-Jacobian product of `op`
    int execute(int op, const void *in, void *out, hipStream_t s, int vjp_mode);
    // loss[t] = sum |W Y_t|^2 (Y_t = DFT(x_t)) and grad = 2 DFT^H (W^2 Y)
    int weighted_power(const void *x, const void *w, void *grad, double *loss, hipStream_t s);
};
