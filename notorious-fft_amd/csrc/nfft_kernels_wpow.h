// nfft_kernels_wpow.h -- weighted spectral power and its gradient:
//     L(x) = sum_k |W_k Y_k|^2,  Y = DFT(x),   grad = 2 DFT^H (W^2 . Y)
// (g = dL/dRe x + i dL/dIm x), per transform of a batch.
//
// k_col_wpow fuses the last pass of the forward transform, the weighting and
// the first pass of the adjoint: a tile of columns of axis 0 (length L,
// stride S) is transformed forward, each bin k multiplied by 2 W_k^2 while
// W_k^2 |Y_k|^2 is summed (in FP64), and transformed back (unnormalised
// inverse) before it is stored: one HBM round trip instead of three.  The
// other axes run as ordinary passes before and after it.
// k_wpow_mask is the same weighting as an elementwise pass, for plans whose
// axis 0 is not a single-pass column length.
// Loss partial sums go to one slot per workgroup; k_wpow_reduce adds each
// transform's slots in a fixed order (deterministic, FP64).
#pragma once
#include "nfft_internal.h"

namespace nfft {

template <int L, int LPB> __host__ __device__ constexpr int wpow_threads() { return Sched<L>::NT * LPB; }

// Sum of v over the block, returned to thread 0 (FP64; red holds one entry per wave).
template <int NTH> __device__ __forceinline__ double block_sum(double v, double *red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NTH / 64; i++) t += red[i];
    return t;
}

template <typename T, int L, int LPB>
__global__ void __launch_bounds__((wpow_threads<L, LPB>())) k_col_wpow(WpowArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int NTH = wpow_threads<L, LPB>();
    static_assert(NTH % 64 == 0, "whole waves");
    constexpr int NLDS = S::S > 1 ? L * LPB : 1;
    constexpr int TWN = S::TWN > 0 ? S::TWN : 1;
    __shared__ C lds[NLDS];
    __shared__ C ltw[TWN];
    __shared__ double red[NTH / 64];
    if constexpr (S::S > 1) stage_tw_to_lds<T, L>(ltw, (const double2 *)a.tw);
    const int b = threadIdx.x % LPB, j = threadIdx.x / LPB;
    const long long tile = blockIdx.x + (long long)gridDim.x * blockIdx.y;  // (transform, column tile)
    const long long t = tile / gridDim.x;
    const int c0 = (int)(tile % gridDim.x) * LPB, c = c0 + b;
    const bool valid = c < a.ncols;
    C *col = (C *)a.buf + t * a.dist + c;
    const T *w = (const T *)a.w + c;
    C reg[S::R];
#pragma unroll
    for (int q = 0; q < S::R; q++) reg[q] = valid ? col[(long long)(j + S::NT * q) * a.ncols] : czero<T>();
    __syncthreads();  // stage table in LDS
    LayCol<LPB> lay{b};
    line_fft_tw<T, L, -1>(reg, lds, lay, j, StageTwLds<T>{ltw});
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < S::R; q++) {
        const T wk = valid ? w[(long long)(j + S::NT * q) * a.ncols] : (T)0;
        const T w2 = wk * wk;
        const C y = reg[q];
        part += (double)w2 * ((double)y.x * y.x + (double)y.y * y.y);
        reg[q] = cscale(y, (T)2 * w2);
    }
    line_fft_tw<T, L, 1>(reg, lds, lay, j, StageTwLds<T>{ltw});
    if (valid) {
#pragma unroll
        for (int q = 0; q < S::R; q++) col[(long long)(j + S::NT * q) * a.ncols] = reg[q];
    }
    const double tot = block_sum<NTH>(part, red);
    if (threadIdx.x == 0) a.partial[tile] = tot;
}

// Elementwise weighting of a spectrum (batch x P complex), with loss partials
// per workgroup: blockIdx.y = transform, blockIdx.x = chunk of P.
template <typename T>
__global__ void __launch_bounds__(256) k_wpow_mask(void *buf, const void *w, double *partial, long long P)
{
    using C = Cx<T>;
    __shared__ double red[4];
    C *y = (C *)buf + (long long)blockIdx.y * P;
    const T *wt = (const T *)w;
    double part = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < P; i += (long long)gridDim.x * 256) {
        const T w2 = wt[i] * wt[i];
        const C v = y[i];
        part += (double)w2 * ((double)v.x * v.x + (double)v.y * v.y);
        y[i] = cscale(v, (T)2 * w2);
    }
    const double tot = block_sum<256>(part, red);
    if (threadIdx.x == 0) partial[(long long)blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

// loss[t] = sum of the `per` partials of transform t, in index order.
__global__ void __launch_bounds__(64) k_wpow_reduce(const double *partial, double *loss, long long per)
{
    __shared__ double red[1];
    const double *p = partial + (long long)blockIdx.x * per;
    double v = 0.0;
    for (long long i = threadIdx.x; i < per; i += 64) v += p[i];
    const double tot = block_sum<64>(v, red);
    if (threadIdx.x == 0) loss[blockIdx.x] = tot;
}

}  // namespace nfft
