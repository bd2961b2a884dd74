// nfft_kernels_rt.h -- real-input and trigonometric transform kernels.
//
// All of them run an L-point complex line (L = N/2) with the pack/unpack and
// pre/post twiddles fused into the same workgroup, the line resident in LDS:
//   k_r2c_line   realdft    (src/07_legacy.h:162-237): pack N reals as N/2
//                complex, FFT, unpack N/2+1 Hermitian bins.
//   k_c2r_line   invrealdft (src/07_legacy.h:239-293): pack, inverse FFT,
//                unnormalised (invrealdft(realdft(x)) = N x).
//   k_t23_line   dct2/dst2  (src/07_legacy.h:757-866,867-949): even/odd
//                reorder on load, realdft, post-twiddle e^{-i pi k/(2N)}.
//                dct3/dst3  (src/07_legacy.h:951-1066): pre-twiddle,
//                invrealdft, un-reorder on store.
//   k_t4_line    dct4/dst4  (src/07_legacy.h:1091-1137,1298-1334): premul
//                e^{-i pi n/N}, N/2-point FFT, postmul e^{-i pi (2n+1)/(4N)}.
// For N/2 above the single-pass limit the same arithmetic runs as separate
// elementwise kernels around the multi-pass complex FFT (k_rt_* below).
#pragma once
#include "nfft_kernels.h"

namespace nfft {

enum TrigOp { T_DCT2 = 0, T_DST2 = 1, T_DCT3 = 2, T_DST3 = 3, T_DCT4 = 4, T_DST4 = 5 };

struct RtArgs {
    const void *in;
    void *out;
    long long nlines;
    long long in_dist, out_dist;  // in units of the element type on each side
    const void *tw;               // stage twiddles for L
    const void *t1;               // realdft: e^{-2 pi i k/N} (k <= L); t23: e^{-i pi k/(2N)} (k < L);
                                  // t4: premul e^{-i pi n/N} (n < L)
    const void *t2;               // t23: realdft twiddles; t4: postmul (n < N)
    double scale;
    double w0;                    // weight of the DC term (vjp weightings)
    int half;                     // c2r: weight 1/2 on bins 0 < k < N/2 (realdft vjp)
    int halves;                   // realdft / invrealdft: one line per workgroup, exchanges in two LDS halves
};

// Waves per SIMD the FP32 single-pass real / trig line kernels are compiled
// for (wpe_hint: as many as their LDS allows, at most 4, i.e. <= 128 VGPRs):
// above 128 VGPRs an 8192-point line kernel held one workgroup per CU.
template <typename T, int L> __host__ __device__ constexpr int rt_wpe()
{
    return wpe_hint<T>(Sched<L>::NT * line_lpb<L>(), (long)line_lds<T, L, line_lpb<L>()>() * (long)sizeof(Cx<T>));
}

// real view of a padded complex LDS line: real index i -> slot i>>1, comp i&1
template <typename T, int L> __device__ __forceinline__ T lds_real(const Cx<T> *line, int i)
{
    const Cx<T> v = line[lpad<T>(i >> 1)];
    return (i & 1) ? v.y : v.x;
}

// X[k] of the length-2L real DFT from the L-point complex DFT Z of the
// packed sequence: E = (Zk + conj Zm)/2, O = -i (Zk - conj Zm)/2,
// X[k] = E + w O (w = e^{-2 pi i k/(2L)}), X[L] from k = 0 as E - O.
template <typename C> __device__ __forceinline__ void unpack_pair(C zk, C zm, C w, C &xk, C &xl)
{
    using T = decltype(C::x);
    const T h = (T)0.5;
    C e = cmk<C>(h * (zk.x + zm.x), h * (zk.y - zm.y));
    C o = cmk<C>(h * (zk.y + zm.y), -h * (zk.x - zm.x));
    xk = cadd(e, cmul(w, o));
    xl = csub(e, o);
}

// Z[k] = (Xk + conj Xm) + i conj(w) (Xk - conj Xm), w = e^{-2 pi i k/(2L)}:
// inverse FFT of Z gives N * (y[2n] + i y[2n+1]).
template <typename C> __device__ __forceinline__ C pack_pair(C xk, C xm, C w)
{
    C s = cmk<C>(xk.x + xm.x, xk.y - xm.y);
    C d = cmk<C>(xk.x - xm.x, xk.y + xm.y);
    C t = cmulc(d, w);
    return cmk<C>(s.x - t.y, s.y + t.x);
}

// ---------------------------------------------------------------------------
template <typename T, int L>
__global__ void __launch_bounds__(Sched<L>::NT *line_lpb<L>(), (rt_wpe<T, L>())) k_r2c_line(RtArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int LPB = line_lpb<L>();
    __shared__ C lds[line_lds<T, L, LPB>()];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const long long line0 = (long long)blockIdx.x * LPB;
    const long long rem = a.nlines - line0;
    const int nrows = rem < LPB ? (int)rem : LPB;
    const C *src = (const C *)((const T *)a.in + line0 * a.in_dist);
    C reg[S::R];
    tile_load<T, L, LPB>(reg, lds, src, a.in_dist / 2, L, nrows, b, j);
    LayLine<T, L> lay{b * lpadlen<T, L>()};
    line_fft<T, L, -1>(reg, lds, lay, j, (const double2 *)a.tw);
    C *line = lds + b * lpadlen<T, L>();
#pragma unroll
    for (int q = 0; q < S::R; q++) line[lpad<T>(j + S::NT * q)] = reg[q];
    __syncthreads();
    if (b >= nrows) return;
    C *dst = (C *)a.out + (line0 + b) * a.out_dist;
    const C *w = (const C *)a.t1;
    const C wj = w[j];  // w[k] = w[j] w[NT q] (the second workgroup-uniform)
    const T sc = (T)a.scale, sk = a.half ? (T)2 * sc : sc;  // half: adjoint weighting 2 on 0 < k < L
#pragma unroll
    for (int q = 0; q < S::R; q++) {
        const int k = j + S::NT * q;
        const C zk = line[lpad<T>(k)], zm = line[lpad<T>((L - k) & (L - 1))];
        C xk, xl;
        unpack_pair(zk, zm, q == 0 ? wj : cmul(wj, w[S::NT * q]), xk, xl);
        dst[k] = cscale(xk, k == 0 ? sc : sk);
        if (k == 0) dst[L] = cscale(xl, sc);
    }
}

// k_r2c_line_h / k_c2r_line_h: realdft / invrealdft of 2L reals with the
// L-point complex line of k_line_h (one line per workgroup, NT / 2 threads
// holding two sets of 16 points, exchanges through half a line of LDS).
// The unpack's partner Z[L - k] of a thread's point k comes through the same
// half buffer in two steps: with the upper half (points of q >= 8) in LDS
// the threads read the partners of their points 0 < k < L/2, then with the
// lower half those of L/2 < k < L; k = 0 and k = L/2 are their own partners.
// Each step's outputs are stored before the next step's partners are read.
// The inverse reads both X[k] and X[L - k] from global memory, so it needs no
// extra exchange.  Arithmetic per point is that of k_r2c_line / k_c2r_line.
template <typename T, int L>
__global__ void __launch_bounds__(Sched<L>::NT / 2, (line_h_wpe<T, L>())) k_r2c_line_h(RtArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int NH = S::NT / 2, H = L / 2;
    __shared__ C lds[lpadlen<T, L / 2>()];
    const long long line = blockIdx.x;
    const int j = threadIdx.x;
    const int jj[2] = {j, j + NH};
    const uint32_t off = (uint32_t)j * (uint32_t)sizeof(C);
    C reg[2][16];
    {
        const auto rs = buf_rsrc((const T *)a.in + line * a.in_dist);
#pragma unroll
        for (int l = 0; l < 2; l++)
#pragma unroll
            for (int q = 0; q < 16; q++) reg[l][q] = buf_ld<C>(rs, off, (uint32_t)((l * NH + S::NT * q) * sizeof(C)));
    }
    line_fft_halves<T, L, -1>(reg, lds, j, (const double2 *)a.tw);
    const C *w = (const C *)a.t1;
    const C wj[2] = {w[jj[0]], w[jj[1]]};  // w[k] = w[jj] w[NT q] (the second workgroup-uniform)
    const T sc = (T)a.scale, sk = a.half ? (T)2 * sc : sc;  // half: adjoint weighting 2 on 0 < k < L
    C *dst = (C *)a.out + line * a.out_dist;
    // h = 0: upper half in LDS, outputs of q < 8; h = 1: lower half, q >= 8
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int qs = h == 0 ? 8 : 0;  // the points written (their half is in LDS)
        const int qo = h == 0 ? 0 : 8;  // the points whose outputs this step gives
#pragma unroll
        for (int l = 0; l < 2; l++)
#pragma unroll
            for (int q = 0; q < 8; q++) lds[lpad<T>(jj[l] + S::NT * q)] = reg[l][q + qs];  // point - (1-h) H
        __syncthreads();
#pragma unroll
        for (int l = 0; l < 2; l++)
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = jj[l] + S::NT * (q + qo);
                const int p = (L - k) & (L - 1);        // in the half held now unless k is 0 or L/2
                const int pl = p - (h == 0 ? H : 0);
                const C v = lds[lpad<T>(pl < 0 ? 0 : (pl >= H ? H - 1 : pl))];
                const C zm = (k == 0 || k == H) ? reg[l][q + qo] : v;
                C xk, xl;
                unpack_pair(reg[l][q + qo], zm, q + qo == 0 ? wj[l] : cmul(wj[l], w[S::NT * (q + qo)]), xk, xl);
                dst[k] = cscale(xk, k == 0 ? sc : sk);
                if (k == 0) dst[L] = cscale(xl, sc);
            }
        if (h == 0) __syncthreads();  // the buffer is rewritten for h = 1
    }
}

template <typename T, int L>
__global__ void __launch_bounds__(Sched<L>::NT / 2, (line_h_wpe<T, L>())) k_c2r_line_h(RtArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int NH = S::NT / 2;
    __shared__ C lds[lpadlen<T, L / 2>()];
    const long long line = blockIdx.x;
    const int j = threadIdx.x;
    const int jj[2] = {j, j + NH};
    const C *src = (const C *)a.in + line * a.in_dist;
    const C *w = (const C *)a.t1;
    const C wj[2] = {w[jj[0]], w[jj[1]]};  // w[k] = w[jj] w[NT q] (the second workgroup-uniform)
    const T hw = a.half ? (T)0.5 : (T)1;
    C reg[2][16];
#pragma unroll
    for (int l = 0; l < 2; l++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int k = jj[l] + S::NT * q;
            C xk = src[k], xm = src[L - k];
            if (k == 0) {  // bins 0 and N/2 are real (imaginary parts ignored)
                xk.y = 0;
                xm.y = 0;
            } else {
                xk = cscale(xk, hw);
                xm = cscale(xm, hw);
            }
            reg[l][q] = pack_pair(xk, xm, q == 0 ? wj[l] : cmul(wj[l], w[S::NT * q]));
        }
    line_fft_halves<T, L, 1>(reg, lds, j, (const double2 *)a.tw);
    const T sc = (T)a.scale;
    const auto rd = buf_rsrc((T *)a.out + line * a.out_dist);
    const uint32_t off = (uint32_t)j * (uint32_t)sizeof(C);
#pragma unroll
    for (int l = 0; l < 2; l++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            C v = reg[l][q];
            if (sc != (T)1) v = cscale(v, sc);
            buf_st<C>(rd, off, (uint32_t)((l * NH + S::NT * q) * sizeof(C)), v);
        }
}

template <typename T, int L>
__global__ void __launch_bounds__(Sched<L>::NT *line_lpb<L>(), (rt_wpe<T, L>())) k_c2r_line(RtArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int LPB = line_lpb<L>();
    __shared__ C lds[line_lds<T, L, LPB>()];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const long long line0 = (long long)blockIdx.x * LPB;
    const long long rem = a.nlines - line0;
    const int nrows = rem < LPB ? (int)rem : LPB;
    const C *src = (const C *)a.in + (line0 + (b < nrows ? b : 0)) * a.in_dist;
    const C *w = (const C *)a.t1;
    const C wj = w[j];  // w[k] = w[j] w[NT q] (the second workgroup-uniform)
    C reg[S::R];
    const T hw = a.half ? (T)0.5 : (T)1;
#pragma unroll
    for (int q = 0; q < S::R; q++) {
        const int k = j + S::NT * q;
        C xk = src[k], xm = src[L - k];
        if (k == 0) {  // bins 0 and N/2 are real (imaginary parts ignored)
            xk.y = 0;
            xm.y = 0;
        } else {
            xk = cscale(xk, hw);
            xm = cscale(xm, hw);
        }
        reg[q] = pack_pair(xk, xm, q == 0 ? wj : cmul(wj, w[S::NT * q]));
    }
    LayLine<T, L> lay{b * lpadlen<T, L>()};
    line_fft<T, L, 1>(reg, lds, lay, j, (const double2 *)a.tw);
    const T sc = (T)a.scale;
    if (sc != (T)1) {
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = cscale(reg[q], sc);
    }
    C *dst = (C *)((T *)a.out + line0 * a.out_dist);
    tile_store<T, L, LPB>(reg, lds, dst, a.out_dist / 2, L, nrows, b, j);
}

// DCT/DST types 2 and 3 -----------------------------------------------------
// (DST-3 keeps the compiler's register choice: at 128 VGPRs it spills)
template <typename T, int L, int OP>
__global__ void __launch_bounds__(Sched<L>::NT *line_lpb<L>(), (OP == T_DST3 ? 1 : rt_wpe<T, L>())) k_t23_line(RtArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int LPB = line_lpb<L>();
    constexpr int NTH = S::NT * LPB;
    constexpr int N = 2 * L;
    __shared__ C lds[line_lds<T, L, LPB>()];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const long long line0 = (long long)blockIdx.x * LPB;
    const long long rem = a.nlines - line0;
    const int nrows = rem < LPB ? (int)rem : LPB;
    const C *src = (const C *)((const T *)a.in + line0 * a.in_dist);
    // lines of >= 64 threads (one or more whole waves per line) read their
    // permuted input straight from global memory, and DCT2/DST2 write their
    // output straight back; shorter lines stage the line through LDS
    constexpr bool DIRECT = S::NT >= 64;
    if constexpr (!DIRECT) {
        stage_in_rows<T, L, LPB, NTH>(lds, src, a.in_dist / 2, L, nrows);
        __syncthreads();
    }
    const C *line = lds + b * lpadlen<T, L>();
    C reg[S::R];
    LayLine<T, L> lay{b * lpadlen<T, L>()};
    const T sq2 = (T)1.41421356237309504880;
    if constexpr (OP == T_DCT2 || OP == T_DST2) {
        constexpr T sg = OP == T_DCT2 ? (T)1 : (T)-1;
        // v[i] = x[2i] (i < L), v[L+i] = sg * x[N-1-2i]; z[m] = (v[2m], v[2m+1])
        if constexpr (DIRECT) {
            // z[m] = (x[4m], x[4m+2]) for m < L/2, else (sg x[4c+3], sg x[4c+1])
            // with c = L-1-m: two complex-pair loads of the real line each, or
            // (FP32 lines on 16-byte boundaries) one 16-byte load
            const T *xr = (const T *)a.in + (line0 + b) * a.in_dist;
            const C *xl = (const C *)xr;
            if (sizeof(T) == 4 && (((uintptr_t)xr) & 15) == 0) {  // uniform per wave (NT >= 64)
                const float4 *x4 = (const float4 *)xr;
#pragma unroll
                for (int q = 0; q < S::R; q++) {
                    const int m = j + S::NT * q;
                    C v = czero<T>();
                    if (b < nrows) {
                        if (q < S::R / 2) {
                            const float4 p = x4[m];
                            v = cmk<C>((T)p.x, (T)p.z);
                        } else {
                            const float4 p = x4[L - 1 - m];
                            v = cmk<C>(sg * (T)p.w, sg * (T)p.y);
                        }
                    }
                    reg[q] = v;
                }
            } else {
#pragma unroll
                for (int q = 0; q < S::R; q++) {
                    const int m = j + S::NT * q;
                    C v = czero<T>();
                    if (b < nrows) {
                        if (q < S::R / 2) {
                            const C p0 = xl[2 * m], p1 = xl[2 * m + 1];
                            v = cmk<C>(p0.x, p1.x);
                        } else {
                            const int c = L - 1 - m;
                            const C p0 = xl[2 * c], p1 = xl[2 * c + 1];
                            v = cmk<C>(sg * p1.y, sg * p0.y);
                        }
                    }
                    reg[q] = v;
                }
            }
        } else {
            auto xv = [&](int i) -> T {
                return i < L ? lds_real<T, L>(line, 2 * i) : sg * lds_real<T, L>(line, N - 1 - 2 * (i - L));
            };
#pragma unroll
            for (int q = 0; q < S::R; q++) {
                const int m = j + S::NT * q;
                reg[q] = cmk<C>(xv(2 * m), xv(2 * m + 1));
            }
            __syncthreads();
        }
        line_fft<T, L, -1>(reg, lds, lay, j, (const double2 *)a.tw);
        C *ln = lds + b * lpadlen<T, L>();
#pragma unroll
        for (int q = 0; q < S::R; q++) ln[lpad<T>(j + S::NT * q)] = reg[q];
        __syncthreads();
        // outputs for k = j + NT q: (y[k], y[N-k]) (DCT2) from X[k]; X[0] and X[L] special
        const C *rw = (const C *)a.t2;   // e^{-2 pi i k/N}
        const C *dw = (const C *)a.t1;   // e^{-i pi k/(2N)}
        // the tables at k = j + NT q as w(j) w(NT q): two per-lane loads and
        // 2 (R - 1) workgroup-uniform ones instead of 2 R per-lane loads
        const C rwj = rw[j], dwj = dw[j];
        T ya[S::R], yb[S::R];
        T yl = 0;
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int k = j + S::NT * q;
            const C zk = ln[lpad<T>(k)], zm = ln[lpad<T>((L - k) & (L - 1))];
            const C rwk = q == 0 ? rwj : cmul(rwj, rw[S::NT * q]);
            C xk, xl;
            unpack_pair(zk, zm, rwk, xk, xl);
            if (k == 0) {
                ya[q] = (T)2 * xk.x * (T)a.w0;
                yb[q] = 0;
                yl = sq2 * xl.x;
            } else {
                const C dwk = q == 0 ? dwj : cmul(dwj, dw[S::NT * q]);
                const C p = cmul(xk, dwk);
                ya[q] = (T)2 * p.x;
                yb[q] = (T)-2 * p.y;
            }
        }
        const T sc = (T)a.scale;
        if constexpr (DIRECT) {
            // y[k] and y[N-k] (DCT2) / y[k-1] and y[N-k-1] (DST2): each store
            // instruction of a wave writes one contiguous run
            if (b < nrows) {
                T *y = (T *)a.out + (line0 + b) * a.out_dist;
#pragma unroll
                for (int q = 0; q < S::R; q++) {
                    const int k = j + S::NT * q;
                    int ia, ib;
                    if constexpr (OP == T_DCT2) {
                        ia = k;
                        ib = N - k;
                    } else {
                        ia = k == 0 ? N - 1 : k - 1;
                        ib = N - k - 1;
                    }
                    const T va = OP == T_DCT2 ? ya[q] : (k == 0 ? ya[q] : yb[q]);
                    const T vb = OP == T_DCT2 ? yb[q] : ya[q];
                    y[ia] = va * sc;
                    if (k != 0) y[ib] = vb * sc;
                    else y[OP == T_DCT2 ? L : L - 1] = yl * sc;
                }
            }
            return;
        }
        __syncthreads();
        // scatter y into the LDS line (real view), then a coalesced store
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int k = j + S::NT * q;
            int ia, ib;
            if constexpr (OP == T_DCT2) {
                ia = k;
                ib = N - k;
            } else {
                ia = k == 0 ? N - 1 : k - 1;
                ib = N - k - 1;
            }
            const T va = OP == T_DCT2 ? ya[q] : (k == 0 ? ya[q] : yb[q]);
            const T vb = OP == T_DCT2 ? yb[q] : ya[q];
            auto put = [&](int i, T v) {
                C *slot = ln + lpad<T>(i >> 1);
                if (i & 1) slot->y = v * sc;
                else slot->x = v * sc;
            };
            put(ia, va);
            if (k != 0) put(ib, vb);
            else put(OP == T_DCT2 ? L : L - 1, yl);
        }
        __syncthreads();
        C *dst = (C *)((T *)a.out + line0 * a.out_dist);
        stage_out_rows<T, L, LPB, NTH>(lds, dst, a.out_dist / 2, L, nrows);
    } else {
        // DCT3: X[0] = w0 x0, X[k] = conj((x[k] + i x[N-k]) w_k), X[L] = sqrt2 x[L]
        // DST3: X[0] = w0 x[N-1], X[k] = conj((x[N-1-k] + i x[k-1]) w_k), X[L] = sqrt2 x[L-1]
        const C *dw = (const C *)a.t1;
        const C *rw = (const C *)a.t2;
        // rows past nrows read row 0 (always present); tile_store drops them
        const T *xg = (const T *)a.in + (line0 + (b < nrows ? b : 0)) * a.in_dist;
        auto rx = [&](int i) -> T {
            if constexpr (DIRECT) return xg[i];
            else return lds_real<T, L>(line, i);
        };
        // X(k) with its twiddle w = dw[k] (unused at k = 0 and k = L)
        auto X = [&](int k, C w) -> C {
            if (k == 0) {
                const T v = OP == T_DCT3 ? rx(0) : rx(N - 1);
                return cmk<C>(v * (T)a.w0, (T)0);
            }
            if (k == L) {
                const T v = OP == T_DCT3 ? rx(L) : rx(L - 1);
                return cmk<C>(sq2 * v, (T)0);
            }
            T xr, xi;
            if constexpr (OP == T_DCT3) {
                xr = rx(k);
                xi = rx(N - k);
            } else {
                xr = rx(N - 1 - k);
                xi = rx(k - 1);
            }
            return cconj(cmul(cmk<C>(xr, xi), w));
        };
        // tables at k = j + NT q as w(j) w(NT q), and dw[L - k] = e^{-i pi/4} conj(dw[k])
        const C rwj = rw[j], dwj = dw[j];
        const C e8 = cmk<C>((T)0.70710678118654752440, (T)-0.70710678118654752440);
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int k = j + S::NT * q;
            const C dwk = q == 0 ? dwj : cmul(dwj, dw[S::NT * q]);
            const C rwk = q == 0 ? rwj : cmul(rwj, rw[S::NT * q]);
            reg[q] = pack_pair(X(k, dwk), X(L - k, cmul(e8, cconj(dwk))), rwk);
        }
        if constexpr (!DIRECT) __syncthreads();
        line_fft<T, L, 1>(reg, lds, lay, j, (const double2 *)a.tw);
        C *ln = lds + b * lpadlen<T, L>();
#pragma unroll
        for (int q = 0; q < S::R; q++) ln[lpad<T>(j + S::NT * q)] = reg[q];
        __syncthreads();
        // temp = real view of ln; y pair m = (temp[m], sg * temp[N-1-m])
        constexpr T sg = OP == T_DCT3 ? (T)1 : (T)-1;
        const T sc = (T)a.scale;
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int m = j + S::NT * q;
            reg[q] = cmk<C>(lds_real<T, L>(ln, m) * sc, sg * lds_real<T, L>(ln, N - 1 - m) * sc);
        }
        __syncthreads();
        C *dst = (C *)((T *)a.out + line0 * a.out_dist);
        tile_store<T, L, LPB>(reg, lds, dst, a.out_dist / 2, L, nrows, b, j);
    }
}

// DCT/DST type 4 -------------------------------------------------------------
template <typename T, int L, int OP>
__global__ void __launch_bounds__(Sched<L>::NT *line_lpb<L>(), (rt_wpe<T, L>())) k_t4_line(RtArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int LPB = line_lpb<L>();
    constexpr int NTH = S::NT * LPB;
    constexpr int N = 2 * L;
    __shared__ C lds[line_lds<T, L, LPB>()];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const long long line0 = (long long)blockIdx.x * LPB;
    const long long rem = a.nlines - line0;
    const int nrows = rem < LPB ? (int)rem : LPB;
    const C *src = (const C *)((const T *)a.in + line0 * a.in_dist);
    const C *pre = (const C *)a.t1;
    const C *post = (const C *)a.t2;
    // tables at n = j + NT q as w(j) w(NT q): pre[n] = e^{-i pi n/N}, and
    // post[2n + e] = post[2j + e] pre[NT q]
    const C prej = pre[j];
    auto pre_at = [&](int q) -> C { return q == 0 ? prej : cmul(prej, pre[S::NT * q]); };
    C reg[S::R];
    if constexpr (S::NT >= 64) {
        // x[2n] = pair n .x, x[N-1-2n] = pair L-1-n .y, straight from global memory
        const C *xl = src + (long long)b * (a.in_dist / 2);
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int n = j + S::NT * q;
            T xe = 0, xo = 0;
            if (b < nrows) {
                xe = xl[n].x;
                xo = xl[L - 1 - n].y;
            }
            if constexpr (OP == T_DCT4) reg[q] = cmul(pre_at(q), cmk<C>(xe, xo));
            else reg[q] = cneg(cmul(pre_at(q), cmk<C>(xe, -xo)));
        }
    } else {
        stage_in_rows<T, L, LPB, NTH>(lds, src, a.in_dist / 2, L, nrows);
        __syncthreads();
        const C *line = lds + b * lpadlen<T, L>();
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int n = j + S::NT * q;
            const T xe = lds_real<T, L>(line, 2 * n), xo = lds_real<T, L>(line, N - 1 - 2 * n);
            if constexpr (OP == T_DCT4) reg[q] = cmul(pre_at(q), cmk<C>(xe, xo));
            else reg[q] = cneg(cmul(pre_at(q), cmk<C>(xe, -xo)));
        }
        __syncthreads();
    }
    LayLine<T, L> lay{b * lpadlen<T, L>()};
    line_fft<T, L, -1>(reg, lds, lay, j, (const double2 *)a.tw);
    C *ln = lds + b * lpadlen<T, L>();
#pragma unroll
    for (int q = 0; q < S::R; q++) ln[lpad<T>(j + S::NT * q)] = reg[q];
    __syncthreads();
    const T sc = (T)a.scale;
    const C post0 = post[2 * j], post1 = post[2 * j + 1];
#pragma unroll
    for (int q = 0; q < S::R; q++) {
        const int k = j + S::NT * q;
        const C tk = ln[lpad<T>(k)], tm = cconj(ln[lpad<T>(L - 1 - k)]);
        const C s = pre[S::NT * q];  // workgroup-uniform
        const C w0 = q == 0 ? post0 : cmul(post0, s), w1 = q == 0 ? post1 : cmul(post1, s);
        const C p0 = cmul(w0, tk), p1 = cmul(w1, tm);
        if constexpr (OP == T_DCT4) reg[q] = cmk<C>((T)2 * p0.x * sc, (T)2 * p1.x * sc);
        else reg[q] = cmk<C>((T)2 * p0.y * sc, (T)2 * p1.y * sc);
    }
    __syncthreads();
    C *dst = (C *)((T *)a.out + line0 * a.out_dist);
    tile_store<T, L, LPB>(reg, lds, dst, a.out_dist / 2, L, nrows, b, j);
}

// ---------------------------------------------------------------------------
// Elementwise kernels for the multi-pass (large N) real and trig paths and
// for N == 1.  One thread per output index; grid.y = line.
// ---------------------------------------------------------------------------
struct EwArgs {
    const void *in;
    void *out;
    long long nlines;
    long long in_dist, out_dist;
    int L;             // N/2
    const void *t1;
    const void *t2;
    double scale;
    double w0;
    int half;
    int op;
};

// realdft unpack: Z (L complex) -> X (L+1 complex); thread k < L.
template <typename T> __global__ void k_ew_unpack(EwArgs a)
{
    using C = Cx<T>;
    const int L = a.L;
    const long long t = (long long)blockIdx.y + (long long)blockIdx.z * gridDim.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L || t >= a.nlines) return;
    const C *z = (const C *)a.in + t * a.in_dist;
    C *x = (C *)a.out + t * a.out_dist;
    C xk, xl;
    unpack_pair(z[k], z[(L - k) & (L - 1)], ((const C *)a.t1)[k], xk, xl);
    const T sc = (T)a.scale, sk = a.half ? (T)2 * sc : sc;
    x[k] = cscale(xk, k == 0 ? sc : sk);
    if (k == 0) x[L] = cscale(xl, sc);
}

// invrealdft pack: X (L+1 complex) -> Z (L complex); thread k < L.
template <typename T> __global__ void k_ew_pack(EwArgs a)
{
    using C = Cx<T>;
    const int L = a.L;
    const long long t = (long long)blockIdx.y + (long long)blockIdx.z * gridDim.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L || t >= a.nlines) return;
    const C *x = (const C *)a.in + t * a.in_dist;
    C *z = (C *)a.out + t * a.out_dist;
    C xk = x[k], xm = x[L - k];
    const T hw = a.half ? (T)0.5 : (T)1;
    if (k == 0) {
        xk.y = 0;
        xm.y = 0;
    } else {
        xk = cscale(xk, hw);
        xm = cscale(xm, hw);
    }
    z[k] = pack_pair(xk, xm, ((const C *)a.t1)[k]);
}

// Trig pre-processing into a complex buffer of L (dct2/dst2: packed reorder,
// dct4/dst4: premul) or L+1 (dct3/dst3: pre-twiddled half spectrum) entries.
template <typename T> __global__ void k_ew_trig_pre(EwArgs a)
{
    using C = Cx<T>;
    const int L = a.L, N = 2 * L;
    const long long t = (long long)blockIdx.y + (long long)blockIdx.z * gridDim.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nlines) return;
    const T *x = (const T *)a.in + t * a.in_dist;
    C *z = (C *)a.out + t * a.out_dist;
    const T sq2 = (T)1.41421356237309504880;
    if (a.op == T_DCT2 || a.op == T_DST2) {
        if (k >= L) return;
        const T sg = a.op == T_DCT2 ? (T)1 : (T)-1;
        auto xv = [&](int i) -> T { return i < L ? x[2 * i] : sg * x[N - 1 - 2 * (i - L)]; };
        z[k] = cmk<C>(xv(2 * k), xv(2 * k + 1));
    } else if (a.op == T_DCT3 || a.op == T_DST3) {
        if (k > L) return;
        C v;
        if (k == 0) v = cmk<C>((a.op == T_DCT3 ? x[0] : x[N - 1]) * (T)a.w0, (T)0);
        else if (k == L) v = cmk<C>(sq2 * (a.op == T_DCT3 ? x[L] : x[L - 1]), (T)0);
        else {
            const T xr = a.op == T_DCT3 ? x[k] : x[N - 1 - k];
            const T xi = a.op == T_DCT3 ? x[N - k] : x[k - 1];
            v = cconj(cmul(cmk<C>(xr, xi), ((const C *)a.t1)[k]));
        }
        z[k] = v;
    } else {
        if (k >= L) return;
        const C pre = ((const C *)a.t1)[k];
        const T xe = x[2 * k], xo = x[N - 1 - 2 * k];
        z[k] = a.op == T_DCT4 ? cmul(pre, cmk<C>(xe, xo)) : cneg(cmul(pre, cmk<C>(xe, -xo)));
    }
}

// Trig post-processing from a complex buffer to N reals.
//   dct2/dst2: in = X (L+1 bins of the realdft of the reorder)
//   dct3/dst3: in = temp (N reals, the invrealdft output)
//   dct4/dst4: in = t (L complex after the forward FFT)
template <typename T> __global__ void k_ew_trig_post(EwArgs a)
{
    using C = Cx<T>;
    const int L = a.L, N = 2 * L;
    const long long t = (long long)blockIdx.y + (long long)blockIdx.z * gridDim.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L || t >= a.nlines) return;
    T *y = (T *)a.out + t * a.out_dist;
    const T sc = (T)a.scale;
    const T sq2 = (T)1.41421356237309504880;
    if (a.op == T_DCT2 || a.op == T_DST2) {
        const C *X = (const C *)a.in + t * a.in_dist;
        if (k == 0) {
            const T d = (T)2 * X[0].x * (T)a.w0, l = sq2 * X[L].x;
            if (a.op == T_DCT2) { y[0] = d * sc; y[L] = l * sc; }
            else { y[N - 1] = d * sc; y[L - 1] = l * sc; }
            return;
        }
        const C p = cmul(X[k], ((const C *)a.t1)[k]);
        if (a.op == T_DCT2) { y[k] = (T)2 * p.x * sc; y[N - k] = (T)-2 * p.y * sc; }
        else { y[k - 1] = (T)-2 * p.y * sc; y[N - k - 1] = (T)2 * p.x * sc; }
    } else if (a.op == T_DCT3 || a.op == T_DST3) {
        const T *tmp = (const T *)a.in + t * a.in_dist;
        const T sg = a.op == T_DCT3 ? (T)1 : (T)-1;
        y[2 * k] = tmp[k] * sc;
        y[N - 1 - 2 * k] = sg * tmp[L + k] * sc;
    } else {
        const C *tt = (const C *)a.in + t * a.in_dist;
        const C *post = (const C *)a.t2;
        const C p0 = cmul(post[2 * k], tt[k]), p1 = cmul(post[2 * k + 1], cconj(tt[L - 1 - k]));
        if (a.op == T_DCT4) { y[2 * k] = (T)2 * p0.x * sc; y[2 * k + 1] = (T)2 * p1.x * sc; }
        else { y[2 * k] = (T)2 * p0.y * sc; y[2 * k + 1] = (T)2 * p1.y * sc; }
    }
}

// N == 1 for every op: y = f * x (complex or real), one thread per line.
template <typename T> __global__ void k_ew_n1(EwArgs a)
{
    using C = Cx<T>;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nlines) return;
    const T f = (T)a.scale;
    switch (a.op) {
    case 0: {  // complex copy * scale
        ((C *)a.out)[t * a.out_dist] = cscale(((const C *)a.in)[t * a.in_dist], f);
        break;
    }
    case 1: {  // real -> complex (realdft N=1)
        ((C *)a.out)[t * a.out_dist] = cmk<C>(((const T *)a.in)[t * a.in_dist] * f, (T)0);
        break;
    }
    case 2: {  // complex -> real (invrealdft N=1)
        ((T *)a.out)[t * a.out_dist] = ((const C *)a.in)[t * a.in_dist].x * f;
        break;
    }
    default: {  // real -> real
        ((T *)a.out)[t * a.out_dist] = ((const T *)a.in)[t * a.in_dist] * f;
        break;
    }
    }
}

}  // namespace nfft

namespace nfft {

// ---------------------------------------------------------------------------
// k_rowT_r2c: the row pass of the four-step for the N/2-point complex FFT of
// a realdft, with the realdft unpack fused in.  Z (Lc = n1 * L points per
// transform, row r = L contiguous points) -> row FFTs -> X[k] for k <= Lc
// (FFTW layout, Lc + 1 complex).  Output k = k2 * n1 + r needs
// Z[Lc - k], which lies in row n1 - r (column L - 1 - k2), or in row 0
// (column -k2 mod L) for r = 0 and in row n1/2 itself for r = n1/2: so
// each workgroup transforms row pairs (p, n1 - p), with pair 0 = rows
// (0, n1/2), and unpacks from its own LDS.  Replaces the separate
// elementwise unpack pass (one full read + write of the spectrum).
// ---------------------------------------------------------------------------
struct RowR2CArgs {
    const void *in;    // Z: Lc complex per transform
    void *out;         // X: out_dist complex per transform (>= Lc + 1)
    int nrows;         // n1 (even)
    long long in_dist, out_dist;
    long long ntrans;
    const void *tw;    // stage table of L
    const void *w;     // e^{-2 pi i k / (2 Lc)}, k < Lc
    double scale;
    int dbl;           // adjoint of invrealdft: bins 0 < k < Lc times 2
    int lpb;           // rows per workgroup (0: the default of rowT_lpb)
};

template <typename T, int L, int LPB>
__global__ void __launch_bounds__(Sched<L>::NT *LPB) k_rowT_r2c(RowR2CArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    static_assert(LPB % 2 == 0, "row pairs");
    constexpr int NTH = S::NT * LPB, PADL = lpadlen<T, L>();
    // unpack twiddle w[k2 n1 + r] = w[k2 n1] w[r]: the L column factors and
    // the LPB row factors of this workgroup in LDS (where they fit beside the
    // tile) instead of one global load per output point
    constexpr bool WLDS = (size_t)(LPB * PADL + L + LPB) * sizeof(C) <= 160 * 1024;
    // stage twiddles from an LDS copy too where that still fits
    constexpr int TWN = Sched<L>::TWN > 0 ? Sched<L>::TWN : 1;
    constexpr bool TWL = S::S > 1 && WLDS && (size_t)(LPB * PADL + L + LPB + TWN) * sizeof(C) <= 160 * 1024;
    __shared__ C lds[LPB * PADL];
    __shared__ C wtab[WLDS ? L + LPB : 1];
    __shared__ C ltw[TWL ? TWN : 1];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const long long lin = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    const long long lg = xcd_logical(lin, (long long)gridDim.x * gridDim.y * gridDim.z);
    const int pt = (int)(lg % gridDim.x);
    const long long t = lg / gridDim.x;
    if (t >= a.ntrans) return;  // uniform per workgroup
    const int n1 = a.nrows, npairs = n1 / 2;
    // row of a slot (-1: past the last pair)
    auto row_of = [&](int slot) {
        const int p = pt * (LPB / 2) + slot / 2;
        if (p >= npairs) return -1;
        return (slot & 1) ? (p == 0 ? npairs : n1 - p) : p;
    };
    const C *src = (const C *)a.in + t * a.in_dist;
    const C *w = (const C *)a.w;
    if constexpr (TWL) stage_tw_to_lds<T, L>(ltw, (const double2 *)a.tw);
    if constexpr (WLDS) {
        for (int i = threadIdx.x; i < L; i += NTH) wtab[i] = w[(long long)i * n1];
        if (threadIdx.x < LPB) {
            const int r = row_of(threadIdx.x);
            wtab[L + threadIdx.x] = w[r >= 0 ? r : 0];
        }
    }
    C reg[S::R];
    if constexpr (rows_direct<T, L>()) {
        const int r = row_of(b);
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = r >= 0 ? src[(long long)r * L + j + S::NT * q] : czero<T>();
    } else {
        for (int idx = threadIdx.x; idx < LPB * L; idx += NTH) {
            const int slot = idx / L, p = idx & (L - 1);
            const int r = row_of(slot);
            lds[slot * PADL + lpad<T>(p)] = r >= 0 ? src[(long long)r * L + p] : czero<T>();
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = lds[b * PADL + lpad<T>(j + S::NT * q)];
        __syncthreads();
    }
    LayLine<T, L> lay{b * PADL};
    if constexpr (TWL) line_fft_tw<T, L, -1>(reg, lds, lay, j, StageTwLds<T>{ltw});
    else line_fft<T, L, -1>(reg, lds, lay, j, (const double2 *)a.tw);
#pragma unroll
    for (int q = 0; q < S::R; q++) lds[b * PADL + lpad<T>(j + S::NT * q)] = reg[q];
    __syncthreads();
    const long long Lc = (long long)n1 * L;
    C *dst = (C *)a.out + t * a.out_dist;
    const T sc = (T)a.scale, sk = a.dbl ? (T)2 * sc : sc;
    // Output k = k2 n1 + r.  Thread (slot, k0) keeps its row slot and walks
    // k2 = k0 + KS i: row, partner slot and row factor are per-thread
    // constants, and consecutive lanes store consecutive rows of one k2
    // (runs of LPB / 2 points on each side of the pair).
    constexpr int KS = NTH / LPB;
    static_assert(NTH % LPB == 0 && L % KS == 0, "output sweep");
    const int slot = threadIdx.x % LPB, k0 = threadIdx.x / LPB;
    const int r = row_of(slot);
    if (r < 0) return;  // past the last pair (no barrier follows)
    // partner Z[Lc - k]: row n1 - r (slot ^ 1) at column L - 1 - k2; row 0
    // pairs with itself at column -k2 mod L, row n1/2 at column L - 1 - k2
    const int ps = (r == 0 || r == npairs) ? slot : (slot ^ 1);
    const int pmask = r == 0 ? L - 1 : -1, pbase = r == 0 ? L : L - 1;
    C wr = czero<T>();
    if constexpr (WLDS) wr = wtab[L + slot];
    const C *zrow = lds + slot * PADL, *prow = lds + ps * PADL;
    const bool fast = off_fits<C>(Lc + 1);
    const auto rd = buf_rsrc(dst);
    const uint32_t voff = (uint32_t)((long long)k0 * n1 + r) * (uint32_t)sizeof(C);
    const uint32_t ist = (uint32_t)((long long)KS * n1) * (uint32_t)sizeof(C);
#pragma unroll 4
    for (int i = 0; i < L / KS; i++) {
        const int k2 = k0 + KS * i;
        const C zk = zrow[lpad<T>(k2)], zm = prow[lpad<T>((pbase - k2) & pmask)];
        C wk;
        if constexpr (WLDS) wk = cmul(wtab[k2], wr);
        else wk = w[(long long)k2 * n1 + r];
        C xk, xl;
        unpack_pair(zk, zm, wk, xk, xl);
        const bool k_is_0 = r == 0 && k2 == 0;
        xk = cscale(xk, k_is_0 ? sc : sk);
        if (fast) buf_st(rd, voff, (uint32_t)i * ist, xk);
        else dst[(long long)k2 * n1 + r] = xk;
        if (k_is_0) dst[Lc] = cscale(xl, sc);
    }
}

}  // namespace nfft
