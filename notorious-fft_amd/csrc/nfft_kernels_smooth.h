// nfft_kernels_smooth.h -- mixed-radix (2, 3, 4, 5, 7, 8) transforms for
// lengths N = 2^a 3^b 5^c 7^d.
//
// The reference sends every length that is not a power of two through
// Bluestein (src/05_algorithms.h:292-344: three power-of-two FFTs of
// M >= 2N - 1 plus chirp passes).  Here a 7-smooth length is transformed
// directly by a self-sorting Stockham sequence of radix-r stages held in LDS
// (ping-pong buffers), one runtime schedule per length: stage s with radix r
// and Ns = r_0 ... r_{s-1} maps butterfly j (k = j mod Ns) from inputs
// x[j + t L/r] to outputs y[(j / Ns) Ns r + k + t Ns] after the twiddles
// w_{Ns r}^{t k}, read from a table of e^{-2 pi i q / L} generated in FP64.
// Each stage dispatches once (uniformly) on its radix to a compile-time
// butterfly, so the per-thread values stay in registers.
//
// Used for whole lines (k_smooth_line, several lines per workgroup) and as
// the column / row passes of a four-step for lengths above the LDS limit.
#pragma once
#include "nfft_kernels.h"
#include "nfft_kernels_persist.h"  // coh_ld / coh_st

namespace nfft {

constexpr int kSmoothMaxStages = 24;
// Largest radix of the standalone mixed-radix kernels (k_smooth_line,
// k_smooth_col, k_smooth_rowT): 10 in FP32 (N = 100 as 10 x 10: 0.78 -> 0.70
// ms per 2^27 points, N = 10^4 1.85 -> 1.61 ms), 8 in FP64, where the radix-10
// butterfly's registers (92 -> 118 VGPRs) cost more occupancy than the stages
// it saves (N = 100 1.21 -> 1.26 ms).
template <typename T> __host__ __device__ constexpr int smooth_radix() { return sizeof(T) == 4 ? 10 : 8; }

// Stage schedule of a 7-smooth length (radices in execution order).
struct SmoothPlan {
    int L;
    int ns;
    unsigned char radix[kSmoothMaxStages];
};

// Odd-radix DFTs from the symmetric form: with a_k = x_k + x_{P-k},
// b_k = x_k - x_{P-k}, y_m = x_0 + sum_k cos(2 pi k m / P) a_k
//                                -/+ i sum_k sin(2 pi k m / P) b_k (forward / inverse),
// y_{P-m} with the opposite sign of the imaginary part.
template <int P> struct OddTab;
template <> struct OddTab<3> {
    static constexpr double c[1][1] = {{-0.5}};
    static constexpr double s[1][1] = {{0.86602540378443864676}};
};
template <> struct OddTab<5> {
    static constexpr double c[2][2] = {{0.30901699437494742410, -0.80901699437494742410},
                                       {-0.80901699437494742410, 0.30901699437494742410}};
    static constexpr double s[2][2] = {{0.95105651629515357212, 0.58778525229247312917},
                                       {0.58778525229247312917, -0.95105651629515357212}};
};
template <> struct OddTab<7> {
    // c[m-1][k-1] = cos(2 pi k m / 7), s[m-1][k-1] = sin(2 pi k m / 7)
    static constexpr double c[3][3] = {
        {0.62348980185873353053, -0.22252093395631440429, -0.90096886790241912624},
        {-0.22252093395631440429, -0.90096886790241912624, 0.62348980185873353053},
        {-0.90096886790241912624, 0.62348980185873353053, -0.22252093395631440429}};
    static constexpr double s[3][3] = {
        {0.78183148246802980871, 0.97492791218182360702, 0.43388373911755812048},
        {0.97492791218182360702, -0.43388373911755812048, -0.78183148246802980871},
        {0.43388373911755812048, -0.78183148246802980871, 0.97492791218182360702}};
};

template <int P, int DIR, typename C> __device__ __forceinline__ void dft_odd(C *v)
{
    using R = decltype(C::x);
    constexpr int H = P / 2;
    C a[H], b[H];
#pragma unroll
    for (int k = 0; k < H; k++) {
        a[k] = cadd(v[k + 1], v[P - 1 - k]);
        b[k] = csub(v[k + 1], v[P - 1 - k]);
    }
    C y0 = v[0];
#pragma unroll
    for (int k = 0; k < H; k++) y0 = cadd(y0, a[k]);
#pragma unroll
    for (int m = 0; m < H; m++) {
        C t = v[0], u = cmk<C>((R)0, (R)0);
#pragma unroll
        for (int k = 0; k < H; k++) {
            const R cc = (R)OddTab<P>::c[m][k], ss = (R)OddTab<P>::s[m][k];
            t = cmk<C>(t.x + cc * a[k].x, t.y + cc * a[k].y);
            u = cmk<C>(u.x + ss * b[k].x, u.y + ss * b[k].y);
        }
        // forward: y_m = t - i u, y_{P-m} = t + i u; inverse swaps them
        const C iu = cmk<C>(-u.y, u.x);
        if constexpr (DIR < 0) {
            v[m + 1] = csub(t, iu);
            v[P - 1 - m] = cadd(t, iu);
        } else {
            v[m + 1] = cadd(t, iu);
            v[P - 1 - m] = csub(t, iu);
        }
    }
    v[0] = y0;
}

// n / d for 0 <= n < 2^19 and 1 <= d < 2^24 from inv = rcp_up(d): the FP32
// reciprocal raised by 2^-20 (relative) puts the FP32 product n * inv in
// [n/d, n/d + n 2^-19 / d), below the next integer (at least 1/d above
// n/d), so truncation gives the quotient with no correction step.
__device__ __forceinline__ float rcp_up(int d)
{
    return (1.0f / (float)d) * 1.00000095367431640625f;
}
__device__ __forceinline__ int fdiv(int n, float inv)
{
    return (int)((float)n * inv);
}

// cos and sin of 2 pi e / P, e = 0 .. P-1 (FP64, correctly rounded), for the
// inner twiddles of the two-level codelets below.
template <int P> struct RootTab;
template <> struct RootTab<6> {
    static constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
    static constexpr double s[6] = {0.0, 0.8660254037844386, 0.8660254037844386, 0.0, -0.8660254037844386, -0.8660254037844386};
};
template <> struct RootTab<9> {
    static constexpr double c[9] = {1.0, 0.766044443118978, 0.17364817766693036, -0.5, -0.9396926207859084, -0.9396926207859084, -0.5, 0.17364817766693036, 0.766044443118978};
    static constexpr double s[9] = {0.0, 0.6427876096865394, 0.984807753012208, 0.8660254037844386, 0.3420201433256687, -0.3420201433256687, -0.8660254037844386, -0.984807753012208, -0.6427876096865394};
};
template <> struct RootTab<10> {
    static constexpr double c[10] = {1.0, 0.8090169943749475, 0.30901699437494745, -0.30901699437494745, -0.8090169943749475, -1.0, -0.8090169943749475, -0.30901699437494745, 0.30901699437494745, 0.8090169943749475};
    static constexpr double s[10] = {0.0, 0.5877852522924731, 0.9510565162951535, 0.9510565162951535, 0.5877852522924731, 0.0, -0.5877852522924731, -0.9510565162951535, -0.9510565162951535, -0.5877852522924731};
};
template <> struct RootTab<12> {
    static constexpr double c[12] = {1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5, 0.0, 0.5, 0.8660254037844386};
    static constexpr double s[12] = {0.0, 0.5, 0.8660254037844386, 1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5};
};
template <> struct RootTab<15> {
    static constexpr double c[15] = {1.0, 0.9135454576426009, 0.6691306063588582, 0.30901699437494745, -0.10452846326765347, -0.5, -0.8090169943749475, -0.9781476007338057, -0.9781476007338057, -0.8090169943749475, -0.5, -0.10452846326765347, 0.30901699437494745, 0.6691306063588582, 0.9135454576426009};
    static constexpr double s[15] = {0.0, 0.4067366430758002, 0.7431448254773942, 0.9510565162951535, 0.9945218953682733, 0.8660254037844386, 0.5877852522924731, 0.20791169081775934, -0.20791169081775934, -0.5877852522924731, -0.8660254037844386, -0.9945218953682733, -0.9510565162951535, -0.7431448254773942, -0.4067366430758002};
};
template <> struct RootTab<20> {
    static constexpr double c[20] = {1.0, 0.9510565162951535, 0.8090169943749475, 0.5877852522924731, 0.30901699437494745, 0.0, -0.30901699437494745, -0.5877852522924731, -0.8090169943749475, -0.9510565162951535, -1.0, -0.9510565162951535, -0.8090169943749475, -0.5877852522924731, -0.30901699437494745, 0.0, 0.30901699437494745, 0.5877852522924731, 0.8090169943749475, 0.9510565162951535};
    static constexpr double s[20] = {0.0, 0.30901699437494745, 0.5877852522924731, 0.8090169943749475, 0.9510565162951535, 1.0, 0.9510565162951535, 0.8090169943749475, 0.5877852522924731, 0.30901699437494745, 0.0, -0.30901699437494745, -0.5877852522924731, -0.8090169943749475, -0.9510565162951535, -1.0, -0.9510565162951535, -0.8090169943749475, -0.5877852522924731, -0.30901699437494745};
};
template <> struct RootTab<25> {
    static constexpr double c[25] = {1.0, 0.9685831611286311, 0.8763066800438636, 0.7289686274214116, 0.5358267949789967, 0.30901699437494745, 0.06279051952931337, -0.18738131458572463, -0.42577929156507266, -0.6374239897486897, -0.8090169943749475, -0.9297764858882515, -0.9921147013144779, -0.9921147013144779, -0.9297764858882515, -0.8090169943749475, -0.6374239897486897, -0.42577929156507266, -0.18738131458572463, 0.06279051952931337, 0.30901699437494745, 0.5358267949789967, 0.7289686274214116, 0.8763066800438636, 0.9685831611286311};
    static constexpr double s[25] = {0.0, 0.2486898871648548, 0.48175367410171527, 0.6845471059286887, 0.8443279255020151, 0.9510565162951535, 0.9980267284282716, 0.9822872507286887, 0.9048270524660196, 0.7705132427757893, 0.5877852522924731, 0.368124552684678, 0.12533323356430426, -0.12533323356430426, -0.368124552684678, -0.5877852522924731, -0.7705132427757893, -0.9048270524660196, -0.9822872507286887, -0.9980267284282716, -0.9510565162951535, -0.8443279255020151, -0.6845471059286887, -0.48175367410171527, -0.2486898871648548};
};

// a * e^{DIR 2 pi i E / P}, E and P compile-time.
template <int DIR, int P, int E, typename C> __device__ __forceinline__ C twp(C a)
{
    constexpr int e = E % P;
    using R = decltype(C::x);
    if constexpr (e == 0) return a;
    else {
        constexpr double c = RootTab<P>::c[e];
        constexpr double s = DIR * RootTab<P>::s[e];
        return cmk<C>(a.x * (R)c - a.y * (R)s, a.x * (R)s + a.y * (R)c);
    }
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dft_any(C *v);

// A DFT of P = R1 R2 points in registers, natural order in and out: with
// n = R2 n1 + n2 and k = k1 + R1 k2, R2 DFTs of length R1 (over n1), the
// inner twiddles w_P^{n2 k1}, then R1 DFTs of length R2 (over n2).  One stage
// of radix P replaces a stage of R1 and a stage of R2: half the LDS round
// trips, index arithmetic and table twiddles of the two.
template <int R1, int R2, int DIR, typename C> __device__ __forceinline__ void dft_two(C *v)
{
    C a[R2][R1];
    sfor<0, R2>([&](auto n2c) {
        constexpr int n2 = decltype(n2c)::value;
        C x[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) x[n1] = v[R2 * n1 + n2];
        dft_any<R1, DIR>(x);
        sfor<0, R1>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            a[n2][k1] = twp<DIR, R1 * R2, n2 * k1>(x[k1]);
        });
    });
    sfor<0, R1>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        C y[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) y[n2] = a[n2][k1];
        dft_any<R2, DIR>(y);
#pragma unroll
        for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = y[k2];
    });
}

// Radices of the mixed-radix stages: 2, 3, 4, 5, 7, 8, 16, and the two-level
// 6 = 2 x 3, 9 = 3 x 3, 10 = 2 x 5, 12 = 4 x 3, 15 = 3 x 5, 20 = 4 x 5, 25 = 5 x 5.
template <int R, int DIR, typename C> __device__ __forceinline__ void dft_any(C *v)
{
    if constexpr (R == 3 || R == 5 || R == 7) dft_odd<R, DIR>(v);
    else if constexpr (R == 6) dft_two<2, 3, DIR>(v);
    else if constexpr (R == 9) dft_two<3, 3, DIR>(v);
    else if constexpr (R == 10) dft_two<2, 5, DIR>(v);
    else if constexpr (R == 12) dft_two<4, 3, DIR>(v);
    else if constexpr (R == 15) dft_two<3, 5, DIR>(v);
    else if constexpr (R == 20) dft_two<4, 5, DIR>(v);
    else if constexpr (R == 25) dft_two<5, 5, DIR>(v);
    else dftR<R, DIR>(v);
}

// One Stockham stage of radix R over G lines of length L (line g at
// src/dst + g * L), all threads of the workgroup sharing the G*L/R butterflies.
// TW: Ns > 1 (the first stage, Ns = 1, has no twiddles).
template <typename T, int R, int DIR, bool TW>
__device__ __forceinline__ void smooth_stage(const Cx<T> *src, Cx<T> *dst, int G, int L, int Ns, const Cx<T> *tw)
{
    using C = Cx<T>;
    const int nb = L / R;
    const int tws = TW ? L / (Ns * R) : 0;  // w_{Ns R}^{t k} = tw[t k tws], t k tws < L
    const float inb = rcp_up(nb), ins = rcp_up(Ns);
    for (int idx = threadIdx.x; idx < G * nb; idx += blockDim.x) {
        const int g = fdiv(idx, inb), j = idx - __mul24(g, nb);
        int jq = j, k = 0;
        if constexpr (TW) {
            jq = fdiv(j, ins);
            k = j - __mul24(jq, Ns);
        }
        const C *x = src + __mul24(g, L) + j;
        C v[R];
#pragma unroll
        for (int t = 0; t < R; t++) v[t] = x[t * nb];
        if constexpr (TW) {
            const unsigned kt = (unsigned)__mul24(k, tws);
#pragma unroll
            for (int t = 1; t < R; t++) v[t] = ctw<DIR>(v[t], tw[t * kt]);
        }
        dft_any<R, DIR>(v);
        C *y = dst + __mul24(g, L) + __mul24(jq, Ns * R) + k;
#pragma unroll
        for (int t = 0; t < R; t++) y[t * Ns] = v[t];
    }
}

template <typename T, int R, int DIR>
__device__ __forceinline__ void smooth_stage_any(const Cx<T> *src, Cx<T> *dst, int G, int L, int Ns,
                                                 const Cx<T> *tw)
{
    if (Ns == 1) smooth_stage<T, R, DIR, false>(src, dst, G, L, 1, tw);
    else smooth_stage<T, R, DIR, true>(src, dst, G, L, Ns, tw);
}

// All stages of G lines held in LDS buffer `a` (buffer `b` is the other half
// of the ping-pong); returns the buffer holding the result.  All threads of
// the workgroup must call it.  MAXR is the largest radix compiled in: 8 (2-8),
// 10 (also 6, 9, 10), 25 (also 12, 15, 16, 20, 25); the schedule must come
// from smooth_factor(..., maxr) with the same bound.  The larger butterflies
// raise the register allocation of the whole switch (FP32: 59-62 VGPRs at 8,
// 103 at 25; FP64: 78-92 and 170), which kernels without a fixed register
// budget pay for in occupancy.
template <typename T, int DIR, int MAXR = 8>
__device__ __forceinline__ Cx<T> *smooth_lines(Cx<T> *a, Cx<T> *b, int G, const SmoothPlan &sp, const Cx<T> *tw)
{
    Cx<T> *src = a, *dst = b;
    int Ns = 1;
    for (int s = 0; s < sp.ns; s++) {
        const int r = sp.radix[s];
        switch (r) {
        case 2: smooth_stage_any<T, 2, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 3: smooth_stage_any<T, 3, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 4: smooth_stage_any<T, 4, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 5: smooth_stage_any<T, 5, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 7: smooth_stage_any<T, 7, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 8: smooth_stage_any<T, 8, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 6: if constexpr (MAXR >= 10) smooth_stage_any<T, 6, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 9: if constexpr (MAXR >= 10) smooth_stage_any<T, 9, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 10: if constexpr (MAXR >= 10) smooth_stage_any<T, 10, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 12: if constexpr (MAXR >= 25) smooth_stage_any<T, 12, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 15: if constexpr (MAXR >= 25) smooth_stage_any<T, 15, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 16: if constexpr (MAXR >= 25) smooth_stage_any<T, 16, DIR>(src, dst, G, sp.L, Ns, tw); break;
        case 20: if constexpr (MAXR >= 25) smooth_stage_any<T, 20, DIR>(src, dst, G, sp.L, Ns, tw); break;
        default: if constexpr (MAXR >= 25) smooth_stage_any<T, 25, DIR>(src, dst, G, sp.L, Ns, tw); break;
        }
        __syncthreads();
        Cx<T> *t = src;
        src = dst;
        dst = t;
        Ns *= r;
    }
    return src;
}

// ---------------------------------------------------------------------------
// k_smooth_line: G lines of a 7-smooth length L per workgroup, contiguous
// lines (line l at in + l * in_dist), out scaled by `scale`.
// ---------------------------------------------------------------------------
struct SmoothArgs {
    const void *in;
    void *out;
    long long nlines;
    long long in_dist, out_dist;
    int G;            // lines per workgroup
    SmoothPlan sp;
    const void *tw;   // L entries e^{-2 pi i q / L}
    double scale;
};

// TWL: the L twiddles are copied into LDS after the two line buffers (the
// launcher gives up lines per workgroup for them, keeping the LDS size).
template <typename T, int DIR, bool TWL = false> __global__ void __launch_bounds__(256) k_smooth_line(SmoothArgs a)
{
    using C = Cx<T>;
    extern __shared__ char smem[];
    C *A = (C *)smem;
    const int L = a.sp.L, G = a.G;
    C *B = A + G * L;
    const long long l0 = (long long)blockIdx.x * G;
    const int g_here = (int)((a.nlines - l0) < G ? (a.nlines - l0) : G);
    const C *src = (const C *)a.in;
    const C *twp = (const C *)a.tw;
    if constexpr (TWL) {
        C *tl = B + G * L;
        for (int i = threadIdx.x; i < L; i += blockDim.x) tl[i] = twp[i];
        twp = tl;
    }
    const float iL = rcp_up(L);
    for (int idx = threadIdx.x; idx < G * L; idx += blockDim.x) {
        const int g = fdiv(idx, iL), p = idx - __mul24(g, L);
        A[idx] = g < g_here ? src[(l0 + g) * a.in_dist + p] : cmk<C>((T)0, (T)0);
    }
    __syncthreads();
    const C *res = smooth_lines<T, DIR, smooth_radix<T>()>(A, B, G, a.sp, twp);
    C *dst = (C *)a.out;
    const T sc = (T)a.scale;
    for (int idx = threadIdx.x; idx < G * L; idx += blockDim.x) {
        const int g = fdiv(idx, iL), p = idx - __mul24(g, L);
        if (g < g_here) {
            C v = res[idx];
            if (sc != (T)1) v = cscale(v, sc);
            dst[(l0 + g) * a.out_dist + p] = v;
        }
    }
}

}  // namespace nfft

namespace nfft {

// ---------------------------------------------------------------------------
// Mixed-radix four-step tiles for N = N1 x N2 (both 7-smooth): element n of
// a transform is x[n1 * N2 + n2].
//   column tile: columns [c0, c0 + w) of the N1 x N2 view, FFT_N1 down each
//     column, times w_N^{k1 c}, stored at scr[k1 * N2 + c];
//   row tile: rows [r0, r0 + w) of scr (N2 contiguous points each), FFT_N2,
//     stored transposed at out[k2 * N1 + r] (natural order of the result).
// COH selects agent-scope (sc1) accesses for the intermediate (tiles that
// other workgroups consume within the same launch).
// ---------------------------------------------------------------------------
template <typename T, int DIR, bool COH, int MAXR = 8>
__device__ __forceinline__ void smooth_col_tile(const Cx<T> *in, Cx<T> *scr, const SmoothPlan &sp1, int N2, int c0, int w,
                                const Cx<T> *tw1, const Tw2 &tw2, Cx<T> *A, Cx<T> *B)
{
    using C = Cx<T>;
    const int N1 = sp1.L;
    const float iw = rcp_up(w);
    for (int idx = threadIdx.x; idx < w * N1; idx += blockDim.x) {
        const int p = fdiv(idx, iw), b = idx - __mul24(p, w);
        A[b * N1 + p] = in[(long long)p * N2 + c0 + b];
    }
    __syncthreads();
    const C *res = smooth_lines<T, DIR, MAXR>(A, B, w, sp1, tw1);
    for (int idx = threadIdx.x; idx < w * N1; idx += blockDim.x) {
        const int k1 = fdiv(idx, iw), b = idx - __mul24(k1, w);
        const uint32_t e = (uint32_t)k1 * (uint32_t)(c0 + b);
        const C v = ctw64<DIR>(res[b * N1 + k1], tw2_get(tw2, e));
        const long long di = (long long)k1 * N2 + c0 + b;
        if constexpr (COH) coh_st(scr, (uint32_t)di, v);
        else scr[di] = v;
    }
    __syncthreads();
}

template <typename T, int DIR, bool COH, int MAXR = 8>
__device__ __forceinline__ void smooth_row_tile(const Cx<T> *scr, Cx<T> *out, const SmoothPlan &sp2, int N1, int r0, int w,
                                const Cx<T> *tw2t, T scale, Cx<T> *A, Cx<T> *B)
{
    using C = Cx<T>;
    const int N2 = sp2.L;
    for (int idx = threadIdx.x; idx < w * N2; idx += blockDim.x) {
        const long long si = (long long)r0 * N2 + idx;
        if constexpr (COH) A[idx] = coh_ld(scr, (uint32_t)si);
        else A[idx] = scr[si];
    }
    __syncthreads();
    const C *res = smooth_lines<T, DIR, MAXR>(A, B, w, sp2, tw2t);
    const float iw = rcp_up(w);
    for (int idx = threadIdx.x; idx < w * N2; idx += blockDim.x) {
        const int k2 = fdiv(idx, iw), b = idx - __mul24(k2, w);
        C v = res[b * N2 + k2];
        if (scale != (T)1) v = cscale(v, scale);
        out[(long long)k2 * N1 + r0 + b] = v;
    }
    __syncthreads();
}

// Standalone passes over a batch (grid.x: tiles of one transform, grid.y x
// grid.z: transforms), the intermediate in a scratch of N per transform.
struct Smooth4Args {
    const void *in;
    void *out;
    void *scratch;
    long long ntrans;
    long long in_dist, out_dist;
    SmoothPlan sp1, sp2;      // N1, N2
    int w1, w2;               // columns / rows per tile
    const void *tw1, *tw2t;   // e^{-2 pi i q / N1}, e^{-2 pi i q / N2}
    Tw2 tw2;                  // e^{-2 pi i e / N}
    double scale;
};

template <typename T, int DIR> __global__ void __launch_bounds__(256) k_smooth_col(Smooth4Args a)
{
    using C = Cx<T>;
    extern __shared__ char smem[];
    C *A = (C *)smem;
    C *B = A + a.w1 * a.sp1.L;
    const long long t = (long long)blockIdx.y + (long long)gridDim.y * blockIdx.z;
    if (t >= a.ntrans) return;  // uniform per workgroup
    const int N2 = a.sp2.L;
    const long long N = (long long)a.sp1.L * N2;
    const int c0 = blockIdx.x * a.w1;
    const int w = (N2 - c0) < a.w1 ? (N2 - c0) : a.w1;
    smooth_col_tile<T, DIR, false, smooth_radix<T>()>((const C *)a.in + t * a.in_dist, (C *)a.scratch + t * N, a.sp1, N2, c0, w,
                                   (const C *)a.tw1, a.tw2, A, B);
}

template <typename T, int DIR> __global__ void __launch_bounds__(256) k_smooth_rowT(Smooth4Args a)
{
    using C = Cx<T>;
    extern __shared__ char smem[];
    C *A = (C *)smem;
    C *B = A + a.w2 * a.sp2.L;
    const long long t = (long long)blockIdx.y + (long long)gridDim.y * blockIdx.z;
    if (t >= a.ntrans) return;
    const int N1 = a.sp1.L;
    const long long N = (long long)N1 * a.sp2.L;
    const int r0 = blockIdx.x * a.w2;
    const int w = (N1 - r0) < a.w2 ? (N1 - r0) : a.w2;
    smooth_row_tile<T, DIR, false, smooth_radix<T>()>((const C *)a.scratch + t * N, (C *)a.out + t * a.out_dist, a.sp2, N1, r0, w,
                                   (const C *)a.tw2t, (T)a.scale, A, B);
}

}  // namespace nfft
