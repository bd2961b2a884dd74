// nfft_device.h -- device-side building blocks for the CDNA4 (gfx950) FFT
// kernels: complex helpers, register radix-2/4/8/16 codelets and the
// LDS-exchanged Stockham "line" FFT that every transform kernel is built on.
//
// A line of L = 2^p points is transformed by NT = L/16 threads (L >= 16), each
// holding 16 points in VGPRs: reg[q] is point (j + NT*q) of the line, both on
// entry (input order) and on exit (natural output order).  Stages are radix-16
// except the first (radix L/16^(S-1) in {2,4,8,16}); between stages the line
// is exchanged through LDS in Stockham (self-sorting) order, so no
// bit-reversal pass exists.  Stage twiddles come from a per-L table generated
// in FP64 on the device (one 16-entry row per butterfly index k).
//
// The algorithm replaces the reference's depth-first split-radix DIF
// (src/05_algorithms.h:358-717) and radix-2 DIT (src/05_algorithms.h:28-287):
// same transform, different factorisation chosen for 64-wide wavefronts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace nfft {

template <typename T> struct CxT;
template <> struct CxT<float> { using V = float2; };
template <> struct CxT<double> { using V = double2; };
template <typename T> using Cx = typename CxT<T>::V;

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b)
{
    C r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cneg(C a) { return cmk<C>(-a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmk<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b)
{
    return cmk<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// a * w where w is a forward twiddle; DIR > 0 uses conj(w).
template <int DIR, typename C> __device__ __forceinline__ C ctw(C a, C w)
{
    if constexpr (DIR < 0) return cmul(a, w);
    else return cmulc(a, w);
}
// multiply by DIR*i  (DIR = -1: forward e^{-i pi/2}; +1: inverse)
template <int DIR, typename C> __device__ __forceinline__ C mulj(C a)
{
    if constexpr (DIR < 0) return cmk<C>(a.y, -a.x);
    else return cmk<C>(-a.y, a.x);
}

// cos(pi*m/8), sin(pi*m/8), m = 0..15
constexpr double kCos16[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
constexpr double kSin16[16] = {
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

// a * e^{DIR * 2 pi i * E / R}, E and R compile-time, R divides 16.
template <int DIR, int R, int E, typename C> __device__ __forceinline__ C twc(C a)
{
    constexpr int e = ((E % R) + R) % R;
    constexpr int m = e * (16 / R);  // angle = pi*m/8
    using T = decltype(C::x);
    if constexpr (m == 0) return a;
    else if constexpr (m == 4) return mulj<DIR>(a);
    else if constexpr (m == 8) return cneg(a);
    else if constexpr (m == 12) return mulj<-DIR>(a);
    else {
        constexpr double c = kCos16[m];
        constexpr double s = DIR * kSin16[m];
        if constexpr (m == 2 || m == 6 || m == 10 || m == 14) {
            // |c| == |s| == 1/sqrt2: (a.x c - a.y s, a.x s + a.y c) with one product each
            constexpr T h = (T)0.70710678118654752440;
            constexpr T sc = c > 0 ? 1 : -1, ss = s > 0 ? 1 : -1;
            return cmk<C>(h * (sc * a.x - ss * a.y), h * (ss * a.x + sc * a.y));
        } else {
            return cmk<C>(a.x * (T)c - a.y * (T)s, a.x * (T)s + a.y * (T)c);
        }
    }
}

template <int I, int N, typename F> __device__ __forceinline__ void sfor(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

// ---------------------------------------------------------------------------
// Register codelets: v[0..R-1] natural order in and out, X_k = sum v_n w^{nk},
// w = e^{DIR 2 pi i / R}.
// ---------------------------------------------------------------------------
template <int DIR, typename C> __device__ __forceinline__ void dft2(C &a, C &b)
{
    C t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int DIR, typename C> __device__ __forceinline__ void dft4(C &a0, C &a1, C &a2, C &a3)
{
    C s0 = cadd(a0, a2), d0 = csub(a0, a2);
    C s1 = cadd(a1, a3), d1 = mulj<DIR>(csub(a1, a3));
    a0 = cadd(s0, s1);
    a2 = csub(s0, s1);
    a1 = cadd(d0, d1);
    a3 = csub(d0, d1);
}

template <int DIR, typename C> __device__ __forceinline__ void dft8(C *v)
{
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    o1 = twc<DIR, 8, 1>(o1);
    o2 = twc<DIR, 8, 2>(o2);
    o3 = twc<DIR, 8, 3>(o3);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

// 16 = 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2.
template <int DIR, typename C> __device__ __forceinline__ void dft16(C *v)
{
    C a[16];
    sfor<0, 4>([&](auto n2c) {
        constexpr int n2 = decltype(n2c)::value;
        C x0 = v[n2], x1 = v[4 + n2], x2 = v[8 + n2], x3 = v[12 + n2];
        dft4<DIR>(x0, x1, x2, x3);
        a[n2 * 4 + 0] = x0;
        a[n2 * 4 + 1] = twc<DIR, 16, n2 * 1>(x1);
        a[n2 * 4 + 2] = twc<DIR, 16, n2 * 2>(x2);
        a[n2 * 4 + 3] = twc<DIR, 16, n2 * 3>(x3);
    });
    sfor<0, 4>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        C y0 = a[0 + k1], y1 = a[4 + k1], y2 = a[8 + k1], y3 = a[12 + k1];
        dft4<DIR>(y0, y1, y2, y3);
        v[k1] = y0;
        v[k1 + 4] = y1;
        v[k1 + 8] = y2;
        v[k1 + 12] = y3;
    });
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dftR(C *v)
{
    if constexpr (R == 1) {
    } else if constexpr (R == 2) dft2<DIR>(v[0], v[1]);
    else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8<DIR>(v);
    else if constexpr (R == 16) dft16<DIR>(v);
    else static_assert(R == 16, "unsupported radix");
}

// ---------------------------------------------------------------------------
// Stage schedule of an L-point line
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

template <int L> struct Sched {
    static_assert(L >= 1 && (L & (L - 1)) == 0, "power of two");
    static constexpr int LG = ilog2c(L);
    static constexpr int S = LG <= 4 ? 1 : (LG + 3) / 4;  // stages
    static constexpr int R = L >= 16 ? 16 : L;            // points per thread
    static constexpr int NT = L / R;                      // threads per line
    static constexpr int R0 = L >> (4 * (S - 1));         // first radix
    static constexpr int radix(int s) { return s == 0 ? R0 : 16; }
    static constexpr int ns(int s) { return s == 0 ? 1 : R0 << (4 * (s - 1)); }
    // twiddle table (FP64): stage s >= 1 holds ns(s) rows of the 4 base
    // powers [w, w^2, w^4, w^8] of w = e^{-2 pi i k / (16 ns(s))}
    static constexpr int twoff(int s) { return s <= 1 ? 0 : twoff(s - 1) + ns(s - 1) * 4; }
    static constexpr int TWN = S <= 1 ? 0 : twoff(S);
};

// LDS padding for line-major layouts: one element per 16 (FP32) / 8 (FP64)
// keeps the strided Stockham writes of the first stages off a single bank.
template <typename T> __host__ __device__ constexpr int lpad(int p)
{
    return sizeof(T) == 4 ? p + (p >> 4) : p + (p >> 3);
}
template <typename T, int L> __host__ __device__ constexpr int lpadlen()
{
    return lpad<T>(L - 1) + 1;
}

// Line-major LDS layout: line b owns [b*PADL, (b+1)*PADL).
template <typename T, int L> struct LayLine {
    int base;
    __device__ __forceinline__ int operator()(int p) const { return base + lpad<T>(p); }
};
// Column layout: point p of line b at p*LPB + b (b fastest).
template <int LPB> struct LayCol {
    int b;
    __device__ __forceinline__ int operator()(int p) const { return p * LPB + b; }
};

// ---------------------------------------------------------------------------
// FP64 complex helpers for twiddle arithmetic
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 dmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ Cx<T> to_cx(double2 w)
{
    return cmk<Cx<T>>((T)w.x, (T)w.y);
}
// a * w (DIR < 0) or a * conj(w) (DIR > 0) with w in FP64: FP64 kernels
// multiply in FP64, FP32 kernels round w once and multiply in FP32.
template <int DIR, typename C> __device__ __forceinline__ C ctw64(C a, double2 w)
{
    return ctw<DIR>(a, cmk<C>((decltype(C::x))w.x, (decltype(C::x))w.y));
}

// Radix-16 stage twiddles w^r (r = 1..15) from the base powers
// b = [w, w^2, w^4, w^8] (FP64 tables): 11 products, each power at most 3
// products away from correctly rounded bases.  FP64 lines multiply in FP64;
// FP32 lines round the four bases once and form the powers in FP32 (a few
// FP32 ulps per twiddle, at half the cost of FP64 products and without the
// per-twiddle FP64 -> FP32 conversions).
template <int DIR> __device__ __forceinline__ void apply_stage_tw32(Cx<float> *a, const float2 (&b)[4])
{
    const Cx<float> w1 = b[0], w2 = b[1], w4 = b[2], w8 = b[3];
    const Cx<float> w3 = cmul(w2, w1), w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w4, w3);
    a[1] = ctw<DIR>(a[1], w1);
    a[2] = ctw<DIR>(a[2], w2);
    a[3] = ctw<DIR>(a[3], w3);
    a[4] = ctw<DIR>(a[4], w4);
    a[5] = ctw<DIR>(a[5], w5);
    a[6] = ctw<DIR>(a[6], w6);
    a[7] = ctw<DIR>(a[7], w7);
    a[8] = ctw<DIR>(a[8], w8);
    a[9] = ctw<DIR>(a[9], cmul(w8, w1));
    a[10] = ctw<DIR>(a[10], cmul(w8, w2));
    a[11] = ctw<DIR>(a[11], cmul(w8, w3));
    a[12] = ctw<DIR>(a[12], cmul(w8, w4));
    a[13] = ctw<DIR>(a[13], cmul(w8, w5));
    a[14] = ctw<DIR>(a[14], cmul(w8, w6));
    a[15] = ctw<DIR>(a[15], cmul(w8, w7));
}

template <int DIR> __device__ __forceinline__ void apply_stage_tw(Cx<float> *a, const float2 (&b)[4])
{
    apply_stage_tw32<DIR>(a, b);
}
template <int DIR, typename C> __device__ __forceinline__ void apply_stage_tw(C *a, const double2 (&b)[4])
{
    if constexpr (sizeof(a[0].x) == 4) {
        const float2 f[4] = {to_cx<float>(b[0]), to_cx<float>(b[1]), to_cx<float>(b[2]), to_cx<float>(b[3])};
        apply_stage_tw32<DIR>(a, f);
        return;
    }
    // ordered so that at most two FP64 temporaries are live at a time
    a[1] = ctw64<DIR>(a[1], b[0]);
    a[2] = ctw64<DIR>(a[2], b[1]);
    a[4] = ctw64<DIR>(a[4], b[2]);
    a[8] = ctw64<DIR>(a[8], b[3]);
    a[9] = ctw64<DIR>(a[9], dmul(b[3], b[0]));
    a[10] = ctw64<DIR>(a[10], dmul(b[3], b[1]));
    a[12] = ctw64<DIR>(a[12], dmul(b[3], b[2]));
    {
        const double2 p3 = dmul(b[1], b[0]);
        a[3] = ctw64<DIR>(a[3], p3);
        a[11] = ctw64<DIR>(a[11], dmul(b[3], p3));
        const double2 p7 = dmul(b[2], p3);
        a[7] = ctw64<DIR>(a[7], p7);
        a[15] = ctw64<DIR>(a[15], dmul(b[3], p7));
    }
    {
        const double2 p5 = dmul(b[2], b[0]);
        a[5] = ctw64<DIR>(a[5], p5);
        a[13] = ctw64<DIR>(a[13], dmul(b[3], p5));
    }
    {
        const double2 p6 = dmul(b[2], b[1]);
        a[6] = ctw64<DIR>(a[6], p6);
        a[14] = ctw64<DIR>(a[14], dmul(b[3], p6));
    }
}

// Where the line FFT takes the base powers [w, w^2, w^4, w^8] of its
// twiddled stages: the FP64 table in global memory (requested one stage
// ahead), or a copy in LDS (StageTwLds, Cx<T> entries: FP32 kernels store the
// bases rounded once, exactly the values they would convert).  The LDS copy
// keeps a kernel's FFT free of global loads, so loads it issued for its next
// tile stay in flight through the whole transform.
struct StageTwGlobal {
    const double2 *tw;
    using B = double2;
    static constexpr bool kAhead = true;  // global latency: fetch one stage early
    __device__ __forceinline__ void load(int off, B (&bw)[4]) const
    {
#pragma unroll
        for (int e = 0; e < 4; e++) bw[e] = tw[off + e];
    }
};
template <typename T> struct StageTwLds {
    const Cx<T> *tw;  // LDS
    using B = Cx<T>;
    static constexpr bool kAhead = false;
    __device__ __forceinline__ void load(int off, B (&bw)[4]) const
    {
#pragma unroll
        for (int e = 0; e < 4; e++) bw[e] = tw[off + e];
    }
};
// Copy the stage bases of length L into LDS (all threads of the block take
// part; the caller synchronises before use).
template <typename T, int L> __device__ __forceinline__ void stage_tw_to_lds(Cx<T> *dst, const double2 *tw)
{
    for (int i = threadIdx.x; i < Sched<L>::TWN; i += blockDim.x) dst[i] = to_cx<T>(tw[i]);
}

// ---------------------------------------------------------------------------
// The line FFT.  All threads of the block must call it (it contains block
// barriers when S > 1).  reg[q] <-> point j + NT*q.  With global bases, those
// of stage s+1 are requested at the start of stage s so that their latency
// overlaps the butterflies and the LDS exchange.
// ---------------------------------------------------------------------------
template <typename T, int L, int DIR, class Lay, class Tw>
__device__ __forceinline__ void line_fft_tw(Cx<T> (&reg)[Sched<L>::R], Cx<T> *lds, const Lay &lay, int j,
                                            const Tw &twsrc)
{
    using S = Sched<L>;
    using C = Cx<T>;
    typename Tw::B bw[4];  // base powers of the current / next twiddled stage
    sfor<0, S::S>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int Rs = S::radix(s), Ns = S::ns(s), m = S::R / Rs;
        if constexpr (s == 0 && S::S > 1 && Tw::kAhead) twsrc.load(S::twoff(1) + (j & (S::ns(1) - 1)) * 4, bw);
        if constexpr (s > 0 && !Tw::kAhead) twsrc.load(S::twoff(s) + (j & (S::ns(s) - 1)) * 4, bw);
#pragma unroll
        for (int i = 0; i < m; i++) {
            C a[Rs];
#pragma unroll
            for (int r = 0; r < Rs; r++) a[r] = reg[i + r * m];
            if constexpr (s > 0) {
                static_assert(Rs == 16 && m == 1, "twiddled stages are radix 16");
                apply_stage_tw<DIR>(a, bw);
            }
            dftR<Rs, DIR>(a);
#pragma unroll
            for (int r = 0; r < Rs; r++) reg[i + r * m] = a[r];
        }
        if constexpr (s + 1 < S::S) {
            // request the next stage's bases; the exchange hides the latency
            if constexpr (s > 0 && Tw::kAhead) twsrc.load(S::twoff(s + 1) + (j & (S::ns(s + 1) - 1)) * 4, bw);
#pragma unroll
            for (int i = 0; i < m; i++) {
                const int v = j + S::NT * i;
                const int base = (v / Ns) * Ns * Rs + (v & (Ns - 1));
#pragma unroll
                for (int r = 0; r < Rs; r++) lds[lay(base + r * Ns)] = reg[i + r * m];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < S::R; q++) reg[q] = lds[lay(j + S::NT * q)];
            __syncthreads();
        }
    });
}
template <typename T, int L, int DIR, class Lay>
__device__ __forceinline__ void line_fft(Cx<T> (&reg)[Sched<L>::R], Cx<T> *lds, const Lay &lay, int j,
                                         const double2 *__restrict__ tw)
{
    line_fft_tw<T, L, DIR>(reg, lds, lay, j, StageTwGlobal{tw});
}

// NL independent lines per thread (line l: registers reg[l], LDS layout
// lay[l], thread index j[l] within its line), transformed together so that
// they share the block barriers of each exchange: a workgroup of NT * LPB / NL
// threads covers a tile of LPB lines, each thread holding NL x 16 points.
// Same arithmetic, in the same order per line, as line_fft_tw.
template <typename T, int L, int DIR, int NL, class Lay, class Tw>
__device__ __forceinline__ void line_fft_tw_n(Cx<T> (&reg)[NL][Sched<L>::R], Cx<T> *lds, const Lay (&lay)[NL],
                                              const int (&j)[NL], const Tw &twsrc)
{
    using S = Sched<L>;
    using C = Cx<T>;
    typename Tw::B bw[NL][4];
    sfor<0, S::S>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int Rs = S::radix(s), Ns = S::ns(s), m = S::R / Rs;
#pragma unroll
        for (int l = 0; l < NL; l++) {
            if constexpr (s == 0 && S::S > 1 && Tw::kAhead)
                twsrc.load(S::twoff(1) + (j[l] & (S::ns(1) - 1)) * 4, bw[l]);
            if constexpr (s > 0 && !Tw::kAhead) twsrc.load(S::twoff(s) + (j[l] & (S::ns(s) - 1)) * 4, bw[l]);
#pragma unroll
            for (int i = 0; i < m; i++) {
                C a[Rs];
#pragma unroll
                for (int r = 0; r < Rs; r++) a[r] = reg[l][i + r * m];
                if constexpr (s > 0) {
                    static_assert(Rs == 16 && m == 1, "twiddled stages are radix 16");
                    apply_stage_tw<DIR>(a, bw[l]);
                }
                dftR<Rs, DIR>(a);
#pragma unroll
                for (int r = 0; r < Rs; r++) reg[l][i + r * m] = a[r];
            }
            // one set's butterflies at a time: interleaved, the sets' temporaries
            // would be live together
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (s + 1 < S::S) {
#pragma unroll
            for (int l = 0; l < NL; l++) {
                if constexpr (s > 0 && Tw::kAhead)
                    twsrc.load(S::twoff(s + 1) + (j[l] & (S::ns(s + 1) - 1)) * 4, bw[l]);
#pragma unroll
                for (int i = 0; i < m; i++) {
                    const int v = j[l] + S::NT * i;
                    const int base = (v / Ns) * Ns * Rs + (v & (Ns - 1));
#pragma unroll
                    for (int r = 0; r < Rs; r++) lds[lay[l](base + r * Ns)] = reg[l][i + r * m];
                }
            }
            __syncthreads();
#pragma unroll
            for (int l = 0; l < NL; l++)
#pragma unroll
                for (int q = 0; q < S::R; q++) reg[l][q] = lds[lay[l](j[l] + S::NT * q)];
            __syncthreads();
        }
    });
}

// Twiddle for the four-step / Bluestein inter-pass products: w_N^{e},
// e < N, from two FP64 tables (hi: e >> sh, lo: e & mask), product in FP64.
struct Tw2 {
    const double2 *hi;
    const double2 *lo;
    int sh;
    int mask;
};
// The same two-level table rounded to FP32 (float2 entries).
struct Tw2f {
    const void *hi;
    const void *lo;
    int sh;
    int mask;
};
__device__ __forceinline__ double2 tw2_get(const Tw2 &t, uint32_t e)
{
    double2 h = t.hi[e >> t.sh], l = t.lo[e & t.mask];
    double2 r;
    r.x = h.x * l.x - h.y * l.y;
    r.y = h.x * l.y + h.y * l.x;
    return r;
}
// reg[q] *= w^{base_e + step_e * q} (conj for DIR > 0), q = 0..R-1: two
// table lookups, then a recurrence.  FP64 lines recur in FP64 (error ~R ulps
// of FP64).  FP32 lines form w^base, w^(base+step) and w^(2 step) in FP64,
// round them once and run two interleaved FP32 chains of R/2 products each
// (error ~R/2 FP32 ulps on the last twiddles).
template <int DIR, typename C, int R>
__device__ __forceinline__ void apply_tw2_from(C (&reg)[R], double2 w, const double2 st)
{
    if constexpr (sizeof(reg[0].x) == 4 && R % 2 == 0) {
        Cx<float> e = to_cx<float>(w), o = to_cx<float>(dmul(w, st));
        const Cx<float> st2 = to_cx<float>(dmul(st, st));
#pragma unroll
        for (int q = 0; q < R; q += 2) {
            reg[q] = ctw<DIR>(reg[q], e);
            reg[q + 1] = ctw<DIR>(reg[q + 1], o);
            if (q + 2 < R) {
                e = cmul(e, st2);
                o = cmul(o, st2);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < R; q++) {
            reg[q] = ctw64<DIR>(reg[q], w);
            if (q + 1 < R) w = dmul(w, st);
        }
    }
}

template <int DIR, typename C, int R>
__device__ __forceinline__ void apply_tw2_series(C (&reg)[R], const Tw2 &t, uint32_t base_e, uint32_t step_e)
{
    apply_tw2_from<DIR>(reg, tw2_get(t, base_e), tw2_get(t, step_e));
}

// Global access as a workgroup-uniform base plus a 32-bit per-lane byte
// offset: the compiler then addresses with the base in scalar registers
// instead of 64-bit per-lane arithmetic on every access.  Callers take this
// form only when off_fits(largest lane offset in elements) holds.
template <typename C> __device__ __forceinline__ bool off_fits(long long max_elems)
{
    return max_elems >= 0 && max_elems * (long long)sizeof(C) < (1LL << 32);
}
template <typename C> __device__ __forceinline__ C ld_off(const C *base, uint32_t off)
{
    return *(const C *)((const char *)base + off);
}
template <typename C> __device__ __forceinline__ void st_off(C *base, uint32_t off, C v)
{
    *(C *)((char *)base + off) = v;
}

// Raw buffer access (gfx9 resource, no stride, range check wide open):
// address = base + soff (scalar register, one per access) + voff (the lane's
// 32-bit byte offset).  Strided tiles use it so that each access costs no
// per-lane address arithmetic; callers keep soff + voff below 2^32.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, -1, 0x00020000);
}
// The same view, or (live == false) one with no records: every load through
// it returns zero without a memory access.  Lets a prefetch be issued
// unconditionally (no branch, so the wait counts after it stay exact).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc_if(const void *base, bool live)
{
    // the arguments are workgroup-uniform: pin them to scalar registers so
    // that the descriptor needs no per-lane handling
    const uint64_t p = (uint64_t)base;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)p);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(p >> 32));
    const int n = __builtin_amdgcn_readfirstlane(live ? -1 : 0);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(((uint64_t)hi << 32) | lo), (short)0, n, 0x00020000);
}
// CP: cache-policy bits of the access (0: default; 2: non-temporal).
template <typename C, int CP = 0>
__device__ __forceinline__ C buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    if constexpr (sizeof(C) == 8)
        return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, CP));
    else
        return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, CP));
}
// 16-byte stores take the whole offset in the lane register (scalar offset
// 0): the compiler's wait state between a store of more than 8 bytes and a
// following write of its data registers is only inserted for that form, and
// FP64 results written through the scalar-offset form were intermittently
// corrupted when other kernels ran at the same time.
template <typename C, int CP = 0>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, C v)
{
    if constexpr (sizeof(C) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r,
                                              (int)voff, (int)soff, CP);
    else
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v),
                                               r, (int)(voff + soff), 0, CP);
}

// Logical workgroup index for XCD-aware tiling: workgroups are dispatched
// round-robin over the 8 XCDs (linear id i runs on XCD i % 8), so giving
// each XCD a contiguous range of logical tiles keeps tiles that share
// 128-byte lines (adjacent column groups) in the same L2.
__device__ __forceinline__ long long xcd_logical(long long lin, long long total)
{
    const long long q = total >> 3, r = total & 7;
    const long long x = lin & 7, k = lin >> 3;
    return x * q + (x < r ? x : r) + k;
}

}  // namespace nfft