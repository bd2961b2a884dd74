// nfft_kernels.h -- kernel templates for the CDNA4 FFT engine.
//
// Families (each instantiated per precision and line length in the .hip files):
//   k_line      single-pass batched C2C: one L-point line per L/16 threads,
//               LPB lines per workgroup, the whole line resident in VGPRs/LDS.
//   k_col       column pass: LPB adjacent columns of stride S (coalesced across
//               columns), used for the first four-step pass (with the
//               inter-pass twiddle), N-D axes and the Bluestein passes.
//   k_rowT      row pass with transposed store: second four-step pass.
//   k_blue_line single-kernel Bluestein for M <= Lmax: chirp, FFT_M, x H,
//               IFFT_M, chirp, all in one workgroup.
//   k_blue_row  middle pass of the three-pass Bluestein (FFT_M2, x H, IFFT_M2).
//
// Replaces the reference's per-call CPU loops: split-radix DIF
// (src/05_algorithms.h:358-736), Bluestein (src/05_algorithms.h:742-896) and
// the strided N-D passes (src/07_legacy.h:81-139).
#pragma once
#include "nfft_device.h"

namespace nfft {

template <typename T> __device__ __forceinline__ Cx<T> czero() { return cmk<Cx<T>>((T)0, (T)0); }

// Coalesced copy of a tile of rows (row r at src + r*dist, `len` valid points,
// rows >= nrows or points >= len read as zero) into the line-major LDS
// layout of lines of length L.
template <typename T, int L, int LPB, int NTH>
__device__ __forceinline__ void stage_in_rows(Cx<T> *lds, const Cx<T> *src, long long dist, int len, int nrows)
{
    constexpr int PADL = lpadlen<T, L>();
    for (int idx = threadIdx.x; idx < LPB * L; idx += NTH) {
        const int b = idx / L, p = idx & (L - 1);
        Cx<T> v = czero<T>();
        if (b < nrows && p < len) v = src[b * dist + p];
        lds[b * PADL + lpad<T>(p)] = v;
    }
}

template <typename T, int L, int LPB, int NTH>
__device__ __forceinline__ void stage_out_rows(const Cx<T> *lds, Cx<T> *dst, long long dist, int len, int nrows)
{
    constexpr int PADL = lpadlen<T, L>();
    for (int idx = threadIdx.x; idx < LPB * L; idx += NTH) {
        const int b = idx / L, p = idx & (L - 1);
        if (b < nrows && p < len) dst[b * dist + p] = lds[b * PADL + lpad<T>(p)];
    }
}

// LDS elements needed by a line-major kernel of LPB lines.
template <typename T, int L, int LPB> __host__ __device__ constexpr int line_lds()
{
    return LPB * lpadlen<T, L>();
}

// Lines per workgroup for the single-pass kernels: ~256 threads, >= 1 line.
template <int L> __host__ __device__ constexpr int line_lpb()
{
    return Sched<L>::NT >= 256 ? 1 : 256 / Sched<L>::NT;
}

// ---------------------------------------------------------------------------
// Load / store of a tile of contiguous lines into reg[] (line-major mapping).
// Direct global access when the NT threads of a line cover at least one
// 128-byte segment per access (a wave then reads 64 / NT runs of whole
// segments), LDS-staged coalesced copy for shorter lines.
// ---------------------------------------------------------------------------
// WIDE: direct access only for lines of a whole wave (NT >= 64), the rule
// of the one-workgroup-per-tile row pass (see k_rowT).
template <typename T, int L, bool WIDE = false> __host__ __device__ constexpr bool rows_direct()
{
    return WIDE ? Sched<L>::NT >= 64 : Sched<L>::NT * sizeof(Cx<T>) >= 128;
}

template <typename T, int L, int LPB, bool WIDE = false>
__device__ __forceinline__ void tile_load(Cx<T> (&reg)[Sched<L>::R], Cx<T> *lds, const Cx<T> *src, long long dist,
                                          int len, int nrows, int b, int j)
{
    using S = Sched<L>;
    constexpr int NTH = S::NT * LPB;
    if constexpr (rows_direct<T, L, WIDE>()) {
        if (nrows == LPB && len == L && off_fits<Cx<T>>((long long)(LPB - 1) * dist + L)) {
            // full tile (uniform): unpredicated loads off one base
            const uint32_t off = (uint32_t)(b * dist + j) * (uint32_t)sizeof(Cx<T>);
#pragma unroll
            for (int q = 0; q < S::R; q++) reg[q] = ld_off(src + S::NT * q, off);
        } else {
            const Cx<T> *s = src + b * dist;
#pragma unroll
            for (int q = 0; q < S::R; q++) {
                const int p = j + S::NT * q;
                reg[q] = (b < nrows && p < len) ? s[p] : czero<T>();
            }
        }
    } else {
        stage_in_rows<T, L, LPB, NTH>(lds, src, dist, len, nrows);
        __syncthreads();
        const int base = b * lpadlen<T, L>();
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = lds[base + lpad<T>(j + S::NT * q)];
        __syncthreads();
    }
}

template <typename T, int L, int LPB, bool WIDE = false>
__device__ __forceinline__ void tile_store(const Cx<T> (&reg)[Sched<L>::R], Cx<T> *lds, Cx<T> *dst, long long dist,
                                           int len, int nrows, int b, int j)
{
    using S = Sched<L>;
    constexpr int NTH = S::NT * LPB;
    if constexpr (rows_direct<T, L, WIDE>()) {
        if (nrows == LPB && len == L && off_fits<Cx<T>>((long long)(LPB - 1) * dist + L)) {
            const uint32_t off = (uint32_t)(b * dist + j) * (uint32_t)sizeof(Cx<T>);
#pragma unroll
            for (int q = 0; q < S::R; q++) st_off(dst + S::NT * q, off, reg[q]);
        } else if (b < nrows) {
            Cx<T> *d = dst + b * dist;
#pragma unroll
            for (int q = 0; q < S::R; q++) {
                const int p = j + S::NT * q;
                if (p < len) d[p] = reg[q];
            }
        }
    } else {
        const int base = b * lpadlen<T, L>();
#pragma unroll
        for (int q = 0; q < S::R; q++) lds[base + lpad<T>(j + S::NT * q)] = reg[q];
        __syncthreads();
        stage_out_rows<T, L, LPB, NTH>(lds, dst, dist, len, nrows);
    }
}

// ---------------------------------------------------------------------------
// k_line: batched contiguous C2C lines (in-place allowed).
// ---------------------------------------------------------------------------
struct LineArgs {
    const void *in;
    void *out;
    long long nlines;
    long long in_dist, out_dist;  // elements between consecutive lines
    const void *tw;               // stage twiddles for L
    double scale;
    int halves;                   // one line per workgroup, exchanges in two LDS halves (k_line_h)
};

template <typename T, int L, int DIR>
__global__ void __launch_bounds__(Sched<L>::NT *line_lpb<L>()) k_line(LineArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int LPB = line_lpb<L>();
    constexpr int NLDS = (S::NT >= 64 && S::S == 1) ? 1 : line_lds<T, L, LPB>();
    __shared__ C lds[NLDS];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const long long line0 = (long long)blockIdx.x * LPB;
    const long long rem = a.nlines - line0;
    const int nrows = rem < LPB ? (int)rem : LPB;
    const C *src = (const C *)a.in + line0 * a.in_dist;
    C *dst = (C *)a.out + line0 * a.out_dist;
    C reg[S::R];
    tile_load<T, L, LPB>(reg, lds, src, a.in_dist, L, nrows, b, j);
    LayLine<T, L> lay{b * lpadlen<T, L>()};
    line_fft<T, L, DIR>(reg, lds, lay, j, (const double2 *)a.tw);
    const T sc = (T)a.scale;
    if (sc != (T)1) {
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = cscale(reg[q], sc);
    }
    tile_store<T, L, LPB>(reg, lds, dst, a.out_dist, L, nrows, b, j);
}

// ---------------------------------------------------------------------------
// Lines with their exchanges through half a line of LDS: twice the
// LDS-resident length (FP32 L = 2^15, FP64 L = 2^14: 256 KiB), and 8192 /
// FP32 16384 points, where the smaller buffer lets more lines share a CU.
// One line per workgroup of NT / 2 threads, held in VGPRs as NL = 2 sets of
// 16 points (set l of thread j: points j + l NT/2 + NT q), with each
// Stockham exchange made through an LDS buffer of half a line in two steps.
// Every value a stage writes lands in a half known at compile time:
//   - stage 0 (Ns = 1, radix R0, m = 16 / R0 >= 2 butterflies per set):
//     butterfly i < m / 2 of either set writes points < L / 2;
//   - a twiddled stage s < S - 1 (Ns <= L / 256, so 16 Ns divides L / 2):
//     set 0 (j < NT / 2) writes the lower half, set 1 the upper;
//   - the reads of q < 8 (points < 8 NT = L / 2) come from the lower half.
// Step h writes the values of half h, then every thread reads its points of
// that half; the lower-half reads go to a side array until the upper-half
// values have been written.  Arithmetic per point is that of line_fft_tw.
// ---------------------------------------------------------------------------
template <typename T, int L> __host__ __device__ constexpr bool line_h_ok()
{
    using S = Sched<L>;
    return S::S >= 2 && S::R == 16 && S::R0 <= 8 && (S::NT / 2) * 2 == S::NT && S::NT / 2 <= 1024;
}

// Waves per SIMD the FP32 half-LDS kernels are compiled for: as many as the
// workgroups that fit in the LDS can use, capped at 4 (<= 128 VGPRs; at 131
// the 16384-point line held one workgroup per CU instead of two).
template <typename T, int L> __host__ __device__ constexpr int line_h_wpe()
{
    if (sizeof(T) != 4) return 1;
    const long lds = (long)lpadlen<T, L / 2>() * (long)sizeof(Cx<T>);
    const long wg = (160L * 1024) / lds;
    const long w = wg * (Sched<L>::NT / 2 / 64) / 4;
    return w < 1 ? 1 : (w > 4 ? 4 : (int)w);
}

template <typename T, int L, int DIR>
__device__ __forceinline__ void line_fft_halves(Cx<T> (&reg)[2][16], Cx<T> *lds, int j0, const double2 *__restrict__ tw)
{
    using S = Sched<L>;
    using C = Cx<T>;
    static_assert(line_h_ok<T, L>(), "half-buffer line shape");
    constexpr int H = L / 2, NH = S::NT / 2;
    const int jj[2] = {j0, j0 + NH};
    double2 bw[2][4];
    sfor<0, S::S>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int Rs = S::radix(s), Ns = S::ns(s), m = S::R / Rs;
        // bases of stage s' for both sets: the same row unless ns(s') exceeds NT / 2
        auto bases = [&](auto sn) {
            constexpr int s2 = decltype(sn)::value;
            constexpr bool same = NH % S::ns(s2) == 0;
#pragma unroll
            for (int l = 0; l < (same ? 1 : 2); l++)
#pragma unroll
                for (int e = 0; e < 4; e++) bw[l][e] = tw[S::twoff(s2) + (jj[l] & (S::ns(s2) - 1)) * 4 + e];
            if constexpr (same) {
#pragma unroll
                for (int e = 0; e < 4; e++) bw[1][e] = bw[0][e];
            }
        };
        if constexpr (s == 0) bases(std::integral_constant<int, 1>{});
#pragma unroll
        for (int l = 0; l < 2; l++) {
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
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (s + 1 < S::S) {
            static_assert(s > 0 || m >= 2, "stage 0 splits by butterfly");
            static_assert(s == 0 || (H % (16 * Ns) == 0 && NH % Ns == 0), "twiddled stages split by set");
            if constexpr (s > 0) bases(std::integral_constant<int, s + 1>{});
            C lo[2][8];
#pragma unroll
            for (int h = 0; h < 2; h++) {
#pragma unroll
                for (int l = 0; l < 2; l++)
#pragma unroll
                    for (int i = 0; i < m; i++) {
                        if ((s == 0 ? (2 * i >= m ? 1 : 0) : l) != h) continue;
                        const int v = jj[l] + S::NT * i;
                        const int base = (v / Ns) * Ns * Rs + (v & (Ns - 1)) - h * H;
#pragma unroll
                        for (int r = 0; r < Rs; r++) lds[lpad<T>(base + r * Ns)] = reg[l][i + r * m];
                    }
                __syncthreads();
#pragma unroll
                for (int l = 0; l < 2; l++)
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const C v = lds[lpad<T>(jj[l] + S::NT * q)];  // point jj + NT (q + 8 h) - h H
                        if (h == 0) lo[l][q] = v;
                        else reg[l][q + 8] = v;
                    }
                __syncthreads();
            }
#pragma unroll
            for (int l = 0; l < 2; l++)
#pragma unroll
                for (int q = 0; q < 8; q++) reg[l][q] = lo[l][q];
        }
    });
}

// k_line_h: one such line per workgroup, read and written straight from
// global memory (a wave covers 64 consecutive points per access).
template <typename T, int L, int DIR>
__global__ void __launch_bounds__(Sched<L>::NT / 2, (line_h_wpe<T, L>())) k_line_h(LineArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int NH = S::NT / 2;
    __shared__ C lds[lpadlen<T, L / 2>()];
    const long long line = blockIdx.x;
    const int j = threadIdx.x;
    const auto rs = buf_rsrc((const C *)a.in + line * a.in_dist);
    const uint32_t off = (uint32_t)j * (uint32_t)sizeof(C);
    C reg[2][16];
#pragma unroll
    for (int l = 0; l < 2; l++)
#pragma unroll
        for (int q = 0; q < 16; q++) reg[l][q] = buf_ld<C>(rs, off, (uint32_t)((l * NH + S::NT * q) * sizeof(C)));
    line_fft_halves<T, L, DIR>(reg, lds, j, (const double2 *)a.tw);
    const T sc = (T)a.scale;
    const auto rd = buf_rsrc((C *)a.out + line * a.out_dist);
#pragma unroll
    for (int l = 0; l < 2; l++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            C v = reg[l][q];
            if (sc != (T)1) v = cscale(v, sc);
            buf_st<C>(rd, off, (uint32_t)((l * NH + S::NT * q) * sizeof(C)), v);
        }
}

// ---------------------------------------------------------------------------
// k_col: LPB adjacent columns per workgroup; element (t, c, p) lives at
// base + t*dist + c*cstride + p*S with cstride = 1 for coalescing.
// MODE selects the prologue/epilogue:
//   COL_PLAIN     in -> FFT -> out
//   COL_TWIDDLE   in -> FFT -> * w_N^{DIR' c p} -> out   (four-step pass 1)
//   COL_BLUE_IN   x[n] * chirp[n] (n < Nv, else 0) -> FFT -> * w_M^{c p} -> out
//   COL_BLUE_OUT  in -> IFFT -> y[n] = conj?(v * chirp[n]) * scale, n < Nv
// ---------------------------------------------------------------------------
enum ColMode { COL_PLAIN = 0, COL_TWIDDLE = 1, COL_BLUE_IN = 2, COL_BLUE_OUT = 3 };

struct ColArgs {
    const void *in;
    void *out;
    int ncols;                 // columns per transform
    long long S;               // point stride (elements)
    long long in_dist, out_dist;  // per-transform distance
    long long ntrans;          // transforms (grid.y * gridDim.z covers them)
    const void *tw;            // stage twiddles for L
    Tw2 tw2;                   // inter-pass twiddle (TWIDDLE / BLUE_IN)
    const void *chirp;         // Bluestein chirp (length Nv)
    long long nv;              // Bluestein valid length
    int conj;                  // Bluestein: conjugate input / output
    double scale;
    int lpb;                   // columns per workgroup (0: default for L)
};

template <int L, int LPB> __host__ __device__ constexpr int col_threads() { return Sched<L>::NT * LPB; }

// Minimum waves per SIMD the FP32 four-step kernels are compiled for: as many
// as the workgroups that fit in the LDS can use, capped at 4 (<= 128 VGPRs);
// without the hint the twiddle pass lands just above 128 VGPRs and loses a
// quarter of its resident waves.  FP64 kernels keep the compiler's choice.
template <typename T> __host__ __device__ constexpr int wpe_hint(int threads, long lds_bytes)
{
    if (sizeof(T) != 4) return 1;
    long wg = lds_bytes > 0 ? (160L * 1024) / lds_bytes : 8;
    if (wg > 8) wg = 8;
    const long w = wg * (threads / 64) / 4;
    return w < 1 ? 1 : (w > 4 ? 4 : (int)w);
}

template <typename T, int L, int LPB, int DIR, int MODE>
__global__ void __launch_bounds__((col_threads<L, LPB>()),
                                  (MODE == COL_BLUE_IN || MODE == COL_BLUE_OUT
                                       ? 1
                                       : wpe_hint<T>(col_threads<L, LPB>(), (long)L * LPB * sizeof(Cx<T>))))
k_col(ColArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int NLDS = S::S > 1 ? L * LPB : 1;
    // plain / twiddle columns read their stage twiddles from an LDS copy
    // (rounded once to T) where it fits beside the tile without lowering the
    // number of workgroups the LDS admits per CU (1 KiB allocation steps)
    constexpr int TWN = Sched<L>::TWN > 0 ? Sched<L>::TWN : 1;
    constexpr size_t kB0 = ((size_t)NLDS * sizeof(C) + 1023) / 1024 * 1024;
    constexpr size_t kB1 = ((size_t)(NLDS + TWN) * sizeof(C) + 1023) / 1024 * 1024;
    constexpr bool TWL = S::S > 1 && (MODE == COL_PLAIN || MODE == COL_TWIDDLE) && kB1 <= 160 * 1024 &&
                         (160 * 1024) / kB1 == (160 * 1024) / kB0;
    __shared__ C lds[NLDS];
    __shared__ C ltw[TWL ? TWN : 1];
    if constexpr (TWL) stage_tw_to_lds<T, L>(ltw, (const double2 *)a.tw);
    const int b = threadIdx.x % LPB, j = threadIdx.x / LPB;
    const long long lin = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    const long long lg = xcd_logical(lin, (long long)gridDim.x * gridDim.y * gridDim.z);
    const int c0 = (int)(lg % gridDim.x) * LPB, c = c0 + b;
    const long long t = lg / gridDim.x;
    const bool valid = c < a.ncols && t < a.ntrans;
    // uniform: every column of the tile exists and tile offsets fit 32 bits
    const bool full = c0 + LPB <= a.ncols && t < a.ntrans && off_fits<C>((long long)L * a.S);
    const uint32_t qst = (uint32_t)(S::NT * a.S) * (uint32_t)sizeof(C);  // byte step of q
    const uint32_t loff = (uint32_t)(j * a.S + b) * (uint32_t)sizeof(C);
    C reg[S::R];
    if constexpr (MODE == COL_BLUE_IN) {
        const C *src = (const C *)a.in + t * a.in_dist;
        const C *ch = (const C *)a.chirp;
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const long long n = (long long)(j + S::NT * q) * a.S + c;
            C v = czero<T>();
            if (valid && n < a.nv) {
                v = src[n];
                if (a.conj) v = cconj(v);
                v = cmul(v, ch[n]);
            }
            reg[q] = v;
        }
    } else if (full) {
        // whole tile in range: unpredicated loads, uniform base per q
        const auto rs = buf_rsrc((const C *)a.in + t * a.in_dist + c0);
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = buf_ld<C>(rs, loff, q * qst);
    } else {
        const C *src = (const C *)a.in + t * a.in_dist + c;
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = valid ? src[(long long)(j + S::NT * q) * a.S] : czero<T>();
    }
    LayCol<LPB> lay{b};
    if constexpr (TWL) line_fft_tw<T, L, DIR>(reg, lds, lay, j, StageTwLds<T>{ltw});
    else line_fft<T, L, DIR>(reg, lds, lay, j, (const double2 *)a.tw);
    if constexpr (MODE == COL_TWIDDLE || MODE == COL_BLUE_IN) {
        const uint32_t cc = valid ? (uint32_t)c : 0u;
        apply_tw2_series<DIR>(reg, a.tw2, cc * (uint32_t)j, cc * (uint32_t)S::NT);
    }
    if (!valid) return;
    if constexpr (MODE == COL_BLUE_OUT) {
        C *dst = (C *)a.out + t * a.out_dist;
        const C *ch = (const C *)a.chirp;
        const T sc = (T)a.scale;
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const long long n = (long long)(j + S::NT * q) * a.S + c;
            if (n < a.nv) {
                C v = cmul(reg[q], ch[n]);
                if (a.conj) v = cconj(v);
                dst[n] = cscale(v, sc);
            }
        }
    } else if (full) {
        const auto rd = buf_rsrc((C *)a.out + t * a.out_dist + c0);
        const T sc = (T)a.scale;
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            C v = reg[q];
            if (sc != (T)1) v = cscale(v, sc);
            buf_st(rd, loff, q * qst, v);
        }
    } else {
        C *dst = (C *)a.out + t * a.out_dist + c;
        const T sc = (T)a.scale;
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            C v = reg[q];
            if (sc != (T)1) v = cscale(v, sc);
            dst[(long long)(j + S::NT * q) * a.S] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// k_rowT: second four-step pass.  Row r (< nrows) of transform t holds L
// contiguous points; output point k of row r goes to r + nrows*k (natural
// order of the full transform), written through an LDS transpose so that
// LPB consecutive rows form one contiguous store segment.
// ---------------------------------------------------------------------------
struct RowTArgs {
    const void *in;
    void *out;
    int nrows;
    long long in_dist, out_dist;  // per-transform distances
    long long ntrans;
    const void *tw;
    double scale;
    int lpb;                      // rows per workgroup (0: default for L)
    int pf;                       // full tiles through k_rowT_pf (resident, prefetching)
    Tw2f twf;                     // TWIN kernels: inter-pass twiddle w_N^{r p} of row r, point p
};

// TWIN: the row pass applies the four-step's inter-pass twiddle to its input
// (row r = k1 of the column FFT, point p = n2) instead of the column pass to
// its output.  The column pass of a 1024 x 1024 FP32 four-step is the slower
// of the two, and the prefetching row grid hides the twiddle work under its
// loads.  Rows of 1024 FP32 points only: a thread's base and step exponents
// r j and r 64 stay below N / 16, so the row grid keeps the used part of the
// two-level table (hi[0, 64), lo[0, 2048)) in LDS, and no global load joins
// the prefetch in the in-order memory counter.
constexpr int kTwinHi = 64, kTwinLo = 2048;
template <int DIR, int R>
__device__ __forceinline__ void apply_twin(Cx<float> (&reg)[R], const Cx<float> *hi, const Cx<float> *lo, int sh,
                                           int mask, uint32_t base_e, uint32_t step_e)
{
    const Cx<float> w = cmul(hi[base_e >> sh], lo[base_e & mask]);
    const Cx<float> st = cmul(hi[step_e >> sh], lo[step_e & mask]);
    Cx<float> e = w, o = cmul(w, st);
    const Cx<float> st2 = cmul(st, st);
#pragma unroll
    for (int q = 0; q < R; q += 2) {
        reg[q] = ctw<DIR>(reg[q], e);
        reg[q + 1] = ctw<DIR>(reg[q + 1], o);
        if (q + 2 < R) {
            e = cmul(e, st2);
            o = cmul(o, st2);
        }
    }
}

template <typename T, int L, int LPB> __host__ __device__ constexpr int rowT_lds()
{
    return (L * (LPB + 1)) > line_lds<T, L, LPB>() ? L * (LPB + 1) : line_lds<T, L, LPB>();
}

template <typename T, int L, int LPB, int DIR, bool TWIN = false>
__global__ void __launch_bounds__(Sched<L>::NT *LPB,
                                  (wpe_hint<T>(Sched<L>::NT * LPB, (long)rowT_lds<T, L, LPB>() * sizeof(Cx<T>))))
k_rowT(RowTArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int NTH = S::NT * LPB;
    __shared__ C lds[rowT_lds<T, L, LPB>()];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const long long lin = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    const long long lg = xcd_logical(lin, (long long)gridDim.x * gridDim.y * gridDim.z);
    const int r0 = (int)(lg % gridDim.x) * LPB;
    const long long t = lg / gridDim.x;
    if (t >= a.ntrans) return;  // uniform per workgroup
    const int nrows = (a.nrows - r0) < LPB ? (a.nrows - r0) : LPB;
    const C *src = (const C *)a.in + t * a.in_dist + (long long)r0 * L;
    C reg[S::R];
    // rows shorter than a wave are staged through LDS here (WIDE): production
    // runs them on the prefetching grid (k_rowT_pf), and this pass is the one
    // that grid is compared with bit for bit (built with direct loads, its
    // FP64 256-point rows differed from the grid's in the last bits)
    tile_load<T, L, LPB, true>(reg, lds, src, L, L, nrows, b, j);
    if constexpr (TWIN) {
        // the table read from global memory; the same arithmetic as k_rowT_pf
        const uint32_t rr = b < nrows ? (uint32_t)(r0 + b) : 0u;
        apply_twin<DIR>(reg, (const Cx<float> *)a.twf.hi, (const Cx<float> *)a.twf.lo, a.twf.sh, a.twf.mask,
                        rr * (uint32_t)j, rr * (uint32_t)S::NT);
    }
    LayLine<T, L> lay{b * lpadlen<T, L>()};
    line_fft<T, L, DIR>(reg, lds, lay, j, (const double2 *)a.tw);
    const T sc = (T)a.scale;
#pragma unroll
    for (int q = 0; q < S::R; q++) {
        C v = reg[q];
        if (sc != (T)1) v = cscale(v, sc);
        lds[(j + S::NT * q) * (LPB + 1) + b] = v;
    }
    __syncthreads();
    C *dst = (C *)a.out + t * a.out_dist + r0;
    constexpr int KS = NTH / LPB;  // output points k advanced per sweep
    if (nrows == LPB && off_fits<C>((long long)L * a.nrows)) {
        // full tile (uniform): L/KS independent LDS reads and stores per thread
        const int bb = threadIdx.x % LPB, k0 = threadIdx.x / LPB;
        const uint32_t off = (uint32_t)(k0 * a.nrows + bb) * (uint32_t)sizeof(C);
        const uint32_t ist = (uint32_t)(KS * a.nrows) * (uint32_t)sizeof(C);
        const auto rd = buf_rsrc(dst);
#pragma unroll
        for (int i = 0; i < L / KS; i++) buf_st(rd, off, i * ist, lds[(k0 + KS * i) * (LPB + 1) + bb]);
        return;
    }
    for (int idx = threadIdx.x; idx < LPB * L; idx += NTH) {
        const int bb = idx % LPB, k = idx / LPB;
        if (bb < nrows) dst[(long long)k * a.nrows + bb] = lds[k * (LPB + 1) + bb];
    }
}

template <typename T, int L> __host__ __device__ constexpr int stage_tw_lds()
{
    return Sched<L>::TWN > 0 ? Sched<L>::TWN : 1;
}

// ---------------------------------------------------------------------------
// k_rowT_pf: k_rowT over full tiles (nrows % LPB == 0) as a resident grid.
// The tiles (tile = t * (nrows / LPB) + row group) are split into 8
// contiguous shares, one per XCD (workgroup w runs on XCD w % 8); the
// m = gridDim.x / 8 workgroups of an XCD walk its share m apart, so the
// output segments written at any time are adjacent (the same DRAM pages).
// Each workgroup issues the loads of its next tile before it transforms the
// current one: with one or two workgroups per CU (the transpose fills the
// LDS) the loads then overlap the FFT work instead of alternating with it.
// The FFT reads its stage twiddles from an LDS copy, so no global load is
// queued behind the prefetch.  Lanes read their points straight from global
// memory (a wave covers 64 / NT rows of NT consecutive points: segments of
// >= 128 bytes for L >= 256).  The host guarantees L * nrows elements and
// ntiles fit the 32-bit offsets and indices used here.
// ---------------------------------------------------------------------------
template <typename T, int L, int LPB> __host__ __device__ constexpr int rowT_pf_lds()
{
    return rowT_lds<T, L, LPB>() + stage_tw_lds<T, L>();
}

template <typename T, int L, int LPB, int DIR, bool TWIN = false>
__global__ void __launch_bounds__(Sched<L>::NT *LPB,
                                  (wpe_hint<T>(Sched<L>::NT * LPB, (long)rowT_pf_lds<T, L, LPB>() * sizeof(Cx<T>))))
k_rowT_pf(RowTArgs a, long long ntiles)
{
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int NTH = S::NT * LPB;
    constexpr int KS = NTH / LPB;  // output points k advanced per sweep
    __shared__ C lds[rowT_lds<T, L, LPB>()];
    __shared__ C ltw[stage_tw_lds<T, L>()];
    __shared__ C thi[TWIN ? kTwinHi : 1], tlo[TWIN ? kTwinLo : 1];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const unsigned gpt = (unsigned)(a.nrows / LPB);
    const unsigned x = blockIdx.x & 7, m = gridDim.x >> 3;
    const unsigned q8 = (unsigned)(ntiles >> 3), r8 = (unsigned)(ntiles & 7);
    const unsigned lo = x * q8 + (x < r8 ? x : r8), hi = lo + q8 + (x < r8 ? 1 : 0);
    const uint32_t loff = (uint32_t)(b * L + j) * (uint32_t)sizeof(C);
    const int bb = threadIdx.x % LPB, k0 = threadIdx.x / LPB;
    const uint32_t soff = (uint32_t)(k0 * a.nrows + bb) * (uint32_t)sizeof(C);
    const uint32_t ist = (uint32_t)(KS * a.nrows) * (uint32_t)sizeof(C);
    const T sc = (T)a.scale;
    unsigned tl = lo + (blockIdx.x >> 3);
    if (tl >= hi) return;  // uniform: no tile for this workgroup
    stage_tw_to_lds<T, L>(ltw, (const double2 *)a.tw);
    if constexpr (TWIN) {
        for (int i = threadIdx.x; i < kTwinHi; i += NTH) thi[i] = ((const C *)a.twf.hi)[i];
        for (int i = threadIdx.x; i <= a.twf.mask; i += NTH) tlo[i] = ((const C *)a.twf.lo)[i];
    }
    C reg[S::R], nxt[S::R];
    {
        const unsigned t = tl / gpt;
        const auto rs = buf_rsrc_if((const C *)a.in + t * a.in_dist + (tl - t * gpt) * LPB * (long long)L, true);
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = buf_ld<C>(rs, loff, q * S::NT * (uint32_t)sizeof(C));
    }
    // Complete the first tile's loads here: the loop-top wait then covers
    // only the prefetch, not the stores issued after it.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // Each iteration starts with its tile in reg[]; the hand-over from nxt[]
    // sits at the end of the previous iteration, after its stores.
    for (; tl < hi; tl += m) {
        const unsigned t = tl / gpt;
        const int r0 = (int)(tl - t * gpt) * LPB;
        {
            const unsigned tn = tl + m < hi ? tl + m : tl;
            const unsigned u = tn / gpt;
            const auto rs =
                buf_rsrc_if((const C *)a.in + u * a.in_dist + (tn - u * gpt) * LPB * (long long)L, tl + m < hi);
#pragma unroll
            for (int q = 0; q < S::R; q++) nxt[q] = buf_ld<C>(rs, loff, q * S::NT * (uint32_t)sizeof(C));
        }
        if constexpr (TWIN) {
            const uint32_t rr = (uint32_t)(r0 + b);
            apply_twin<DIR>(reg, thi, tlo, a.twf.sh, a.twf.mask, rr * (uint32_t)j, rr * (uint32_t)S::NT);
        }
        LayLine<T, L> lay{b * lpadlen<T, L>()};
        line_fft_tw<T, L, DIR>(reg, lds, lay, j, StageTwLds<T>{ltw});
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            C v = reg[q];
            if (sc != (T)1) v = cscale(v, sc);
            lds[(j + S::NT * q) * (LPB + 1) + b] = v;
        }
        __syncthreads();
        const auto rd = buf_rsrc_if((C *)a.out + t * a.out_dist + r0, true);
#pragma unroll
        for (int i = 0; i < L / KS; i++) buf_st(rd, soff, i * ist, lds[(k0 + KS * i) * (LPB + 1) + bb]);
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = nxt[q];
        __syncthreads();  // the next tile's exchanges overwrite the transpose
    }
}

// ---------------------------------------------------------------------------
// k_blue_line: Bluestein for M <= Lmax in one workgroup per line group.
// y[k] = conj?( a[k] * IFFT_M( FFT_M(pad(a * conj?(x))) * Hs )[k] ) * scale
// with a = e^{-i pi k^2/N} and Hs = FFT_M(h)/M (h the wrapped conj chirp).
// ---------------------------------------------------------------------------
struct BlueArgs {
    const void *in;
    void *out;
    long long nlines;
    long long in_dist, out_dist;  // line distances
    int n;                 // transform length N
    const void *tw;        // stage twiddles for M
    const void *chirp;     // N entries
    const void *H;         // M entries (natural order), includes 1/M
    int conj;
    double scale;
};

template <typename T, int M>
__global__ void __launch_bounds__(Sched<M>::NT *line_lpb<M>()) k_blue_line(BlueArgs a)
{
    using S = Sched<M>;
    using C = Cx<T>;
    constexpr int LPB = line_lpb<M>();
    __shared__ C lds[line_lds<T, M, LPB>()];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const long long line0 = (long long)blockIdx.x * LPB;
    const long long rem = a.nlines - line0;
    const int nrows = rem < LPB ? (int)rem : LPB;
    const C *src = (const C *)a.in + line0 * a.in_dist;
    C *dst = (C *)a.out + line0 * a.out_dist;
    const C *ch = (const C *)a.chirp;
    const C *H = (const C *)a.H;
    C reg[S::R];
    tile_load<T, M, LPB>(reg, lds, src, a.in_dist, a.n, nrows, b, j);
#pragma unroll
    for (int q = 0; q < S::R; q++) {
        const int p = j + S::NT * q;
        if (p < a.n) {
            C v = reg[q];
            if (a.conj) v = cconj(v);
            reg[q] = cmul(v, ch[p]);
        }
    }
    LayLine<T, M> lay{b * lpadlen<T, M>()};
    line_fft<T, M, -1>(reg, lds, lay, j, (const double2 *)a.tw);
#pragma unroll
    for (int q = 0; q < S::R; q++) reg[q] = cmul(reg[q], H[j + S::NT * q]);
    line_fft<T, M, 1>(reg, lds, lay, j, (const double2 *)a.tw);
    const T sc = (T)a.scale;
#pragma unroll
    for (int q = 0; q < S::R; q++) {
        const int p = j + S::NT * q;
        C v = p < a.n ? cmul(reg[q], ch[p]) : czero<T>();
        if (a.conj) v = cconj(v);
        reg[q] = cscale(v, sc);
    }
    if constexpr (S::NT < 64) __syncthreads();  // staged store reuses LDS
    tile_store<T, M, LPB>(reg, lds, dst, a.out_dist, a.n, nrows, b, j);
}

// ---------------------------------------------------------------------------
// k_blue_row: middle pass of the three-pass Bluestein on an M = M1 x M2
// buffer (row k1 holds M2 contiguous points): FFT over n2, x Hp[k1][k2],
// inverse FFT over k2, x w_M^{+n2 k1}; in place.
// ---------------------------------------------------------------------------
struct BlueRowArgs {
    void *buf;
    int nrows;            // M1
    long long dist;       // M per transform
    long long ntrans;
    const void *tw;       // stage twiddles for M2
    const void *Hp;       // M entries, [k1][k2] order, includes 1/M
    Tw2 tw2;              // w_M
};

template <typename T, int L, int LPB>
__global__ void __launch_bounds__(Sched<L>::NT *LPB) k_blue_row(BlueRowArgs a)
{
    using S = Sched<L>;
    using C = Cx<T>;
    __shared__ C lds[line_lds<T, L, LPB>()];
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const int r0 = blockIdx.x * LPB;
    const long long t = (long long)blockIdx.y + (long long)blockIdx.z * gridDim.y;
    if (t >= a.ntrans) return;
    const int nrows = (a.nrows - r0) < LPB ? (a.nrows - r0) : LPB;
    C *buf = (C *)a.buf + t * a.dist + (long long)r0 * L;
    C reg[S::R];
    tile_load<T, L, LPB>(reg, lds, buf, L, L, nrows, b, j);
    LayLine<T, L> lay{b * lpadlen<T, L>()};
    line_fft<T, L, -1>(reg, lds, lay, j, (const double2 *)a.tw);
    const int k1 = r0 + (b < nrows ? b : 0);
    const C *Hp = (const C *)a.Hp + (long long)k1 * L;
#pragma unroll
    for (int q = 0; q < S::R; q++) reg[q] = cmul(reg[q], Hp[j + S::NT * q]);
    line_fft<T, L, 1>(reg, lds, lay, j, (const double2 *)a.tw);
    apply_tw2_series<1>(reg, a.tw2, (uint32_t)k1 * (uint32_t)j, (uint32_t)k1 * (uint32_t)S::NT);
    if constexpr (S::NT < 64) __syncthreads();
    tile_store<T, L, LPB>(reg, lds, buf, L, L, nrows, b, j);
}

}  // namespace nfft

