// nfft_ragged.hip -- ragged batches of mixed-size 1D complex transforms in
// one persistent kernel launch, recorded into a HIP graph.
//
// The reference has no batch API at all (src/07_legacy.h:29-57); a caller
// with many transforms of different sizes loops over plans on the CPU.  Here
// the whole ragged batch is one launch:
//   - the host sorts the transforms into work items: a group of up to LPB
//     same-size short transforms (power of two <= 4096, or Bluestein with
//     M <= 4096) processed as LDS-resident lines, or one tile of a long
//     transform (power of two up to 65536: four-step, 2 phases; Bluestein
//     up to M = 131072: 3 phases of column / row / column tiles) whose
//     intermediate lives in one of a set of rotating scratch slots;
//   - the long transforms' tiles come first, in a software-pipelined order
//     (phase 1 of transform s, phase 2 of s - D, phase 3 of s - 2D), then
//     the short groups longest-first; an atomic counter hands items out;
//   - a tile waits (bounded, with an error flag) for the previous phase of
//     its transform, a phase-1 tile for the last phase of its slot's
//     previous user; both were handed out earlier, so the queue cannot
//     deadlock;
//   - the graph holds [counter reset kernel, persistent kernel]; there is no
//     memset node.
// Inverse transforms follow the reference's normalisation: unnormalised for
// powers of two, 1/N for Bluestein lengths (src/05_algorithms.h:315-333).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/notorious_fft_hip.h"
#include "nfft_internal.h"

namespace nfft {

constexpr int RG_THREADS = 256;
enum RagType { RG_LINE = 0, RG_BLUE1 = 1, RG_FOUR = 2, RG_BLUE3 = 3, RG_SMOOTH = 4, RG_SMOOTH4 = 5 };
// longest mixed-radix line a ragged workgroup holds (two LDS buffers within
// the kernel's dynamic LDS) and the points of one mixed-radix tile
constexpr int RG_SMOOTH_MAX = 2048;

// A work item: a group of short transforms (LINE, BLUE1, SMOOTH), or one
// tile of a long transform (FOUR, SMOOTH4: phase 1 column tiles, phase 2 row
// tiles; BLUE3: phases 1-3 = column, row, column tiles), covering column/row
// groups [g0, g1) of that phase.
struct RagItem {
    int type;
    int L, L2;     // LINE: n; BLUE1: M; FOUR: n1 x n2; BLUE3: m1 x m2
    int n;
    int count;     // transforms in this item (tiles: 1)
    int first;     // first entry in the id list
    int tab;       // table descriptor
    int phase;     // tiles: 1..3
    int g0, g1;    // tiles: group range
    int slot;      // tiles: intermediate slot
    int done;      // tiles: counter bumped on completion
    int wait;      // tiles: counter to wait on (-1: none)
    int need;      // tiles: count it must reach
};

struct RagTab {
    const void *tw_a, *tw_b;
    const void *chirp, *H;
    Tw2 tw2;
    SmoothPlan sp1, sp2;   // mixed radix: line (sp1) or four-step N1 (sp1) x N2 (sp2)
    int w1, w2;            // mixed-radix four-step: columns / rows per group
};

struct RagArgs {
    const void *in;
    void *out;
    const RagItem *items;
    int nitems;
    const int *ids;
    const long long *offs;
    const RagTab *tabs;
    unsigned *ctr;          // [0] queue head, [1] error, [2] sticky error, [3 + k] tile counters
    void *scratch;          // slots x slot_elems intermediate elements
    long long slot_elems;
    long long timeout;      // wall-clock ticks a dependency wait may take
};

template <int L> constexpr int rg_lines() { return RG_THREADS / Sched<L>::NT; }

// ---------------------------------------------------------------------------
// groups of short lines
// ---------------------------------------------------------------------------
template <typename T, int L, int DIR, bool BLUE>
__device__ __noinline__ void rg_line(const RagArgs &a, const RagItem &it, Cx<T> *, long long *bases)
{
    // the LDS through the kernel's own symbol (address space known: ds
    // instructions rather than flat ones)
    extern __shared__ char smem[];
    Cx<T> *lds = (Cx<T> *)smem;
    using S = Sched<L>;
    using C = Cx<T>;
    constexpr int LPB = rg_lines<L>();
    constexpr int PADL = lpadlen<T, L>();
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    const RagTab tb = a.tabs[it.tab];
    const int len = BLUE ? it.n : L;
    if (threadIdx.x < LPB) bases[threadIdx.x] = threadIdx.x < it.count ? a.offs[a.ids[it.first + threadIdx.x]] : 0;
    __syncthreads();
    const C *in = (const C *)a.in;
    C *out = (C *)a.out;
    C reg[S::R];
    if constexpr (S::NT >= 64) {
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int p = j + S::NT * q;
            reg[q] = (b < it.count && p < len) ? in[bases[b] + p] : czero<T>();
        }
    } else {
        for (int idx = threadIdx.x; idx < LPB * L; idx += RG_THREADS) {
            const int bb = idx / L, p = idx & (L - 1);
            lds[bb * PADL + lpad<T>(p)] = (bb < it.count && p < len) ? in[bases[bb] + p] : czero<T>();
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = lds[b * PADL + lpad<T>(j + S::NT * q)];
        __syncthreads();
    }
    LayLine<T, L> lay{b * PADL};
    T scale = (T)1;
    if constexpr (BLUE) {
        const C *ch = (const C *)tb.chirp;
        const C *H = (const C *)tb.H;
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int p = j + S::NT * q;
            if (p < len) reg[q] = cmul(DIR > 0 ? cconj(reg[q]) : reg[q], ch[p]);
        }
        line_fft<T, L, -1>(reg, lds, lay, j, (const double2 *)tb.tw_a);
#pragma unroll
        for (int q = 0; q < S::R; q++) reg[q] = cmul(reg[q], H[j + S::NT * q]);
        line_fft<T, L, 1>(reg, lds, lay, j, (const double2 *)tb.tw_a);
        if (DIR > 0) scale = (T)1 / (T)len;
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int p = j + S::NT * q;
            C v = p < len ? cmul(reg[q], ch[p]) : czero<T>();
            if (DIR > 0) v = cconj(v);
            reg[q] = cscale(v, scale);
        }
    } else {
        line_fft<T, L, DIR>(reg, lds, lay, j, (const double2 *)tb.tw_a);
    }
    if constexpr (S::NT >= 64) {
        if (b < it.count) {
#pragma unroll
            for (int q = 0; q < S::R; q++) {
                const int p = j + S::NT * q;
                if (p < len) out[bases[b] + p] = reg[q];
            }
        }
    } else {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < S::R; q++) lds[b * PADL + lpad<T>(j + S::NT * q)] = reg[q];
        __syncthreads();
        for (int idx = threadIdx.x; idx < LPB * L; idx += RG_THREADS) {
            const int bb = idx / L, p = idx & (L - 1);
            if (bb < it.count && p < len) out[bases[bb] + p] = lds[bb * PADL + lpad<T>(p)];
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// one long transform per workgroup
// ---------------------------------------------------------------------------
// Scratch accesses of tiles go through agent-scope (sc1) loads and stores,
// coherent across XCDs without cache maintenance (see nfft_kernels_persist.h).
// Element i of a workgroup-uniform base (i below 2^28 when COH).
template <bool COH, typename C> __device__ __forceinline__ C rg_ld(const C *base, long long i)
{
    if constexpr (COH) return coh_ld(base, (uint32_t)i);
    else return base[i];
}
template <bool COH, typename C> __device__ __forceinline__ void rg_st(C *base, long long i, C v)
{
    if constexpr (COH) coh_st(base, (uint32_t)i, v);
    else base[i] = v;
}

// column pass over an L1 x L2 matrix (row-major, L2 columns), column groups
// [g0, g1) of LPB columns each:
// MODE 0: src -> FFT_L1 -> * tw2 (c*p, DIR) -> dst (scratch)
// MODE 1: x[n] * chirp (n < nv, conj if cj) -> FFT -> * tw2 -> dst (scratch; Bluestein in)
// MODE 2: src (scratch) -> IFFT -> y[n] = conj?(v * chirp[n]) * sc, n < nv (Bluestein out)
template <typename T, int L1, int L2, int DIR, int MODE>
__device__ __forceinline__ void rg_cols(const Cx<T> *src, Cx<T> *dst, const RagTab &tb, int nv, int cj, T sc, Cx<T> *lds, int g0,
                        int g1)
{
    using S = Sched<L1>;
    using C = Cx<T>;
    constexpr int LPB = RG_THREADS / S::NT;
    static_assert(L2 % LPB == 0, "column tiles must cover the columns");
    const int b = threadIdx.x % LPB, j = threadIdx.x / LPB;
    for (int g = g0; g < g1; g++) {
        const int c = g * LPB + b;
        C reg[S::R];
#pragma unroll
        for (int q = 0; q < S::R; q++) {
            const int p = j + S::NT * q;
            const int n = p * L2 + c;
            if constexpr (MODE == 1) {
                C v = czero<T>();
                if (n < nv) {
                    v = src[n];
                    if (cj) v = cconj(v);
                    v = cmul(v, ((const C *)tb.chirp)[n]);
                }
                reg[q] = v;
            } else {
                reg[q] = rg_ld<MODE == 2>(src, n);
            }
        }
        LayCol<LPB> lay{b};
        line_fft<T, L1, DIR>(reg, lds, lay, j, (const double2 *)tb.tw_a);
        if constexpr (MODE == 2) {
#pragma unroll
            for (int q = 0; q < S::R; q++) {
                const int p = j + S::NT * q;
                const int n = p * L2 + c;
                if (n < nv) {
                    C v = cmul(reg[q], ((const C *)tb.chirp)[n]);
                    if (cj) v = cconj(v);
                    dst[n] = cscale(v, sc);
                }
            }
        } else {
            apply_tw2_series<DIR>(reg, tb.tw2, (uint32_t)c * (uint32_t)j, (uint32_t)c * (uint32_t)S::NT);
#pragma unroll
            for (int q = 0; q < S::R; q++) rg_st<true>(dst, (j + S::NT * q) * L2 + c, reg[q]);
        }
    }
}

// row pass: rows of L2 points, row groups [g0, g1) of LPB rows each.
// MODE 0: scratch -> FFT_L2 (DIR) -> transposed store dst[r + L1*k] (four-step pass 2)
// MODE 1: scratch -> FFT, * H[r][k], IFFT, * conj(tw2(r*n2)) -> scratch in place (Bluestein middle)
template <typename T, int L1, int L2, int DIR, int MODE>
__device__ __forceinline__ void rg_rows(const Cx<T> *src, Cx<T> *dst, const RagTab &tb, Cx<T> *lds, int g0, int g1)
{
    using S = Sched<L2>;
    using C = Cx<T>;
    constexpr int LPB = RG_THREADS / S::NT;
    constexpr int PADL = lpadlen<T, L2>();
    const int b = threadIdx.x / S::NT, j = threadIdx.x % S::NT;
    for (int g = g0; g < g1; g++) {
        const int r0 = g * LPB;
        const int r = r0 + b;
        C reg[S::R];
        if constexpr (S::NT >= 64) {
#pragma unroll
            for (int q = 0; q < S::R; q++) reg[q] = rg_ld<true>(src, (long long)r * L2 + j + S::NT * q);
        } else {
            for (int idx = threadIdx.x; idx < LPB * L2; idx += RG_THREADS) {
                const int bb = idx / L2, p = idx % L2;
                lds[bb * PADL + lpad<T>(p)] = rg_ld<true>(src, (long long)(r0 + bb) * L2 + p);
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < S::R; q++) reg[q] = lds[b * PADL + lpad<T>(j + S::NT * q)];
            __syncthreads();
        }
        LayLine<T, L2> lay{b * PADL};
        if constexpr (MODE == 0) {
            line_fft<T, L2, DIR>(reg, lds, lay, j, (const double2 *)tb.tw_b);
#pragma unroll
            for (int q = 0; q < S::R; q++) lds[(j + S::NT * q) * (LPB + 1) + b] = reg[q];
            __syncthreads();
            for (int idx = threadIdx.x; idx < LPB * L2; idx += RG_THREADS) {
                const int bb = idx % LPB, k = idx / LPB;
                dst[(long long)k * L1 + r0 + bb] = lds[k * (LPB + 1) + bb];
            }
            __syncthreads();
        } else {
            line_fft<T, L2, -1>(reg, lds, lay, j, (const double2 *)tb.tw_b);
            const C *Hp = (const C *)tb.H + (long long)r * L2;
#pragma unroll
            for (int q = 0; q < S::R; q++) reg[q] = cmul(reg[q], Hp[j + S::NT * q]);
            line_fft<T, L2, 1>(reg, lds, lay, j, (const double2 *)tb.tw_b);
            apply_tw2_series<1>(reg, tb.tw2, (uint32_t)r * (uint32_t)j, (uint32_t)r * (uint32_t)S::NT);
            if constexpr (S::NT >= 64) {
#pragma unroll
                for (int q = 0; q < S::R; q++) rg_st<true>(dst, (long long)r * L2 + j + S::NT * q, reg[q]);
            } else {
#pragma unroll
                for (int q = 0; q < S::R; q++) lds[b * PADL + lpad<T>(j + S::NT * q)] = reg[q];
                __syncthreads();
                for (int idx = threadIdx.x; idx < LPB * L2; idx += RG_THREADS) {
                    const int bb = idx / L2, p = idx % L2;
                    rg_st<true>(dst, (long long)(r0 + bb) * L2 + p, lds[bb * PADL + lpad<T>(p)]);
                }
                __syncthreads();
            }
        }
    }
}

// one tile of a long power-of-two transform (four-step N1 x N2)
template <typename T, int N1, int N2, int DIR>
__device__ __noinline__ void rg_four(const RagArgs &a, const RagItem &it, Cx<T> *)
{
    // the LDS through the kernel's own symbol (address space known: ds
    // instructions rather than flat ones)
    extern __shared__ char smem[];
    Cx<T> *lds = (Cx<T> *)smem;
    using C = Cx<T>;
    const RagTab tb = a.tabs[it.tab];
    const long long base = a.offs[a.ids[it.first]];
    C *scr = (C *)a.scratch + (long long)it.slot * a.slot_elems;
    if (it.phase == 1) rg_cols<T, N1, N2, DIR, 0>((const C *)a.in + base, scr, tb, 0, 0, (T)1, lds, it.g0, it.g1);
    else rg_rows<T, N1, N2, DIR, 0>(scr, (C *)a.out + base, tb, lds, it.g0, it.g1);
}

// one tile of a long Bluestein transform (three passes over M1 x M2)
template <typename T, int M1, int M2, int DIR>
__device__ __noinline__ void rg_blue3(const RagArgs &a, const RagItem &it, Cx<T> *)
{
    // the LDS through the kernel's own symbol (address space known: ds
    // instructions rather than flat ones)
    extern __shared__ char smem[];
    Cx<T> *lds = (Cx<T> *)smem;
    using C = Cx<T>;
    const RagTab tb = a.tabs[it.tab];
    const long long base = a.offs[a.ids[it.first]];
    C *scr = (C *)a.scratch + (long long)it.slot * a.slot_elems;
    const int cj = DIR > 0 ? 1 : 0;
    const T sc = DIR > 0 ? (T)1 / (T)it.n : (T)1;
    if (it.phase == 1)
        rg_cols<T, M1, M2, -1, 1>((const C *)a.in + base, scr, tb, it.n, cj, (T)1, lds, it.g0, it.g1);
    else if (it.phase == 2)
        rg_rows<T, M1, M2, -1, 1>(scr, scr, tb, lds, it.g0, it.g1);
    else
        rg_cols<T, M1, M2, 1, 2>(scr, (C *)a.out + base, tb, it.n, cj, sc, lds, it.g0, it.g1);
}

// a group of short mixed-radix transforms (same length), LDS-resident
template <typename T, int DIR>
__device__ __noinline__ void rg_smooth_group(const RagArgs &a, const RagItem &it, Cx<T> *, long long *bases)
{
    // the LDS through the kernel's own symbol (address space known: ds
    // instructions rather than flat ones)
    extern __shared__ char smem[];
    Cx<T> *lds = (Cx<T> *)smem;
    using C = Cx<T>;
    const RagTab &tb = a.tabs[it.tab];
    const int L = tb.sp1.L, G = it.count;
    if (threadIdx.x < G) bases[threadIdx.x] = a.offs[a.ids[it.first + threadIdx.x]];
    __syncthreads();
    C *A = lds, *B = lds + G * L;
    const float iL = rcp_up(L);
    for (int idx = threadIdx.x; idx < G * L; idx += RG_THREADS) {
        const int g = fdiv(idx, iL), p = idx - __mul24(g, L);
        A[idx] = ((const C *)a.in)[bases[g] + p];
    }
    __syncthreads();
    const C *res = smooth_lines<T, DIR, 25>(A, B, G, tb.sp1, (const C *)tb.tw_a);
    const T sc = DIR > 0 ? (T)1 / (T)L : (T)1;  // inverse: 1/N for non-powers of two
    for (int idx = threadIdx.x; idx < G * L; idx += RG_THREADS) {
        const int g = fdiv(idx, iL), p = idx - __mul24(g, L);
        ((C *)a.out)[bases[g] + p] = cscale(res[idx], sc);
    }
    __syncthreads();
}

// one tile of a long mixed-radix transform (four-step N1 x N2)
template <typename T, int DIR>
__device__ __noinline__ void rg_smooth4(const RagArgs &a, const RagItem &it, Cx<T> *)
{
    // the LDS through the kernel's own symbol (address space known: ds
    // instructions rather than flat ones)
    extern __shared__ char smem[];
    Cx<T> *lds = (Cx<T> *)smem;
    using C = Cx<T>;
    const RagTab &tb = a.tabs[it.tab];
    const long long base = a.offs[a.ids[it.first]];
    C *scr = (C *)a.scratch + (long long)it.slot * a.slot_elems;
    const int N1 = tb.sp1.L, N2 = tb.sp2.L;
    if (it.phase == 1) {
        for (int g = it.g0; g < it.g1; g++) {
            const int c0 = g * tb.w1, w = (N2 - c0) < tb.w1 ? (N2 - c0) : tb.w1;
            smooth_col_tile<T, DIR, true, 25>((const C *)a.in + base, scr, tb.sp1, N2, c0, w, (const C *)tb.tw_a, tb.tw2,
                                          lds, lds + w * N1);
        }
    } else {
        const T sc = DIR > 0 ? (T)1 / (T)((long long)N1 * N2) : (T)1;
        for (int g = it.g0; g < it.g1; g++) {
            const int r0 = g * tb.w2, w = (N1 - r0) < tb.w2 ? (N1 - r0) : tb.w2;
            smooth_row_tile<T, DIR, true, 25>(scr, (C *)a.out + base, tb.sp2, N1, r0, w, (const C *)tb.tw_b, sc, lds,
                                          lds + w * N2);
        }
    }
}

// ---------------------------------------------------------------------------
// persistent kernel
// ---------------------------------------------------------------------------
// The item bodies are separate (not inlined) functions: inlined into one
// switch they shared a register allocation of 395 (FP32) / 512 (FP64) VGPRs,
// which left a single workgroup per CU.  Called, each keeps its own, and the
// kernel fits 4 (FP32) / 2 (FP64) workgroups per CU.
template <typename T, int DIR>
__global__ void __launch_bounds__(RG_THREADS, (sizeof(T) == 4 ? 4 : 2)) k_ragged(RagArgs a)
{
    extern __shared__ char smem[];
    Cx<T> *lds = (Cx<T> *)smem;
    __shared__ long long bases[RG_THREADS];
    __shared__ int cur;
    unsigned *head = a.ctr, *err = a.ctr + 1, *tiles = a.ctr + 3;
    // Lane 0 of wave 0 publishes the finished tile and fetches the next item
    // in one region behind a wave-uniform branch (a second lane-divergent
    // region between this loop's barriers lets the compiler reorder lane 0's
    // work past the next barrier).  Tiles wait only for items earlier in the
    // queue, held by running workgroups, so the queue cannot deadlock; each
    // wait is bounded and raises the error flag instead of spinning.
    const bool wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x) == 0;
    int done = -1;  // lane 0: counter of the tile just finished
    for (long long iter = 0;; iter++) {
        if (iter > a.nitems) {  // cannot happen: a workgroup takes at most nitems items
            if (threadIdx.x == 0) atomicOr(err, 2u);
            break;
        }
        if (wave0) {
            if (threadIdx.x == 0) {
                if (done >= 0) __hip_atomic_fetch_add(tiles + done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const int i = (int)atomicAdd(head, 1u);
                done = -1;
                if (i < a.nitems) {
                    const RagItem &q = a.items[i];
                    if (q.wait >= 0) persist_wait(tiles + q.wait, (unsigned)q.need, err, a.timeout);
                    done = q.done;
                }
                cur = i;
            }
        }
        __syncthreads();
        const int i = __builtin_amdgcn_readfirstlane(cur);
        if (i >= a.nitems) break;  // uniform: the queue is drained
        const RagItem it = a.items[i];
        switch (it.type) {
        case RG_LINE:
            switch (it.L) {
#define NFFT_RG(LL) case LL: rg_line<T, LL, DIR, false>(a, it, lds, bases); break;
                NFFT_RG(1) NFFT_RG(2) NFFT_RG(4) NFFT_RG(8) NFFT_RG(16) NFFT_RG(32) NFFT_RG(64) NFFT_RG(128)
                NFFT_RG(256) NFFT_RG(512) NFFT_RG(1024) NFFT_RG(2048) NFFT_RG(4096)
#undef NFFT_RG
            }
            break;
        case RG_BLUE1:
            switch (it.L) {
#define NFFT_RG(LL) case LL: rg_line<T, LL, DIR, true>(a, it, lds, bases); break;
                NFFT_RG(4) NFFT_RG(8) NFFT_RG(16) NFFT_RG(32) NFFT_RG(64) NFFT_RG(128) NFFT_RG(256) NFFT_RG(512)
                NFFT_RG(1024) NFFT_RG(2048) NFFT_RG(4096)
#undef NFFT_RG
            }
            break;
        case RG_FOUR:
            switch (it.L * 100000 + it.L2) {
            case 64 * 100000 + 128: rg_four<T, 64, 128, DIR>(a, it, lds); break;
            case 128 * 100000 + 128: rg_four<T, 128, 128, DIR>(a, it, lds); break;
            case 128 * 100000 + 256: rg_four<T, 128, 256, DIR>(a, it, lds); break;
            case 256 * 100000 + 256: rg_four<T, 256, 256, DIR>(a, it, lds); break;
            }
            break;
        case RG_SMOOTH: rg_smooth_group<T, DIR>(a, it, lds, bases); break;
        case RG_SMOOTH4: rg_smooth4<T, DIR>(a, it, lds); break;
        case RG_BLUE3:
            switch (it.L * 100000 + it.L2) {
            case 64 * 100000 + 128: rg_blue3<T, 64, 128, DIR>(a, it, lds); break;
            case 128 * 100000 + 128: rg_blue3<T, 128, 128, DIR>(a, it, lds); break;
            case 128 * 100000 + 256: rg_blue3<T, 128, 256, DIR>(a, it, lds); break;
            case 256 * 100000 + 256: rg_blue3<T, 256, 256, DIR>(a, it, lds); break;
            case 512 * 100000 + 256: rg_blue3<T, 512, 256, DIR>(a, it, lds); break;
            }
            break;
        }
        // tiles that wrote the intermediate: each wave's stores are complete
        // (acknowledged at agent scope) before the barrier that precedes
        // lane 0's counter increment
        if (((it.type == RG_FOUR || it.type == RG_SMOOTH4) && it.phase == 1) || (it.type == RG_BLUE3 && it.phase < 3))
            __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
}

// LDS bytes of the largest path: 256 threads x 16 points (+ padding / the
// transposed tile's extra column)
static size_t rg_lds_bytes(int prec)
{
    const size_t e = prec == F32 ? 8 : 16;
    const size_t pts = (size_t)RG_THREADS * 16;
    const size_t pad = prec == F32 ? pts / 16 : pts / 8;
    return (pts + std::max(pad, (size_t)256 + 16)) * e;
}

}  // namespace nfft

// ===========================================================================
using namespace nfft;

namespace {
bool is_pow2(long long n) { return n > 0 && (n & (n - 1)) == 0; }
int ilog2i(long long n)
{
    int l = 0;
    while ((1LL << l) < n) l++;
    return l;
}
int env_int_rg(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}
// Tile size (column/row groups per tile) and pipeline-lag scale (quarters of
// the default lag) of the batch being created: set by the autotuner around
// its candidate builds and from a tuning record (0: the defaults).
struct RagShape {
    int gpt = 0, lag4 = 0;
};
thread_local RagShape t_rag_shape;
// splits of the long transforms (see the kernel's dispatch table)
bool four_split(int n, int *n1, int *n2)
{
    switch (n) {
    case 8192: *n1 = 64; *n2 = 128; return true;
    case 16384: *n1 = 128; *n2 = 128; return true;
    case 32768: *n1 = 128; *n2 = 256; return true;
    case 65536: *n1 = 256; *n2 = 256; return true;
    case 131072: *n1 = 512; *n2 = 256; return true;
    }
    return false;
}

struct GraphKey {
    const void *in;
    void *out;
    int inv;
    bool operator<(const GraphKey &o) const
    {
        if (in != o.in) return in < o.in;
        if (out != o.out) return out < o.out;
        return inv < o.inv;
    }
};
}  // namespace

struct nfft_ragged {
    int prec = 0;
    int count = 0;
    long long total = 0;
    std::vector<C2C> engines;  // per distinct size (tables)
    void *d_items = nullptr, *d_ids = nullptr, *d_offs = nullptr, *d_tabs = nullptr;
    unsigned *ctr = nullptr;   // queue head, error flags, tile counters
    long long ctr_words = 0;
    void *scratch = nullptr;   // rotating intermediate slots of the long transforms
    long long slot_elems = 0;
    long long timeout = 0;
    int nitems = 0;
    int grid = 0;
    size_t lds = 0;
    hipStream_t cap = nullptr;
    std::map<GraphKey, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    std::vector<void *> owned;  // per-size Bluestein H buffers
};

static int launch_ragged(nfft_ragged *r, int inverse, const void *in, void *out, hipStream_t s)
{
    RagArgs a{in, out, (const RagItem *)r->d_items, r->nitems, (const int *)r->d_ids,
              (const long long *)r->d_offs, (const RagTab *)r->d_tabs, r->ctr, r->scratch,
              r->slot_elems, r->timeout};
    if (launch_persist_reset(r->ctr, r->ctr_words, s) != hipSuccess) return NFFT_E_HIP;
    if (r->prec == F32) {
        if (inverse) hipLaunchKernelGGL((k_ragged<float, 1>), dim3(r->grid), dim3(RG_THREADS), r->lds, s, a);
        else hipLaunchKernelGGL((k_ragged<float, -1>), dim3(r->grid), dim3(RG_THREADS), r->lds, s, a);
    } else {
        if (inverse) hipLaunchKernelGGL((k_ragged<double, 1>), dim3(r->grid), dim3(RG_THREADS), r->lds, s, a);
        else hipLaunchKernelGGL((k_ragged<double, -1>), dim3(r->grid), dim3(RG_THREADS), r->lds, s, a);
    }
    return hipGetLastError() == hipSuccess ? NFFT_OK : NFFT_E_HIP;
}

static void ragged_free(nfft_ragged *r)
{
    for (auto &kv : r->graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    for (auto &e : r->engines) e.release();
    for (void *p : r->owned) (void)hipFree(p);
    void *ptrs[] = {r->d_items, r->d_ids, r->d_offs, r->d_tabs, r->ctr, r->scratch};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (r->cap) (void)hipStreamDestroy(r->cap);
    delete r;
}

// Workgroups per CU for the persistent grid: from the tuning cache file
// (NFFT_TUNE_CACHE) when present, else 4 (FP32) / 2 (FP64).
static std::string tune_key(int prec, const std::vector<int> &sizes)
{
    unsigned long long h = 1469598103934665603ULL;
    for (int v : sizes) h = (h ^ (unsigned long long)v) * 1099511628211ULL;
    char b[96];
    snprintf(b, sizeof b, "ragged p%d c%zu h%016llx", prec, sizes.size(), h);
    return b;
}

static int tune_lookup(const std::string &key)
{
    const char *path = getenv("NFFT_TUNE_CACHE");
    if (!path || !*path) return 0;
    FILE *f = fopen(path, "r");
    if (!f) return 0;
    char line[256];
    int found = 0;
    const size_t kl = key.size();
    while (fgets(line, sizeof line, f)) {
        if (strncmp(line, key.c_str(), kl) == 0 && line[kl] == ' ') found = atoi(line + kl + 1);
    }
    fclose(f);
    return found;
}

static void tune_store(const std::string &key, int val)
{
    const char *path = getenv("NFFT_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %d\n", key.c_str(), val);
    fclose(f);
}

// Status of a failed table / engine setup: NFFT_E_ALLOC when a device
// allocation failed since `fails0` (alloc_failures()), else NFFT_E_HIP.
static int setup_fail(long long fails0) { return alloc_failures() > fails0 ? NFFT_E_ALLOC : NFFT_E_HIP; }

static int ragged_create_impl(nfft_ragged **out, int prec, int count, const int *sizes, int flags)
{
    clear_error();
    const long long fails0 = alloc_failures();
    if (!out || count < 1 || !sizes || (prec != NFFT_F32 && prec != NFFT_F64)) {
        set_error("invalid ragged arguments");
        return NFFT_E_INVALID;
    }
    *out = nullptr;
    nfft_ragged *r = new (std::nothrow) nfft_ragged();
    if (!r) return NFFT_E_ALLOC;
    r->prec = prec;
    r->count = count;
    const size_t cb = cx_bytes(prec);
    std::vector<long long> offs(count);
    std::map<int, int> size_tab;  // size -> table index
    std::map<int, int> size_kind; // size -> item type
    std::vector<RagTab> tabs;
    // every size is checked before any device work
    for (int i = 0; i < count; i++) {
        const int n = sizes[i];
        if (n < 1) {
            set_error("ragged sizes must be >= 1");
            ragged_free(r);
            return NFFT_E_SIZE;
        }
        long long m = 1;
        while (m < 2LL * n - 1) m <<= 1;
        if ((is_pow2(n) && n > 65536) || (!is_pow2(n) && m > 131072)) {
            set_error("ragged batches support N <= 65536");
            ragged_free(r);
            return NFFT_E_UNSUPPORTED;
        }
    }
    for (int i = 0; i < count; i++) {
        const int n = sizes[i];
        offs[i] = r->total;
        r->total += n;
        if (size_tab.count(n)) continue;
        // tables for this size (twiddles, chirp, H) come from a plain engine
        C2C e;
        const bool p2 = is_pow2(n);
        long long m = 1;
        while (m < 2LL * n - 1) m <<= 1;
        if ((p2 && n > 65536) || (!p2 && m > 131072)) {
            set_error("ragged batches support N <= 65536");
            ragged_free(r);
            return NFFT_E_UNSUPPORTED;
        }
        RagTab t{};
        SmoothPlan sp1{}, sp2{};
        const bool smooth_line = !p2 && n <= RG_SMOOTH_MAX && smooth_factor(n, &sp1, 25);
        const bool smooth_four = !p2 && !smooth_line && smooth_split(n, RG_SMOOTH_MAX, &sp1, &sp2, 25);
        if (smooth_line) {
            size_kind[n] = RG_SMOOTH;
            t.sp1 = sp1;
            t.tw_a = phase_table(prec, n, 1, 0, n);
            if (!t.tw_a) { ragged_free(r); return setup_fail(fails0); }
        } else if (smooth_four) {
            size_kind[n] = RG_SMOOTH4;
            t.sp1 = sp1;
            t.sp2 = sp2;
            auto balanced = [](int len, int wmax) {  // near-equal tile widths <= wmax
                wmax = std::max(1, std::min(len, wmax));
                const int tiles = (len + wmax - 1) / wmax;
                return (len + tiles - 1) / tiles;
            };
            t.w1 = balanced(sp2.L, RG_SMOOTH_MAX / sp1.L);
            t.w2 = balanced(sp1.L, RG_SMOOTH_MAX / sp2.L);
            t.tw_a = phase_table(prec, sp1.L, 1, 0, sp1.L);
            t.tw_b = phase_table(prec, sp2.L, 1, 0, sp2.L);
            if (!t.tw_a || !t.tw_b || !tw2_table(n, &t.tw2)) { ragged_free(r); return setup_fail(fails0); }
        } else if (p2 && n <= 4096) {
            size_kind[n] = RG_LINE;
            t.tw_a = stage_table(prec, n);
            if (!t.tw_a) { ragged_free(r); return setup_fail(fails0); }
        } else if (p2) {
            size_kind[n] = RG_FOUR;
            int n1, n2;
            four_split(n, &n1, &n2);
            t.tw_a = stage_table(prec, n1);
            t.tw_b = stage_table(prec, n2);
            if (!t.tw_a || !t.tw_b || !tw2_table(n, &t.tw2)) { ragged_free(r); return setup_fail(fails0); }
        } else {
            // Bluestein: the engine builds chirp and H in the layout it uses
            // for M <= line limit (natural) or M1 x M2 (permuted); the ragged
            // kernel needs natural order for M <= 4096 and the [k1][k2] order
            // for its own M1 x M2 split, so H is rebuilt here when they differ.
            if (!e.init(n, prec, 0)) { ragged_free(r); return setup_fail(fails0); }
            t.chirp = e.chirp;
            size_kind[n] = m <= 4096 ? RG_BLUE1 : RG_BLUE3;
            if (m <= 4096) {
                t.tw_a = stage_table(prec, (int)m);
                if (e.algo != ALG_BLUE1) { set_error("unexpected Bluestein layout"); ragged_free(r); return setup_fail(fails0); }
                t.H = e.H;
            } else {
                int m1, m2;
                four_split((int)m, &m1, &m2);
                t.tw_a = stage_table(prec, m1);
                t.tw_b = stage_table(prec, m2);
                if (!t.tw_a || !t.tw_b || !tw2_table(m, &t.tw2)) { ragged_free(r); return setup_fail(fails0); }
                // permute H (natural or engine-permuted) into [k1][k2] for m1 x m2
                void *Hn = nullptr;
                if (dev_malloc(&Hn, (size_t)m * cb) != hipSuccess) { ragged_free(r); return NFFT_E_ALLOC; }
                r->owned.push_back(Hn);
                // recompute H in FP64 from the definition with this split
                C2C h64;
                void *hb = nullptr, *Hb = nullptr, *sb = nullptr;
                const double2 *chirp64 = nullptr;
                chirp_table(prec, n, &chirp64);
                hipStream_t st = nullptr;
                (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
                bool ok = h64.init((int)m, F64) && dev_malloc(&hb, (size_t)m * 16) == hipSuccess &&
                          dev_malloc(&Hb, (size_t)m * 16) == hipSuccess;
                if (ok && h64.scratch_per_line) ok = dev_malloc(&sb, (size_t)h64.scratch_per_line * 16) == hipSuccess;
                ok = ok && launch_blue_filter((double2 *)hb, chirp64, n, (int)m, st) == hipSuccess &&
                     h64.run(hb, Hb, 1, m, m, -1, 1.0, sb, 1, st) == hipSuccess &&
                     launch_blue_hprep((const double2 *)Hb, Hn, prec, (int)m, m1, m2, 1.0 / (double)m, st) ==
                         hipSuccess &&
                     hipStreamSynchronize(st) == hipSuccess;
                if (hb) (void)hipFree(hb);
                if (Hb) (void)hipFree(Hb);
                if (sb) (void)hipFree(sb);
                h64.release();
                if (st) (void)hipStreamDestroy(st);
                if (!ok) { set_error("ragged Bluestein table setup failed"); ragged_free(r); return setup_fail(fails0); }
                t.H = Hn;
            }
            r->engines.push_back(e);  // keeps chirp/H alive
        }
        size_tab[n] = (int)tabs.size();
        tabs.push_back(t);
    }
    // work items: groups of short transforms, and tiles of the long ones
    // (FOUR, BLUE3), each long transform split into phases of column/row
    // tiles that many workgroups share
    struct Long {
        int id, type, L, L2, n, tab, phases;
        int groups[3], tiles[3];  // per phase
        long long elems;
        double cost;
    };
    std::vector<Long> longs;
    std::map<std::pair<int, int>, std::vector<int>> groups;  // (type, size) -> ids
    for (int i = 0; i < count; i++) {
        const int n = sizes[i];
        const int kind = size_kind[n];
        long long m = 1;
        while (m < 2LL * n - 1) m <<= 1;
        Long l{};
        l.id = i;
        l.type = kind;
        l.n = n;
        l.tab = size_tab[n];
        if (kind == RG_FOUR) {
            four_split(n, &l.L, &l.L2);
            l.phases = 2;
            l.groups[0] = l.groups[1] = n / (RG_THREADS * 16);
            l.elems = n;
            l.cost = 4.0 * n;
            longs.push_back(l);
        } else if (kind == RG_BLUE3) {
            four_split((int)m, &l.L, &l.L2);
            l.phases = 3;
            l.groups[0] = l.groups[1] = l.groups[2] = (int)(m / (RG_THREADS * 16));
            l.elems = m;
            l.cost = 6.0 * m;
            longs.push_back(l);
        } else if (kind == RG_SMOOTH4) {
            const RagTab &t = tabs[size_tab[n]];
            l.L = t.sp1.L;
            l.L2 = t.sp2.L;
            l.phases = 2;
            l.groups[0] = (t.sp2.L + t.w1 - 1) / t.w1;
            l.groups[1] = (t.sp1.L + t.w2 - 1) / t.w2;
            l.elems = n;
            l.cost = 4.0 * n;
            longs.push_back(l);
        } else {
            groups[{kind, n}].push_back(i);
        }
    }
    std::stable_sort(longs.begin(), longs.end(), [](const Long &a, const Long &b) { return a.cost > b.cost; });
    // column/row groups per tile: 16 measured best on the 10,000-size batch
    // (FP32 4 / 8 / 16 / 32 groups: 2.65 / 2.55 / 2.52 / 2.61 ms)
    const int gpt = std::max(1, env_int_rg("NFFT_RAGGED_GPT", t_rag_shape.gpt > 0 ? t_rag_shape.gpt : 16));
    long long slot_elems = 0;
    for (Long &l : longs) {
        for (int ph = 0; ph < l.phases; ph++) l.tiles[ph] = (l.groups[ph] + gpt - 1) / gpt;
        slot_elems = std::max(slot_elems, l.elems);
    }
    const long long B = (long long)longs.size();
    // rotating slots within a memory budget; the pipeline lag D (in long
    // transforms) must stay below slots/2 so that a slot's next user waits
    // for work handed out well before it
    long long slots = 0;
    int lag = 0;
    if (B > 0) {
        // Far enough that a later phase seldom finds the earlier one still
        // running: 3/4 of the workgroups the default grid keeps in flight
        // (measured on the 10,000-size batch: FP32 lag 150 -> 768 took it
        // from 3.42 to 2.65 ms, FP64 6.6 -> 6.1 ms at 384; no gain beyond).
        // The default slot budget gives the slots four times the lag, so
        // that a phase-1 tile seldom waits for its slot's previous user
        // (with slots near 2 x lag it did, and lost the gain).
        int dev0 = 0, ncu0 = 256;
        (void)hipGetDevice(&dev0);
        (void)hipDeviceGetAttribute(&ncu0, hipDeviceAttributeMultiprocessorCount, dev0);
        int dflt = std::max(16, ncu0 * (prec == F32 ? 4 : 2) * 3 / 4);
        if (t_rag_shape.lag4 > 0) dflt = std::max(16, dflt * t_rag_shape.lag4 / 4);
        const long long slot_bytes = slot_elems * (long long)cb;
        const long long want_mb = std::max(512LL, std::min(4096LL, (4LL * dflt * slot_bytes) >> 20));
        const long long budget = (long long)std::max(16, env_int_rg("NFFT_RAGGED_MB", (int)want_mb)) << 20;
        slots = std::max(1LL, std::min(B, budget / slot_bytes));
        lag = std::min<long long>(std::max(0, env_int_rg("NFFT_RAGGED_LAG", dflt)), (slots - 1) / 2);
    }
    std::vector<RagItem> sorted;
    std::vector<int> ids;
    std::vector<int> first(B);
    for (long long b = 0; b < B; b++) {
        first[b] = (int)ids.size();
        ids.push_back(longs[b].id);
    }
    auto emit = [&](long long b, int phase) {
        const Long &l = longs[b];
        for (int t = 0; t < l.tiles[phase - 1]; t++) {
            RagItem it{};
            it.type = l.type;
            it.L = l.L;
            it.L2 = l.L2;
            it.n = l.n;
            it.count = 1;
            it.first = first[b];
            it.tab = l.tab;
            it.phase = phase;
            it.g0 = t * gpt;
            it.g1 = std::min(l.groups[phase - 1], (t + 1) * gpt);
            it.slot = (int)(b % slots);
            it.done = (int)(3 * b + phase - 1);
            if (phase > 1) {
                it.wait = (int)(3 * b + phase - 2);
                it.need = l.tiles[phase - 2];
            } else if (b >= slots) {
                const Long &prev = longs[b - slots];
                it.wait = (int)(3 * (b - slots) + prev.phases - 1);
                it.need = prev.tiles[prev.phases - 1];
            } else {
                it.wait = -1;
            }
            sorted.push_back(it);
        }
    };
    for (long long st = 0; st < B + 2LL * lag; st++) {
        if (st < B) emit(st, 1);
        if (st - lag >= 0 && st - lag < B) emit(st - lag, 2);
        if (st - 2 * lag >= 0 && st - 2 * lag < B && longs[st - 2 * lag].phases == 3) emit(st - 2 * lag, 3);
    }
    std::vector<RagItem> short_items;
    std::vector<double> cost;
    for (auto &kv : groups) {
        const int type = kv.first.first, n = kv.first.second;
        const std::vector<int> &v = kv.second;
        long long m = 1;
        while (m < 2LL * n - 1) m <<= 1;
        const int L = type == RG_BLUE1 ? (int)m : n;
        int per = RG_THREADS / std::max(1, L / 16);
        if (L < 16) per = RG_THREADS;
        if (type == RG_SMOOTH) per = std::max(1, std::min(RG_THREADS, RG_SMOOTH_MAX / n));
        const double c1 = type == RG_BLUE1 ? 2.0 * n + 0.5 * m : 2.0 * n;
        for (size_t k = 0; k < v.size(); k += per) {
            RagItem it{};
            it.type = type;
            it.L = L;
            it.n = n;
            it.count = (int)std::min<size_t>(per, v.size() - k);
            it.first = (int)ids.size();
            it.tab = size_tab[n];
            it.done = -1;
            it.wait = -1;
            for (size_t q = k; q < k + it.count; q++) ids.push_back(v[q]);
            short_items.push_back(it);
            cost.push_back(c1 * it.count);
        }
    }
    // short groups after the long transforms' tiles, longest first
    std::vector<int> order(short_items.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    for (int o : order) sorted.push_back(short_items[o]);
    r->nitems = (int)sorted.size();
    int dev = 0, ncu = 256, khz = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    r->timeout = (long long)(2.0 * 1000.0 * std::max(khz, 1000));
    r->lds = rg_lds_bytes(prec);
    {
        const int lds = (int)r->lds;
        hipError_t e1, e2;
        if (prec == F32) {
            e1 = hipFuncSetAttribute((const void *)k_ragged<float, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            e2 = hipFuncSetAttribute((const void *)k_ragged<float, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        } else {
            e1 = hipFuncSetAttribute((const void *)k_ragged<double, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            e2 = hipFuncSetAttribute((const void *)k_ragged<double, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        }
        (void)e1;
        (void)e2;
        (void)hipGetLastError();
    }
    const std::string key = tune_key(prec, std::vector<int>(sizes, sizes + count));
    int per_cu = tune_lookup(key);
    const bool recorded = per_cu > 0;
    if (per_cu <= 0) per_cu = prec == F32 ? 4 : 2;
    r->grid = std::max(1, std::min(r->nitems, ncu * per_cu));
    r->slot_elems = slot_elems;
    r->ctr_words = persist_ctr_words(0) + 3 * B;
    bool ok = dev_malloc(&r->d_items, sorted.size() * sizeof(RagItem)) == hipSuccess &&
              dev_malloc(&r->d_ids, ids.size() * sizeof(int)) == hipSuccess &&
              dev_malloc(&r->d_offs, offs.size() * sizeof(long long)) == hipSuccess &&
              dev_malloc(&r->d_tabs, tabs.size() * sizeof(RagTab)) == hipSuccess &&
              dev_malloc((void **)&r->ctr, (size_t)r->ctr_words * sizeof(unsigned)) == hipSuccess &&
              hipMemset(r->ctr, 0, (size_t)r->ctr_words * sizeof(unsigned)) == hipSuccess;
    if (ok && B > 0) ok = dev_malloc(&r->scratch, (size_t)slots * slot_elems * cb) == hipSuccess;
    ok = ok && hipMemcpy(r->d_items, sorted.data(), sorted.size() * sizeof(RagItem), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(r->d_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(r->d_offs, offs.data(), offs.size() * sizeof(long long), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(r->d_tabs, tabs.data(), tabs.size() * sizeof(RagTab), hipMemcpyHostToDevice) == hipSuccess &&
         hipStreamCreateWithFlags(&r->cap, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        set_error("ragged setup: device allocation or copy failed");
        ragged_free(r);
        return NFFT_E_ALLOC;
    }
    if ((flags & 1) && !recorded) {
        // project-written autotuner: time the persistent launch for a few
        // workgroups-per-CU choices on a scratch batch and keep the fastest
        void *tin = nullptr, *tout = nullptr;
        if (dev_malloc(&tin, (size_t)r->total * cb) == hipSuccess && dev_malloc(&tout, (size_t)r->total * cb) == hipSuccess) {
            (void)launch_fill_uniform(tin, 2 * r->total, 7, 0, prec, r->cap);
            hipEvent_t e0, e1;
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            float best = 1e30f;
            int best_pc = per_cu;
            // at most what the kernel's register budget keeps resident
            const int cands[] = {1, 2, 3, 4};
            for (int pc : cands) {
                if (prec == F64 && pc > 2) continue;
                r->grid = std::max(1, std::min(r->nitems, ncu * pc));
                (void)launch_ragged(r, 0, tin, tout, r->cap);
                (void)hipEventRecord(e0, r->cap);
                for (int k = 0; k < 3; k++) (void)launch_ragged(r, 0, tin, tout, r->cap);
                (void)hipEventRecord(e1, r->cap);
                (void)hipEventSynchronize(e1);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms < best) { best = ms; best_pc = pc; }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            per_cu = best_pc;
            r->grid = std::max(1, std::min(r->nitems, ncu * per_cu));
            tune_store(key, per_cu);
        }
        if (tin) (void)hipFree(tin);
        if (tout) (void)hipFree(tout);
    }
    *out = r;
    return NFFT_OK;
}

// Project-written autotuner for the tile size and the pipeline lag: builds
// the batch with each candidate (16 groups per tile and the default lag, and
// the neighbours measured on the 10,000-size batch), times four launches of
// each on a scratch batch and keeps the fastest ("ragged-gl" record).
static int ragged_tune_shape(int prec, int count, const int *sizes, const std::string &key, RagShape *best)
{
    const size_t cb = cx_bytes(prec);
    long long total = 0;
    for (int i = 0; i < count; i++) total += sizes[i];
    void *tin = nullptr, *tout = nullptr;
    if (dev_malloc(&tin, (size_t)total * cb) != hipSuccess || dev_malloc(&tout, (size_t)total * cb) != hipSuccess) {
        if (tin) (void)hipFree(tin);
        return NFFT_E_ALLOC;
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
              hipEventCreate(&e1) == hipSuccess && launch_fill_uniform(tin, 2 * total, 7, 0, prec, st) == hipSuccess;
    float best_ms = 1e30f;
    int fail = NFFT_E_HIP;
    const RagShape cands[] = {{16, 4}, {8, 4}, {32, 4}, {16, 2}, {16, 8}};
    for (const RagShape &c : cands) {
        if (!ok) break;
        t_rag_shape = c;
        nfft_ragged *r = nullptr;
        const int rc = ragged_create_impl(&r, prec, count, sizes, 0);
        t_rag_shape = RagShape{};
        if (rc != NFFT_OK) {
            if (rc == NFFT_E_ALLOC) continue;  // this shape's slots did not fit
            fail = rc;
            ok = false;
            break;
        }
        float ms = 0;
        unsigned ew[2] = {1, 1};
        ok = launch_ragged(r, 0, tin, tout, st) == NFFT_OK && hipEventRecord(e0, st) == hipSuccess;
        for (int k = 0; k < 4 && ok; k++) ok = launch_ragged(r, 0, tin, tout, st) == NFFT_OK;
        ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
             hipMemcpy(ew, r->ctr + 1, sizeof ew, hipMemcpyDeviceToHost) == hipSuccess && ew[0] == 0 && ew[1] == 0;
        ragged_free(r);
        if (ok && ms < best_ms) {
            best_ms = ms;
            *best = c;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(tin);
    (void)hipFree(tout);
    if (!ok || best_ms >= 1e30f) {
        (void)hipGetLastError();
        return fail;
    }
    tune_store(key, best->gpt * 100 + best->lag4);
    return NFFT_OK;
}

extern "C" {

int nfft_ragged_create(nfft_ragged **out, int prec, int count, const int *sizes, int flags)
{
    clear_error();
    if (!out || count < 1 || !sizes || (prec != NFFT_F32 && prec != NFFT_F64))
        return ragged_create_impl(out, prec, count, sizes, flags);  // reports the argument error
    // tile size and lag: a tuning record, else measured with flags bit 0
    RagShape shape;
    const std::string key = "ragged-gl" + tune_key(prec, std::vector<int>(sizes, sizes + count)).substr(6);
    const int rec = tune_lookup(key);
    if (rec > 0) {
        shape.gpt = rec / 100;
        shape.lag4 = rec % 100;
    } else if (flags & 1) {
        // the sizes are validated by the first candidate build
        const int rc = ragged_tune_shape(prec, count, sizes, key, &shape);
        if (rc == NFFT_E_SIZE || rc == NFFT_E_UNSUPPORTED || rc == NFFT_E_INVALID) return rc;
        if (rc != NFFT_OK) shape = RagShape{};  // nothing measured: the defaults
        clear_error();
    }
    t_rag_shape = shape;
    const int rc = ragged_create_impl(out, prec, count, sizes, flags);
    t_rag_shape = RagShape{};
    return rc;
}

int nfft_ragged_execute(nfft_ragged *r, int inverse, const void *in, void *out, void *stream)
{
    clear_error();
    if (!r || !in || !out) {
        set_error("NULL ragged handle or buffer");
        return NFFT_E_INVALID;
    }
    GraphKey k{in, out, inverse ? 1 : 0};
    auto it = r->graphs.find(k);
    if (it == r->graphs.end()) {
        if (r->graphs.size() >= 8) {
            (void)hipGraphExecDestroy(r->graphs.begin()->second.second);
            (void)hipGraphDestroy(r->graphs.begin()->second.first);
            r->graphs.erase(r->graphs.begin());
        }
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        if (hipStreamBeginCapture(r->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            set_error("graph capture failed to start");
            return NFFT_E_HIP;
        }
        const int st = launch_ragged(r, k.inv, in, out, r->cap);
        if (hipStreamEndCapture(r->cap, &g) != hipSuccess || st != NFFT_OK ||
            hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            set_error("graph capture of the ragged launch failed");
            return NFFT_E_HIP;
        }
        it = r->graphs.emplace(k, std::make_pair(g, ge)).first;
    }
    if (hipGraphLaunch(it->second.second, (hipStream_t)stream) != hipSuccess) {
        set_error("graph launch failed");
        return NFFT_E_HIP;
    }
    return NFFT_OK;
}

int nfft_ragged_destroy(nfft_ragged *r)
{
    if (r) ragged_free(r);
    return NFFT_OK;
}

long long nfft_ragged_total(const nfft_ragged *r) { return r ? r->total : 0; }

int nfft_ragged_check(nfft_ragged *r, void *stream)
{
    clear_error();
    if (!r) {
        set_error("NULL ragged handle");
        return NFFT_E_INVALID;
    }
    unsigned w[2] = {0, 0};
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess ||
        hipMemcpy(w, r->ctr + 1, sizeof w, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("hipStreamSynchronize / hipMemcpy failed");
        return NFFT_E_HIP;
    }
    if (w[0] | w[1]) {
        set_error("ragged batch: a tile dependency wait timed out");
        return NFFT_E_HIP;
    }
    return NFFT_OK;
}

}  // extern "C"
