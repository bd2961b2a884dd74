// nfft_kernels_persist.h -- the whole four-step of a batch in one launch.
//
// The two-launch four-step (k_col with twiddle, then k_rowT) streams the
// N-point intermediate of every transform through HBM twice: its ceiling is
// half the HBM bandwidth.  k_four_persist keeps the intermediate of only a
// few transforms alive at a time, in rotating scratch slots small enough to
// stay in the 256 MB Infinity Cache, so that HBM sees one read of the input
// and one write of the output.
//
// Work items are handed out in pipeline order by an atomic head counter (a
// workgroup takes `grab` consecutive items per access and runs them in order):
// step s holds the A column tiles of transform s (pass 1) followed by the B
// row tiles of transform s - D (pass 2).  A pass-2 tile of transform t waits
// until all A pass-1 tiles of t are published; a pass-1 tile of t waits until
// the B pass-2 tiles of t - S (the previous user of its slot, S > D slots)
// are done.  Both waits then reach back several queue steps (D and S - D - 1),
// so that in steady state they are already satisfied: with D = S - 1 a pass-1
// tile would wait on pass-2 tiles handed out one step earlier and the queue
// would advance at one item latency per step.
// Every wait points to items handed out earlier, held by workgroups that are
// already running and that only wait on items earlier still, so the chain
// always ends at an item that can proceed: the queue cannot deadlock, whether
// or not the whole grid is resident.  A wait also gives up after a wall-clock
// timeout (and as soon as another wait has given up), raising an error flag
// that the host reads back, so a fault in this scheme never hangs the device.
//
// Publication without cache maintenance: the intermediate is written and
// read with agent-scope (sc1) accesses, which are coherent across the XCDs
// by themselves (no L2 write-back or invalidate needed, and they meet in the
// memory-side cache).  A pass-1 item ends with every wave waiting for its own
// stores to complete, then the workgroup barrier, then lane 0's counter
// increment; a pass-2 item's counter only guards the slot's reuse, and its
// loads are consumed before the barrier.  The consumer's lane 0 observes the
// count before the workgroup's first load of the slot is issued.
//
// Tile shapes are those of k_col/k_rowT: a pass-1 tile is LPB1 adjacent
// columns of the N1 x N2 view (column FFT of length N1, then the inter-pass
// twiddle w_N^{c p}), a pass-2 tile is LPB2 rows of N2 points written back
// transposed into natural order (point k2*N1 + r of the transform).
#pragma once
#include "nfft_kernels.h"

namespace nfft {

struct PersistArgs {
    const void *in;
    void *out;
    long long ntrans;
    long long in_dist, out_dist;  // per-transform distances (elements)
    void *scratch;                // slots * N elements
    int slots;                    // rotating intermediate slots S
    int lag;                      // D < S: pass 2 of transform t runs in queue step t + D
    unsigned *ctr;                // see persist_ctr_words
    const void *tw_a, *tw_b;      // stage tables of N1 and N2
    Tw2 tw2;                      // e^{-2 pi i e / N}
    double scale;
    long long timeout;            // wall-clock ticks one wait may take
    int grab;                     // consecutive items taken per queue access (>= 1)
};

// Counter block: [0] queue head, [1] error flag of this launch, [2] sticky
// error (previous launches), [3, 3+T) pass-1 tiles done per transform,
// [3+T, 3+2T) pass-2 tiles done per transform.
__host__ __device__ constexpr long long persist_ctr_words(long long ntrans) { return 3 + 2 * ntrans; }

// Agent-scope (sc1) store / load of one complex element.
__device__ __forceinline__ void coh_store(Cx<float> *p, Cx<float> v)
{
    __hip_atomic_store((unsigned long long *)p, __builtin_bit_cast(unsigned long long, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void coh_store(Cx<double> *p, Cx<double> v)
{
    unsigned long long *q = (unsigned long long *)p;
    __hip_atomic_store(q, __builtin_bit_cast(unsigned long long, v.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, __builtin_bit_cast(unsigned long long, v.y), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ Cx<float> coh_load(const Cx<float> *p)
{
    return __builtin_bit_cast(Cx<float>, __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ Cx<double> coh_load(const Cx<double> *p)
{
    const unsigned long long *q = (const unsigned long long *)p;
    const unsigned long long x = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long y = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return cmk<Cx<double>>(__builtin_bit_cast(double, x), __builtin_bit_cast(double, y));
}

// Agent-scope (sc1) store / load of element idx of a workgroup-uniform base
// (byte offsets below 2^32).  An FP64 element takes one 16-byte buffer access
// with the sc1 cache-policy bit, the access an agent-scope atomic of that
// width would be; the atomic form above splits it into two 8-byte halves, and
// each half writes through to memory on its own, so every line of an
// intermediate is written twice.
constexpr int kCpSc1 = 16;
__device__ __forceinline__ void coh_st(Cx<float> *base, uint32_t idx, Cx<float> v) { coh_store(base + idx, v); }
__device__ __forceinline__ void coh_st(Cx<double> *base, uint32_t idx, Cx<double> v)
{
    buf_st<Cx<double>, kCpSc1>(buf_rsrc_if(base, true), idx * (uint32_t)sizeof(Cx<double>), 0u, v);
}
__device__ __forceinline__ Cx<float> coh_ld(const Cx<float> *base, uint32_t idx) { return coh_load(base + idx); }
__device__ __forceinline__ Cx<double> coh_ld(const Cx<double> *base, uint32_t idx)
{
    return buf_ld<Cx<double>, kCpSc1>(buf_rsrc_if(base, true), idx * (uint32_t)sizeof(Cx<double>), 0u);
}

// OCC: workgroups per CU the register allocation must allow (1 = no limit
// beyond one workgroup).
template <typename T, int N1, int N2, int LPB1, int LPB2, int OCC = 1> struct PShape {
    static constexpr int NT1 = Sched<N1>::NT, NT2 = Sched<N2>::NT;
    static constexpr int BT = NT1 * LPB1;  // threads per workgroup
    static constexpr int WPE = OCC * BT / 256 > 1 ? OCC * BT / 256 : 1;  // waves per SIMD
    static constexpr int A = N2 / LPB1;    // pass-1 tiles per transform
    static constexpr int B = N1 / LPB2;    // pass-2 tiles per transform
    static constexpr int L1 = Sched<N1>::S > 1 ? N1 * LPB1 : 1;
    static constexpr int L2 = rowT_lds<T, N2, LPB2>();
    static constexpr int LDS = L1 > L2 ? L1 : L2;
    static constexpr bool ok = NT2 * LPB2 == BT && BT <= 1024 && N2 % LPB1 == 0 && N1 % LPB2 == 0 &&
                               (long long)OCC * ((long long)LDS * (long long)sizeof(Cx<T>) + 64) <= 160 * 1024;
};

// Item index -> (kind 1: pass-1 tile, 2: pass-2 tile; transform; tile);
// false once the queue is drained.
__device__ __forceinline__ bool persist_decode(long long i, long long T, long long D, long long A, long long B,
                                               int *kind, long long *t, int *tile)
{
    auto cum = [&](long long s) {  // items in steps [0, s)
        const long long p1 = s < T ? s : T;
        long long p2 = s - D;
        p2 = p2 < 0 ? 0 : (p2 > T ? T : p2);
        return A * p1 + B * p2;
    };
    if (i >= cum(T + D)) return false;
    long long lo = 0, hi = T + D;  // cum(lo) <= i < cum(hi)
    while (hi - lo > 1) {
        const long long mid = (lo + hi) >> 1;
        if (cum(mid) <= i) lo = mid;
        else hi = mid;
    }
    const long long w = i - cum(lo);
    const long long p1 = lo < T ? A : 0;
    if (w < p1) {
        *kind = 1;
        *t = lo;
        *tile = (int)w;
    } else {
        *kind = 2;
        *t = lo - D;
        *tile = (int)(w - p1);
    }
    return true;
}

// Bounded spin of one thread until *c >= need.
__device__ __forceinline__ void persist_wait(const unsigned *c, unsigned need, unsigned *err, long long timeout)
{
    const unsigned long long t0 = wall_clock64();
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        if ((long long)(wall_clock64() - t0) > timeout) {
            atomicOr(err, 1u);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

template <typename T, int N1, int N2, int LPB1, int LPB2, int OCC, int DIR>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, PShape<T, N1, N2, LPB1, LPB2, OCC>::BT),
                               amdgpu_waves_per_eu(PShape<T, N1, N2, LPB1, LPB2, OCC>::WPE)))
k_four_persist(PersistArgs a)
{
    using C = Cx<T>;
    using Sh = PShape<T, N1, N2, LPB1, LPB2, OCC>;
    using S1 = Sched<N1>;
    using S2 = Sched<N2>;
    static_assert(Sh::ok, "persistent four-step shape");
    constexpr long long N = (long long)N1 * N2;
    __shared__ C lds[Sh::LDS];
    __shared__ int s_item[3];  // kind, transform, tile
    const long long ntr = a.ntrans;
    unsigned *head = a.ctr, *err = a.ctr + 1, *done1 = a.ctr + 3, *done2 = a.ctr + 3 + ntr;
    // Lane 0 of wave 0 publishes the finished item and fetches the next one
    // in a single region at the top of the loop, entered through a
    // wave-uniform branch: a second lane-divergent region between the
    // barriers of this loop lets the compiler split the loop into nests in
    // which the other lanes reach the next barrier before lane 0 has run.
    const bool wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x) == 0;
    int kind = 0, tile = 0, t = 0;
    long long base = 0;  // lane 0: current run of `grab` items from the queue
    int sub = a.grab;
    // A workgroup cannot take more items than the queue holds: the bound
    // only matters if the loop misbehaves, and then it still ends.
    const long long max_iter = ntr * (Sh::A + Sh::B) + 1;
    for (long long iter = 0;; iter++) {
        if (iter > max_iter) {
            if (threadIdx.x == 0) atomicOr(err, 2u);
            break;
        }
        if (wave0) {
            if (threadIdx.x == 0) {
                if (kind != 0)  // completion of the previous item
                    __hip_atomic_fetch_add(kind == 1 ? done1 + t : done2 + t, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                int nk = 0, ntile = 0;
                long long nt = 0;
                if (sub == a.grab) {
                    base = (long long)atomicAdd(head, (unsigned)a.grab);
                    sub = 0;
                }
                const long long i = base + sub++;
                if (persist_decode(i, ntr, a.lag, Sh::A, Sh::B, &nk, &nt, &ntile)) {
                    if (nk == 2) persist_wait(done1 + nt, (unsigned)Sh::A, err, a.timeout);
                    else if (nt >= a.slots) persist_wait(done2 + (nt - a.slots), (unsigned)Sh::B, err, a.timeout);
                }
                s_item[0] = nk;
                s_item[1] = (int)nt;
                s_item[2] = ntile;
            }
        }
        __syncthreads();
        kind = __builtin_amdgcn_readfirstlane(s_item[0]);
        t = __builtin_amdgcn_readfirstlane(s_item[1]);
        tile = __builtin_amdgcn_readfirstlane(s_item[2]);
        if (kind == 0) break;  // uniform: the queue is drained
        C *slot = (C *)a.scratch + (t % a.slots) * N;
        if (kind == 1) {
            const int b = threadIdx.x % LPB1, j = threadIdx.x / LPB1;
            const int c = tile * LPB1 + b;
            const C *src = (const C *)a.in + t * a.in_dist + c;
            C reg[S1::R];
#pragma unroll
            for (int q = 0; q < S1::R; q++) reg[q] = src[(long long)(j + S1::NT * q) * N2];
            LayCol<LPB1> lay{b};
            line_fft<T, N1, DIR>(reg, lds, lay, j, (const double2 *)a.tw_a);
            apply_tw2_series<DIR>(reg, a.tw2, (uint32_t)c * (uint32_t)j, (uint32_t)c * (uint32_t)S1::NT);
#pragma unroll
            for (int q = 0; q < S1::R; q++) coh_st(slot, (uint32_t)((j + S1::NT * q) * N2 + c), reg[q]);
            // the slot's stores of this wave are complete (acknowledged at
            // agent scope) before the barrier that precedes the counter
            __builtin_amdgcn_s_waitcnt(0);
        } else {
            const int b = threadIdx.x / S2::NT, j = threadIdx.x % S2::NT;
            const int r0 = tile * LPB2;
            C reg[S2::R];
            const uint32_t r0off = (uint32_t)r0 * (uint32_t)N2;  // the slot's row r0
            if constexpr (S2::NT >= 64) {
#pragma unroll
                for (int q = 0; q < S2::R; q++) reg[q] = coh_ld(slot, r0off + (uint32_t)(b * N2 + j + S2::NT * q));
            } else {  // rows shorter than a wavefront: coalesced copy through LDS
                constexpr int PADL = lpadlen<T, N2>();
                for (int idx = threadIdx.x; idx < LPB2 * N2; idx += Sh::BT) {
                    const int bb = idx / N2, p = idx % N2;
                    lds[bb * PADL + lpad<T>(p)] = coh_ld(slot, r0off + (uint32_t)idx);
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < S2::R; q++) reg[q] = lds[b * PADL + lpad<T>(j + S2::NT * q)];
                __syncthreads();
            }
            LayLine<T, N2> lay{b * lpadlen<T, N2>()};
            line_fft<T, N2, DIR>(reg, lds, lay, j, (const double2 *)a.tw_b);
            const T sc = (T)a.scale;  // line_fft leaves the LDS free (barrier after its last read)
#pragma unroll
            for (int q = 0; q < S2::R; q++) {
                C v = reg[q];
                if (sc != (T)1) v = cscale(v, sc);
                lds[(j + S2::NT * q) * (LPB2 + 1) + b] = v;
            }
            __syncthreads();
            C *dst = (C *)a.out + t * a.out_dist + r0;
            for (int idx = threadIdx.x; idx < LPB2 * N2; idx += Sh::BT) {
                const int bb = idx % LPB2, k = idx / LPB2;
                dst[(long long)k * N1 + bb] = lds[k * (LPB2 + 1) + bb];
            }
        }
        __syncthreads();
    }
}

}  // namespace nfft
