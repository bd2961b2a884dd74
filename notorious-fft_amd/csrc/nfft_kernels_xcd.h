// nfft_kernels_xcd.h -- the whole FP32 four-step of a batch in one launch,
// with each transform kept on one XCD.
//
// The two-launch four-step (k_col, then k_rowT_pf) moves every point through
// HBM four times: input read, intermediate write, intermediate read, output
// write, in two launches.  k_four_xcd runs both passes in one launch and
// keeps the intermediate of a transform on the XCD that produces it: the
// column tiles and the row tiles of a transform are run by workgroups of the
// same XCD, through the XCD's own L2 and the memory-side cache behind it.
// An 8 MiB intermediate (N = 2^20) does not fit the 4 MB L2, so it still
// crosses the fabric twice (counted by the L2's FETCH_SIZE / WRITE_SIZE);
// what the single launch removes is the second pass's ramp and the
// inter-launch gap, and the intermediate's slots (two per XCD, 128 MiB) stay
// within the memory-side cache's capacity.
//
// Placement.  A workgroup reads the id of the XCD it runs on from the
// hardware register XCC_ID and takes its work from that XCD's queue, so the
// producer and the consumer of an intermediate always share an L2, whatever
// the dispatcher did.  Within one XCD the L2 is the coherence point: the
// intermediate is written with plain (write-back) stores and read with
// agent-scope loads (cache-policy bit sc1: served by the L2, never by the
// reading CU's L1, as the compiler emits for an agent-scope atomic load); no
// L2 write-back is needed.  Plain loads after a workgroup-scope invalidate
// (buffer_inv sc0) are not enough: that does not drop the CU's L1 lines, so
// a row tile could read lines that an earlier user of the same slot left in
// that L1 (whole rows of stale points, seen when a slot is reused within the
// L1's lifetime: N1 <= 128).  An agent-scope invalidate (buffer_inv sc1) is
// correct but costs 25 % at N = 2^20 (1.63 -> 2.03 ms).  The host maps the XCC ids the device
// reports to queues 0..nq-1 (xcd_topology); a workgroup on an unmapped id
// raises the error flag and does nothing.
//
// Queues.  Queue q owns transforms q, q + nq, q + 2 nq, ... (local step s is
// transform s nq + q).  Its items come off the queue's atomic head counter in
// the order
//     step s:  A column tiles of step s,  then  B row tiles of step s - 1,
// and step s uses intermediate slot s % 2 of the queue (two slots each).
//   - a row tile of step s - 1 loads the slot once the A column tiles of that
//     step have stored their results (they were handed out a full step
//     earlier);
//   - a column tile of step s, before its stores (after its loads and column
//     FFT), waits until the B row tiles of step s - 2, the slot's previous
//     user, have loaded it (handed out earlier still).
// A workgroup claims one item ahead and issues the loads of that item before
// it transforms the current one (for a row tile only if its column tiles are
// already complete; otherwise it waits and loads once the current item is
// stored), so that loads overlap the FFT work as in k_rowT_pf.  It only ever
// blocks on behalf of its current item, and every such wait points to items
// handed out earlier than it; the earliest unfinished item therefore always
// belongs to a workgroup that can run it, and the queues cannot deadlock,
// whether or not the whole grid is resident.  A wait also gives up after a
// wall-clock timeout (and as soon as another wait has given up), raising an
// error flag that the host reads back, so a fault never hangs the device.
// The last workgroup to leave checks that every tile of every transform was
// done (error flag otherwise), so a transform whose queue received no
// workgroup cannot go unnoticed.
//
// Tiles: a column tile is LPB1 adjacent columns of the N1 x 1024 view
// (column FFT of length N1, stored to the slot, row-major N1 x 1024); a row
// tile is 16 rows of 1024 points: inter-pass twiddle w_N^{r p}, row FFT and
// the transposed store into natural order (point k2 * N1 + r), as k_rowT_pf.
// TWROW = false moves the inter-pass twiddle into the column tiles (FP64
// table).  The arithmetic per point is that of k_col + k_rowT_pf, so the
// results are bit-identical to the two-launch four-step.
#pragma once
#include "nfft_kernels.h"

namespace nfft {

constexpr int kXcdQueues = 8;   // queues (XCDs) at most
constexpr int kCpAgent = 16;    // cache-policy bit sc1: agent-scope (L1-bypassing) access
constexpr int kXcdSlots = 2;    // intermediate slots per queue
constexpr int kXcdIds = 16;     // XCC_ID values the map covers

struct XcdArgs {
    const void *in;
    void *out;
    long long ntrans;
    long long in_dist, out_dist;  // per-transform distances (elements)
    void *scratch;                // nq queues x 2 slots x N elements
    unsigned *ctr;                // counter block (xcd_ctr_words)
    const void *tw_a, *tw_b;      // FP64 stage tables of N1 and N2
    Tw2f twf;                     // FP32 two-level table of w_N^e (twiddle in the row tiles)
    Tw2 tw2;                      // FP64 two-level table of w_N^e (twiddle in the column tiles)
    double scale;
    long long timeout;            // wall-clock ticks one wait may take
    int nq;                       // queues in use (1..kXcdQueues)
    signed char qmap[kXcdIds];    // XCC_ID -> queue (-1: not expected)
};

// Steps of a queue: transforms per queue rounded up, plus the drain step.
__host__ __device__ constexpr long long xcd_steps(long long ntrans, int nq) { return (ntrans + nq - 1) / nq + 2; }
// Counter block (32-bit words): [1] error flag of this launch, [2] sticky
// error (earlier launches), [4, 6) one 64-bit completion word (bits 0-19
// workgroups finished, 20-41 column tiles done, 42-63 row tiles done),
// [6, 6 + kXcdQueues) queue heads, then per queue: column tiles stored per
// step, row tiles loaded per step.
constexpr int kXcdHead = 6;
__host__ __device__ constexpr long long xcd_ctr_words(long long ntrans)
{
    return kXcdHead + kXcdQueues + (long long)kXcdQueues * 2 * xcd_steps(ntrans, 1);
}
// Launch limits the completion word can count.
constexpr long long kXcdMaxGrid = 1 << 20, kXcdMaxTiles = 1LL << 22;

// Tile shapes: 512 threads, each holding two lines' worth (NL = 2 sets of
// 16 points) of the current tile and, prefetched, two of the next one: two
// waves per SIMD leave 256 VGPRs per thread for them and the FFT.
template <int N1> struct XShape {
    static constexpr int N2 = 1024, LPB2 = 16, NL = 2;
    static constexpr int NT1 = Sched<N1>::NT, NT2 = Sched<N2>::NT;
    static constexpr int BT = NT2 * LPB2 / NL;     // 512 threads
    static constexpr int LPB1 = BT * NL / NT1;     // adjacent columns per column tile
    static constexpr int A = N2 / LPB1;            // column tiles per transform
    static constexpr int B = N1 / LPB2;            // row tiles per transform
    static constexpr int LDS = rowT_lds<float, N2, LPB2>() > N1 * LPB1 ? rowT_lds<float, N2, LPB2>() : N1 * LPB1;
    static constexpr int TWA = N1 == N2 ? 1 : stage_tw_lds<float, N1>();
    static constexpr int TWB = stage_tw_lds<float, N2>();
    // (upper bound: with the row-tile twiddle tables)
    static constexpr long long LDS_BYTES =
        (long long)(LDS + TWA + TWB + kTwinHi + kTwinLo) * (long long)sizeof(Cx<float>) + 64;
    static constexpr bool ok = NT1 % NL == 0 && LPB2 % NL == 0 && BT % 64 == 0 && LPB1 >= 4 &&
                               N2 % LPB1 == 0 && N1 % LPB2 == 0 && LDS_BYTES <= 160 * 1024;
};

// Bounded spin of one thread until *c >= need (false: gave up, error raised).
__device__ __forceinline__ bool xcd_wait_ge(const unsigned *c, unsigned need, unsigned *err, long long timeout)
{
    const unsigned long long t0 = wall_clock64();
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if ((long long)(wall_clock64() - t0) > timeout) {
            atomicOr(err, 1u);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// Id of the XCD this wave runs on: hwreg(HW_REG_XCC_ID, 0, 4).
__device__ __forceinline__ unsigned xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20); }

// Item kinds.
enum { XI_STOP = 0, XI_COL = 1, XI_ROW = 2, XI_SKIP = 3 };

// Item i of queue q -> kind, transform, tile, step (slot user).
__device__ __forceinline__ void xcd_decode(unsigned i, int q, int nq, long long T, int A, int B, int *kind,
                                           long long *t, int *tile, long long *s)
{
    const long long sq = q < T ? (T - q + nq - 1) / nq : 0;  // valid steps of this queue
    const long long st = i / (unsigned)(A + B);
    const int w = (int)(i % (unsigned)(A + B));
    *kind = XI_SKIP;
    *t = 0;
    *tile = 0;
    *s = st;
    if (st > sq) {
        *kind = XI_STOP;
    } else if (w < A) {
        if (st < sq) {
            *kind = XI_COL;
            *t = st * nq + q;
            *tile = w;
        }
    } else if (st >= 1) {
        *kind = XI_ROW;
        *s = st - 1;
        *t = (st - 1) * nq + q;
        *tile = w - A;
    }
}

// CPS, CPI: cache-policy bits of the streamed accesses (input loads, output
// stores) and of the intermediate's accesses.  Measured at N = 2^20 x 256:
// non-temporal streams 1.65 ms against 1.73 ms with the default policy;
// non-temporal intermediate accesses 1.81 ms, both 2.05 ms.
template <int N1, int DIR, bool TWROW, int CPS = 2, int CPI = 0>
__global__ void __launch_bounds__(XShape<N1>::BT, 1) k_four_xcd(XcdArgs a)
{
    using C = Cx<float>;
    using Sh = XShape<N1>;
    using S1 = Sched<N1>;
    using S2 = Sched<Sh::N2>;
    static_assert(Sh::ok, "XCD four-step shape");
    static_assert(S1::R == S2::R, "both tile kinds hold 16 points per line and thread");
    constexpr int N2 = Sh::N2, LPB1 = Sh::LPB1, LPB2 = Sh::LPB2, A = Sh::A, B = Sh::B, BT = Sh::BT, NL = Sh::NL;
    constexpr int R = S1::R;
    constexpr long long N = (long long)N1 * N2;
    constexpr int KS = BT / LPB2;  // output points k advanced per sweep of the transposed store
    __shared__ C lds[Sh::LDS];
    __shared__ C ltwb[Sh::TWB];
    __shared__ C ltwa[Sh::TWA];
    __shared__ C thi[TWROW ? kTwinHi : 1], tlo[TWROW ? kTwinLo : 1];
    __shared__ int s_item[5];  // kind, transform, tile, step, ready
    const long long T = a.ntrans;
    const int nq = a.nq;
    const long long SM = xcd_steps(T, 1);
    unsigned *err = a.ctr + 1;
    const unsigned id = xcc_id();
    const int q = id < (unsigned)kXcdIds ? (int)a.qmap[id] : -1;
    const bool lane0 = threadIdx.x == 0;
    const bool wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x) == 0;
    if (q < 0 || q >= nq) {  // an XCD the host did not see when it mapped the device
        if (lane0) atomicOr(err, 4u);
        return;
    }
    unsigned *head = a.ctr + kXcdHead + q;
    unsigned *done1 = a.ctr + kXcdHead + kXcdQueues + (long long)q * 2 * SM;
    unsigned *done2 = done1 + SM;
    // tables into LDS (made visible by the first barrier below)
    stage_tw_to_lds<float, N2>(ltwb, (const double2 *)a.tw_b);
    if constexpr (N1 != N2) stage_tw_to_lds<float, N1>(ltwa, (const double2 *)a.tw_a);
    if constexpr (TWROW) {
        for (int i = threadIdx.x; i < kTwinHi; i += BT) thi[i] = ((const C *)a.twf.hi)[i];
        for (int i = threadIdx.x; i <= a.twf.mask; i += BT) tlo[i] = ((const C *)a.twf.lo)[i];
    }
    const C *twa = N1 == N2 ? ltwb : ltwa;
    const float sc = (float)a.scale;
    // Column tile: thread (cb, cj) holds points cj + l NT1 / 2 (l < NL) of
    // column cb.  Row tile: thread (rb, rj) holds point rj of rows rb + l 8.
    // Consecutive lanes cover consecutive columns / points, so every access
    // of a wave is a run of whole 128-byte segments.
    constexpr int CJ = S1::NT / NL;      // column tile: line threads per set
    constexpr int RBS = LPB2 / NL;       // row tile: rows per set
    constexpr uint32_t cqst = (uint32_t)(S1::NT * N2) * (uint32_t)sizeof(C);
    constexpr uint32_t clst = (uint32_t)(CJ * N2) * (uint32_t)sizeof(C);
    constexpr uint32_t rqst = (uint32_t)(S2::NT) * (uint32_t)sizeof(C);
    constexpr uint32_t rlst = (uint32_t)(RBS * N2) * (uint32_t)sizeof(C);
    // The thread roles are derived from `tid`, which each loop iteration
    // makes opaque to the compiler: hoisted out of the loop, the lane offsets
    // and LDS addresses of both tile kinds would stay live through the whole
    // loop.
    int tid = threadIdx.x;

    // lane 0: claim the next item of the queue and publish it in s_item
    auto claim = [&]() {
        int kind, tile;
        long long t, s;
        xcd_decode(atomicAdd(head, 1u), q, nq, T, A, B, &kind, &t, &tile, &s);
        int ready = 0;
        if (kind == XI_ROW)
            ready = __hip_atomic_load(done1 + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (unsigned)A;
        s_item[0] = kind;
        s_item[1] = (int)t;
        s_item[2] = tile;
        s_item[3] = (int)s;
        s_item[4] = ready;
    };
    // the loads of an item into v (row tiles: the caller has checked that
    // the slot is complete)
    auto load = [&](C (&v)[NL][R], int kind, long long t, int tile, long long s) {
        if (kind == XI_COL) {
            const int cb = tid % LPB1, cj = tid / LPB1;
            const uint32_t off = (uint32_t)(cj * N2 + cb) * (uint32_t)sizeof(C);
            const auto rs = buf_rsrc((const C *)a.in + t * a.in_dist + tile * LPB1);
#pragma unroll
            for (int l = 0; l < NL; l++)
#pragma unroll
                for (int i = 0; i < R; i++) v[l][i] = buf_ld<C, CPS>(rs, off, l * clst + i * cqst);
        } else if (kind == XI_ROW) {
            const int rb = tid / S2::NT, rj = tid % S2::NT;
            const uint32_t off = (uint32_t)(rb * N2 + rj) * (uint32_t)sizeof(C);
            const C *slot = (const C *)a.scratch + ((long long)q * kXcdSlots + (s % kXcdSlots)) * N;
            const auto rs = buf_rsrc(slot + (long long)tile * LPB2 * N2);
#pragma unroll
            for (int l = 0; l < NL; l++)
#pragma unroll
                for (int i = 0; i < R; i++) v[l][i] = buf_ld<C, CPI | kCpAgent>(rs, off, l * rlst + i * rqst);
        }
    };

    // prologue: the first item, loaded; the next one claimed
    if (wave0 && lane0) claim();
    __syncthreads();
    int kind = __builtin_amdgcn_readfirstlane(s_item[0]);
    long long t = __builtin_amdgcn_readfirstlane(s_item[1]);
    int tile = __builtin_amdgcn_readfirstlane(s_item[2]);
    long long s = __builtin_amdgcn_readfirstlane(s_item[3]);
    if (kind == XI_ROW && !__builtin_amdgcn_readfirstlane(s_item[4])) {
        if (wave0 && lane0) (void)xcd_wait_ge(done1 + s, (unsigned)A, err, a.timeout);
        __syncthreads();
    }
    C reg[NL][R], nxt[NL][R];
    load(reg, kind, t, tile, s);
    __syncthreads();  // s_item is read
    if (wave0 && lane0) claim();
    __syncthreads();
    int nkind = __builtin_amdgcn_readfirstlane(s_item[0]);
    long long nt = __builtin_amdgcn_readfirstlane(s_item[1]);
    int ntile = __builtin_amdgcn_readfirstlane(s_item[2]);
    long long ns = __builtin_amdgcn_readfirstlane(s_item[3]);
    int nready = __builtin_amdgcn_readfirstlane(s_item[4]);
    unsigned ncol = 0, nrow = 0;  // tiles done by this workgroup
    // a workgroup cannot take more items than the queue holds
    const long long max_iter = SM * (A + B) + 2;
    for (long long iter = 0; kind != XI_STOP; iter++) {
        if (iter > max_iter) {
            if (lane0) atomicOr(err, 2u);
            break;
        }
        asm volatile("" : "+v"(tid));
        // the current item's loads are complete (the prefetch of the previous
        // iteration or the loads of the hand-over)
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (kind == XI_ROW && lane0)  // the slot is free for its next user
            __hip_atomic_fetch_add(done2 + s, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the next item's loads, where its data is ready, overlap this item
        const bool pf = nkind == XI_COL || (nkind == XI_ROW && nready);
        if (pf) load(nxt, nkind, nt, ntile, ns);
        if (kind == XI_COL) {
            const int cb = tid % LPB1, cj = tid / LPB1;
            const uint32_t off = (uint32_t)(cj * N2 + cb) * (uint32_t)sizeof(C);
            const LayCol<LPB1> lay[NL] = {LayCol<LPB1>{cb}, LayCol<LPB1>{cb}};
            const int jj[NL] = {cj, cj + CJ};
            line_fft_tw_n<float, N1, DIR, NL>(reg, lds, lay, jj, StageTwLds<float>{twa});
            if constexpr (!TWROW) {
                const uint32_t c = (uint32_t)(tile * LPB1 + cb);
#pragma unroll
                for (int l = 0; l < NL; l++)
                    apply_tw2_series<DIR>(reg[l], a.tw2, c * (uint32_t)jj[l], c * (uint32_t)S1::NT);
            }
            // the slot's previous user (row tiles of step s - 2) has loaded it
            if (s >= kXcdSlots && lane0) (void)xcd_wait_ge(done2 + s - kXcdSlots, (unsigned)B, err, a.timeout);
            __syncthreads();
            C *slot = (C *)a.scratch + ((long long)q * kXcdSlots + (s % kXcdSlots)) * N;
            const auto rd = buf_rsrc(slot + tile * LPB1);
#pragma unroll
            for (int l = 0; l < NL; l++)
#pragma unroll
                for (int i = 0; i < R; i++) buf_st<C, CPI>(rd, off, l * clst + i * cqst, reg[l][i]);
            // this wave's stores have reached the L2 before the barrier that
            // precedes the count
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (lane0) __hip_atomic_fetch_add(done1 + s, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ncol++;
        } else if (kind == XI_ROW) {
            const int rb = tid / S2::NT, rj = tid % S2::NT;
            const int r0 = tile * LPB2;
            if constexpr (TWROW) {
#pragma unroll
                for (int l = 0; l < NL; l++) {
                    const uint32_t rr = (uint32_t)(r0 + rb + l * RBS);
                    apply_twin<DIR>(reg[l], thi, tlo, a.twf.sh, a.twf.mask, rr * (uint32_t)rj,
                                    rr * (uint32_t)S2::NT);
                }
            }
            const LayLine<float, N2> lay[NL] = {LayLine<float, N2>{rb * lpadlen<float, N2>()},
                                                LayLine<float, N2>{(rb + RBS) * lpadlen<float, N2>()}};
            const int jj[NL] = {rj, rj};
            line_fft_tw_n<float, N2, DIR, NL>(reg, lds, lay, jj, StageTwLds<float>{ltwb});
#pragma unroll
            for (int l = 0; l < NL; l++)
#pragma unroll
                for (int i = 0; i < R; i++) {
                    C v = reg[l][i];
                    if (sc != 1.0f) v = cscale(v, sc);
                    lds[(rj + S2::NT * i) * (LPB2 + 1) + rb + l * RBS] = v;
                }
            __syncthreads();
            const int bb = tid % LPB2, k0 = tid / LPB2;
            const uint32_t soff = (uint32_t)(k0 * N1 + bb) * (uint32_t)sizeof(C);
            constexpr uint32_t ist = (uint32_t)(KS * N1) * (uint32_t)sizeof(C);
            const auto rd = buf_rsrc((C *)a.out + t * a.out_dist + r0);
#pragma unroll
            for (int i = 0; i < N2 / KS; i++)
                buf_st<C, CPS>(rd, soff, i * ist, lds[(k0 + KS * i) * (LPB2 + 1) + bb]);
            nrow++;
        }
        // hand over: the item after next is claimed, the next becomes current
        __syncthreads();  // LDS and s_item are free
        if (wave0 && lane0) claim();
        __syncthreads();
        if (pf) {
#pragma unroll
            for (int l = 0; l < NL; l++)
#pragma unroll
                for (int i = 0; i < R; i++) reg[l][i] = nxt[l][i];
        } else {
            if (nkind == XI_ROW) {  // its column tiles were not complete at claim time
                if (wave0 && lane0) (void)xcd_wait_ge(done1 + ns, (unsigned)A, err, a.timeout);
                __syncthreads();
            }
            load(reg, nkind, nt, ntile, ns);
        }
        kind = nkind;
        t = nt;
        tile = ntile;
        s = ns;
        nkind = __builtin_amdgcn_readfirstlane(s_item[0]);
        nt = __builtin_amdgcn_readfirstlane(s_item[1]);
        ntile = __builtin_amdgcn_readfirstlane(s_item[2]);
        ns = __builtin_amdgcn_readfirstlane(s_item[3]);
        nready = __builtin_amdgcn_readfirstlane(s_item[4]);
    }
    // completion: one atomic adds this workgroup's exit and tile counts; the
    // last workgroup out checks that every tile of every transform was done
    if (lane0) {
        const unsigned long long mine = 1ull | ((unsigned long long)ncol << 20) | ((unsigned long long)nrow << 42);
        const unsigned long long all =
            __hip_atomic_fetch_add((unsigned long long *)(a.ctr + 4), mine, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) + mine;
        constexpr unsigned long long m20 = (1ull << 20) - 1, m22 = (1ull << 22) - 1;
        if ((all & m20) == gridDim.x &&
            (((all >> 20) & m22) != (unsigned long long)(T * A) || ((all >> 42) & m22) != (unsigned long long)(T * B)))
            atomicOr(err, 8u);
    }
}

}  // namespace nfft
