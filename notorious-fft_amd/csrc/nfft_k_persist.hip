// nfft_k_persist.hip -- instantiations and launcher of the persistent
// four-step kernel (k_four_persist) for both precisions.
//
// Each power-of-two length has a list of shapes (N1 x N2 split and tile
// widths); entry 0 is the default, the others are alternatives a caller can
// select for tuning.
#include "nfft_internal.h"

namespace nfft {

// Clears the counter block before a launch, folding the last launch's error
// flag into the sticky word (a fill kernel rather than a memset, so the pair
// can be captured in a graph).
__global__ void k_persist_reset(unsigned *ctr, long long words)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 1) {
        ctr[2] |= ctr[1];
        ctr[1] = 0u;
    } else if (i != 2 && i < words) {
        ctr[i] = 0u;
    }
}

namespace {

template <typename T, int N1, int N2, int LPB1, int LPB2, int OCC>
hipError_t persist_go(int dir, const PersistArgs &a, int grid, hipStream_t s)
{
    constexpr int BT = PShape<T, N1, N2, LPB1, LPB2, OCC>::BT;
    if (dir < 0)
        hipLaunchKernelGGL((k_four_persist<T, N1, N2, LPB1, LPB2, OCC, -1>), dim3(grid), dim3(BT), 0, s, a);
    else
        hipLaunchKernelGGL((k_four_persist<T, N1, N2, LPB1, LPB2, OCC, 1>), dim3(grid), dim3(BT), 0, s, a);
    return hipGetLastError();
}

template <typename T, int N1, int N2, int LPB1, int LPB2, int OCC> int persist_occ()
{
    constexpr int BT = PShape<T, N1, N2, LPB1, LPB2, OCC>::BT;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_four_persist<T, N1, N2, LPB1, LPB2, OCC, -1>, BT,
                                                     0) != hipSuccess)
        return 0;
    return nb;
}

struct Entry {
    int prec, n1, n2, lpb1, lpb2, threads;
    hipError_t (*go)(int, const PersistArgs &, int, hipStream_t);
    int (*occ)();
};

#define NFFT_PERSIST(T, P, N1, N2, LPB1, LPB2, OCC)                                                          \
    Entry{P, N1, N2, LPB1, LPB2, PShape<T, N1, N2, LPB1, LPB2, OCC>::BT,                                      \
          &persist_go<T, N1, N2, LPB1, LPB2, OCC>, &persist_occ<T, N1, N2, LPB1, LPB2, OCC>}

// Order matters: the first entry of a length is its default shape.
const Entry kShapes[] = {
    NFFT_PERSIST(float, F32, 256, 128, 32, 64, 1),
    NFFT_PERSIST(float, F32, 256, 256, 32, 32, 1),
    NFFT_PERSIST(float, F32, 512, 256, 16, 32, 1),
    NFFT_PERSIST(float, F32, 512, 512, 16, 16, 1),
    NFFT_PERSIST(float, F32, 512, 512, 16, 16, 2),
    NFFT_PERSIST(float, F32, 1024, 512, 8, 16, 1),
    NFFT_PERSIST(float, F32, 1024, 512, 8, 16, 2),
    NFFT_PERSIST(float, F32, 1024, 1024, 8, 8, 1),
    NFFT_PERSIST(float, F32, 1024, 1024, 16, 16, 1),
    NFFT_PERSIST(float, F32, 512, 2048, 16, 4, 1),
    NFFT_PERSIST(float, F32, 2048, 512, 4, 16, 1),
    NFFT_PERSIST(float, F32, 1024, 1024, 8, 8, 2),
    NFFT_PERSIST(float, F32, 2048, 512, 4, 16, 2),
    NFFT_PERSIST(float, F32, 2048, 1024, 4, 8, 1),
    NFFT_PERSIST(float, F32, 2048, 2048, 4, 4, 1),
    NFFT_PERSIST(float, F32, 4096, 2048, 2, 4, 1),
    NFFT_PERSIST(double, F64, 128, 128, 32, 32, 1),
    NFFT_PERSIST(double, F64, 256, 128, 16, 32, 1),
    NFFT_PERSIST(double, F64, 256, 256, 16, 16, 1),
    NFFT_PERSIST(double, F64, 512, 256, 8, 16, 1),
    NFFT_PERSIST(double, F64, 512, 512, 8, 8, 1),
    NFFT_PERSIST(double, F64, 1024, 512, 4, 8, 1),
    NFFT_PERSIST(double, F64, 1024, 1024, 4, 4, 1),
    NFFT_PERSIST(double, F64, 1024, 1024, 8, 8, 1),
    NFFT_PERSIST(double, F64, 2048, 1024, 2, 4, 1),
    NFFT_PERSIST(double, F64, 4096, 1024, 1, 4, 1),
};

}  // namespace

bool persist_find(int prec, long long n, int variant, PersistCfg *cfg)
{
    int k = 0;
    for (const Entry &e : kShapes) {
        if (e.prec != prec || (long long)e.n1 * e.n2 != n) continue;
        if (k++ != variant) continue;
        cfg->n1 = e.n1;
        cfg->n2 = e.n2;
        cfg->lpb1 = e.lpb1;
        cfg->lpb2 = e.lpb2;
        cfg->threads = e.threads;
        cfg->per_cu = e.occ();
        cfg->id = (int)(&e - kShapes);
        return cfg->per_cu > 0;
    }
    return false;
}

hipError_t launch_four_persist(const PersistCfg &c, int dir, const PersistArgs &a, int grid, hipStream_t s)
{
    const Entry &e = kShapes[c.id];
    const int prec = e.prec;
    const int slot = prof_on() ? prof_begin(s) : -1;
    hipError_t err = e.go(dir, a, grid, s);
    if (slot >= 0)
        prof_end(slot, s, "k_four_persist", prec, e.n1 * e.n2,
                 2.0 * (double)a.ntrans * e.n1 * e.n2 * (prec == F32 ? 8.0 : 16.0));
    return err;
}

hipError_t launch_persist_reset(unsigned *ctr, long long words, hipStream_t s)
{
    const int bt = 256;
    const long long g = (words + bt - 1) / bt;
    hipLaunchKernelGGL(k_persist_reset, dim3((unsigned)g), dim3(bt), 0, s, ctr, words);
    return hipGetLastError();
}

}  // namespace nfft
