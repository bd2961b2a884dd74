// nfft_k_xcd.hip -- instantiations and launcher of the XCD-affine FP32
// four-step (k_four_xcd): N = N1 x 1024 with N1 in {64, ..., 2048}, and the
// one-time probe of the XCC ids a device runs workgroups on.
#include <algorithm>
#include <mutex>
#include <vector>

#include "nfft_internal.h"
#include "nfft_kernels_xcd.h"

namespace nfft {

// One workgroup per slot records the XCC_ID it ran on.
__global__ void k_xcc_probe(int *ids)
{
    if (threadIdx.x == 0) ids[blockIdx.x] = (int)xcc_id();
}

namespace {

struct Topo {
    std::once_flag once;
    bool ok = false;
    int nq = 0;
    signed char qmap[kXcdIds];
};
Topo g_topo[64];

void topo_probe(Topo &tp)
{
    int ncu = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return;
    const int g = std::max(64, 8 * ncu);
    int *d = nullptr;
    // (not dev_malloc: a failed probe would switch the path off for the process)
    if (hipMalloc(&d, (size_t)g * sizeof(int)) != hipSuccess) return;
    std::vector<int> h((size_t)g, -1);
    hipStream_t st = nullptr;
    bool good = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    if (good) {
        hipLaunchKernelGGL(k_xcc_probe, dim3(g), dim3(64), 0, st, d);
        good = hipGetLastError() == hipSuccess &&
               hipMemcpyAsync(h.data(), d, (size_t)g * sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
               hipStreamSynchronize(st) == hipSuccess;
        (void)hipStreamDestroy(st);
    }
    (void)hipFree(d);
    if (!good) return;
    for (int i = 0; i < kXcdIds; i++) tp.qmap[i] = -1;
    int nq = 0;
    for (int v : h) {
        if (v < 0 || v >= kXcdIds) return;
        if (tp.qmap[v] < 0) {
            if (nq == kXcdQueues) return;
            tp.qmap[v] = (signed char)nq++;
        }
    }
    // queues numbered in XCC_ID order
    nq = 0;
    for (int i = 0; i < kXcdIds; i++)
        if (tp.qmap[i] >= 0) tp.qmap[i] = (signed char)nq++;
    tp.nq = nq;
    tp.ok = nq > 0;
}

template <int N1, bool TWROW> hipError_t xcd_go(int dir, const XcdArgs &a, int grid, hipStream_t s)
{
    if (dir < 0) hipLaunchKernelGGL((k_four_xcd<N1, -1, TWROW>), dim3(grid), dim3(XShape<N1>::BT), 0, s, a);
    else hipLaunchKernelGGL((k_four_xcd<N1, 1, TWROW>), dim3(grid), dim3(XShape<N1>::BT), 0, s, a);
    return hipGetLastError();
}

template <int N1> int xcd_occ()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_four_xcd<N1, -1, true>, XShape<N1>::BT, 0) != hipSuccess)
        return 0;
    return nb;
}

}  // namespace

bool xcd_topology(int *nq, signed char *qmap)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
    Topo &tp = g_topo[dev];
    std::call_once(tp.once, [&] { topo_probe(tp); });
    if (!tp.ok) return false;
    *nq = tp.nq;
    for (int i = 0; i < kXcdIds; i++) qmap[i] = tp.qmap[i];
    return true;
}

bool xcd_supported(long long n, int *n1, int *per_cu)
{
    int occ = 0;
    switch (n) {
    case 1LL << 16: *n1 = 64; occ = xcd_occ<64>(); break;
    case 1LL << 17: *n1 = 128; occ = xcd_occ<128>(); break;
    case 1LL << 18: *n1 = 256; occ = xcd_occ<256>(); break;
    case 1LL << 19: *n1 = 512; occ = xcd_occ<512>(); break;
    case 1LL << 20: *n1 = 1024; occ = xcd_occ<1024>(); break;
    case 1LL << 21: *n1 = 2048; occ = xcd_occ<2048>(); break;
    default: return false;
    }
    *per_cu = occ;
    return occ > 0;
}

hipError_t launch_four_xcd(int n1, int dir, bool twrow, const XcdArgs &a, int grid, hipStream_t s)
{
    const int slot = prof_on() ? prof_begin(s) : -1;
    hipError_t e = hipErrorInvalidValue;
    switch (n1 * 2 + (twrow ? 1 : 0)) {
    case 128: e = xcd_go<64, false>(dir, a, grid, s); break;
    case 129: e = xcd_go<64, true>(dir, a, grid, s); break;
    case 256: e = xcd_go<128, false>(dir, a, grid, s); break;
    case 257: e = xcd_go<128, true>(dir, a, grid, s); break;
    case 512: e = xcd_go<256, false>(dir, a, grid, s); break;
    case 513: e = xcd_go<256, true>(dir, a, grid, s); break;
    case 1024: e = xcd_go<512, false>(dir, a, grid, s); break;
    case 1025: e = xcd_go<512, true>(dir, a, grid, s); break;
    case 2048: e = xcd_go<1024, false>(dir, a, grid, s); break;
    case 2049: e = xcd_go<1024, true>(dir, a, grid, s); break;
    case 4096: e = xcd_go<2048, false>(dir, a, grid, s); break;
    case 4097: e = xcd_go<2048, true>(dir, a, grid, s); break;
    default: break;
    }
    if (slot >= 0) prof_end(slot, s, "k_four_xcd", F32, n1 * 1024, 2.0 * (double)a.ntrans * n1 * 1024 * 8.0);
    return e;
}

}  // namespace nfft
