// nfft_k_wpow.hip -- instantiations and launchers of the weighted spectral
// power kernels (nfft_kernels_wpow.h) for both precisions.
#include "nfft_launch_impl.h"
#include "nfft_kernels_wpow.h"

namespace nfft {

namespace {

template <typename T, int L> constexpr bool wpow_fits()
{
    constexpr int lpb = col_lpb<T, L>();
    return col_fits_v<T, L, lpb>() && wpow_threads<L, lpb>() % 64 == 0;
}

template <typename T> hipError_t col_wpow_impl(int L, const WpowArgs &a, hipStream_t s, long long *per)
{
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (wpow_fits<T, LL>()) {                                                                   \
            constexpr int lpb = col_lpb<T, LL>();                                                             \
            const unsigned gx = (unsigned)((a.ncols + lpb - 1) / lpb);                                        \
            *per = gx;                                                                                        \
            if (!s && !a.buf) return hipSuccess; /* size query */                                           \
            hipLaunchKernelGGL((k_col_wpow<T, LL, lpb>), dim3(gx, (unsigned)a.ntrans), dim3((wpow_threads<LL, lpb>())), \
                               0, s, a);                                                                      \
            return hipGetLastError();                                                                         \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

}  // namespace

bool col_wpow_supported(int prec, int L)
{
    long long per = 0;
    WpowArgs a{};
    a.ncols = 1;
    return (prec == F32 ? col_wpow_impl<float>(L, a, nullptr, &per) : col_wpow_impl<double>(L, a, nullptr, &per)) ==
           hipSuccess;
}

long long col_wpow_tiles(int prec, int L, int ncols)
{
    long long per = 0;
    WpowArgs a{};
    a.ncols = ncols;
    if ((prec == F32 ? col_wpow_impl<float>(L, a, nullptr, &per) : col_wpow_impl<double>(L, a, nullptr, &per)) !=
        hipSuccess)
        return 0;
    return per;
}

hipError_t launch_col_wpow(int prec, int L, const WpowArgs &a, hipStream_t s)
{
    if (a.ntrans < 1 || a.ntrans > 65535 || !a.buf) return hipErrorInvalidValue;
    long long per = 0;
    const int slot = prof_on() ? prof_begin(s) : -1;
    const hipError_t e = prec == F32 ? col_wpow_impl<float>(L, a, s, &per) : col_wpow_impl<double>(L, a, s, &per);
    if (slot >= 0)
        prof_end(slot, s, "k_col_wpow", prec, L,
                 (double)a.ntrans * a.dist * (prec == F32 ? 16.0 : 32.0) + (double)a.dist * (prec == F32 ? 4 : 8));
    return e;
}

long long wpow_mask_chunks(long long P) { return std::max(1LL, std::min(1024LL, (P + 2047) / 2048)); }

hipError_t launch_wpow_mask(int prec, void *buf, const void *w, double *partial, long long P, long long ntrans,
                            hipStream_t s)
{
    if (ntrans < 1 || ntrans > 65535) return hipErrorInvalidValue;
    const dim3 g((unsigned)wpow_mask_chunks(P), (unsigned)ntrans);
    if (prec == F32) hipLaunchKernelGGL(k_wpow_mask<float>, g, dim3(256), 0, s, buf, w, partial, P);
    else hipLaunchKernelGGL(k_wpow_mask<double>, g, dim3(256), 0, s, buf, w, partial, P);
    return hipGetLastError();
}

hipError_t launch_wpow_reduce(const double *partial, double *loss, long long per, long long ntrans, hipStream_t s)
{
    if (ntrans < 1) return hipSuccess;
    hipLaunchKernelGGL(k_wpow_reduce, dim3((unsigned)ntrans), dim3(64), 0, s, partial, loss, per);
    return hipGetLastError();
}

}  // namespace nfft
