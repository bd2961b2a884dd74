// nfft_k_smooth.hip -- launchers of the mixed-radix kernels (both precisions).
#include "nfft_internal.h"

#include <algorithm>
#include <cstdlib>

namespace nfft {

namespace {
template <typename T, int DIR, bool TWL> void smooth_attr()
{
    // one attribute call per instantiation and process (thread-safe static init)
    static const bool done = [] {
        (void)hipFuncSetAttribute((const void *)k_smooth_line<T, DIR, TWL>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipGetLastError();
        return true;
    }();
    (void)done;
}

template <typename T, bool TWL> hipError_t smooth_line_launch(int dir, const SmoothArgs &a, hipStream_t s)
{
    const long long grid = (a.nlines + a.G - 1) / a.G;
    const size_t lds = (2 * (size_t)a.G + (TWL ? 1 : 0)) * a.sp.L * sizeof(Cx<T>);
    if (grid < 1) return hipSuccess;
    if (grid > 0x7fffffffLL || lds > 160 * 1024) return hipErrorInvalidValue;
    if (dir < 0) {
        smooth_attr<T, -1, TWL>();
        hipLaunchKernelGGL((k_smooth_line<T, -1, TWL>), dim3((unsigned)grid), dim3(256), lds, s, a);
    } else {
        smooth_attr<T, 1, TWL>();
        hipLaunchKernelGGL((k_smooth_line<T, 1, TWL>), dim3((unsigned)grid), dim3(256), lds, s, a);
    }
    return hipGetLastError();
}

// FP32 lines read their twiddles from LDS when a workgroup can give up lines
// for them within the same LDS size (2 G L >= 2 G' L + L with G' >= 1):
// measured N = 100 0.815 -> 0.78 ms, N = 480 0.958 -> 0.879 ms per 2^27
// points.  FP64 lines lost with it (N = 100 1.23 -> 1.31 ms) and keep the
// global table.
template <typename T> hipError_t smooth_line_impl(int dir, const SmoothArgs &a, hipStream_t s)
{
    const int L = a.sp.L;
    const int g2 = (2 * a.G * L - L) / (2 * L);
    if (sizeof(T) == 4 && g2 >= 1) {
        SmoothArgs b = a;
        b.G = g2;
        if constexpr (sizeof(T) == 4) return smooth_line_launch<T, true>(dir, b, s);
    }
    return smooth_line_launch<T, false>(dir, a, s);
}
template <typename T, int DIR> void smooth4_attr()
{
    static const bool done = [] {
        (void)hipFuncSetAttribute((const void *)k_smooth_col<T, DIR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void *)k_smooth_rowT<T, DIR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipGetLastError();
        return true;
    }();
    (void)done;
}

template <typename T> hipError_t smooth4_impl(int dir, const Smooth4Args &a, hipStream_t s)
{
    if (a.ntrans < 1) return hipSuccess;
    const long long gy = a.ntrans < 65535 ? a.ntrans : 65535;
    const long long gz = (a.ntrans + gy - 1) / gy;
    const int N1 = a.sp1.L, N2 = a.sp2.L;
    const dim3 g1((unsigned)((N2 + a.w1 - 1) / a.w1), (unsigned)gy, (unsigned)gz);
    const dim3 g2((unsigned)((N1 + a.w2 - 1) / a.w2), (unsigned)gy, (unsigned)gz);
    const size_t l1 = 2 * (size_t)a.w1 * N1 * sizeof(Cx<T>), l2 = 2 * (size_t)a.w2 * N2 * sizeof(Cx<T>);
    if (gz > 65535 || l1 > 160 * 1024 || l2 > 160 * 1024) return hipErrorInvalidValue;
    if (dir < 0) {
        smooth4_attr<T, -1>();
        hipLaunchKernelGGL((k_smooth_col<T, -1>), g1, dim3(256), l1, s, a);
        hipLaunchKernelGGL((k_smooth_rowT<T, -1>), g2, dim3(256), l2, s, a);
    } else {
        smooth4_attr<T, 1>();
        hipLaunchKernelGGL((k_smooth_col<T, 1>), g1, dim3(256), l1, s, a);
        hipLaunchKernelGGL((k_smooth_rowT<T, 1>), g2, dim3(256), l2, s, a);
    }
    return hipGetLastError();
}
}  // namespace

// the standalone kernels run schedules of radices up to smooth_radix<T>() only
static bool small_radices(const SmoothPlan &sp, int prec)
{
    for (int i = 0; i < sp.ns; i++)
        if (sp.radix[i] > smooth_radix_of(prec)) return false;
    return true;
}

hipError_t launch_smooth4(int prec, int dir, const Smooth4Args &a, hipStream_t s)
{
    if (!small_radices(a.sp1, prec) || !small_radices(a.sp2, prec)) return hipErrorInvalidValue;
    const int slot = prof_on() ? prof_begin(s) : -1;
    hipError_t e = prec == F32 ? smooth4_impl<float>(dir, a, s) : smooth4_impl<double>(dir, a, s);
    if (slot >= 0)
        prof_end(slot, s, "k_smooth_col+rowT", prec, a.sp1.L * a.sp2.L,
                 4.0 * (double)a.ntrans * a.sp1.L * a.sp2.L * (prec == F32 ? 8.0 : 16.0));
    return e;
}

bool smooth_split(long long n, int maxlen, SmoothPlan *sp1, SmoothPlan *sp2, int maxr)
{
    // N1 = the divisor of n nearest sqrt(n) with both factors <= maxlen
    long long best = 0;
    for (long long d = 2; d * d <= n * 4 && d <= maxlen; d++) {
        if (n % d) continue;
        const long long e = n / d;
        if (e > maxlen || e < 2) continue;
        if (!best || std::llabs(d * d - n) < std::llabs(best * best - n)) best = d;
    }
    if (!best) return false;
    return smooth_factor(best, sp1, maxr) && smooth_factor(n / best, sp2, maxr);
}

hipError_t launch_smooth_line(int prec, int dir, const SmoothArgs &a, hipStream_t s)
{
    if (!small_radices(a.sp, prec)) return hipErrorInvalidValue;
    const int slot = prof_on() ? prof_begin(s) : -1;
    hipError_t e = prec == F32 ? smooth_line_impl<float>(dir, a, s) : smooth_line_impl<double>(dir, a, s);
    if (slot >= 0)
        prof_end(slot, s, "k_smooth_line", prec, a.sp.L, 2.0 * (double)a.nlines * a.sp.L * (prec == F32 ? 8.0 : 16.0));
    return e;
}

int smooth_line_max(int prec)
{
    const char *v = getenv("NFFT_SMOOTH_LINE_MAX");
    if (v && *v && atoi(v) >= 2) return std::min(atoi(v), prec == F32 ? 8192 : 4096);
    return prec == F32 ? 2048 : 1024;
}

bool smooth_factor(long long n, SmoothPlan *sp, int maxr)
{
    if (n < 2 || n > (1LL << 30)) return false;
    SmoothPlan p{};
    p.L = (int)n;
    int ns = 0;
    auto push = [&](int r) {
        if (ns >= kSmoothMaxStages) return false;
        p.radix[ns++] = (unsigned char)r;
        return true;
    };
    int a = 0, b = 0, c = 0, d = 0;
    long long m = n;
    for (; m % 2 == 0; m /= 2) a++;
    for (; m % 3 == 0; m /= 3) b++;
    for (; m % 5 == 0; m /= 5) c++;
    for (; m % 7 == 0; m /= 7) d++;
    if (m != 1) return false;
    const char *env = getenv("NFFT_SMOOTH_RADIX");
    if (env && *env) maxr = std::min(maxr, std::max(8, atoi(env)));
    bool ok = true;
    if (maxr >= 25) {
        // fewest stages: 16 and the squares 25, 9 first, then one stage for
        // the leftover 5, 3 and powers of two together (20, 10, 15, 12, 6)
        for (; a >= 4; a -= 4) ok = ok && push(16);
        for (; c >= 2; c -= 2) ok = ok && push(25);
        for (; b >= 2; b -= 2) ok = ok && push(9);
        if (c == 1) {
            int r = 5;
            if (a >= 2) r = 20, a -= 2;
            else if (a == 1) r = 10, a = 0;
            else if (b == 1) r = 15, b = 0;
            ok = ok && push(r);
        }
        if (b == 1) {
            int r = 3;
            if (a >= 2) r = 12, a -= 2;
            else if (a == 1) r = 6, a = 0;
            ok = ok && push(r);
        }
        if (a > 0) ok = ok && push(1 << a);
    } else if (maxr >= 10) {
        // radices up to 10: each 5 with a 2 (10), 3s in pairs (9) or with a
        // 2 (6), then 8 / 4 / 2 and the remaining 5s and 3
        for (; c >= 1 && a >= 1; c--, a--) ok = ok && push(10);
        for (; b >= 2; b -= 2) ok = ok && push(9);
        if (b == 1 && a >= 1 && a % 3 != 0) ok = ok && push(6), b = 0, a--;
        for (; a >= 3; a -= 3) ok = ok && push(8);
        if (a > 0) ok = ok && push(1 << a);
        for (; c > 0; c--) ok = ok && push(5);
        if (b == 1) ok = ok && push(3);
    } else {
        // radices 2-8: 8s, 5s, a 4, 3s, 7s, a 2
        for (; a >= 3; a -= 3) ok = ok && push(8);
        for (; c > 0; c--) ok = ok && push(5);
        if (a == 2) ok = ok && push(4);
        for (; b > 0; b--) ok = ok && push(3);
        for (; d > 0; d--) ok = ok && push(7);
        if (a == 1) ok = ok && push(2);
        a = 0;
    }
    for (; d > 0; d--) ok = ok && push(7);
    if (!ok) return false;
    p.ns = ns;
    *sp = p;
    return true;
}

}  // namespace nfft
