// nfft_launch_impl.h -- size dispatch from runtime lengths to the compiled
// kernel instantiations.  Each nfft_k_*.hip translation unit includes this
// and explicitly instantiates one family for one precision.
#pragma once
#include "nfft_internal.h"

namespace nfft {

constexpr size_t kLdsLimit = 160 * 1024;

#define NFFT_POW2_LIST(X)                                                                                     \
    X(1) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192) X(16384)

template <typename T, int L> constexpr bool line_fits()
{
    return (size_t)line_lds<T, L, line_lpb<L>()>() * sizeof(Cx<T>) <= kLdsLimit &&
           Sched<L>::NT * line_lpb<L>() <= 1024;
}

inline dim3 grid_lines(long long nlines, int lpb)
{
    return dim3((unsigned)((nlines + lpb - 1) / lpb));
}

// transforms split over grid.y (<= 65535) and grid.z
inline void grid_trans(long long ntrans, unsigned &gy, unsigned &gz)
{
    long long y = ntrans < 65535 ? ntrans : 65535;
    if (y < 1) y = 1;
    gy = (unsigned)y;
    gz = (unsigned)((ntrans + y - 1) / y);
    if (gz < 1) gz = 1;
}

// lines of twice the LDS-resident length (k_line_h)
template <typename T, int L> constexpr bool line_h_fits()
{
    return L >= 4096 && line_h_ok<T, L>() && (size_t)lpadlen<T, L / 2>() * sizeof(Cx<T>) <= kLdsLimit;
}
template <typename T, int L> hipError_t line_h_go(int dir, const LineArgs &a, hipStream_t s)
{
    if constexpr (line_h_fits<T, L>()) {
        if (a.nlines > 0x7fffffffLL) return hipErrorInvalidValue;
        const dim3 g((unsigned)a.nlines), b(Sched<L>::NT / 2);
        if (dir < 0) hipLaunchKernelGGL((k_line_h<T, L, -1>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_line_h<T, L, 1>), g, b, 0, s, a);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
template <typename T> hipError_t line_impl(int L, int dir, const LineArgs &a, hipStream_t s)
{
    if (a.halves) {
        switch (L) {
        case 8192: return line_h_go<T, 8192>(dir, a, s);
        case 16384: return line_h_go<T, 16384>(dir, a, s);
        case 32768: return line_h_go<T, 32768>(dir, a, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (line_fits<T, LL>()) {                                                                   \
            constexpr int lpb = line_lpb<LL>();                                                               \
            if (dir < 0) hipLaunchKernelGGL((k_line<T, LL, -1>), grid_lines(a.nlines, lpb),                   \
                                            dim3(Sched<LL>::NT * lpb), 0, s, a);                              \
            else hipLaunchKernelGGL((k_line<T, LL, 1>), grid_lines(a.nlines, lpb), dim3(Sched<LL>::NT * lpb), \
                                    0, s, a);                                                                 \
            return hipGetLastError();                                                                         \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

template <typename T> hipError_t blue_line_impl(int M, const BlueArgs &a, hipStream_t s)
{
    switch (M) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (LL >= 2 && line_fits<T, LL>()) {                                                        \
            constexpr int lpb = line_lpb<LL>();                                                               \
            hipLaunchKernelGGL((k_blue_line<T, LL>), grid_lines(a.nlines, lpb), dim3(Sched<LL>::NT * lpb), 0, \
                               s, a);                                                                         \
            return hipGetLastError();                                                                         \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
template <typename T, int L, int LPB> constexpr bool col_fits_v()
{
    constexpr size_t lds = Sched<L>::S > 1 ? (size_t)L * LPB * sizeof(Cx<T>) : 0;
    return L >= 2 && L <= 4096 && lds <= kLdsLimit && Sched<L>::NT * LPB <= 1024 && Sched<L>::NT * LPB >= 64;
}
template <typename T, int L> constexpr bool col_fits() { return col_fits_v<T, L, col_lpb<T, L>()>(); }

template <typename T, int L, int LPB>
hipError_t col_launch(int dir, int mode, const ColArgs &a, hipStream_t s)
{
    if constexpr (col_fits_v<T, L, LPB>()) {
        unsigned gy, gz;
        grid_trans(a.ntrans, gy, gz);
        dim3 g((unsigned)((a.ncols + LPB - 1) / LPB), gy, gz), b(Sched<L>::NT * LPB);
        if (mode == COL_PLAIN && dir < 0) hipLaunchKernelGGL((k_col<T, L, LPB, -1, COL_PLAIN>), g, b, 0, s, a);
        else if (mode == COL_PLAIN) hipLaunchKernelGGL((k_col<T, L, LPB, 1, COL_PLAIN>), g, b, 0, s, a);
        else if (mode == COL_TWIDDLE && dir < 0) hipLaunchKernelGGL((k_col<T, L, LPB, -1, COL_TWIDDLE>), g, b, 0, s, a);
        else if (mode == COL_TWIDDLE) hipLaunchKernelGGL((k_col<T, L, LPB, 1, COL_TWIDDLE>), g, b, 0, s, a);
        else if (mode == COL_BLUE_IN) hipLaunchKernelGGL((k_col<T, L, LPB, -1, COL_BLUE_IN>), g, b, 0, s, a);
        else if (mode == COL_BLUE_OUT) hipLaunchKernelGGL((k_col<T, L, LPB, 1, COL_BLUE_OUT>), g, b, 0, s, a);
        else return hipErrorInvalidValue;
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

// Tile widths selectable at run time for the lengths the four-step passes
// use; other lengths always take the default width.
template <int L> constexpr bool tunable_len() { return L >= 128 && L <= 2048; }

template <typename T> hipError_t col_impl(int L, int dir, int mode, const ColArgs &a, hipStream_t s)
{
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (col_fits<T, LL>()) {                                                                    \
            if constexpr (tunable_len<LL>()) {                                                                \
                switch (a.lpb) {                                                                              \
                case 4: if constexpr (col_fits_v<T, LL, 4>()) return col_launch<T, LL, 4>(dir, mode, a, s); break;                                         \
                case 8: if constexpr (col_fits_v<T, LL, 8>()) return col_launch<T, LL, 8>(dir, mode, a, s); break;                                         \
                case 16: if constexpr (col_fits_v<T, LL, 16>()) return col_launch<T, LL, 16>(dir, mode, a, s); break;                                       \
                case 32: if constexpr (col_fits_v<T, LL, 32>()) return col_launch<T, LL, 32>(dir, mode, a, s); break;                                       \
                default: break;                                                                               \
                }                                                                                             \
            }                                                                                                 \
            return col_launch<T, LL, col_lpb<T, LL>()>(dir, mode, a, s);                                      \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

template <typename T, int L, int LPB> constexpr bool rowT_fits_v()
{
    return L >= 2 && L <= (sizeof(T) == 4 ? 2048 : 1024) &&
           (size_t)rowT_lds<T, L, LPB>() * sizeof(Cx<T>) <= kLdsLimit && Sched<L>::NT * LPB <= 1024 &&
           Sched<L>::NT * LPB >= 64;
}
template <typename T, int L> constexpr bool rowT_fits() { return rowT_fits_v<T, L, rowT_lpb<T, L>()>(); }

// The resident prefetching grid is used where it measured faster: rows of
// 256..512 points, and longer rows with tiles of 16 or more rows (with 8-row
// tiles of >= 1024 points the one-workgroup-per-tile pass was faster: FP32
// 2048 x 8 961 vs 1084 us, 1024 x 8 899 vs 1118 us; FP64 1024 x 8 1581 vs
// 1669 us, on 2^28 / 2^27 points).
template <typename T, int L, int LPB> constexpr bool rowT_pf_fits()
{
    return Sched<L>::S > 1 && L >= 256 && (L <= 512 || LPB >= 16) &&
           (size_t)rowT_pf_lds<T, L, LPB>() * sizeof(Cx<T>) <= kLdsLimit;
}

template <typename T, int L, int LPB, int DIR, bool TWIN> bool rowT_pf_try(const RowTArgs &a, hipStream_t s)
{
    if (a.nrows % LPB != 0 || (long long)L * a.nrows * (long long)sizeof(Cx<T>) >= (1LL << 32)) return false;
    static const long long resident = [] {
        int dev = 0, ncu = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)k_rowT_pf<T, L, LPB, DIR, TWIN>,
                                                         Sched<L>::NT * LPB, 0) != hipSuccess)
            return 0LL;
        return (long long)n * ncu;
    }();
    const long long ntiles = a.ntrans * (a.nrows / LPB);
    if (ntiles >= (1LL << 31)) return false;  // 32-bit tile indices
    const long long g = std::min(resident, (ntiles + 7) / 8 * 8) / 8 * 8;
    if (g < 8) return false;
    hipLaunchKernelGGL((k_rowT_pf<T, L, LPB, DIR, TWIN>), dim3((unsigned)g), dim3(Sched<L>::NT * LPB), 0, s, a, ntiles);
    return true;
}

// Rows that take the inter-pass twiddle (RowTArgs::twf set): FP32 rows of
// 1024 points, the row pass of the 2^14..2^21-point four-steps.
template <typename T, int L> constexpr bool rowT_twin_v() { return sizeof(T) == 4 && L == 1024; }

template <typename T, int L, int LPB, bool TWIN> constexpr bool rowT_pf_fits_t()
{
    return rowT_pf_fits<T, L, LPB>() &&
           (!TWIN || (size_t)(rowT_pf_lds<T, L, LPB>() + kTwinHi + kTwinLo) * sizeof(Cx<T>) <= kLdsLimit);
}

template <typename T, int L, int LPB, bool TWIN> hipError_t rowT_launch_t(int dir, const RowTArgs &a, hipStream_t s)
{
    if (TWIN && a.twf.mask >= kTwinLo) return hipErrorInvalidValue;
    if constexpr (rowT_pf_fits_t<T, L, LPB, TWIN>()) {
        if (a.pf && (dir < 0 ? rowT_pf_try<T, L, LPB, -1, TWIN>(a, s) : rowT_pf_try<T, L, LPB, 1, TWIN>(a, s)))
            return hipGetLastError();
    }
    unsigned gy, gz;
    grid_trans(a.ntrans, gy, gz);
    dim3 g((unsigned)((a.nrows + LPB - 1) / LPB), gy, gz), b(Sched<L>::NT * LPB);
    if (dir < 0) hipLaunchKernelGGL((k_rowT<T, L, LPB, -1, TWIN>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_rowT<T, L, LPB, 1, TWIN>), g, b, 0, s, a);
    return hipGetLastError();
}

template <typename T, int L, int LPB> hipError_t rowT_launch(int dir, const RowTArgs &a, hipStream_t s)
{
    if constexpr (rowT_fits_v<T, L, LPB>()) {
        if (a.twf.hi) {
            if constexpr (rowT_twin_v<T, L>()) return rowT_launch_t<T, L, LPB, true>(dir, a, s);
            return hipErrorInvalidValue;
        }
        return rowT_launch_t<T, L, LPB, false>(dir, a, s);
    }
    return hipErrorInvalidValue;
}

template <typename T> hipError_t rowT_impl(int L, int dir, const RowTArgs &a, hipStream_t s)
{
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (rowT_fits<T, LL>()) {                                                                   \
            if constexpr (tunable_len<LL>()) {                                                                \
                switch (a.lpb) {                                                                              \
                case 4: if constexpr (rowT_fits_v<T, LL, 4>()) return rowT_launch<T, LL, 4>(dir, a, s); break;                                              \
                case 8: if constexpr (rowT_fits_v<T, LL, 8>()) return rowT_launch<T, LL, 8>(dir, a, s); break;                                              \
                case 16: if constexpr (rowT_fits_v<T, LL, 16>()) return rowT_launch<T, LL, 16>(dir, a, s); break;                                            \
                case 32: if constexpr (rowT_fits_v<T, LL, 32>()) return rowT_launch<T, LL, 32>(dir, a, s); break;                                            \
                default: break;                                                                               \
                }                                                                                             \
            }                                                                                                 \
            return rowT_launch<T, LL, rowT_lpb<T, LL>()>(dir, a, s);                                          \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

template <typename T, int L> constexpr bool rowr2c_fits()
{
    return rowT_fits<T, L>() && rowT_lpb<T, L>() % 2 == 0;
}

// twice the rows of rowT_lpb where the row pairs' tile still fits the LDS
// and 1024 threads (longer output runs on both sides of each pair)
template <typename T, int L, int LPB> constexpr bool rowr2c_wide_fits()
{
    return Sched<L>::NT * LPB <= 1024 && Sched<L>::S > 1 &&
           (size_t)(LPB * lpadlen<T, L>() + L + LPB) * sizeof(Cx<T>) <= 150 * 1024;
}

template <typename T, int L, int LPB> hipError_t rowT_r2c_go(const RowR2CArgs &a, hipStream_t s)
{
    unsigned gy, gz;
    grid_trans(a.ntrans, gy, gz);
    dim3 g((unsigned)((a.nrows / 2 + LPB / 2 - 1) / (LPB / 2)), gy, gz), b(Sched<L>::NT * LPB);
    hipLaunchKernelGGL((k_rowT_r2c<T, L, LPB>), g, b, 0, s, a);
    return hipGetLastError();
}

template <typename T> hipError_t rowT_r2c_impl(int L, const RowR2CArgs &a, hipStream_t s)
{
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (rowr2c_fits<T, LL>()) {                                                                 \
            constexpr int LPB = rowT_lpb<T, LL>();                                                            \
            if constexpr (LPB >= 4) {                                                                         \
                if (a.lpb == LPB / 2) return rowT_r2c_go<T, LL, LPB / 2>(a, s);                               \
            }                                                                                                 \
            if constexpr (rowr2c_wide_fits<T, LL, 2 * LPB>()) {                                               \
                if (a.lpb == 2 * LPB) return rowT_r2c_go<T, LL, 2 * LPB>(a, s);                               \
            }                                                                                                 \
            return rowT_r2c_go<T, LL, LPB>(a, s);                                                             \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

template <typename T, int L> constexpr bool bluerow_fits()
{
    constexpr int lpb = bluerow_lpb<T, L>();
    return L >= 2 && L <= (sizeof(T) == 4 ? 4096 : 2048) &&
           (size_t)line_lds<T, L, lpb>() * sizeof(Cx<T>) <= kLdsLimit && Sched<L>::NT * lpb <= 1024;
}

template <typename T> hipError_t blue_row_impl(int L, const BlueRowArgs &a, hipStream_t s)
{
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (bluerow_fits<T, LL>()) {                                                                \
            constexpr int lpb = bluerow_lpb<T, LL>();                                                         \
            unsigned gy, gz;                                                                                  \
            grid_trans(a.ntrans, gy, gz);                                                                     \
            dim3 g((unsigned)((a.nrows + lpb - 1) / lpb), gy, gz), b(Sched<LL>::NT * lpb);                    \
            hipLaunchKernelGGL((k_blue_row<T, LL, lpb>), g, b, 0, s, a);                                      \
            return hipGetLastError();                                                                         \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

template <typename T, int L> hipError_t r2c_h_go(bool inverse, const RtArgs &a, hipStream_t s)
{
    if constexpr (line_h_fits<T, L>()) {
        if (a.nlines > 0x7fffffffLL) return hipErrorInvalidValue;
        const dim3 g((unsigned)a.nlines), b(Sched<L>::NT / 2);
        if (!inverse) hipLaunchKernelGGL((k_r2c_line_h<T, L>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_c2r_line_h<T, L>), g, b, 0, s, a);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
template <typename T> hipError_t r2c_impl(int L, bool inverse, const RtArgs &a, hipStream_t s)
{
    if (a.halves) {
        switch (L) {
        case 8192: return r2c_h_go<T, 8192>(inverse, a, s);
        case 16384: return r2c_h_go<T, 16384>(inverse, a, s);
        case 32768: return r2c_h_go<T, 32768>(inverse, a, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (line_fits<T, LL>()) {                                                                   \
            constexpr int lpb = line_lpb<LL>();                                                               \
            if (!inverse) hipLaunchKernelGGL((k_r2c_line<T, LL>), grid_lines(a.nlines, lpb),                  \
                                             dim3(Sched<LL>::NT * lpb), 0, s, a);                             \
            else hipLaunchKernelGGL((k_c2r_line<T, LL>), grid_lines(a.nlines, lpb),                           \
                                    dim3(Sched<LL>::NT * lpb), 0, s, a);                                      \
            return hipGetLastError();                                                                         \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

template <typename T> hipError_t t23_impl(int L, int op, const RtArgs &a, hipStream_t s)
{
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (line_fits<T, LL>()) {                                                                   \
            constexpr int lpb = line_lpb<LL>();                                                               \
            dim3 g = grid_lines(a.nlines, lpb), b(Sched<LL>::NT * lpb);                                       \
            switch (op) {                                                                                     \
            case T_DCT2: hipLaunchKernelGGL((k_t23_line<T, LL, T_DCT2>), g, b, 0, s, a); break;             \
            case T_DST2: hipLaunchKernelGGL((k_t23_line<T, LL, T_DST2>), g, b, 0, s, a); break;             \
            case T_DCT3: hipLaunchKernelGGL((k_t23_line<T, LL, T_DCT3>), g, b, 0, s, a); break;             \
            case T_DST3: hipLaunchKernelGGL((k_t23_line<T, LL, T_DST3>), g, b, 0, s, a); break;             \
            default: return hipErrorInvalidValue;                                                             \
            }                                                                                                 \
            return hipGetLastError();                                                                         \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

template <typename T> hipError_t t4_impl(int L, int op, const RtArgs &a, hipStream_t s)
{
    switch (L) {
#define NFFT_CASE(LL)                                                                                         \
    case LL:                                                                                                  \
        if constexpr (line_fits<T, LL>()) {                                                                   \
            constexpr int lpb = line_lpb<LL>();                                                               \
            dim3 g = grid_lines(a.nlines, lpb), b(Sched<LL>::NT * lpb);                                       \
            if (op == T_DCT4) hipLaunchKernelGGL((k_t4_line<T, LL, T_DCT4>), g, b, 0, s, a);                 \
            else if (op == T_DST4) hipLaunchKernelGGL((k_t4_line<T, LL, T_DST4>), g, b, 0, s, a);            \
            else return hipErrorInvalidValue;                                                                 \
            return hipGetLastError();                                                                         \
        }                                                                                                     \
        break;
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return hipErrorInvalidValue;
}

// explicit instantiations live in the nfft_k_*.hip files
extern template hipError_t line_impl<float>(int, int, const LineArgs &, hipStream_t);
extern template hipError_t line_impl<double>(int, int, const LineArgs &, hipStream_t);
extern template hipError_t blue_line_impl<float>(int, const BlueArgs &, hipStream_t);
extern template hipError_t blue_line_impl<double>(int, const BlueArgs &, hipStream_t);
extern template hipError_t col_impl<float>(int, int, int, const ColArgs &, hipStream_t);
extern template hipError_t col_impl<double>(int, int, int, const ColArgs &, hipStream_t);
extern template hipError_t rowT_impl<float>(int, int, const RowTArgs &, hipStream_t);
extern template hipError_t rowT_impl<double>(int, int, const RowTArgs &, hipStream_t);
extern template hipError_t rowT_r2c_impl<float>(int, const RowR2CArgs &, hipStream_t);
extern template hipError_t rowT_r2c_impl<double>(int, const RowR2CArgs &, hipStream_t);
extern template hipError_t blue_row_impl<float>(int, const BlueRowArgs &, hipStream_t);
extern template hipError_t blue_row_impl<double>(int, const BlueRowArgs &, hipStream_t);
extern template hipError_t r2c_impl<float>(int, bool, const RtArgs &, hipStream_t);
extern template hipError_t r2c_impl<double>(int, bool, const RtArgs &, hipStream_t);
extern template hipError_t t23_impl<float>(int, int, const RtArgs &, hipStream_t);
extern template hipError_t t23_impl<double>(int, int, const RtArgs &, hipStream_t);
extern template hipError_t t4_impl<float>(int, int, const RtArgs &, hipStream_t);
extern template hipError_t t4_impl<double>(int, int, const RtArgs &, hipStream_t);

}  // namespace nfft
