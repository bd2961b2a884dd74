// nfft_k_misc.hip -- FP64 table generation on the device, Bluestein filter
// preparation, elementwise real/trig passes, tiled transposes, the synthetic
// input generator and the precision dispatch of every launcher.
#include "nfft_launch_impl.h"

namespace nfft {

// ---------------------------------------------------------------------------
// precision dispatch
// ---------------------------------------------------------------------------
// Each dispatcher records its launch with the kernel timer when it is on.
#define NFFT_TIMED(FAMILY, L, BYTES, CALL)                                                                    \
    do {                                                                                                      \
        const int slot_ = prof_on() ? prof_begin(s) : -1;                                                     \
        hipError_t e_ = (CALL);                                                                               \
        if (slot_ >= 0) prof_end(slot_, s, FAMILY, prec, L, (BYTES));                                         \
        return e_;                                                                                            \
    } while (0)

static double cxb(int prec) { return prec == F32 ? 8.0 : 16.0; }

template <typename T> static bool line_h_sup(int L)
{
    switch (L) {
    case 8192: return line_h_fits<T, 8192>();
    case 16384: return line_h_fits<T, 16384>();
    case 32768: return line_h_fits<T, 32768>();
    default: return false;
    }
}
bool line_h_supported(int prec, int L) { return prec == F32 ? line_h_sup<float>(L) : line_h_sup<double>(L); }

hipError_t launch_line(int prec, int L, int dir, const LineArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_line", L, 2.0 * a.nlines * L * cxb(prec),
               prec == F32 ? line_impl<float>(L, dir, a, s) : line_impl<double>(L, dir, a, s));
}
hipError_t launch_blue_line(int prec, int M, const BlueArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_blue_line", M, 2.0 * a.nlines * a.n * cxb(prec),
               prec == F32 ? blue_line_impl<float>(M, a, s) : blue_line_impl<double>(M, a, s));
}
hipError_t launch_col(int prec, int L, int dir, int mode, const ColArgs &a, hipStream_t s)
{
    static const char *names[4] = {"k_col_plain", "k_col_twiddle", "k_col_blue_in", "k_col_blue_out"};
    const double tile = (double)a.ntrans * a.ncols * L * cxb(prec);
    const double nvb = (double)a.ntrans * a.nv * cxb(prec);
    const double bytes = mode == COL_BLUE_IN || mode == COL_BLUE_OUT ? tile + nvb : 2.0 * tile;
    NFFT_TIMED(names[mode & 3], L, bytes,
               prec == F32 ? col_impl<float>(L, dir, mode, a, s) : col_impl<double>(L, dir, mode, a, s));
}
hipError_t launch_rowT(int prec, int L, int dir, const RowTArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_rowT", L, 2.0 * a.ntrans * a.nrows * L * cxb(prec),
               prec == F32 ? rowT_impl<float>(L, dir, a, s) : rowT_impl<double>(L, dir, a, s));
}
hipError_t launch_rowT_r2c(int prec, int L, const RowR2CArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_rowT_r2c", L,
               (double)a.ntrans * a.nrows * L * cxb(prec) + (double)a.ntrans * (a.nrows * (double)L + 1) * cxb(prec),
               prec == F32 ? rowT_r2c_impl<float>(L, a, s) : rowT_r2c_impl<double>(L, a, s));
}
hipError_t launch_blue_row(int prec, int L, const BlueRowArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_blue_row", L, 2.0 * a.ntrans * a.nrows * L * cxb(prec),
               prec == F32 ? blue_row_impl<float>(L, a, s) : blue_row_impl<double>(L, a, s));
}
hipError_t launch_r2c(int prec, int L, const RtArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_r2c_line", L, (double)a.nlines * (2.0 * L + 1) * cxb(prec),
               prec == F32 ? r2c_impl<float>(L, false, a, s) : r2c_impl<double>(L, false, a, s));
}
hipError_t launch_c2r(int prec, int L, const RtArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_c2r_line", L, (double)a.nlines * (2.0 * L + 1) * cxb(prec),
               prec == F32 ? r2c_impl<float>(L, true, a, s) : r2c_impl<double>(L, true, a, s));
}
hipError_t launch_t23(int prec, int L, int op, const RtArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_t23_line", L, 2.0 * a.nlines * L * cxb(prec),
               prec == F32 ? t23_impl<float>(L, op, a, s) : t23_impl<double>(L, op, a, s));
}
hipError_t launch_t4(int prec, int L, int op, const RtArgs &a, hipStream_t s)
{
    NFFT_TIMED("k_t4_line", L, 2.0 * a.nlines * L * cxb(prec),
               prec == F32 ? t4_impl<float>(L, op, a, s) : t4_impl<double>(L, op, a, s));
}

template <int L> static int col_lpb_of(int prec) { return prec == F32 ? col_lpb<float, L>() : col_lpb<double, L>(); }
template <int L> static int rowT_lpb_of(int prec) { return prec == F32 ? rowT_lpb<float, L>() : rowT_lpb<double, L>(); }
template <int L> static int bluerow_lpb_of(int prec)
{
    return prec == F32 ? bluerow_lpb<float, L>() : bluerow_lpb<double, L>();
}
int col_lpb_rt(int prec, int L)
{
    switch (L) {
#define NFFT_CASE(LL) case LL: return col_lpb_of<LL>(prec);
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return 1;
}
int rowT_lpb_rt(int prec, int L)
{
    switch (L) {
#define NFFT_CASE(LL) case LL: return rowT_lpb_of<LL>(prec);
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return 1;
}
int bluerow_lpb_rt(int prec, int L)
{
    switch (L) {
#define NFFT_CASE(LL) case LL: return bluerow_lpb_of<LL>(prec);
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return 1;
}
int line_lpb_rt(int L)
{
    switch (L) {
#define NFFT_CASE(LL) case LL: return line_lpb<LL>();
        NFFT_POW2_LIST(NFFT_CASE)
#undef NFFT_CASE
    }
    return 1;
}

// ---------------------------------------------------------------------------
// elementwise real / trig passes
// ---------------------------------------------------------------------------
hipError_t launch_ew(int prec, int kind, const EwArgs &a, long long nx, hipStream_t s)
{
    unsigned gy, gz;
    grid_trans(a.nlines, gy, gz);
    dim3 b(256), g((unsigned)((nx + 255) / 256), gy, gz);
    if (kind == EW_N1) g = dim3((unsigned)((a.nlines + 255) / 256));
#define NFFT_EW(K)                                                                                            \
    if (prec == F32) hipLaunchKernelGGL((K<float>), g, b, 0, s, a);                                           \
    else hipLaunchKernelGGL((K<double>), g, b, 0, s, a);
    switch (kind) {
    case EW_UNPACK: NFFT_EW(k_ew_unpack) break;
    case EW_PACK: NFFT_EW(k_ew_pack) break;
    case EW_TRIG_PRE: NFFT_EW(k_ew_trig_pre) break;
    case EW_TRIG_POST: NFFT_EW(k_ew_trig_post) break;
    case EW_N1: NFFT_EW(k_ew_n1) break;
    default: return hipErrorInvalidValue;
    }
#undef NFFT_EW
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// twiddle / chirp tables: e^{-2 pi i m/den}, evaluated in FP64 from the exact
// integer phase m (reduced mod den), so the argument of sincospi is the
// correctly rounded rational 2m/den in (-1, 1].
// ---------------------------------------------------------------------------
__global__ void k_gen_phase(double2 *o64, float2 *o32, long long count, int mode, long long a, long long c,
                            long long den)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const unsigned long long ud = (unsigned long long)den;
    unsigned long long m;
    if (mode == 0) m = ((unsigned long long)a * (unsigned long long)i + (unsigned long long)c) % ud;
    else if (mode == 1) m = ((unsigned long long)(i >> 4) * (unsigned long long)(i & 15)) % ud;
    else if (mode == 3) m = ((unsigned long long)(i >> 2) << (i & 3)) % ud;
    else m = ((unsigned long long)i * (unsigned long long)i) % ud;
    double x = 2.0 * (double)m / (double)den;  // in [0, 2)
    if (x > 1.0) x -= 2.0;                     // in (-1, 1]
    double sv, cv;
    sincospi(x, &sv, &cv);
    if (o64) o64[i] = make_double2(cv, -sv);
    if (o32) o32[i] = make_float2((float)cv, (float)-sv);
}

hipError_t launch_gen_phase(double2 *out64, float2 *out32, long long count, int mode, long long a, long long c,
                            long long den, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_phase, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, out64, out32, count,
                       mode, a, c, den);
    return hipGetLastError();
}

// h[m] = conj(a[m]) for m < n, conj(a[M-m]) for M-m < n, else 0 (wrapped
// symmetric filter of the chirp a[k] = e^{-i pi k^2/n}).
__global__ void k_blue_filter(double2 *h, const double2 *a, int n, int M)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    double2 v = make_double2(0.0, 0.0);
    if (m < n) v = make_double2(a[m].x, -a[m].y);
    else if (M - m < n) v = make_double2(a[M - m].x, -a[M - m].y);
    h[m] = v;
}

hipError_t launch_blue_filter(double2 *h, const double2 *chirp64, int n, int m, hipStream_t s)
{
    hipLaunchKernelGGL(k_blue_filter, dim3((m + 255) / 256), dim3(256), 0, s, h, chirp64, n, m);
    return hipGetLastError();
}

// out = scale * H, natural (m1 == 0) or [k1][k2] = H[k1 + m1 k2] order.
__global__ void k_blue_hprep(const double2 *H, void *out, int prec, int m, int m1, int m2, double scale)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m) return;
    int src = idx;
    if (m1 > 0) {
        const int k1 = idx / m2, k2 = idx % m2;
        src = k1 + m1 * k2;
    }
    const double2 v = make_double2(H[src].x * scale, H[src].y * scale);
    if (prec == F32) ((float2 *)out)[idx] = make_float2((float)v.x, (float)v.y);
    else ((double2 *)out)[idx] = v;
}

hipError_t launch_blue_hprep(const double2 *H, void *out, int prec, int m, int m1, int m2, double scale,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_blue_hprep, dim3((m + 255) / 256), dim3(256), 0, s, H, out, prec, m, m1, m2, scale);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// synthetic input: counter-based SplitMix64, uniform [-1, 1)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sm64_uniform(unsigned long long seed, unsigned long long idx)
{
    unsigned long long z = ((seed << 40) | idx) + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

__global__ void k_fill_uniform(void *out, long long count, unsigned long long seed, unsigned long long start,
                               int prec)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const double v = sm64_uniform(seed, start + (unsigned long long)i);
    if (prec == F32) ((float *)out)[i] = (float)v;
    else ((double *)out)[i] = v;
}

hipError_t launch_fill_uniform(void *out, long long count, uint64_t seed, uint64_t start, int prec, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const long long blocks = (count + 255) / 256;
    if (blocks > 0xffffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)blocks), dim3(256), 0, s, out, count, seed, start, prec);
    return hipGetLastError();
}

// Gaussian N(0, 1) by Box-Muller on consecutive uniforms of the same stream:
// value j (j = start + i) belongs to pair p = j / 2 with u1 = U(2p), u2 =
// U(2p + 1), U(k) = ((z_k >> 11) + 1) 2^-53 in (0, 1]; it is
// sqrt(-2 ln u1) cos(2 pi u2) for even j and sqrt(-2 ln u1) sin(2 pi u2) for
// odd j (FP64, rounded once to the output type).
__device__ __forceinline__ double sm64_unit_open0(unsigned long long seed, unsigned long long idx)
{
    unsigned long long z = ((seed << 40) | idx) + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)((z >> 11) + 1) * (1.0 / 9007199254740992.0);
}

__global__ void k_fill_gaussian(void *out, long long count, unsigned long long seed, unsigned long long start,
                                int prec)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const unsigned long long j = start + (unsigned long long)i, p = j >> 1;
    const double u1 = sm64_unit_open0(seed, 2 * p), u2 = sm64_unit_open0(seed, 2 * p + 1);
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    const double v = r * ((j & 1) ? sn : cs);
    if (prec == F32) ((float *)out)[i] = (float)v;
    else ((double *)out)[i] = v;
}

hipError_t launch_fill_gaussian(void *out, long long count, uint64_t seed, uint64_t start, int prec, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const long long blocks = (count + 255) / 256;
    if (blocks > 0xffffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fill_gaussian, dim3((unsigned)blocks), dim3(256), 0, s, out, count, seed, start, prec);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// batched transpose through a 32 x 33 LDS tile: out[b][c][r] = in[b][r][c]
// ---------------------------------------------------------------------------
template <typename E> __global__ void k_transpose(const E *in, E *out, long long rows, long long cols)
{
    __shared__ E tile[32][33];
    const long long b = blockIdx.z;
    const long long r0 = (long long)blockIdx.y * 32, c0 = (long long)blockIdx.x * 32;
    const E *src = in + b * rows * cols;
    E *dst = out + b * rows * cols;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const long long r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = src[r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const long long c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[c * rows + r] = tile[tx][k];
    }
}

hipError_t launch_transpose(const void *in, void *out, long long batch, long long rows, long long cols, int esz,
                            hipStream_t s)
{
    if (batch <= 0 || rows <= 0 || cols <= 0) return hipSuccess;
    if ((cols + 31) / 32 > 0x7fffffffLL || (rows + 31) / 32 > 65535) return hipErrorInvalidValue;
    if (esz != 4 && esz != 8 && esz != 16) return hipErrorInvalidValue;
    // grid.z <= 65535: walk the batch in slices
    for (long long b0 = 0; b0 < batch; b0 += 65535) {
        const long long nb = batch - b0 < 65535 ? batch - b0 : 65535;
        const size_t off = (size_t)b0 * rows * cols * esz;
        const void *src = (const char *)in + off;
        void *dst = (char *)out + off;
        dim3 g((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32), (unsigned)nb), b(32, 8);
        if (esz == 4) hipLaunchKernelGGL((k_transpose<float>), g, b, 0, s, (const float *)src, (float *)dst, rows, cols);
        else if (esz == 8)
            hipLaunchKernelGGL((k_transpose<float2>), g, b, 0, s, (const float2 *)src, (float2 *)dst, rows, cols);
        else hipLaunchKernelGGL((k_transpose<double2>), g, b, 0, s, (const double2 *)src, (double2 *)dst, rows, cols);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Six-step pieces for lengths beyond the four-step (N = N1 x N2, both
// transformed by ordinary engines): out[b][c][r] = in[b][r][c] * w_N^{dir r c}
// * scale through a 32 x 33 LDS tile; the twiddle from the FP64 two-level
// table (exponent r c < N, exact), the product in FP64.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_tw_transpose(const Cx<T> *in, Cx<T> *out, long long rows, long long cols, Tw2 tw, int dir,
                               double scale)
{
    __shared__ Cx<T> tile[32][33];
    const long long b = blockIdx.z;
    const long long r0 = (long long)blockIdx.y * 32, c0 = (long long)blockIdx.x * 32;
    const Cx<T> *src = in + b * rows * cols;
    Cx<T> *dst = out + b * rows * cols;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const long long r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) {
            const Cx<T> v = src[r * cols + c];
            double2 w = tw2_get(tw, (uint32_t)(r * c));
            w.x *= scale;
            w.y *= dir > 0 ? -scale : scale;
            tile[k][tx] = cmk<Cx<T>>((T)((double)v.x * w.x - (double)v.y * w.y), (T)((double)v.x * w.y + (double)v.y * w.x));
        }
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const long long c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[c * rows + r] = tile[tx][k];
    }
}

hipError_t launch_tw_transpose(int prec, const void *in, void *out, long long batch, long long rows, long long cols,
                               const Tw2 &tw, int dir, double scale, hipStream_t s)
{
    if (batch <= 0 || rows <= 0 || cols <= 0) return hipSuccess;
    if ((cols + 31) / 32 > 0x7fffffffLL || (rows + 31) / 32 > 65535 || rows * cols > (1LL << 32))
        return hipErrorInvalidValue;
    const size_t esz = prec == F32 ? 8 : 16;
    for (long long b0 = 0; b0 < batch; b0 += 65535) {
        const long long nb = batch - b0 < 65535 ? batch - b0 : 65535;
        const size_t off = (size_t)b0 * rows * cols * esz;
        dim3 g((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32), (unsigned)nb), bl(32, 8);
        if (prec == F32)
            hipLaunchKernelGGL((k_tw_transpose<float>), g, bl, 0, s, (const float2 *)((const char *)in + off),
                               (float2 *)((char *)out + off), rows, cols, tw, dir, scale);
        else
            hipLaunchKernelGGL((k_tw_transpose<double>), g, bl, 0, s, (const double2 *)((const char *)in + off),
                               (double2 *)((char *)out + off), rows, cols, tw, dir, scale);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Bluestein through whole-length engines (lengths beyond the three-pass
// form): Z = pad(chirp * conj?(x)) over M points; Z *= H; y = conj?(chirp *
// Z[0, n)) * scale.  blockIdx.y = transform.
template <typename T>
__global__ void k_blue_pre(const Cx<T> *x, Cx<T> *z, const Cx<T> *chirp, long long n, long long m, long long in_dist,
                           int conj)
{
    const long long t = blockIdx.y;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (long long)gridDim.x * blockDim.x) {
        Cx<T> v = czero<T>();
        if (k < n) {
            v = x[t * in_dist + k];
            if (conj) v = cconj(v);
            v = cmul(v, chirp[k]);
        }
        z[t * m + k] = v;
    }
}
template <typename T> __global__ void k_blue_mulH(Cx<T> *z, const Cx<T> *H, long long m)
{
    const long long t = blockIdx.y;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (long long)gridDim.x * blockDim.x)
        z[t * m + k] = cmul(z[t * m + k], H[k]);
}
template <typename T>
__global__ void k_blue_post(const Cx<T> *z, Cx<T> *y, const Cx<T> *chirp, long long n, long long m, long long out_dist,
                            int conj, double scale)
{
    const long long t = blockIdx.y;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        Cx<T> v = cmul(z[t * m + k], chirp[k]);
        if (conj) v = cconj(v);
        y[t * out_dist + k] = cscale(v, (T)scale);
    }
}

hipError_t launch_blue_ew(int prec, int stage, const void *src, void *dst, const void *table, long long n, long long m,
                          long long dist, long long ntrans, int conj, double scale, hipStream_t s)
{
    if (ntrans <= 0) return hipSuccess;
    const long long len = stage == 2 ? n : m;
    for (long long t0 = 0; t0 < ntrans; t0 += 65535) {
        const long long nt = ntrans - t0 < 65535 ? ntrans - t0 : 65535;
        const dim3 g((unsigned)std::min<long long>(1024, (len + 255) / 256), (unsigned)nt), b(256);
        const size_t cb = prec == F32 ? 8 : 16;
        const char *sp = (const char *)src;
        char *dp = (char *)dst;
        // offsets of this slice of transforms: src / dst strides by stage
        const long long sstride = stage == 0 ? dist : m, dstride = stage == 2 ? dist : m;
        sp += (size_t)t0 * sstride * cb;
        dp += (size_t)t0 * dstride * cb;
        if (prec == F32) {
            if (stage == 0) hipLaunchKernelGGL(k_blue_pre<float>, g, b, 0, s, (const float2 *)sp, (float2 *)dp, (const float2 *)table, n, m, dist, conj);
            else if (stage == 1) hipLaunchKernelGGL(k_blue_mulH<float>, g, b, 0, s, (float2 *)dp, (const float2 *)table, m);
            else hipLaunchKernelGGL(k_blue_post<float>, g, b, 0, s, (const float2 *)sp, (float2 *)dp, (const float2 *)table, n, m, dist, conj, scale);
        } else {
            if (stage == 0) hipLaunchKernelGGL(k_blue_pre<double>, g, b, 0, s, (const double2 *)sp, (double2 *)dp, (const double2 *)table, n, m, dist, conj);
            else if (stage == 1) hipLaunchKernelGGL(k_blue_mulH<double>, g, b, 0, s, (double2 *)dp, (const double2 *)table, m);
            else hipLaunchKernelGGL(k_blue_post<double>, g, b, 0, s, (const double2 *)sp, (double2 *)dp, (const double2 *)table, n, m, dist, conj, scale);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

__global__ void k_scale_real(void *buf, long long count, int prec, double sc)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (prec == F32) ((float *)buf)[i] *= (float)sc;
    else ((double *)buf)[i] *= sc;
}

hipError_t launch_scale_real(void *buf, long long count, int prec, double sc, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_real, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, buf, count, prec, sc);
    return hipGetLastError();
}

}  // namespace nfft
