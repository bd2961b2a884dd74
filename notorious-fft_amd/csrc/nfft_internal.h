// nfft_internal.h -- host-side internals shared by the planner, the C ABI
// and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <memory>
#include <string>
#include <vector>

#include "nfft_kernels.h"
#include "nfft_kernels_persist.h"
#include "nfft_kernels_xcd.h"
#include "nfft_kernels_smooth.h"
#include "nfft_kernels_rt.h"

namespace nfft {

enum Prec { F32 = 0, F64 = 1 };

inline size_t cx_bytes(int prec) { return prec == F32 ? 8 : 16; }
inline size_t re_bytes(int prec) { return prec == F32 ? 4 : 8; }

// Every device allocation of the library (nfft_plan.cpp).  Failure
// injection for tests: NFFT_FAIL_ALLOC=k[:tag] fails the k-th allocation
// counted from the first one that sees this exact value with
// hipErrorOutOfMemory.  The count restarts only when an allocation sees a
// different value, so a caller re-arming with the same k gives a new tag.
hipError_t dev_malloc_raw(void **p, size_t bytes);
// Device allocations that failed on this thread so far (constructors compare
// it before and after to report NFFT_E_ALLOC rather than NFFT_E_HIP).
long long alloc_failures();
template <typename T> inline hipError_t dev_malloc(T **p, size_t bytes) { return dev_malloc_raw((void **)p, bytes); }

// thread-local error message + status helpers (nfft_plan.cpp)
void set_error(const std::string &msg);
void clear_error();
const char *last_error();

// single-pass limits (LDS: 160 KB per workgroup on gfx950)
constexpr int kLineMax32 = 16384;
constexpr int kLineMax64 = 8192;
inline int line_max(int prec) { return prec == F32 ? kLineMax32 : kLineMax64; }
// column pass (k_col) and transposed row pass (k_rowT) limits
inline int col_max(int /*prec*/) { return 4096; }
inline int rowT_max(int prec) { return prec == F32 ? 2048 : 1024; }
inline int bluerow_max(int prec) { return prec == F32 ? 4096 : 2048; }

// Columns per workgroup of k_col / rows of k_rowT, k_blue_row: powers of two,
// 128-byte global segments where the LDS budget allows, >= 256 threads.
template <typename T, int L> __host__ __device__ constexpr int col_lpb()
{
    const int base = sizeof(T) == 4 ? 16 : 8;
    const int occ = 256 / Sched<L>::NT;
    int v = occ > base ? occ : base;
    while (v > 1 && ((Sched<L>::S > 1 && (long)L * v * (long)sizeof(Cx<T>) > 128 * 1024) ||
                     Sched<L>::NT * v > 1024))
        v /= 2;
    return v;
}
template <typename T, int L> __host__ __device__ constexpr int rowT_lpb()
{
    const int base = sizeof(T) == 4 ? 16 : 8;
    const int occ = 256 / Sched<L>::NT;
    int v = occ > base ? occ : base;
    while (v > 1 && ((long)L * (v + 1) * (long)sizeof(Cx<T>) > 150 * 1024 ||
                     (long)line_lds<T, L, 1>() * v * (long)sizeof(Cx<T>) > 150 * 1024 || Sched<L>::NT * v > 1024))
        v /= 2;
    return v;
}
template <typename T, int L> __host__ __device__ constexpr int bluerow_lpb()
{
    int v = 256 / Sched<L>::NT;
    if (v < 1) v = 1;
    while (v > 1 && ((long)line_lds<T, L, 1>() * v * (long)sizeof(Cx<T>) > 150 * 1024 || Sched<L>::NT * v > 1024))
        v /= 2;
    return v;
}

// runtime mirrors of the above (pow2 L only)
int col_lpb_rt(int prec, int L);
int rowT_lpb_rt(int prec, int L);
int bluerow_lpb_rt(int prec, int L);
int line_lpb_rt(int L);

// ---------------------------------------------------------------------------
// kernel timing (nfft_plan.cpp): while enabled, each launch is bracketed by
// HIP events on its own stream and attributed to a kernel family name with
// its algorithmic bytes (compulsory input + output traffic of that launch).
// ---------------------------------------------------------------------------
bool prof_on();
int prof_begin(hipStream_t s);  // returns a slot or -1
void prof_end(int slot, hipStream_t s, const char *family, int prec, int L, double bytes);

// ---------------------------------------------------------------------------
// launchers (nfft_k_*.hip); return hipSuccess or the launch error
// ---------------------------------------------------------------------------
hipError_t launch_line(int prec, int L, int dir, const LineArgs &a, hipStream_t s);
// whether lines of L points have the half-LDS form (k_line_h)
bool line_h_supported(int prec, int L);
hipError_t launch_blue_line(int prec, int M, const BlueArgs &a, hipStream_t s);
hipError_t launch_col(int prec, int L, int dir, int mode, const ColArgs &a, hipStream_t s);
hipError_t launch_rowT(int prec, int L, int dir, const RowTArgs &a, hipStream_t s);
hipError_t launch_blue_row(int prec, int L, const BlueRowArgs &a, hipStream_t s);
hipError_t launch_rowT_r2c(int prec, int L, const RowR2CArgs &a, hipStream_t s);
hipError_t launch_r2c(int prec, int L, const RtArgs &a, hipStream_t s);
hipError_t launch_c2r(int prec, int L, const RtArgs &a, hipStream_t s);
hipError_t launch_t23(int prec, int L, int op, const RtArgs &a, hipStream_t s);
hipError_t launch_t4(int prec, int L, int op, const RtArgs &a, hipStream_t s);

// persistent four-step (nfft_k_persist.hip): shape `variant` (0 = default)
// of a power-of-two length n; false if that length/variant is not built.
struct PersistCfg {
    int n1 = 0, n2 = 0, lpb1 = 0, lpb2 = 0;
    int threads = 0;  // per workgroup
    int per_cu = 0;   // resident workgroups per CU
    int id = -1;
};
bool persist_find(int prec, long long n, int variant, PersistCfg *cfg);
hipError_t launch_four_persist(const PersistCfg &c, int dir, const PersistArgs &a, int grid, hipStream_t s);
hipError_t launch_persist_reset(unsigned *ctr, long long words, hipStream_t s);

// weighted spectral power (nfft_k_wpow.hip): fused column pass of axis 0
// (length L, ncols columns of stride ncols), its tiles per transform, and the
// elementwise / reduction passes
struct WpowArgs {
    void *buf;                 // in place: batch x prod(dims) complex
    const void *w;             // weights W, prod(dims) reals (shared by the batch)
    double *partial;           // one FP64 partial sum per workgroup
    int ncols;                 // S: columns of axis 0 (prod(dims[1:]))
    long long dist;            // elements per transform (L * S)
    long long ntrans;          // batch
    const void *tw;            // FP64 stage table of L
};
bool col_wpow_supported(int prec, int L);
long long col_wpow_tiles(int prec, int L, int ncols);
hipError_t launch_col_wpow(int prec, int L, const WpowArgs &a, hipStream_t s);
long long wpow_mask_chunks(long long P);
hipError_t launch_wpow_mask(int prec, void *buf, const void *w, double *partial, long long P, long long ntrans,
                            hipStream_t s);
hipError_t launch_wpow_reduce(const double *partial, double *loss, long long per, long long ntrans, hipStream_t s);

// XCD-affine FP32 four-step (nfft_k_xcd.hip): true if length n has an
// instance (N1 x 1024); n1 and the resident workgroups per CU are returned.
bool xcd_supported(long long n, int *n1, int *per_cu);
// The XCC ids on which the current device runs workgroups, as queue numbers
// 0..nq-1 (a one-time probe launch per device); false if it has more XCDs
// than the kernel has queues.
bool xcd_topology(int *nq, signed char *qmap);
// twrow: inter-pass twiddle in the row tiles (FP32 table) or the column tiles
hipError_t launch_four_xcd(int n1, int dir, bool twrow, const XcdArgs &a, int grid, hipStream_t s);

// mixed-radix lines (nfft_k_smooth.hip); smooth_factor() builds the stage
// schedule of a 7-smooth length (false otherwise) with radices up to maxr
// (8, 10 or 25; the bound smooth_lines<..., MAXR> is compiled with).  The
// standalone mixed-radix kernels are compiled for smooth_radix<T>()
// (smooth_radix_of(prec) on the host).
inline int smooth_radix_of(int prec) { return prec == F32 ? smooth_radix<float>() : smooth_radix<double>(); }
bool smooth_factor(long long n, SmoothPlan *sp, int maxr);
hipError_t launch_smooth_line(int prec, int dir, const SmoothArgs &a, hipStream_t s);
// mixed-radix four-step: n = N1 x N2, both 7-smooth and <= maxlen
bool smooth_split(long long n, int maxlen, SmoothPlan *sp1, SmoothPlan *sp2, int maxr);
hipError_t launch_smooth4(int prec, int dir, const Smooth4Args &a, hipStream_t s);
// longest mixed-radix line run by one workgroup (longer 7-smooth lengths
// use the mixed-radix four-step), and the points per workgroup of the line
// and tile kernels (two LDS buffers each: 32 KiB, so that five workgroups
// share a CU)
int smooth_line_max(int prec);
inline int smooth_tile_points(int prec) { return prec == F32 ? 2048 : 1024; }

enum EwKind { EW_UNPACK = 0, EW_PACK = 1, EW_TRIG_PRE = 2, EW_TRIG_POST = 3, EW_N1 = 4 };
hipError_t launch_ew(int prec, int kind, const EwArgs &a, long long nthreads_x, hipStream_t s);

// table generation: out[i] = e^{-2 pi i m(i)/den} in FP64 (out64) and/or
// rounded to FP32 (out32); mode 0: m = (a*i + c) mod den, 1: m = (i>>4)*(i&15),
// 2: m = (i*i) mod den, 3: m = (i>>2) << (i&3) (stage base powers).
hipError_t launch_gen_phase(double2 *out64, float2 *out32, long long count, int mode, long long a,
                            long long c, long long den, hipStream_t s);
// Bluestein filter h (length M): conj chirp wrapped (FP64)
hipError_t launch_blue_filter(double2 *h, const double2 *chirp64, int n, int m, hipStream_t s);
// H (FP64, natural) -> scaled, optionally permuted [k1][k2] = H[k1 + m1*k2], to FP32/FP64
hipError_t launch_blue_hprep(const double2 *H, void *out, int prec, int m, int m1, int m2, double scale,
                             hipStream_t s);
hipError_t launch_fill_uniform(void *out, long long count, uint64_t seed, uint64_t start, int prec,
                               hipStream_t s);
hipError_t launch_fill_gaussian(void *out, long long count, uint64_t seed, uint64_t start, int prec,
                               hipStream_t s);
// strided transpose of a batch of [rows x cols] matrices of elements of
// `esz` bytes (4, 8 or 16): out[b][c][r] = in[b][r][c]
hipError_t launch_transpose(const void *in, void *out, long long batch, long long rows, long long cols, int esz,
                            hipStream_t s);
// elementwise complex ops for the adjoint / loss helpers
hipError_t launch_scale_real(void *buf, long long count, int prec, double s, hipStream_t st);
// six-step pieces: out[b][c][r] = in[b][r][c] * w_N^{dir r c} * scale (tw: the
// two-level FP64 table of N, rows * cols <= N)
hipError_t launch_tw_transpose(int prec, const void *in, void *out, long long batch, long long rows, long long cols,
                               const Tw2 &tw, int dir, double scale, hipStream_t s);
// Bluestein through whole-length engines: stage 0 z = pad(chirp * conj?(x))
// (x strided by dist), 1 z *= H, 2 y = conj?(chirp * z[0, n)) * scale (y
// strided by dist); z holds ntrans x m points
hipError_t launch_blue_ew(int prec, int stage, const void *src, void *dst, const void *table, long long n, long long m,
                          long long dist, long long ntrans, int conj, double scale, hipStream_t s);

// ---------------------------------------------------------------------------
// per-device table cache (FP64-generated twiddles, shared by all plans)
// ---------------------------------------------------------------------------
struct StageTw {
    const void *ptr;  // Sched<L>::TWN entries of Cx<T>
};
const void *stage_table(int prec, int L);                         // nullptr on failure
bool tw2_table(long long n, Tw2 *out);                            // FP64 two-level e^{-2 pi i e/n}
const void *phase_table(int prec, long long count, long long a, long long c, long long den);
const void *chirp_table(int prec, long long n, const double2 **chirp64);

// ---------------------------------------------------------------------------
// 1D complex engine over contiguous lines
// ---------------------------------------------------------------------------
enum C2CAlgo {
    ALG_LINE = 0, ALG_FOUR = 1, ALG_BLUE1 = 2, ALG_BLUE3 = 3, ALG_N1 = 4, ALG_SMOOTH = 5, ALG_SMOOTH4 = 6,
    ALG_BIG = 7,     // power of two beyond the four-step: N1 x N2 over two sub-engines (six-step)
    ALG_BLUEBIG = 8  // Bluestein with M beyond the three-pass form: M-point sub-engine
};
// Longest length a C2C engine accepts (2^30: offsets and exponents stay
// within the 32-bit twiddle exponent and the int size of the C API).
constexpr long long kMaxLen = 1LL << 30;

struct C2C {
    int n = 0;
    int prec = F32;
    int algo = ALG_LINE;
    int L = 0;           // ALG_LINE: n; ALG_BLUE1: m
    int halves = 0;      // ALG_LINE: one line per workgroup, exchanges in two LDS halves
    int n1 = 0, n2 = 0;  // ALG_FOUR: columns FFT length n1 over n2 columns, then rows of n2
    int m = 0, m1 = 0, m2 = 0;
    const void *tw_a = nullptr, *tw_b = nullptr;  // stage tables (L or n1/m1, n2/m2)
    Tw2 tw2{};
    const void *chirp = nullptr;
    void *H = nullptr;   // owned: Bluestein filter spectrum (scaled 1/m)
    long long scratch_per_line = 0;  // elements of Cx per line (ALG_FOUR: n, ALG_BLUE3: m)
    int col_lpb = 0, row_lpb = 0;     // tile widths of the two passes (0: defaults)
    int row_pf = 1;                   // row pass on the resident prefetching grid
    int row_twin = 0;                 // inter-pass twiddle applied by the row pass
    Tw2f tw2f{};                      // its FP32 table
    SmoothPlan sp{};                  // ALG_SMOOTH: stage schedule; ALG_SMOOTH4: N1
    SmoothPlan sp2{};                 // ALG_SMOOTH4: N2
    int sgroup = 1;                   // ALG_SMOOTH: lines per workgroup; ALG_SMOOTH4: columns per tile
    int sgroup2 = 1;                  // ALG_SMOOTH4: rows per tile
    // ALG_BIG: sub1 (length n1) and sub2 (n2); ALG_BLUEBIG: sub1 (length m).
    // Shared, so that copies of an engine (axis lists) use the same tables;
    // release() of the owner releases them.
    std::shared_ptr<C2C> sub1, sub2;
    // ALG_FOUR in one persistent launch (k_four_persist) when pctr is set
    PersistCfg pcfg{};
    unsigned *pctr = nullptr;         // owned: counter block for up to pcap transforms
    long long pcap = 0;
    int pgrid = 0;                    // workgroups of the persistent launch
    long long ptimeout = 0;           // wall-clock ticks a dependency wait may take
    int pgrab = 1;                    // queue items taken per access
    int plag = 0;                     // pass-2 lag in transforms (0: half the slots)
    long long pslots = 1;             // rotating slots at most (Infinity Cache budget)
    // FP32 ALG_FOUR in one XCD-affine launch (k_four_xcd) when xctr is set
    // and a batch has at least kXcdQueues * kXcdSlots transforms
    unsigned *xctr = nullptr;         // owned: counter block for up to xcap transforms
    long long xcap = 0;
    int xgrid = 0;                    // workgroups of the launch
    int xnq = 0;                      // XCD queues and the XCC_ID -> queue map
    signed char xqmap[kXcdIds] = {};

    // Set up tables for length n; returns false with the error set.
    // smooth_mode for 7-smooth non-power-of-two n: 2 = choose (tuned /
    // measured rule), 1 = mixed radix, 0 = Bluestein (callers that need its
    // chirp / filter tables)
    // persist_mode = 0 keeps the two-launch four-step for powers of two
    // (callers that fuse their own work into its row pass)
    bool init(int n, int prec, int smooth_mode = 2, int persist_mode = 2);
    void release();
    std::string describe() const;
    // dir: -1 forward, +1 inverse (unnormalised); scale applied to the output.
    // scratch: scratch_lines * scratch_per_line elements (may be null when
    // scratch_per_line == 0).
    hipError_t run(const void *in, void *out, long long nlines, long long in_dist, long long out_dist, int dir,
                   double scale, void *scratch, long long scratch_lines, hipStream_t s) const;
    // true if a persistent launch of this engine ever gave up a dependency
    // wait (call after the stream has been synchronised)
    bool device_fault() const;
};

// strided complex transform along one axis of a row-major array:
// element (a, i, b) at a*n*B + i*B + b, in place on `buf`, or from `src`
// (same layout) into `buf` when src is not null.
hipError_t c2c_axis(const C2C &e, void *buf, long long A, long long B, int dir, double scale, void *scratch,
                    long long scratch_elems, hipStream_t s, const void *src = nullptr);

}  // namespace nfft
