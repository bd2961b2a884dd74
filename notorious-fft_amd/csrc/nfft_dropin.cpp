// nfft_dropin.cpp -- the reference's 27-function C API (src/07_legacy.h:29-57)
// on top of device-resident plans.  Compiled twice: without
// NOTORIOUS_FFT_SINGLE (double, libnotorious_fft_hip_f64.so) and with it
// (float, libnotorious_fft_hip_f32.so), as the reference switches its element
// type at compile time (src/01_core.h:36-40).
//
// Semantics kept from the reference:
//   - constructors return NULL for N <= 0 and, for realdft/t2t3/t4, for
//     non-power-of-two sizes (src/07_legacy.h:1406,1442,1529,1579); an N-D
//     aux is NULL when any axis is (src/07_legacy.h:1654-1688);
//   - invdft is unnormalised for power-of-two lengths and scaled by 1/N along
//     non-power-of-two axes (src/05_algorithms.h:315-333);
//   - transforms are synchronous, never modify their input, allow x == y,
//     and do nothing when handed a NULL aux; work the caller queued on the
//     device before the call is complete before the transform reads x;
//   - free_aux(NULL) is a no-op (src/07_legacy.h:1690-1691).
// Multi-dimensional realdft, declared but not implemented by the reference
// (src/07_legacy.h:1492-1514), follows FFTW's layout: the last axis is halved
// to N/2+1 complex bins.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/notorious_fft.h"
#include "../../include/notorious_fft_hip.h"

// the core library's device allocation (with its failure injection for tests)
namespace nfft {
hipError_t dev_malloc_raw(void **p, size_t bytes);
}

#ifdef NOTORIOUS_FFT_SINGLE
#define NFFT_DROPIN_PREC NFFT_F32
#else
#define NFFT_DROPIN_PREC NFFT_F64
#endif

namespace {

struct Impl {
    nfft_plan *plan = nullptr;
    int kind = 0;
    std::vector<int> dims;
    long long P = 1;
    hipStream_t st = nullptr;
    void *din = nullptr, *dout = nullptr;
    size_t stage_bytes = 0;
    void *hin = nullptr, *hout = nullptr;  // pinned host staging of small host-buffer calls
    void *hin_d = nullptr, *hout_d = nullptr;  // their device-side addresses
    size_t hstage_bytes = 0;
};

// Host buffers up to this size go through pinned staging buffers of the aux
// (a host copy into and out of them, DMA transfers on the aux's stream):
// lower latency than transfers from pageable memory for small transforms,
// while larger ones keep the direct pageable transfer (no extra host copy).
constexpr size_t kPinnedStageMax = 4u << 20;

// Up to this many bytes a 1-D transform runs on the mapped staging buffers
// directly, without DMA (NFFT_DROPIN_ZC_MAX overrides; 0 turns it off).
size_t zero_copy_max()
{
    const char *e = getenv("NFFT_DROPIN_ZC_MAX");
    return e && *e ? (size_t)strtoull(e, nullptr, 10) : (size_t)(256u << 10);
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

notorious_fft_aux *make(int kind, int d, const int *Ns)
{
    if (d < 1 || !Ns) return nullptr;
    for (int i = 0; i < d; i++) {
        if (Ns[i] <= 0) return nullptr;
        const bool need_pow2 = kind == NFFT_KIND_T2T3 || kind == NFFT_KIND_T4 ||
                               (kind == NFFT_KIND_REALDFT && i == d - 1);
        if (need_pow2 && !is_pow2(Ns[i])) return nullptr;
    }
    Impl *im = new (std::nothrow) Impl();
    if (!im) return nullptr;
    if (nfft_plan_create(&im->plan, kind, NFFT_DROPIN_PREC, d, Ns, 1, NFFT_NORM_REFERENCE) != NFFT_OK) {
        delete im;
        return nullptr;
    }
    if (hipStreamCreateWithFlags(&im->st, hipStreamNonBlocking) != hipSuccess) {
        nfft_plan_destroy(im->plan);
        delete im;
        return nullptr;
    }
    im->kind = kind;
    im->dims.assign(Ns, Ns + d);
    for (int i = 0; i < d; i++) im->P *= Ns[i];
    notorious_fft_aux *a = new (std::nothrow) notorious_fft_aux;
    if (!a) {
        (void)hipStreamDestroy(im->st);
        nfft_plan_destroy(im->plan);
        delete im;
        return nullptr;
    }
    a->N = (int)im->P;
    a->impl = im;
    return a;
}

bool on_device(const void *p)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

// bytes of the input and output of `op` for this aux
void op_bytes(const Impl *im, int op, size_t *ib, size_t *ob)
{
    const size_t r = sizeof(notorious_fft_real), c = 2 * sizeof(notorious_fft_real);
    const long long last = im->dims.back();
    const long long half = im->P / last * (last / 2 + 1);
    switch (op) {
    case NFFT_OP_DFT:
    case NFFT_OP_INVDFT: *ib = *ob = (size_t)im->P * c; break;
    case NFFT_OP_REALDFT: *ib = (size_t)im->P * r; *ob = (size_t)half * c; break;
    case NFFT_OP_INVREALDFT: *ib = (size_t)half * c; *ob = (size_t)im->P * r; break;
    default: *ib = *ob = (size_t)im->P * r; break;
    }
}

void run(const notorious_fft_aux *a, int op, const void *x, void *y)
{
    if (!a || !a->impl || !x || !y) return;
    Impl *im = (Impl *)a->impl;
    size_t ib, ob;
    op_bytes(im, op, &ib, &ob);
    // The reference's calls are synchronous on the caller's thread, so any
    // write of x (or read of y) the caller has already queued on a device
    // stream -- a kernel, an async copy on the null stream or on a stream of
    // its own -- happens before this transform.  The aux's stream is
    // non-blocking, so that order is established here for every pointer kind.
    if (hipDeviceSynchronize() != hipSuccess) return;
    const bool xd = on_device(x), yd = on_device(y);
    if (xd && yd) {
        if (nfft_execute(im->plan, op, x, y, im->st) == NFFT_OK) (void)hipStreamSynchronize(im->st);
        return;
    }
    const size_t need = ib > ob ? ib : ob;
    bool pin = need <= kPinnedStageMax;
    if (pin && im->hstage_bytes < need) {
        if (im->hin) (void)hipHostFree(im->hin);
        if (im->hout) (void)hipHostFree(im->hout);
        im->hin = im->hout = im->hin_d = im->hout_d = nullptr;
        im->hstage_bytes = 0;
        const unsigned fl = hipHostMallocMapped | hipHostMallocCoherent;
        if (hipHostMalloc(&im->hin, need, fl) == hipSuccess && hipHostMalloc(&im->hout, need, fl) == hipSuccess &&
            hipHostGetDevicePointer(&im->hin_d, im->hin, 0) == hipSuccess &&
            hipHostGetDevicePointer(&im->hout_d, im->hout, 0) == hipSuccess) {
            im->hstage_bytes = need;
        } else {
            (void)hipGetLastError();
            if (im->hin) (void)hipHostFree(im->hin);
            if (im->hout) (void)hipHostFree(im->hout);
            im->hin = im->hout = im->hin_d = im->hout_d = nullptr;
            pin = false;  // pageable transfers instead
        }
    }
    if (pin && need <= zero_copy_max() && im->dims.size() == 1) {
        // The transform reads and writes the mapped staging buffers in place
        // over the bus: no DMA transfers, whose fixed cost (about 10 us each
        // for 64 KiB) exceeds a one-pass transform at these sizes.  N-D plans
        // make several passes over their output and keep the DMA transfers
        // (each pass over the bus would cost more than the two copies).
        const void *src = x;
        if (!xd) {
            memcpy(im->hin, x, ib);
            src = im->hin_d;
        }
        if (nfft_execute(im->plan, op, src, yd ? y : im->hout_d, im->st) != NFFT_OK) return;
        if (hipStreamSynchronize(im->st) != hipSuccess) return;
        if (!yd) memcpy(y, im->hout, ob);
        return;
    }
    if (im->stage_bytes < need) {
        if (im->din) (void)hipFree(im->din);
        if (im->dout) (void)hipFree(im->dout);
        im->din = im->dout = nullptr;
        im->stage_bytes = 0;
        if (nfft::dev_malloc_raw(&im->din, need) != hipSuccess || nfft::dev_malloc_raw(&im->dout, need) != hipSuccess)
            return;
        im->stage_bytes = need;
    }
    const void *src = x;
    void *dst = yd ? y : im->dout;
    if (!xd) {
        if (pin) {
            memcpy(im->hin, x, ib);
            if (hipMemcpyAsync(im->din, im->hin, ib, hipMemcpyHostToDevice, im->st) != hipSuccess) return;
        } else if (hipMemcpyAsync(im->din, x, ib, hipMemcpyHostToDevice, im->st) != hipSuccess) {
            return;
        }
        src = im->din;
    }
    if (nfft_execute(im->plan, op, src, dst, im->st) != NFFT_OK) return;
    if (!yd) {
        if (pin) {
            if (hipMemcpyAsync(im->hout, im->dout, ob, hipMemcpyDeviceToHost, im->st) != hipSuccess) return;
            if (hipStreamSynchronize(im->st) != hipSuccess) return;
            memcpy(y, im->hout, ob);
            return;
        }
        (void)hipMemcpyAsync(y, im->dout, ob, hipMemcpyDeviceToHost, im->st);
    }
    (void)hipStreamSynchronize(im->st);
}

}  // namespace

extern "C" {

void notorious_fft_dft(notorious_fft_cmpl *x, notorious_fft_cmpl *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_DFT, x, y);
}
void notorious_fft_invdft(notorious_fft_cmpl *x, notorious_fft_cmpl *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_INVDFT, x, y);
}
void notorious_fft_realdft(notorious_fft_real *x, notorious_fft_cmpl *z, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_REALDFT, x, z);
}
void notorious_fft_invrealdft(notorious_fft_cmpl *z, notorious_fft_real *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_INVREALDFT, z, y);
}
void notorious_fft_dct2(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_DCT2, x, y);
}
void notorious_fft_dst2(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_DST2, x, y);
}
void notorious_fft_dct3(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_DCT3, x, y);
}
void notorious_fft_dst3(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_DST3, x, y);
}
void notorious_fft_dct4(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_DCT4, x, y);
}
void notorious_fft_dst4(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a)
{
    run(a, NFFT_OP_DST4, x, y);
}

notorious_fft_aux *notorious_fft_mkaux_dft_1d(int N) { return make(NFFT_KIND_DFT, 1, &N); }
notorious_fft_aux *notorious_fft_mkaux_dft_2d(int N1, int N2)
{
    int Ns[2] = {N1, N2};
    return make(NFFT_KIND_DFT, 2, Ns);
}
notorious_fft_aux *notorious_fft_mkaux_dft_3d(int N1, int N2, int N3)
{
    int Ns[3] = {N1, N2, N3};
    return make(NFFT_KIND_DFT, 3, Ns);
}
notorious_fft_aux *notorious_fft_mkaux_dft(int d, int *Ns) { return make(NFFT_KIND_DFT, d, Ns); }

notorious_fft_aux *notorious_fft_mkaux_realdft_1d(int N) { return make(NFFT_KIND_REALDFT, 1, &N); }
notorious_fft_aux *notorious_fft_mkaux_realdft_2d(int N1, int N2)
{
    int Ns[2] = {N1, N2};
    return make(NFFT_KIND_REALDFT, 2, Ns);
}
notorious_fft_aux *notorious_fft_mkaux_realdft_3d(int N1, int N2, int N3)
{
    int Ns[3] = {N1, N2, N3};
    return make(NFFT_KIND_REALDFT, 3, Ns);
}
notorious_fft_aux *notorious_fft_mkaux_realdft(int d, int *Ns) { return make(NFFT_KIND_REALDFT, d, Ns); }

notorious_fft_aux *notorious_fft_mkaux_t2t3_1d(int N) { return make(NFFT_KIND_T2T3, 1, &N); }
notorious_fft_aux *notorious_fft_mkaux_t2t3_2d(int N1, int N2)
{
    int Ns[2] = {N1, N2};
    return make(NFFT_KIND_T2T3, 2, Ns);
}
notorious_fft_aux *notorious_fft_mkaux_t2t3_3d(int N1, int N2, int N3)
{
    int Ns[3] = {N1, N2, N3};
    return make(NFFT_KIND_T2T3, 3, Ns);
}
notorious_fft_aux *notorious_fft_mkaux_t2t3(int d, int *Ns) { return make(NFFT_KIND_T2T3, d, Ns); }

notorious_fft_aux *notorious_fft_mkaux_t4_1d(int N) { return make(NFFT_KIND_T4, 1, &N); }
notorious_fft_aux *notorious_fft_mkaux_t4_2d(int N1, int N2)
{
    int Ns[2] = {N1, N2};
    return make(NFFT_KIND_T4, 2, Ns);
}
notorious_fft_aux *notorious_fft_mkaux_t4_3d(int N1, int N2, int N3)
{
    int Ns[3] = {N1, N2, N3};
    return make(NFFT_KIND_T4, 3, Ns);
}
notorious_fft_aux *notorious_fft_mkaux_t4(int d, int *Ns) { return make(NFFT_KIND_T4, d, Ns); }

void notorious_fft_free_aux(notorious_fft_aux *a)
{
    if (!a) return;
    Impl *im = (Impl *)a->impl;
    if (im) {
        if (im->st) (void)hipStreamSynchronize(im->st);
        nfft_plan_destroy(im->plan);
        if (im->din) (void)hipFree(im->din);
        if (im->dout) (void)hipFree(im->dout);
        if (im->hin) (void)hipHostFree(im->hin);
        if (im->hout) (void)hipHostFree(im->hout);
        if (im->st) (void)hipStreamDestroy(im->st);
        delete im;
    }
    delete a;
}

}  // extern "C"
