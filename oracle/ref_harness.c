/*
 * ref_harness.c -- ctypes harness around the reference library, compiled
 * from the reference's own header where it lies (see oracle/Makefile).
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement, to write the
 * golden vectors under tests/golden/, and as the CPU baseline of bench.py.
 * It calls only the reference's public C API (mkaux_* / transforms /
 * free_aux) and adds an OpenMP loop over a batch with one aux per thread,
 * because an aux holds scratch and is not thread-safe.
 */
#define NOTORIOUS_FFT_IMPLEMENTATION
#include "notorious_fft.h"
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef notorious_fft_real real_t;

/* kind: 0 dft, 1 realdft, 2 t2t3, 3 t4 (same numbering as the oracle) */
notorious_fft_aux *ref_mkaux(int kind, int d, const int *Ns)
{
    int tmp[16];
    if (d < 1 || d > 16) return NULL;
    for (int i = 0; i < d; i++) tmp[i] = Ns[i];
    if (d == 1) {
        switch (kind) {
        case 0: return notorious_fft_mkaux_dft_1d(tmp[0]);
        case 1: return notorious_fft_mkaux_realdft_1d(tmp[0]);
        case 2: return notorious_fft_mkaux_t2t3_1d(tmp[0]);
        case 3: return notorious_fft_mkaux_t4_1d(tmp[0]);
        }
        return NULL;
    }
    switch (kind) {
    case 0: return notorious_fft_mkaux_dft(d, tmp);
    case 2: return notorious_fft_mkaux_t2t3(d, tmp);
    case 3: return notorious_fft_mkaux_t4(d, tmp);
    }
    return NULL; /* multi-D realdft is not implemented by the reference */
}

void ref_free(notorious_fft_aux *a) { notorious_fft_free_aux(a); }

int ref_real_size(void) { return (int)sizeof(real_t); }

void ref_exec(notorious_fft_aux *a, int op, real_t *x, real_t *y)
{
    switch (op) {
    case 0: notorious_fft_dft((notorious_fft_cmpl *)x, (notorious_fft_cmpl *)y, a); break;
    case 1: notorious_fft_invdft((notorious_fft_cmpl *)x, (notorious_fft_cmpl *)y, a); break;
    case 2: notorious_fft_realdft(x, (notorious_fft_cmpl *)y, a); break;
    case 3: notorious_fft_invrealdft((notorious_fft_cmpl *)x, y, a); break;
    case 4: notorious_fft_dct2(x, y, a); break;
    case 5: notorious_fft_dst2(x, y, a); break;
    case 6: notorious_fft_dct3(x, y, a); break;
    case 7: notorious_fft_dst3(x, y, a); break;
    case 8: notorious_fft_dct4(x, y, a); break;
    case 9: notorious_fft_dst4(x, y, a); break;
    }
}

/* Batched run over contiguous transforms; in_n/out_n are reals per
 * transform. nthreads <= 0 uses the OpenMP default. When exec_s is not NULL
 * it receives the wall time of the transform loop alone (the per-thread aux
 * creation and release are outside it). Returns 0 on success. */
static double ref_now(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
#endif
}

int ref_exec_batch(int kind, int d, const int *Ns, int op, long long batch,
                   long long in_n, long long out_n, real_t *x, real_t *y, int nthreads, double *exec_s)
{
    int err = 0;
    double t0 = 0.0, t1 = 0.0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : err)
#endif
    {
        notorious_fft_aux *a = ref_mkaux(kind, d, Ns);
        if (!a) err = 1;
#ifdef _OPENMP
#pragma omp barrier
#pragma omp master
#endif
        t0 = ref_now();
#ifdef _OPENMP
#pragma omp barrier
#pragma omp for schedule(static)
#endif
        for (long long b = 0; b < batch; b++)
            if (a) ref_exec(a, op, x + b * in_n, y + b * out_n);
#ifdef _OPENMP
#pragma omp master
#endif
        t1 = ref_now();
        if (a) notorious_fft_free_aux(a);
    }
    (void)nthreads;
    if (exec_s) *exec_s = t1 - t0;
    return err;
}

/* A list of 1-D complex transforms of mixed sizes, packed back to back
 * (transform i of sizes[i] complex starts where transform i-1 ends): the
 * auxes are made first, then the transforms run OpenMP-parallel over the
 * list (dynamic schedule); exec_s receives the time of that loop alone.
 * Returns 0 on success. */
int ref_exec_list(int count, const int *sizes, int op, real_t *x, real_t *y, int nthreads, double *exec_s)
{
    notorious_fft_aux **aux = (notorious_fft_aux **)calloc((size_t)count, sizeof(*aux));
    long long *off = (long long *)malloc(((size_t)count + 1) * sizeof(long long));
    int err = 0;
    if (!aux || !off) {
        free(aux);
        free(off);
        return 1;
    }
    off[0] = 0;
    for (int i = 0; i < count; i++) off[i + 1] = off[i] + 2LL * sizes[i];
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic) reduction(| : err)
#endif
    for (int i = 0; i < count; i++) {
        aux[i] = notorious_fft_mkaux_dft_1d(sizes[i]);
        if (!aux[i]) err = 1;
    }
    double t0 = ref_now();
    if (!err) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic)
#endif
        for (int i = 0; i < count; i++) ref_exec(aux[i], op, x + off[i], y + off[i]);
    }
    double t1 = ref_now();
    for (int i = 0; i < count; i++) notorious_fft_free_aux(aux[i]);
    free(aux);
    free(off);
    (void)nthreads;
    if (exec_s) *exec_s = t1 - t0;
    return err;
}
