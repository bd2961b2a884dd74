/*
 * notorious_fft.h -- drop-in C API of the MI355X FFT engine.
 *
 * Declares the 27 functions of the reference library's C API with the same
 * names, argument meanings, FFTW-compatible interleaved layout and error
 * behaviour (constructors return NULL on unsupported sizes; transforms are
 * synchronous and return nothing).  Each declaration below cites the
 * reference prototype it replaces (src/07_legacy.h:29-57).  Declarations
 * only: the implementation is libnotorious_fft_hip_f64.so (double) or
 * libnotorious_fft_hip_f32.so (built with NOTORIOUS_FFT_SINGLE: float), as
 * the reference switches precision at compile time (src/01_core.h:36-40).
 *
 * Pointers may be host or device (HBM) memory; device pointers run in place
 * on the GPU, host pointers are copied through device staging buffers.
 * Input is never modified and x == y is allowed (as in the reference).
 * An aux must not be used by two threads at once (src/07_legacy.h:133).
 */
#ifndef NOTORIOUS_FFT_H
#define NOTORIOUS_FFT_H

#ifdef __cplusplus
extern "C" {
#endif

/* element type: src/01_core.h:36-40 */
#ifdef NOTORIOUS_FFT_SINGLE
typedef float notorious_fft_real;
#else
typedef double notorious_fft_real;
#endif

/* interleaved complex: src/07_legacy.h:19-23 */
#ifdef __cplusplus
typedef struct { notorious_fft_real re, im; } notorious_fft_cmpl;
#else
typedef notorious_fft_real notorious_fft_cmpl[2];
#endif

/* Plan handle.  The reference exposes its struct (src/01_core.h:96-105) but
 * no caller reads past the leading size; here only N is public. */
typedef struct notorious_fft_aux {
    int N;       /* transform size (product of dims for N-D) */
    void *impl;  /* private */
} notorious_fft_aux;

/* transforms: src/07_legacy.h:29-38 */
void notorious_fft_dft(notorious_fft_cmpl *x, notorious_fft_cmpl *y, const notorious_fft_aux *a);
void notorious_fft_invdft(notorious_fft_cmpl *x, notorious_fft_cmpl *y, const notorious_fft_aux *a);
void notorious_fft_realdft(notorious_fft_real *x, notorious_fft_cmpl *z, const notorious_fft_aux *a);
void notorious_fft_invrealdft(notorious_fft_cmpl *z, notorious_fft_real *y, const notorious_fft_aux *a);
void notorious_fft_dct2(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a);
void notorious_fft_dst2(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a);
void notorious_fft_dct3(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a);
void notorious_fft_dst3(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a);
void notorious_fft_dct4(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a);
void notorious_fft_dst4(notorious_fft_real *x, notorious_fft_real *y, const notorious_fft_aux *a);

/* plan construction: src/07_legacy.h:40-55 */
notorious_fft_aux *notorious_fft_mkaux_dft_1d(int N);
notorious_fft_aux *notorious_fft_mkaux_dft_2d(int N1, int N2);
notorious_fft_aux *notorious_fft_mkaux_dft_3d(int N1, int N2, int N3);
notorious_fft_aux *notorious_fft_mkaux_dft(int d, int *Ns);
notorious_fft_aux *notorious_fft_mkaux_realdft_1d(int N);
notorious_fft_aux *notorious_fft_mkaux_realdft_2d(int N1, int N2);
notorious_fft_aux *notorious_fft_mkaux_realdft_3d(int N1, int N2, int N3);
notorious_fft_aux *notorious_fft_mkaux_realdft(int d, int *Ns);
notorious_fft_aux *notorious_fft_mkaux_t2t3_1d(int N);
notorious_fft_aux *notorious_fft_mkaux_t2t3_2d(int N1, int N2);
notorious_fft_aux *notorious_fft_mkaux_t2t3_3d(int N1, int N2, int N3);
notorious_fft_aux *notorious_fft_mkaux_t2t3(int d, int *Ns);
notorious_fft_aux *notorious_fft_mkaux_t4_1d(int N);
notorious_fft_aux *notorious_fft_mkaux_t4_2d(int N1, int N2);
notorious_fft_aux *notorious_fft_mkaux_t4_3d(int N1, int N2, int N3);
notorious_fft_aux *notorious_fft_mkaux_t4(int d, int *Ns);

/* teardown: src/07_legacy.h:57; free_aux(NULL) is a no-op */
void notorious_fft_free_aux(notorious_fft_aux *a);

#ifdef __cplusplus
}
#endif

#endif /* NOTORIOUS_FFT_H */
