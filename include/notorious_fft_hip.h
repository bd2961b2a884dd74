/*
 * notorious_fft_hip.h -- extended C ABI of the MI355X (gfx950) FFT engine.
 *
 * Everything here is new relative to the reference library, whose API is
 * one synchronous transform per call on host memory (src/07_legacy.h:29-57).
 * This header adds, with plain pointers and sizes only:
 *   - device-resident plans ("device-aux"): tables and scratch in HBM;
 *   - uniform batches of contiguous transforms on a caller stream;
 *   - an explicit normalisation flag;
 *   - ragged batches of mixed sizes (one persistent launch, graph-captured);
 *   - vector-Jacobian products (adjoints) of every transform;
 *   - status codes with a thread-local message.
 * Library: libnotorious_fft_hip.so (both precisions; the precision is a plan
 * argument).  Streams are hipStream_t passed as void*; NULL is the default
 * stream.  Calls enqueue work and return; synchronise the stream before
 * reading results.  A plan holds scratch: two executions of one plan must not
 * overlap in time (the same rule as the reference's aux, src/07_legacy.h:133).
 */
#ifndef NOTORIOUS_FFT_HIP_H
#define NOTORIOUS_FFT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nfft_plan nfft_plan;
typedef struct nfft_ragged nfft_ragged;

/* status codes */
enum {
    NFFT_OK = 0,
    NFFT_E_INVALID = 1,     /* bad argument (NULL, wrong op for the plan kind) */
    NFFT_E_SIZE = 2,        /* size not supported (the reference's NULL-aux cases) */
    NFFT_E_ALLOC = 3,       /* device or host allocation failed */
    NFFT_E_HIP = 4,         /* a HIP runtime call or kernel launch failed */
    NFFT_E_UNSUPPORTED = 5  /* valid request outside the library's limits */
};

/* precision */
enum { NFFT_F32 = 0, NFFT_F64 = 1 };

/* plan kinds, as the reference's aux families (mkaux_dft / realdft / t2t3 / t4) */
enum { NFFT_KIND_DFT = 0, NFFT_KIND_REALDFT = 1, NFFT_KIND_T2T3 = 2, NFFT_KIND_T4 = 3 };

/* operations, as the reference's transform entry points */
enum {
    NFFT_OP_DFT = 0, NFFT_OP_INVDFT = 1, NFFT_OP_REALDFT = 2, NFFT_OP_INVREALDFT = 3,
    NFFT_OP_DCT2 = 4, NFFT_OP_DST2 = 5, NFFT_OP_DCT3 = 6, NFFT_OP_DST3 = 7,
    NFFT_OP_DCT4 = 8, NFFT_OP_DST4 = 9
};

/* normalisation of inverse complex transforms:
 *   REFERENCE  the reference's behaviour: invdft unnormalised for power-of-two
 *              lengths, scaled by 1/N along every non-power-of-two axis
 *              (src/05_algorithms.h:315-333);
 *   NONE       never scaled (the adjoint of dft);
 *   INVERSE    invdft scaled by 1/prod(N). */
enum { NFFT_NORM_REFERENCE = 0, NFFT_NORM_NONE = 1, NFFT_NORM_INVERSE = 2 };

/* Create a plan for `batch` contiguous transforms of dims[0..d-1] (row-major,
 * last index fastest, FFTW layout).  Element counts per transform:
 *   DFT      in/out prod(dims) complex (interleaved re,im)
 *   REALDFT  real prod(dims) <-> complex prod(dims[0..d-2]) * (dims[d-1]/2+1)
 *   T2T3/T4  prod(dims) reals
 * Size rules follow the reference (src/07_legacy.h:1405-1634): DFT accepts
 * any N >= 1 per axis; REALDFT, T2T3 and T4 need powers of two.
 * Autotuning: records in the file named by NFFT_TUNE_CACHE choose, per
 * precision and length, mixed radix or Bluestein for 7-smooth lengths and
 * the row length and row pass of power-of-two four-steps; with NFFT_TUNE=1 a
 * length without a record is measured at plan creation and its record
 * appended. */
int nfft_plan_create(nfft_plan **plan, int kind, int prec, int d, const int *dims,
                     long long batch, int norm);
int nfft_plan_destroy(nfft_plan *plan);

/* Run `op` on device buffers.  in == out is allowed except for REALDFT /
 * INVREALDFT with batch > 1.  The input is never modified. */
int nfft_execute(nfft_plan *plan, int op, const void *in, void *out, void *stream);

/* nfft_execute for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): dft / invdft of a DFT plan with
 * every kernel's scratch (four-step intermediates, Bluestein passes, the
 * full-array copy of a strided axis) taken from one static device array of
 * 64 MiB per GPU, held from the library's first use on that GPU until the
 * process ends.  No counterpart in the reference.  The call allocates
 * nothing; results are those of nfft_execute (the batch is worked through in
 * chunks that fit).  Plans needing more than 64 MiB for one chunk (e.g. a
 * strided axis whose full array exceeds it) are rejected with NFFT_E_SIZE.
 * All calls on one GPU share the array: calls that overlap in time on one
 * GPU (other threads, other streams without ordering) race. */
int nfft_execute_noalloc(nfft_plan *plan, int op, const void *in, void *out, void *stream);

/* Vector-Jacobian product of `op`: grad_in = J^T grad_out, where complex
 * gradients use the convention g = dL/dRe + i dL/dIm.  Layouts are those of
 * the op's output (grad_out) and input (grad_in).
 *   dft:        DFT^H g  (unnormalised inverse DFT)
 *   invdft:     DFT g, scaled as the plan's invdft is
 *   realdft:    invrealdft(g * [1, 1/2, ..., 1/2, 1])
 *   invrealdft: realdft(g) with bins 0 < k < N/2 doubled
 *   dct2/dst2:  dct3/dst3 with the DC input doubled
 *   dct3/dst3:  dct2/dst2 with the DC output halved
 *   dct4/dst4:  self-adjoint */
int nfft_vjp(nfft_plan *plan, int op, const void *grad_out, void *grad_in, void *stream);

/* Weighted spectral power of each transform of a DFT plan's batch and its
 * gradient (the adjoint pipeline of loss = sum_k |W_k DFT(x)_k|^2):
 *   loss[t] = sum_k |W_k Y_k|^2 with Y = DFT(x_t) (unnormalised), in FP64
 *             (device array of `batch` doubles; NULL: not written);
 *   grad    = 2 DFT^H (W^2 . Y), the convention of nfft_vjp
 *             (g = dL/dRe x + i dL/dIm x), layout of x (grad == x allowed).
 * w: prod(dims) real weights (the plan's precision), shared by the batch.
 * For d >= 2 with a power-of-two axis 0 of at most 4096 points, the forward
 * transform of axis 0, the weighting and the adjoint of axis 0 run as one
 * fused pass; otherwise forward, elementwise weighting and adjoint run as
 * separate passes.  No counterpart in the reference (whose gradient would be
 * its invdft of the weighted spectrum, src/07_legacy.h:146-152). */
int nfft_dft_weighted_power(nfft_plan *plan, const void *x, const void *w, void *grad, double *loss,
                            void *stream);

/* scratch bytes held by the plan in HBM */
long long nfft_plan_workspace_bytes(const nfft_plan *plan);

/* A short description of the execution strategy ("line", "four-step
 * 1024x1024", "bluestein M=2048", ...) for diagnostics. */
const char *nfft_plan_describe(const nfft_plan *plan);

/* Synchronises `stream` (a hipStream_t, NULL = default) and reports whether
 * an earlier execute of `plan` hit a device-side fault: the persistent
 * four-step kernel bounds each of its inter-workgroup dependency waits by a
 * wall-clock timeout and raises a flag instead of spinning forever.  Returns
 * NFFT_OK, or NFFT_E_HIP with the message set (the outputs of the affected
 * executes are then undefined).  The flag is sticky for the plan's lifetime. */
int nfft_plan_check(nfft_plan *plan, void *stream);

/* Ragged batch: `count` 1D complex transforms of sizes[i] (1 <= N <= 65536,
 * any N), packed back to back (transform i starts at element
 * sizes[0] + ... + sizes[i-1]).  Execution is one persistent kernel launch
 * (after a one-node counter reset), recorded into a HIP graph on first use
 * for each (in, out, direction) and replayed afterwards.  Inverse transforms
 * follow the reference's normalisation (1/N for non-power-of-two N only).
 * flags bit 0: autotune on a scratch batch -- first the tile size (column /
 * row groups per tile) and the pipeline lag of the long transforms over five
 * candidates, then the persistent grid (workgroups per CU); the choices are
 * appended to the file named by NFFT_TUNE_CACHE, if set, and read back (not
 * measured again) by later creates with the same size list. */
int nfft_ragged_create(nfft_ragged **r, int prec, int count, const int *sizes, int flags);
int nfft_ragged_execute(nfft_ragged *r, int inverse, const void *in, void *out, void *stream);
int nfft_ragged_destroy(nfft_ragged *r);
long long nfft_ragged_total(const nfft_ragged *r);

/* Synchronises `stream` and reports whether an earlier execute of `r` hit a
 * device-side fault: the tiles of the long transforms wait for each other
 * with bounded waits that raise a flag instead of spinning forever. Returns
 * NFFT_OK, or NFFT_E_HIP with the message set (outputs then undefined). */
int nfft_ragged_check(nfft_ragged *r, void *stream);

/* Kernel timer.  nfft_prof_enable(1, capacity) clears the records and
 * brackets every subsequent kernel launch of the library with HIP events on
 * the launch stream (up to `capacity` launches); (0, 0) stops recording.
 * nfft_prof_count() gives the number of kernel families seen ("name<prec,L>");
 * nfft_prof_get() returns one family's launch count, summed event time in
 * milliseconds and summed algorithmic bytes (compulsory input + output
 * traffic of each launch); it waits for the recorded events. */
int nfft_prof_enable(int on, int capacity);
int nfft_prof_count(void);
int nfft_prof_get(int idx, char *name, int name_len, long long *launches, double *total_ms,
                  double *bytes);

/* Message for the last error on this thread ("" if none). */
const char *nfft_last_error(void);

/* Device helpers used by the tests and the benchmark: fill `count` values
 * (reals) with the counter-based SplitMix64 uniform [-1, 1) stream, value i
 * = U(seed, start + i); prec selects float or double output. */
int nfft_fill_uniform(void *out, long long count, unsigned long long seed,
                      unsigned long long start, int prec, void *stream);

/* Same stream, Gaussian N(0, 1) by Box-Muller: value i (j = start + i) is
 * sqrt(-2 ln u1) cos(2 pi u2) for even j, sqrt(-2 ln u1) sin(2 pi u2) for odd
 * j, with u1 = V(2 (j/2)), u2 = V(2 (j/2) + 1) and V(k) = ((z_k >> 11) + 1) 2^-53
 * in (0, 1] the SplitMix64 draw k of `seed`, computed in FP64. */
int nfft_fill_gaussian(void *out, long long count, unsigned long long seed,
                       unsigned long long start, int prec, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NOTORIOUS_FFT_HIP_H */
