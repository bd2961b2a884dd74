/*
 * nfft_oracle.c -- CPU restatement of the reference FFT algorithms (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity reference: only tests/, the
 * smoke entry point and the CPU-baseline leg of bench.py may load it.  The
 * product (the HIP library) never links it and has no CPU fallback.
 *
 * It restates, in code of its own, the scalar double-precision path of the
 * reference header (no SIMD branches), keeping the arithmetic in the same
 * order so that, built with -O2 -ffp-contract=off and no vector ISA flags, it
 * agrees bit for bit with the reference built the same way.  Each function
 * cites the reference lines it follows.
 *
 * Pins (see tests/test_oracle.py): bitwise against the reference compiled
 * from its own header (oracle/_ref), the committed golden vectors under
 * tests/golden/, a long-double brute-force DFT, numpy.fft / scipy.fft
 * (FFTW REDFT/RODFT definitions) and closed forms (round trips, Parseval).
 *
 * Inputs/outputs are interleaved (re, im) doubles, the FFTW layout.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793238462643383279502884
#define ORC_2PI 6.283185307179586476925286766559005768
#define ORC_SQRT2 1.414213562373095048801688724209698079
#define ORC_INV_SQRT2 0.707106781186547524400844362104849039

/* Transform kinds (aux families) and operations. */
enum { ORC_KIND_DFT = 0, ORC_KIND_REALDFT = 1, ORC_KIND_T2T3 = 2, ORC_KIND_T4 = 3 };
enum {
    ORC_OP_DFT = 0, ORC_OP_INVDFT = 1, ORC_OP_REALDFT = 2, ORC_OP_INVREALDFT = 3,
    ORC_OP_DCT2 = 4, ORC_OP_DST2 = 5, ORC_OP_DCT3 = 6, ORC_OP_DST3 = 7,
    ORC_OP_DCT4 = 8, ORC_OP_DST4 = 9
};

/* ------------------------------------------------------------------------- */
/* 1D complex plan: power of two (split-radix tables) or Bluestein.          */
/* ------------------------------------------------------------------------- */

typedef struct cplan {
    int n;
    int pow2;
    /* power of two */
    int *bitrev;        /* n entries */
    double *tw_re, *tw_im; /* n/2 entries: e^{-2 pi i k/n} */
    double *sr_e;       /* split-radix levels n, n/2, ..., 16: quads (w1, w3) */
    double *sr_t;       /* 2n reals: copy of input */
    double *work;       /* 2n reals: split-radix temp / iterative work */
    /* Bluestein */
    int m;              /* pow2 >= 2n-1 */
    struct cplan *inner;
    double *chirp_re, *chirp_im;   /* e^{-i pi k^2/n}, k < n */
    double *hf_re, *hf_im;         /* FFT_m of the filter h */
    double *buf_re, *buf_im, *fb_re, *fb_im, *in_re, *in_im;
} cplan;

/* Bit reversal table: 02_math.h:134-151. */
static void orc_bitrev(int *rev, int n)
{
    int lg = 0;
    for (int t = n; t > 1; t >>= 1) lg++;
    for (int i = 0; i < n; i++) {
        int j = 0;
        for (int b = 0; b < lg; b++)
            if ((i >> b) & 1) j |= 1 << (lg - 1 - b);
        rev[i] = j;
    }
}

static void cplan_free(cplan *p)
{
    if (!p) return;
    free(p->bitrev); free(p->tw_re); free(p->tw_im); free(p->sr_e); free(p->sr_t);
    free(p->work);
    cplan_free(p->inner);
    free(p->chirp_re); free(p->chirp_im); free(p->hf_re); free(p->hf_im);
    free(p->buf_re); free(p->buf_im); free(p->fb_re); free(p->fb_im);
    free(p->in_re); free(p->in_im);
    free(p);
}

/* Power-of-two plan tables: 06_plan.h:127-226 (angle expressions kept:
 * tw angle = (-2pi * i) / n; split-radix angles (-2pi*k)/sz and
 * ((-2pi*3)*k)/sz). */
static cplan *cplan_pow2(int n)
{
    cplan *p = (cplan *)calloc(1, sizeof(cplan));
    if (!p) return NULL;
    p->n = n;
    p->pow2 = 1;
    p->bitrev = (int *)malloc(sizeof(int) * (size_t)n);
    p->tw_re = (double *)malloc(sizeof(double) * (size_t)(n / 2 + 1));
    p->tw_im = (double *)malloc(sizeof(double) * (size_t)(n / 2 + 1));
    p->work = (double *)malloc(sizeof(double) * (size_t)(2 * n));
    if (!p->bitrev || !p->tw_re || !p->tw_im || !p->work) { cplan_free(p); return NULL; }
    orc_bitrev(p->bitrev, n);
    for (int i = 0; i < n / 2; i++) {
        double ang = -ORC_2PI * (double)i / (double)n;
        p->tw_re[i] = cos(ang);
        p->tw_im[i] = sin(ang);
    }
    if (n > 8) {
        p->sr_e = (double *)malloc(sizeof(double) * (size_t)(2 * n));
        p->sr_t = (double *)malloc(sizeof(double) * (size_t)(2 * n));
        if (!p->sr_e || !p->sr_t) { cplan_free(p); return NULL; }
        double *ep = p->sr_e;
        for (int sz = n; sz >= 16; sz >>= 1) {
            for (int k = 0; k < sz / 4; k++) {
                *ep++ = cos(-ORC_2PI * (double)k / (double)sz);
                *ep++ = sin(-ORC_2PI * (double)k / (double)sz);
                *ep++ = cos(-ORC_2PI * 3.0 * (double)k / (double)sz);
                *ep++ = sin(-ORC_2PI * 3.0 * (double)k / (double)sz);
            }
        }
    }
    return p;
}

/* Split-radix 2/4 DIF on interleaved data, natural-order output at stride sy.
 * Forward: 05_algorithms.h:358-563; inverse (conj twiddles, u/v swapped):
 * 05_algorithms.h:566-717.  Terminal cases N = 1, 2, 4, 8 are the unrolled
 * ones of :366-427 / :574-636. */
static void orc_sr_dif(int N, double *x, double *t, double *y, int sy, const double *e, int inv)
{
    double *xr = x, *xi = x + 1, *tr = t, *ti = t + 1, *yr = y, *yi = y + 1;
    if (N == 1) { yr[0] = xr[0]; yi[0] = xi[0]; return; }
    if (N == 2) {
        double a_r = xr[0] + xr[2], a_i = xi[0] + xi[2];
        double b_r = xr[0] - xr[2], b_i = xi[0] - xi[2];
        yr[0] = a_r; yi[0] = a_i;
        yr[2 * sy] = b_r; yi[2 * sy] = b_i;
        return;
    }
    if (N == 4) {
        double s0r = xr[0] + xr[4], s0i = xi[0] + xi[4];
        double s1r = xr[2] + xr[6], s1i = xi[2] + xi[6];
        double d0r = xr[0] - xr[4], d0i = xi[0] - xi[4];
        double jr = -xi[2] + xi[6], ji = xr[2] - xr[6];  /* i*(x1 - x3) */
        yr[0] = s0r + s1r; yi[0] = s0i + s1i;
        yr[4 * sy] = s0r - s1r; yi[4 * sy] = s0i - s1i;
        if (!inv) {
            yr[2 * sy] = d0r - jr; yi[2 * sy] = d0i - ji;
            yr[6 * sy] = d0r + jr; yi[6 * sy] = d0i + ji;
        } else {
            yr[2 * sy] = d0r + jr; yi[2 * sy] = d0i + ji;
            yr[6 * sy] = d0r - jr; yi[6 * sy] = d0i - ji;
        }
        return;
    }
    if (N == 8) {
        const double c = ORC_INV_SQRT2;
        double s0r, s0i, s1r, s1i, d0r, d0i, jr, ji;
        double a0r, a0i, a1r, a1i, a2r, a2i, a3r, a3i;
        double b0r, b0i, b1r, b1i, b2r, b2i, b3r, b3i, qr, qi;
        /* even half: x0, x2, x4, x6 */
        s0r = xr[0] + xr[8]; s0i = xi[0] + xi[8];
        s1r = xr[4] + xr[12]; s1i = xi[4] + xi[12];
        d0r = xr[0] - xr[8]; d0i = xi[0] - xi[8];
        jr = -xi[4] + xi[12]; ji = xr[4] - xr[12];
        a0r = s0r + s1r; a0i = s0i + s1i;
        a2r = s0r - s1r; a2i = s0i - s1i;
        if (!inv) {
            a1r = d0r - jr; a1i = d0i - ji;
            a3r = d0r + jr; a3i = d0i + ji;
        } else {
            a1r = d0r + jr; a1i = d0i + ji;
            a3r = d0r - jr; a3i = d0i - ji;
        }
        /* odd half: x1, x3, x5, x7 */
        s0r = xr[2] + xr[10]; s0i = xi[2] + xi[10];
        s1r = xr[6] + xr[14]; s1i = xi[6] + xi[14];
        d0r = xr[2] - xr[10]; d0i = xi[2] - xi[10];
        jr = -xi[6] + xi[14]; ji = xr[6] - xr[14];
        b0r = s0r + s1r; b0i = s0i + s1i;
        if (!inv) {
            qr = d0r - jr; qi = d0i - ji;
            b1r = c * (qr + qi); b1i = c * (qi - qr);
            b2r = s0i - s1i; b2i = -s0r + s1r;
            qr = d0r + jr; qi = d0i + ji;
            b3r = c * (qi - qr); b3i = -c * (qi + qr);
        } else {
            qr = d0r + jr; qi = d0i + ji;
            b1r = c * (qr - qi); b1i = c * (qr + qi);
            b2r = -s0i + s1i; b2i = s0r - s1r;
            qr = d0r - jr; qi = d0i - ji;
            b3r = -c * (qr + qi); b3i = c * (qr - qi);
        }
        yr[0] = a0r + b0r; yi[0] = a0i + b0i;
        yr[2 * sy] = a1r + b1r; yi[2 * sy] = a1i + b1i;
        yr[4 * sy] = a2r + b2r; yi[4 * sy] = a2i + b2i;
        yr[6 * sy] = a3r + b3r; yi[6 * sy] = a3i + b3i;
        yr[8 * sy] = a0r - b0r; yi[8 * sy] = a0i - b0i;
        yr[10 * sy] = a1r - b1r; yi[10 * sy] = a1i - b1i;
        yr[12 * sy] = a2r - b2r; yi[12 * sy] = a2i - b2i;
        yr[14 * sy] = a3r - b3r; yi[14 * sy] = a3i - b3i;
        return;
    }
    int q = N / 4, h = N / 2;
    for (int k = 0; k < q; k++) {
        double p0r = xr[2 * k], p0i = xi[2 * k];
        double p1r = xr[2 * (k + h)], p1i = xi[2 * (k + h)];
        double p2r = xr[2 * (k + q)], p2i = xi[2 * (k + q)];
        double p3r = xr[2 * (k + 3 * q)], p3i = xi[2 * (k + 3 * q)];
        double sr = p0r + p1r, si = p0i + p1i;
        double s2r = p2r + p3r, s2i = p2i + p3i;
        double dr = p0r - p1r, di = p0i - p1i;
        double jr = p3i - p2i, ji = p2r - p3r;   /* i*(p2 - p3) */
        double ur, ui, vr, vi;
        if (!inv) { ur = dr - jr; ui = di - ji; vr = dr + jr; vi = di + ji; }
        else      { ur = dr + jr; ui = di + ji; vr = dr - jr; vi = di - ji; }
        const double *w = e + 4 * k;
        tr[2 * k] = sr; ti[2 * k] = si;
        tr[2 * (k + q)] = s2r; ti[2 * (k + q)] = s2i;
        if (!inv) {
            tr[2 * (k + h)] = ur * w[0] - ui * w[1];
            ti[2 * (k + h)] = ur * w[1] + ui * w[0];
            tr[2 * (k + 3 * q)] = vr * w[2] - vi * w[3];
            ti[2 * (k + 3 * q)] = vr * w[3] + vi * w[2];
        } else {
            tr[2 * (k + h)] = ur * w[0] + ui * w[1];
            ti[2 * (k + h)] = -ur * w[1] + ui * w[0];
            tr[2 * (k + 3 * q)] = vr * w[2] + vi * w[3];
            ti[2 * (k + 3 * q)] = -vr * w[3] + vi * w[2];
        }
    }
    const double *en = e + N;
    orc_sr_dif(N / 2, t, t, y, 2 * sy, en, inv);
    orc_sr_dif(N / 4, t + N, t + N, y + 2 * sy, 4 * sy, en + N / 2, inv);
    orc_sr_dif(N / 4, t + 3 * (N / 2), t + 3 * (N / 2), y + 6 * sy, 4 * sy, en + N / 2, inv);
}

/* Radix-2 DIT on split re/im arrays after a bit-reversal gather:
 * 05_algorithms.h:28-117 (scalar branch), used by Bluestein. */
static void orc_iterative_split(const cplan *p, const double *xr, const double *xi,
                                double *yr, double *yi, int inv)
{
    int n = p->n;
    double *wr = p->work, *wi = p->work + n;
    for (int i = 0; i < n; i++) { wr[i] = xr[p->bitrev[i]]; wi[i] = xi[p->bitrev[i]]; }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; j++) {
                double cr = p->tw_re[j * step];
                double ci = inv ? -p->tw_im[j * step] : p->tw_im[j * step];
                int a = i + j, b = i + j + half;
                double ur = wr[a], ui = wi[a];
                double vr = wr[b] * cr - wi[b] * ci;
                double vi = wr[b] * ci + wi[b] * cr;
                wr[a] = ur + vr; wi[a] = ui + vi;
                wr[b] = ur - vr; wi[b] = ui - vi;
            }
        }
    }
    memcpy(yr, wr, sizeof(double) * (size_t)n);
    memcpy(yi, wi, sizeof(double) * (size_t)n);
}

/* Bluestein plan: 06_plan.h:28-121.  Filter h[k] = e^{+i pi (k-n+1)^2 / n}
 * for k <= 2n-2 (angle evaluated as ((pi*idx)*idx)/n in double, no modular
 * reduction), FFT_m(h) by the radix-2 DIT, chirp a[k] = e^{-i pi k^2/n}. */
static cplan *cplan_bluestein(int n)
{
    cplan *p = (cplan *)calloc(1, sizeof(cplan));
    if (!p) return NULL;
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    p->n = n;
    p->m = m;
    p->inner = cplan_pow2(m);
    size_t sz = sizeof(double) * (size_t)m;
    p->chirp_re = (double *)malloc(sz); p->chirp_im = (double *)malloc(sz);
    p->hf_re = (double *)malloc(sz); p->hf_im = (double *)malloc(sz);
    p->buf_re = (double *)malloc(sz); p->buf_im = (double *)malloc(sz);
    p->fb_re = (double *)malloc(sz); p->fb_im = (double *)malloc(sz);
    p->in_re = (double *)malloc(sz); p->in_im = (double *)malloc(sz);
    if (!p->inner || !p->chirp_re || !p->chirp_im || !p->hf_re || !p->hf_im || !p->buf_re ||
        !p->buf_im || !p->fb_re || !p->fb_im || !p->in_re || !p->in_im) {
        cplan_free(p);
        return NULL;
    }
    for (int k = 0; k <= 2 * (n - 1); k++) {
        int idx = k - (n - 1);
        double ang = ORC_PI * (double)idx * (double)idx / (double)n;
        p->buf_re[k] = cos(ang);
        p->buf_im[k] = sin(ang);
    }
    for (int k = 2 * n - 1; k < m; k++) { p->buf_re[k] = 0; p->buf_im[k] = 0; }
    orc_iterative_split(p->inner, p->buf_re, p->buf_im, p->hf_re, p->hf_im, 0);
    for (int i = 0; i < n; i++) {
        double ang = -ORC_PI * (double)i * (double)i / (double)n;
        p->chirp_re[i] = cos(ang);
        p->chirp_im[i] = sin(ang);
    }
    return p;
}

static cplan *cplan_new(int n)
{
    if (n <= 0) return NULL;
    if ((n & (n - 1)) == 0) return cplan_pow2(n);
    return cplan_bluestein(n);
}

/* Forward Bluestein on split arrays: 05_algorithms.h:797-896. */
static void orc_bluestein_fwd(const cplan *p, const double *xr, const double *xi,
                              double *yr, double *yi)
{
    int n = p->n, m = p->m;
    for (int i = 0; i < n; i++) {
        p->buf_re[i] = xr[i] * p->chirp_re[i] - xi[i] * p->chirp_im[i];
        p->buf_im[i] = xr[i] * p->chirp_im[i] + xi[i] * p->chirp_re[i];
    }
    for (int i = n; i < m; i++) { p->buf_re[i] = 0; p->buf_im[i] = 0; }
    orc_iterative_split(p->inner, p->buf_re, p->buf_im, p->fb_re, p->fb_im, 0);
    for (int i = 0; i < m; i++) {
        double a = p->fb_re[i] * p->hf_re[i] - p->fb_im[i] * p->hf_im[i];
        double b = p->fb_re[i] * p->hf_im[i] + p->fb_im[i] * p->hf_re[i];
        p->fb_re[i] = a; p->fb_im[i] = b;
    }
    orc_iterative_split(p->inner, p->fb_re, p->fb_im, p->buf_re, p->buf_im, 1);
    double scale = 1.0 / (double)m;
    for (int i = 0; i < n; i++) {
        int j = n - 1 + i;
        yr[i] = (p->buf_re[j] * p->chirp_re[i] - p->buf_im[j] * p->chirp_im[i]) * scale;
        yi[i] = (p->buf_re[j] * p->chirp_im[i] + p->buf_im[j] * p->chirp_re[i]) * scale;
    }
}

/* Dispatch on one contiguous interleaved transform: 05_algorithms.h:292-344.
 * Bluestein: inverse via conj trick and a 1/n scale (the reference's
 * normalised-Bluestein-inverse behaviour); power of two >= 16: split-radix
 * with a copy of the input; n <= 8: split-radix terminals in place. */
static void orc_exec_cx(const cplan *p, const double *x, double *y, int inv)
{
    int n = p->n;
    if (!p->pow2) {
        double *ir = p->in_re, *ii = p->in_im;
        double *orr = p->fb_re, *oi = p->fb_im;
        for (int i = 0; i < n; i++) { ir[i] = x[2 * i]; ii[i] = x[2 * i + 1]; }
        if (inv) for (int i = 0; i < n; i++) ii[i] = -ii[i];
        orc_bluestein_fwd(p, ir, ii, orr, oi);
        if (inv) {
            double s = 1.0 / (double)n;
            for (int i = 0; i < n; i++) { y[2 * i] = orr[i] * s; y[2 * i + 1] = -oi[i] * s; }
        } else {
            for (int i = 0; i < n; i++) { y[2 * i] = orr[i]; y[2 * i + 1] = oi[i]; }
        }
        return;
    }
    if (n >= 16) {
        memcpy(p->sr_t, x, sizeof(double) * (size_t)(2 * n));
        orc_sr_dif(n, p->sr_t, p->work, y, 1, p->sr_e, inv);
        return;
    }
    if (x != y) memcpy(y, x, sizeof(double) * (size_t)(2 * n));
    if (n == 1) return;
    double tmp[16];
    orc_sr_dif(n, y, tmp, y, 1, NULL, inv);
}

/* ------------------------------------------------------------------------- */
/* Aux tree: 07_legacy.h:1389-1700                                           */
/* ------------------------------------------------------------------------- */

typedef struct orc_aux {
    int N;            /* 1D size, or product of dims for a tree node */
    int kind;
    cplan *plan;      /* 1D plan (size N, or N/2 for real/trig kinds) */
    double *e;        /* realdft unpack twiddles (N/4 pairs) or T4 pre/post */
    double *t;        /* scratch */
    double *dtw;      /* DCT-2/3 twiddles e^{-i pi k/(2N)}, N/2 pairs */
    double *dscr;     /* DCT-2/3 scratch: reordered (N) + z (N+2) */
    struct orc_aux *sub1, *sub2;
} orc_aux;

void orc_free_aux(orc_aux *a);

static orc_aux *aux_alloc(void) { return (orc_aux *)calloc(1, sizeof(orc_aux)); }

static orc_aux *mk_dft_1d(int N)
{
    if (N <= 0) return NULL;
    orc_aux *a = aux_alloc();
    if (!a) return NULL;
    a->N = N;
    a->kind = ORC_KIND_DFT;
    a->plan = cplan_new(N);
    if (!a->plan) { free(a); return NULL; }
    return a;
}

/* realdft aux: 07_legacy.h:1441-1486 (pow2 only, plan of N/2 for N >= 4,
 * unpack twiddles angle (-2pi*k)/N for k < N/4). */
static orc_aux *mk_realdft_1d(int N)
{
    if (N <= 0 || (N & (N - 1))) return NULL;
    orc_aux *a = aux_alloc();
    if (!a) return NULL;
    a->N = N;
    a->kind = ORC_KIND_REALDFT;
    if (N >= 4) {
        a->plan = cplan_new(N / 2);
        a->e = (double *)malloc(sizeof(double) * (size_t)(N / 2));
        a->t = (double *)malloc(sizeof(double) * (size_t)N);
        if (!a->plan || !a->e || !a->t) { orc_free_aux(a); return NULL; }
        for (int k = 0; k < N / 4; k++) {
            double ang = -ORC_2PI * (double)k / (double)N;
            a->e[2 * k] = cos(ang);
            a->e[2 * k + 1] = sin(ang);
        }
    }
    return a;
}

/* t2t3 aux: 07_legacy.h:1527-1558 (DCT twiddles angle (-pi*k)/(2N)). */
static orc_aux *mk_t2t3_1d(int N)
{
    orc_aux *a = mk_realdft_1d(N);
    if (!a) return NULL;
    a->kind = ORC_KIND_T2T3;
    if (N >= 4) {
        a->dtw = (double *)malloc(sizeof(double) * (size_t)N);
        a->dscr = (double *)malloc(sizeof(double) * (size_t)(2 * N + 2));
        if (!a->dtw || !a->dscr) { orc_free_aux(a); return NULL; }
        for (int k = 0; k < N / 2; k++) {
            double ang = -ORC_PI * (double)k / (double)(2 * N);
            a->dtw[2 * k] = cos(ang);
            a->dtw[2 * k + 1] = sin(ang);
        }
    }
    return a;
}

/* t4 aux: 07_legacy.h:1578-1634: premul e^{-i pi n/N} (n < N/2), postmul
 * e^{-i pi (2n+1)/(4N)} (n < N), plan of N/2 when N/2 >= 2. */
static orc_aux *mk_t4_1d(int N)
{
    if (N <= 0 || (N & (N - 1))) return NULL;
    orc_aux *a = aux_alloc();
    if (!a) return NULL;
    a->N = N;
    a->kind = ORC_KIND_T4;
    if (N >= 2) {
        int M = N / 2;
        if (M >= 2) {
            a->plan = cplan_new(M);
            if (!a->plan) { free(a); return NULL; }
        }
        a->e = (double *)malloc(sizeof(double) * (size_t)(2 * M + 2 * N));
        a->t = (double *)malloc(sizeof(double) * (size_t)N);
        if (!a->e || !a->t) { orc_free_aux(a); return NULL; }
        double *ep = a->e;
        for (int n = 0; n < M; n++) {
            double ang = -ORC_PI * (double)n / (double)N;
            ep[2 * n] = cos(ang);
            ep[2 * n + 1] = sin(ang);
        }
        ep += 2 * M;
        for (int n = 0; n < N; n++) {
            double ang = -ORC_PI * (double)(2 * n + 1) / (double)(4 * N);
            ep[2 * n] = cos(ang);
            ep[2 * n + 1] = sin(ang);
        }
    }
    return a;
}

static orc_aux *mk_1d(int kind, int N)
{
    switch (kind) {
    case ORC_KIND_DFT: return mk_dft_1d(N);
    case ORC_KIND_REALDFT: return mk_realdft_1d(N);
    case ORC_KIND_T2T3: return mk_t2t3_1d(N);
    case ORC_KIND_T4: return mk_t4_1d(N);
    }
    return NULL;
}

/* Generic N-D tree: 07_legacy.h:1654-1688.  sub2 = 1D aux of the slowest
 * axis Ns[0], sub1 = tree of Ns[1..]; t = scratch of prod(Ns) elements. */
static orc_aux *mk_tree(int kind, int d, const int *Ns, size_t elsz)
{
    long long p = 1;
    for (int i = 0; i < d; i++) p *= Ns[i];
    orc_aux *a = aux_alloc();
    if (!a) return NULL;
    a->N = (int)p;
    a->kind = kind;
    if (d == 1) {
        a->sub2 = mk_1d(kind, Ns[0]);
    } else {
        a->t = (double *)malloc(elsz * (size_t)p);
        a->sub1 = mk_tree(kind, d - 1, Ns + 1, elsz);
        a->sub2 = mk_1d(kind, Ns[0]);
    }
    if ((d > 1 && (!a->sub1 || !a->t)) || !a->sub2) { orc_free_aux(a); return NULL; }
    return a;
}

void orc_free_aux(orc_aux *a)
{
    if (!a) return;
    cplan_free(a->plan);
    free(a->e); free(a->t); free(a->dtw); free(a->dscr);
    orc_free_aux(a->sub1);
    orc_free_aux(a->sub2);
    free(a);
}

/* ------------------------------------------------------------------------- */
/* Complex DFT, 1D and N-D: 07_legacy.h:73-152                               */
/* ------------------------------------------------------------------------- */

static void orc_s_dft(const double *x, double *y, int sy, const orc_aux *a, int inv)
{
    int N = a->N;
    if (!a->plan) return;
    if (sy == 1) { orc_exec_cx(a->plan, x, y, inv); return; }
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    if (!buf) return;
    orc_exec_cx(a->plan, x, buf, inv);
    for (int i = 0; i < N; i++) {
        y[2 * (size_t)i * sy] = buf[2 * i];
        y[2 * (size_t)i * sy + 1] = buf[2 * i + 1];
    }
    free(buf);
}

static void orc_mkcx(const double *x, double *y, int sy, const orc_aux *a, int inv)
{
    if (a->plan) { orc_s_dft(x, y, sy, a, inv); return; }
    if (!a->sub1) { if (a->sub2) orc_s_dft(x, y, sy, a->sub2, inv); return; }
    int N1 = a->sub1->N, N2 = a->sub2->N;
    double *t = a->t;
    for (int n = 0; n < N2; n++)
        orc_mkcx(x + 2 * (size_t)n * N1, t + 2 * (size_t)n, N2, a->sub1, inv);
    for (int n = 0; n < N1; n++)
        orc_s_dft(t + 2 * (size_t)n * N2, y + 2 * (size_t)sy * n, sy * N1, a->sub2, inv);
}

/* ------------------------------------------------------------------------- */
/* Real DFT: 07_legacy.h:162-293                                             */
/* ------------------------------------------------------------------------- */

static void orc_realdft_1d(const double *x, double *z, const orc_aux *a)
{
    int N = a->N;
    if (N == 1) { z[0] = x[0]; z[1] = 0; return; }
    if (N == 2) {
        double s = x[0] + x[1], d = x[0] - x[1];
        z[0] = s; z[1] = 0; z[2] = d; z[3] = 0;
        return;
    }
    int M = N / 2;
    memcpy(a->t, x, sizeof(double) * (size_t)N);
    orc_exec_cx(a->plan, a->t, z, 0);
    const double *tw = a->e;
    double z0r = z[0], z0i = z[1];
    z[0] = z0r + z0i; z[1] = 0;
    z[2 * M] = z0r - z0i; z[2 * M + 1] = 0;
    for (int k = 1; k < N / 4; k++) {
        double ar = z[2 * k], ai = z[2 * k + 1];
        double br = z[2 * (M - k)], bi = z[2 * (M - k) + 1];
        double ur = (ar + br) * 0.5, ui = (ai - bi) * 0.5;
        double dr = ar - br, di = ai + bi;
        double pr = dr * tw[2 * k] - di * tw[2 * k + 1];
        double pi_ = dr * tw[2 * k + 1] + di * tw[2 * k];
        double vr = pi_ * 0.5, vi = -pr * 0.5;
        z[2 * (M - k)] = ur - vr;
        z[2 * (M - k) + 1] = -(ui - vi);
        z[2 * k] = ur + vr;
        z[2 * k + 1] = ui + vi;
    }
    if (N >= 4) z[2 * (N / 4) + 1] = -z[2 * (N / 4) + 1];
}

static void orc_invrealdft_1d(const double *z, double *y, const orc_aux *a)
{
    int N = a->N;
    if (N == 1) { y[0] = z[0]; return; }
    if (N == 2) { double p = z[0], q = z[2]; y[0] = p + q; y[1] = p - q; return; }
    int M = N / 2;
    double *b = a->t;
    const double *tw = a->e;
    b[0] = z[0] + z[2 * M];
    b[1] = z[0] - z[2 * M];
    for (int k = 1; k < N / 4; k++) {
        double ar = z[2 * k], ai = z[2 * k + 1];
        double cr = z[2 * (M - k)], ci = -z[2 * (M - k) + 1];
        double ur = ar + cr, ui = ai + ci, dr = ar - cr, di = ai - ci;
        double wr = tw[2 * k], wi = -tw[2 * k + 1];
        double vr = -di * wr - dr * wi;
        double vi = -di * wi + dr * wr;
        b[2 * k] = ur + vr; b[2 * k + 1] = ui + vi;
        b[2 * (M - k)] = ur - vr; b[2 * (M - k) + 1] = -(ui - vi);
    }
    if (N >= 4) { b[2 * (N / 4)] = 2 * z[2 * (N / 4)]; b[2 * (N / 4) + 1] = -2 * z[2 * (N / 4) + 1]; }
    orc_exec_cx(a->plan, b, y, 1);
}

/* ------------------------------------------------------------------------- */
/* DCT/DST-2/3 (scalar): 07_legacy.h:757-1066                                */
/* ------------------------------------------------------------------------- */

static void orc_dct2_1d(const double *x, double *y, const orc_aux *a)
{
    int N = a->N;
    if (N == 1) { y[0] = 2 * x[0]; return; }
    if (N == 2) { y[0] = 2 * (x[0] + x[1]); y[1] = ORC_SQRT2 * (x[0] - x[1]); return; }
    double *v = a->dscr, *z = a->dscr + N;
    for (int i = 0; i < N / 2; i++) { v[i] = x[2 * i]; v[N / 2 + i] = x[N - 1 - 2 * i]; }
    orc_realdft_1d(v, z, a);
    const double *tw = a->dtw;
    y[0] = 2 * z[0];
    for (int k = 1; k < N / 2; k++) {
        double wr = tw[2 * k], wi = tw[2 * k + 1], zr = z[2 * k], zi = z[2 * k + 1];
        y[k] = 2 * (zr * wr - zi * wi);
        y[N - k] = -2 * (zr * wi + zi * wr);
    }
    y[N / 2] = ORC_SQRT2 * z[N];
}

static void orc_dst2_1d(const double *x, double *y, const orc_aux *a)
{
    int N = a->N;
    if (N == 1) { y[0] = 2 * x[0]; return; }
    if (N == 2) { y[0] = ORC_SQRT2 * (x[0] + x[1]); y[1] = 2 * (x[0] - x[1]); return; }
    double *v = a->dscr, *z = a->dscr + N;
    for (int i = 0; i < N / 2; i++) { v[i] = x[2 * i]; v[N / 2 + i] = -x[N - 1 - 2 * i]; }
    orc_realdft_1d(v, z, a);
    const double *tw = a->dtw;
    y[N - 1] = 2 * z[0];
    for (int k = 1; k < N / 2; k++) {
        double wr = tw[2 * k], wi = tw[2 * k + 1], zr = z[2 * k], zi = z[2 * k + 1];
        y[k - 1] = -2 * (zr * wi + zi * wr);
        y[N - k - 1] = 2 * (zr * wr - zi * wi);
    }
    y[N / 2 - 1] = ORC_SQRT2 * z[N];
}

static void orc_dct3_1d(const double *x, double *y, const orc_aux *a)
{
    int N = a->N;
    if (N == 1) { y[0] = x[0]; return; }
    if (N == 2) { double s = ORC_SQRT2 * x[1]; y[0] = x[0] + s; y[1] = x[0] - s; return; }
    double *tmp = a->dscr, *z = a->dscr + N;
    const double *tw = a->dtw;
    z[0] = x[0]; z[1] = 0;
    for (int k = 1; k < N / 2; k++) {
        double wr = tw[2 * k], wi = tw[2 * k + 1], xr = x[k], xi = x[N - k];
        z[2 * k] = xr * wr - xi * wi;
        z[2 * k + 1] = -xr * wi - xi * wr;
    }
    z[N] = ORC_SQRT2 * x[N / 2]; z[N + 1] = 0;
    orc_invrealdft_1d(z, tmp, a);
    for (int i = 0; i < N / 2; i++) { y[2 * i] = tmp[i]; y[N - 1 - 2 * i] = tmp[N / 2 + i]; }
}

static void orc_dst3_1d(const double *x, double *y, const orc_aux *a)
{
    int N = a->N;
    if (N == 1) { y[0] = x[0]; return; }
    if (N == 2) { double s = ORC_SQRT2 * x[0]; y[0] = s + x[1]; y[1] = s - x[1]; return; }
    double *tmp = a->dscr, *z = a->dscr + N;
    const double *tw = a->dtw;
    z[0] = x[N - 1]; z[1] = 0;
    for (int k = 1; k < N / 2; k++) {
        double wr = tw[2 * k], wi = tw[2 * k + 1], xr = x[N - k - 1], xi = x[k - 1];
        z[2 * k] = xr * wr - xi * wi;
        z[2 * k + 1] = -xr * wi - xi * wr;
    }
    z[N] = ORC_SQRT2 * x[N / 2 - 1]; z[N + 1] = 0;
    orc_invrealdft_1d(z, tmp, a);
    for (int i = 0; i < N / 2; i++) { y[2 * i] = tmp[i]; y[N - 1 - 2 * i] = -tmp[N / 2 + i]; }
}

/* ------------------------------------------------------------------------- */
/* DCT/DST-4: 07_legacy.h:531-549 (pre/post products), :1091-1137 (body),    */
/* :1140-1183 (N = 8 unrolled, radix-2 DIT 4-point DFT), :1298-1334.         */
/* ------------------------------------------------------------------------- */

static void t4_premul(double *tr, double *ti, double er, double ei, double xe, double xo, int dst)
{
    if (!dst) { *tr = er * xe - ei * xo; *ti = er * xo + ei * xe; }
    else      { *tr = -er * xe - ei * xo; *ti = er * xo - ei * xe; }
}

static double t4_post(double fr, double fi, double tr, double ti, int dst)
{
    return dst ? 2 * (fr * ti + fi * tr) : 2 * (fr * tr - fi * ti);
}

static void orc_t4_1d(const double *x, double *y, const orc_aux *a, int dst)
{
    int N = a->N;
    if (N == 1) { y[0] = ORC_SQRT2 * x[0]; return; }
    int M = N / 2;
    const double *pre = a->e, *post = a->e + 2 * M;
    double *t = a->t;
    if (N == 8) {
        double r[4], i_[4];
        int ev[4] = {0, 2, 4, 6}, od[4] = {7, 5, 3, 1};
        for (int n = 0; n < 4; n++) t4_premul(&r[n], &i_[n], pre[2 * n], pre[2 * n + 1], x[ev[n]], x[od[n]], dst);
        double a0r = r[0] + r[2], a0i = i_[0] + i_[2];
        double a1r = r[1] + r[3], a1i = i_[1] + i_[3];
        double a2r = r[0] - r[2], a2i = i_[0] - i_[2];
        double a3r = r[1] - r[3], a3i = i_[1] - i_[3];
        double f[8];
        f[0] = a0r + a1r; f[1] = a0i + a1i;
        f[4] = a0r - a1r; f[5] = a0i - a1i;
        f[2] = a2r + a3i; f[3] = a2i - a3r;
        f[6] = a2r - a3i; f[7] = a2i + a3r;
        for (int k = 0; k < 4; k++) {
            y[2 * k] = t4_post(post[4 * k], post[4 * k + 1], f[2 * k], f[2 * k + 1], dst);
            y[2 * k + 1] = t4_post(post[4 * k + 2], post[4 * k + 3], f[6 - 2 * k], -f[7 - 2 * k], dst);
        }
        return;
    }
    for (int n = 0; n < M; n++)
        t4_premul(&t[2 * n], &t[2 * n + 1], pre[2 * n], pre[2 * n + 1], x[2 * n], x[N - 1 - 2 * n], dst);
    if (a->plan) orc_exec_cx(a->plan, t, t, 0);
    for (int k = 0; k < M; k++) {
        y[2 * k] = t4_post(post[4 * k], post[4 * k + 1], t[2 * k], t[2 * k + 1], dst);
        y[2 * k + 1] = t4_post(post[4 * k + 2], post[4 * k + 3], t[N - 2 - 2 * k], -t[N - 1 - 2 * k], dst);
    }
}

/* ------------------------------------------------------------------------- */
/* Real-valued transforms on trees: 07_legacy.h:822-861 and the analogous   */
/* loops for dst2/dct3/dst3/dct4/dst4 (hyperplanes via sub1, then columns).  */
/* ------------------------------------------------------------------------- */

typedef void (*orc_r1d)(const double *, double *, const orc_aux *);

static void dct4_1d(const double *x, double *y, const orc_aux *a) { orc_t4_1d(x, y, a, 0); }
static void dst4_1d(const double *x, double *y, const orc_aux *a) { orc_t4_1d(x, y, a, 1); }

static void orc_rnd(const double *x, double *y, const orc_aux *a, orc_r1d f)
{
    /* A 1D aux (made by mk_*_1d) is recognised by having no sub-trees. */
    if (!a->sub1 && !a->sub2) { f(x, y, a); return; }
    int N1 = a->sub1 ? a->sub1->N : 1;
    int N2 = a->sub2->N;
    if (N1 == 1) { orc_rnd(x, y, a->sub2, f); return; }
    size_t total = (size_t)N1 * (size_t)N2;
    double *tmp = (double *)malloc(sizeof(double) * total);
    double *ci = (double *)malloc(sizeof(double) * (size_t)N2);
    double *co = (double *)malloc(sizeof(double) * (size_t)N2);
    if (tmp && ci && co) {
        for (int n = 0; n < N2; n++) orc_rnd(x + (size_t)n * N1, tmp + (size_t)n * N1, a->sub1, f);
        for (int n = 0; n < N1; n++) {
            for (int k = 0; k < N2; k++) ci[k] = tmp[n + (size_t)k * N1];
            orc_rnd(ci, co, a->sub2, f);
            for (int k = 0; k < N2; k++) y[n + (size_t)k * N1] = co[k];
        }
    }
    free(tmp); free(ci); free(co);
}

/* ------------------------------------------------------------------------- */
/* Multi-dimensional real DFT.  The reference declares it but leaves it     */
/* unimplemented (07_legacy.h:1492-1514 builds sub1 = 1D realdft of the     */
/* last axis and sub2 = complex DFT tree of the leading axes, and realdft    */
/* ignores both).  This restates the FFTW/minfft semantics for that tree:    */
/* output N1 x ... x N_{d-1} x (N_d/2+1) complex, unnormalised inverse.      */
/* ------------------------------------------------------------------------- */

static void orc_realdft_nd(const double *x, double *z, const orc_aux *a)
{
    if (!a->sub1) { orc_realdft_1d(x, z, a); return; }
    int Nl = a->sub1->N, H = Nl / 2 + 1;
    int P = a->N / Nl;              /* product of leading dims */
    for (int p = 0; p < P; p++) orc_realdft_1d(x + (size_t)p * Nl, z + 2 * (size_t)p * H, a->sub1);
    /* complex DFT over the leading axes for every one of the H columns */
    double *ci = (double *)malloc(sizeof(double) * 2 * (size_t)P);
    double *co = (double *)malloc(sizeof(double) * 2 * (size_t)P);
    if (ci && co) {
        for (int h = 0; h < H; h++) {
            for (int p = 0; p < P; p++) { ci[2 * p] = z[2 * ((size_t)p * H + h)]; ci[2 * p + 1] = z[2 * ((size_t)p * H + h) + 1]; }
            orc_mkcx(ci, co, 1, a->sub2, 0);
            for (int p = 0; p < P; p++) { z[2 * ((size_t)p * H + h)] = co[2 * p]; z[2 * ((size_t)p * H + h) + 1] = co[2 * p + 1]; }
        }
    }
    free(ci); free(co);
}

static void orc_invrealdft_nd(const double *z, double *y, const orc_aux *a)
{
    if (!a->sub1) { orc_invrealdft_1d(z, y, a); return; }
    int Nl = a->sub1->N, H = Nl / 2 + 1;
    int P = a->N / Nl;
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)P * H);
    double *ci = (double *)malloc(sizeof(double) * 2 * (size_t)P);
    double *co = (double *)malloc(sizeof(double) * 2 * (size_t)P);
    if (w && ci && co) {
        for (int h = 0; h < H; h++) {
            for (int p = 0; p < P; p++) { ci[2 * p] = z[2 * ((size_t)p * H + h)]; ci[2 * p + 1] = z[2 * ((size_t)p * H + h) + 1]; }
            orc_mkcx(ci, co, 1, a->sub2, 1);
            for (int p = 0; p < P; p++) { w[2 * ((size_t)p * H + h)] = co[2 * p]; w[2 * ((size_t)p * H + h) + 1] = co[2 * p + 1]; }
        }
        for (int p = 0; p < P; p++) orc_invrealdft_1d(w + 2 * (size_t)p * H, y + (size_t)p * Nl, a->sub1);
    }
    free(w); free(ci); free(co);
}

static orc_aux *mk_realdft_tree(int d, const int *Ns)
{
    if (d == 1) return mk_realdft_1d(Ns[0]);
    orc_aux *a = aux_alloc();
    if (!a) return NULL;
    long long p = 1;
    for (int i = 0; i < d; i++) p *= Ns[i];
    a->N = (int)p;
    a->kind = ORC_KIND_REALDFT;
    a->sub1 = mk_realdft_1d(Ns[d - 1]);
    a->sub2 = mk_tree(ORC_KIND_DFT, d - 1, Ns, 2 * sizeof(double));
    if (!a->sub1 || !a->sub2) { orc_free_aux(a); return NULL; }
    return a;
}

/* ------------------------------------------------------------------------- */
/* Exported C ABI (ctypes)                                                   */
/* ------------------------------------------------------------------------- */

/* Build an aux of the given kind over d dims (d == 1 gives the direct 1D
 * aux, as the reference's mkaux_*_1d does; d > 1 gives the tree). */
orc_aux *orc_mkaux(int kind, int d, const int *Ns)
{
    if (d < 1 || !Ns) return NULL;
    for (int i = 0; i < d; i++) if (Ns[i] <= 0) return NULL;
    if (d == 1) return mk_1d(kind, Ns[0]);
    if (kind == ORC_KIND_REALDFT) return mk_realdft_tree(d, Ns);
    size_t elsz = kind == ORC_KIND_DFT ? 2 * sizeof(double) : sizeof(double);
    return mk_tree(kind, d, Ns, elsz);
}

int orc_aux_n(const orc_aux *a) { return a ? a->N : 0; }

/* Run one transform.  Complex data is interleaved doubles. */
void orc_exec(const orc_aux *a, int op, const double *x, double *y)
{
    if (!a) return;
    switch (op) {
    case ORC_OP_DFT: orc_mkcx(x, y, 1, a, 0); break;
    case ORC_OP_INVDFT: orc_mkcx(x, y, 1, a, 1); break;
    case ORC_OP_REALDFT: orc_realdft_nd(x, y, a); break;
    case ORC_OP_INVREALDFT: orc_invrealdft_nd(x, y, a); break;
    case ORC_OP_DCT2: orc_rnd(x, y, a, orc_dct2_1d); break;
    case ORC_OP_DST2: orc_rnd(x, y, a, orc_dst2_1d); break;
    case ORC_OP_DCT3: orc_rnd(x, y, a, orc_dct3_1d); break;
    case ORC_OP_DST3: orc_rnd(x, y, a, orc_dst3_1d); break;
    case ORC_OP_DCT4: orc_rnd(x, y, a, dct4_1d); break;
    case ORC_OP_DST4: orc_rnd(x, y, a, dst4_1d); break;
    }
}

/* Element counts (in doubles) of one transform's input and output. */
static void op_sizes(int kind, int d, const int *Ns, int op, size_t *in_d, size_t *out_d)
{
    size_t p = 1;
    for (int i = 0; i < d; i++) p *= (size_t)Ns[i];
    size_t half = p / (size_t)Ns[d - 1] * (size_t)(Ns[d - 1] / 2 + 1);
    (void)kind;
    switch (op) {
    case ORC_OP_DFT: case ORC_OP_INVDFT: *in_d = *out_d = 2 * p; break;
    case ORC_OP_REALDFT: *in_d = p; *out_d = 2 * half; break;
    case ORC_OP_INVREALDFT: *in_d = 2 * half; *out_d = p; break;
    default: *in_d = *out_d = p; break;
    }
}

/* Batched execution: OpenMP over the batch with one aux per thread (an aux
 * holds scratch and is not shared between threads).  nthreads <= 0 uses the
 * OpenMP default. Returns 0 on success. */
int orc_exec_batch(int kind, int d, const int *Ns, int op, long long batch,
                   const double *x, double *y, int nthreads)
{
    size_t in_d, out_d;
    op_sizes(kind, d, Ns, op, &in_d, &out_d);
    int err = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : err)
#endif
    {
        orc_aux *a = orc_mkaux(kind, d, Ns);
        if (!a) err = 1;
        else {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
            for (long long b = 0; b < batch; b++)
                orc_exec(a, op, x + (size_t)b * in_d, y + (size_t)b * out_d);
            orc_free_aux(a);
        }
    }
    (void)nthreads;
    return err;
}

/* Long-double O(N^2) DFT with exact angle reduction (k*n mod N), the
 * brute-force check used for small sizes.  inv selects e^{+2 pi i kn/N}
 * (unnormalised). */
void orc_bruteforce_dft(int n, const double *x, double *y, int inv)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n; k++) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; j++) {
            long long m = ((long long)k * j) % n;
            long double ang = (inv ? 2 : -2) * pi * (long double)m / (long double)n;
            long double c = cosl(ang), s = sinl(ang);
            sr += (long double)x[2 * j] * c - (long double)x[2 * j + 1] * s;
            si += (long double)x[2 * j] * s + (long double)x[2 * j + 1] * c;
        }
        y[2 * k] = (double)sr;
        y[2 * k + 1] = (double)si;
    }
}

/* Counter-based SplitMix64 inputs: element idx of stream seed maps to
 * mix(((seed << 40) | idx) + golden), uniform in [-1, 1) from the top 53
 * bits.  The same generator runs on the GPU and in Python. */
static uint64_t orc_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_fill_uniform(uint64_t seed, uint64_t start, long long count, double *out)
{
    for (long long i = 0; i < count; i++) {
        uint64_t key = ((seed << 40) | (start + (uint64_t)i)) + 0x9E3779B97F4A7C15ULL;
        uint64_t z = orc_mix(key);
        out[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
}
