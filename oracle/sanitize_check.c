/*
 * sanitize_check.c -- drives the oracle (nfft_oracle.c) under
 * AddressSanitizer / UndefinedBehaviorSanitizer: every transform kind and
 * operation on 1-D lengths (powers of two, Bluestein lengths, the N <= 8
 * terminals) and N-D shapes (with unit axes), in and out of place, the
 * constructors' rejected sizes, and a numerical sanity check of each
 * transform against the brute-force DFT or its round trip.
 *
 * TEST INFRASTRUCTURE ONLY (built by `make -C oracle sanitize`, run by
 * tests/test_sanitize.py).  Exit status 0 when every check passes; the
 * sanitizers abort on the first memory or UB error.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_aux orc_aux;
orc_aux *orc_mkaux(int kind, int d, const int *Ns);
void orc_free_aux(orc_aux *a);
void orc_exec(const orc_aux *a, int op, const double *x, double *y);
void orc_bruteforce_dft(int n, const double *x, double *y, int inv);
void orc_fill_uniform(uint64_t seed, uint64_t start, long long count, double *out);

static int failures = 0;

static double rel_l2(const double *a, const double *b, long long n)
{
    double num = 0, den = 0;
    for (long long i = 0; i < n; i++) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return den > 0 ? sqrt(num / den) : sqrt(num);
}

static void check(int ok, const char *what, int d, const int *Ns)
{
    if (ok) return;
    failures++;
    fprintf(stderr, "FAIL %s dims", what);
    for (int i = 0; i < d; i++) fprintf(stderr, " %d", Ns[i]);
    fprintf(stderr, "\n");
}

static long long prodn(int d, const int *Ns)
{
    long long p = 1;
    for (int i = 0; i < d; i++) p *= Ns[i];
    return p;
}

/* complex DFT: forward against brute force (1-D), round trip (any d) */
static void run_dft(int d, const int *Ns)
{
    orc_aux *a = orc_mkaux(0, d, Ns);
    check(a != NULL, "mkaux dft", d, Ns);
    if (!a) return;
    const long long p = prodn(d, Ns);
    double *x = malloc(sizeof(double) * 2 * p), *y = malloc(sizeof(double) * 2 * p);
    double *z = malloc(sizeof(double) * 2 * p), *w = malloc(sizeof(double) * 2 * p);
    orc_fill_uniform(7, 0, 2 * p, x);
    memcpy(w, x, sizeof(double) * 2 * p);
    orc_exec(a, 0, x, y);
    check(memcmp(w, x, sizeof(double) * 2 * p) == 0, "input untouched", d, Ns);
    if (d == 1) {
        orc_bruteforce_dft(Ns[0], x, z, 0);
        check(rel_l2(y, z, 2 * p) < 1e-9, "dft vs brute force", d, Ns);
    }
    orc_exec(a, 1, y, z);
    /* invdft is unnormalised along power-of-two axes and scaled by 1/N
       along the others, so the round trip gives x times the power-of-two part */
    double s = 1;
    for (int i = 0; i < d; i++)
        if ((Ns[i] & (Ns[i] - 1)) == 0) s *= Ns[i];
    for (long long i = 0; i < 2 * p; i++) w[i] = s * x[i];
    check(rel_l2(z, w, 2 * p) < 1e-9, "dft round trip", d, Ns);
    memcpy(w, x, sizeof(double) * 2 * p); /* in place */
    orc_exec(a, 0, w, w);
    check(rel_l2(w, y, 2 * p) < 1e-13, "dft in place", d, Ns);
    free(x), free(y), free(z), free(w);
    orc_free_aux(a);
}

/* realdft / invrealdft: round trip invrealdft(realdft(x)) = x times the
   product of the power-of-two lengths */
static void run_real(int d, const int *Ns)
{
    orc_aux *a = orc_mkaux(1, d, Ns);
    check(a != NULL, "mkaux realdft", d, Ns);
    if (!a) return;
    const long long p = prodn(d, Ns);
    const long long h = p / Ns[d - 1] * (Ns[d - 1] / 2 + 1);
    double *x = malloc(sizeof(double) * p), *z = malloc(sizeof(double) * 2 * h), *y = malloc(sizeof(double) * p);
    orc_fill_uniform(8, 0, p, x);
    orc_exec(a, 2, x, z);
    orc_exec(a, 3, z, y);
    /* unnormalised along power-of-two axes, 1/N along leading non-power-of-two
       axes (the reference's invdft rule): x times the power-of-two part */
    double s = 1;
    for (int i = 0; i < d; i++)
        if ((Ns[i] & (Ns[i] - 1)) == 0) s *= Ns[i];
    for (long long i = 0; i < p; i++) y[i] /= s;
    check(rel_l2(y, x, p) < 1e-12, "realdft round trip", d, Ns);
    free(x), free(z), free(y);
    orc_free_aux(a);
}

/* trig: dct3(dct2) = dst3(dst2) = 2N x per axis (products), dct4 / dst4 involutions up to 2N */
static void run_trig(int kind, int d, const int *Ns)
{
    orc_aux *a = orc_mkaux(kind, d, Ns);
    check(a != NULL, "mkaux trig", d, Ns);
    if (!a) return;
    const long long p = prodn(d, Ns);
    double *x = malloc(sizeof(double) * p), *y = malloc(sizeof(double) * p), *z = malloc(sizeof(double) * p);
    orc_fill_uniform(9, 0, p, x);
    /* the reference skips a trailing run of unit axes */
    double s = 1;
    int act = d;
    while (act > 1 && Ns[act - 1] == 1) act--;
    for (int i = 0; i < act; i++) s *= 2.0 * Ns[i];
    const int ops[2][2] = {{4, 6}, {5, 7}};
    for (int k = 0; k < (kind == 2 ? 2 : 1); k++) {
        const int f = kind == 2 ? ops[k][0] : 8 + k, b = kind == 2 ? ops[k][1] : 8 + k;
        for (int t = 0; t < (kind == 3 ? 2 : 1); t++) {
            const int ff = kind == 3 ? 8 + t : f, bb = kind == 3 ? 8 + t : b;
            orc_exec(a, ff, x, y);
            orc_exec(a, bb, y, z);
            for (long long i = 0; i < p; i++) z[i] /= s;
            check(rel_l2(z, x, p) < 1e-10, "trig round trip", d, Ns);
        }
    }
    free(x), free(y), free(z);
    orc_free_aux(a);
}

int main(void)
{
    static const int n1d[] = {1, 2, 3, 4, 5, 7, 8, 12, 16, 17, 31, 64, 100, 127, 256, 1009, 1024, 4096};
    for (size_t i = 0; i < sizeof n1d / sizeof n1d[0]; i++) run_dft(1, &n1d[i]);
    static const int nd[][3] = {{3, 5, 0}, {6, 10, 4}, {8, 8, 8}, {1, 7, 0}, {5, 1, 0}, {16, 12, 0}, {2, 1, 3}};
    static const int ndd[] = {2, 3, 3, 2, 2, 2, 3};
    for (size_t i = 0; i < sizeof ndd / sizeof ndd[0]; i++) run_dft(ndd[i], nd[i]);
    static const int rp[] = {1, 2, 4, 8, 16, 64, 1024, 4096};
    for (size_t i = 0; i < sizeof rp / sizeof rp[0]; i++) run_real(1, &rp[i]);
    static const int rnd[][3] = {{4, 8, 0}, {6, 16, 0}, {3, 5, 8}};
    static const int rndd[] = {2, 2, 3};
    for (size_t i = 0; i < sizeof rndd / sizeof rndd[0]; i++) run_real(rndd[i], rnd[i]);
    static const int tp[] = {1, 2, 4, 8, 16, 256, 1024};
    for (size_t i = 0; i < sizeof tp / sizeof tp[0]; i++) {
        run_trig(2, 1, &tp[i]);
        run_trig(3, 1, &tp[i]);
    }
    static const int tnd[][3] = {{8, 16, 0}, {4, 1, 0}, {2, 4, 8}};
    static const int tndd[] = {2, 2, 3};
    for (size_t i = 0; i < sizeof tndd / sizeof tndd[0]; i++) {
        run_trig(2, tndd[i], tnd[i]);
        run_trig(3, tndd[i], tnd[i]);
    }
    /* constructors reject what the reference rejects */
    static const int bad_real[] = {12}, bad_zero[] = {0}, bad_neg[] = {-4}, bad_trig[] = {12};
    check(orc_mkaux(1, 1, bad_real) == NULL, "realdft of 12 rejected", 1, bad_real);
    check(orc_mkaux(0, 1, bad_zero) == NULL, "dft of 0 rejected", 1, bad_zero);
    check(orc_mkaux(0, 1, bad_neg) == NULL, "dft of -4 rejected", 1, bad_neg);
    check(orc_mkaux(2, 1, bad_trig) == NULL, "t2t3 of 12 rejected", 1, bad_trig);
    orc_free_aux(NULL);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("oracle sanitizer checks passed\n");
    return 0;
}
