// cpp_wrapper_check.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C++17 caller of the reference's wrapper (include/notorious_fft.hpp of the
// reference, used unchanged) compiled against this repository's
// include/notorious_fft.h and linked with libnotorious_fft_hip_f64.so: the
// drop-in claim for the wrapper's callers.  oracle/Makefile builds it into
// oracle/_ref/ (the wrapper is read where it lies under the reference tree);
// tests/test_gpu_cpp_wrapper.py runs it on the GPU.
//
// Each check compares a wrapper call with a long-double direct sum or with a
// round trip of known scale (unnormalised transforms: invdft(dft(x)) = N x
// for power-of-two N, x for other N; invrealdft(realdft(x)) = N x;
// dct3(dct2(x)) = 2N x; dct4(dct4(x)) = 2N x).  Exit status 0 iff all pass.
#include <cmath>
#include <complex>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "notorious_fft.hpp"

namespace nf = notorious_fft;
using cd = std::complex<double>;

static int g_fail = 0, g_pass = 0;

static void report(const char *what, int n, double err, double tol)
{
    const bool ok = err <= tol;
    std::printf("%-28s N=%-6d rel err %.3e (tol %.1e) %s\n", what, n, err, tol, ok ? "ok" : "FAIL");
    (ok ? g_pass : g_fail)++;
}

static double rel(const std::vector<cd> &a, const std::vector<cd> &b)
{
    long double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return (double)std::sqrt(num / (den > 0 ? den : 1));
}

static double rel(const std::vector<double> &a, const std::vector<double> &b, double scale = 1.0)
{
    long double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        const long double d = (long double)a[i] - scale * (long double)b[i];
        num += d * d;
        den += (long double)scale * b[i] * scale * b[i];
    }
    return (double)std::sqrt(num / (den > 0 ? den : 1));
}

// deterministic inputs in [-1, 1)
static double u01(unsigned long long &s)
{
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

static std::vector<cd> direct_dft(const std::vector<cd> &x, int sign)
{
    const size_t n = x.size();
    std::vector<cd> y(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (size_t k = 0; k < n; k++) {
        long double re = 0, im = 0;
        for (size_t j = 0; j < n; j++) {
            const long double ang = sign * 2 * pi * (long double)((k * j) % n) / n;
            re += x[j].real() * cosl(ang) - x[j].imag() * sinl(ang);
            im += x[j].real() * sinl(ang) + x[j].imag() * cosl(ang);
        }
        y[k] = cd((double)re, (double)im);
    }
    return y;
}

int main()
{
    unsigned long long seed = 7;
    for (int n : {1, 2, 16, 100, 1024, 1009, 4096}) {
        std::vector<cd> x(n);
        for (auto &v : x) v = cd(u01(seed), u01(seed));
        const auto y = nf::dft(x);
        report("dft(vector)", n, rel(y, direct_dft(x, -1)), 1e-12);
        const auto z = nf::invdft(y);
        const double s = (n & (n - 1)) == 0 ? (double)n : 1.0;
        std::vector<cd> sx(x);
        for (auto &v : sx) v *= s;
        report("invdft(dft) round trip", n, rel(z, sx), 1e-12);
        nf::dft_plan p(n);
        report("dft_plan::execute", n, rel(p.execute(x), y), 0.0);
        auto w = x;
        p.execute_in_place(w);
        report("dft_plan::execute_in_place", n, rel(w, y), 0.0);
    }
    for (int n : {2, 16, 1024, 8192}) {
        std::vector<double> x(n);
        for (auto &v : x) v = u01(seed);
        const auto z = nf::realdft(x);
        std::vector<cd> xc(x.begin(), x.end());
        auto full = direct_dft(xc, -1);
        full.resize(n / 2 + 1);
        report("realdft(vector)", n, rel(z, full), 1e-12);
        report("invrealdft(realdft)", n, rel(nf::invrealdft(z), x, (double)n), 1e-12);
        nf::realdft_plan rp(n);
        report("realdft_plan::execute", n, rel(rp.execute(x), z), 0.0);
        report("dct3(dct2)", n, rel(nf::dct3(nf::dct2(x)), x, 2.0 * n), 1e-12);
        report("dst3(dst2)", n, rel(nf::dst3(nf::dst2(x)), x, 2.0 * n), 1e-12);
        report("dct4(dct4)", n, rel(nf::dct4(nf::dct4(x)), x, 2.0 * n), 1e-12);
        report("dst4(dst4)", n, rel(nf::dst4(nf::dst4(x)), x, 2.0 * n), 1e-12);
        nf::dct2_plan dp(n);
        report("dct2_plan::execute", n, rel(dp.execute(x), nf::dct2(x)), 0.0);
    }
    {
        const int n1 = 16, n2 = 32;
        std::vector<cd> x(n1 * n2);
        for (auto &v : x) v = cd(u01(seed), u01(seed));
        // row-major 2-D DFT as two passes of the direct sum
        std::vector<cd> t(x.size()), ref(x.size());
        for (int r = 0; r < n1; r++) {
            std::vector<cd> row(x.begin() + r * n2, x.begin() + (r + 1) * n2);
            const auto f = direct_dft(row, -1);
            for (int c = 0; c < n2; c++) t[r * n2 + c] = f[c];
        }
        for (int c = 0; c < n2; c++) {
            std::vector<cd> col(n1);
            for (int r = 0; r < n1; r++) col[r] = t[r * n2 + c];
            const auto f = direct_dft(col, -1);
            for (int r = 0; r < n1; r++) ref[r * n2 + c] = f[r];
        }
        report("dft_2d", n1 * n2, rel(nf::dft_2d(x, n1, n2), ref), 1e-12);
    }
    bool threw = false;
    try {
        nf::dft_plan bad(0);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    report("mkaux(0) throws", 0, threw ? 0.0 : 1.0, 0.0);
    std::printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
}
