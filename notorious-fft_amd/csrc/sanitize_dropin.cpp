// sanitize_dropin.cpp -- the reference's 27-function API through one drop-in
// library (loaded with dlopen), for the host-side sanitizer driver
// (sanitize_args.cpp).  Compiled twice: as is (double) and with
// NOTORIOUS_FFT_SINGLE (float), so that every call goes through a pointer of
// the function's own type.  Checks: constructors reject invalid sizes (the
// reference's rules), free_aux(NULL) and transforms with a NULL aux are
// no-ops, a valid aux (where a device is visible) keeps int N first.
//
// Test infrastructure; not part of the library.
#include <dlfcn.h>
#include <stdio.h>
#include <string.h>

#include "../../include/notorious_fft.h"

#ifdef NOTORIOUS_FFT_SINGLE
#define DROPIN_FN dropin_paths_f32
#else
#define DROPIN_FN dropin_paths_f64
#endif

int DROPIN_FN(const char *path)
{
    int failures = 0;
#define EXPECT(cond)                                                           \
    do {                                                                       \
        if (!(cond)) {                                                         \
            failures++;                                                        \
            fprintf(stderr, "FAIL %s line %d: %s\n", path, __LINE__, #cond);  \
        }                                                                      \
    } while (0)
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "dlopen %s: %s\n", path, dlerror());
        return 1;
    }
#define SYM(name) ((decltype(&name))dlsym(h, #name))
    auto mk1 = {SYM(notorious_fft_mkaux_dft_1d), SYM(notorious_fft_mkaux_realdft_1d), SYM(notorious_fft_mkaux_t2t3_1d),
                SYM(notorious_fft_mkaux_t4_1d)};
    auto mk2 = {SYM(notorious_fft_mkaux_dft_2d), SYM(notorious_fft_mkaux_realdft_2d), SYM(notorious_fft_mkaux_t2t3_2d),
                SYM(notorious_fft_mkaux_t4_2d)};
    auto mk3 = {SYM(notorious_fft_mkaux_dft_3d), SYM(notorious_fft_mkaux_realdft_3d), SYM(notorious_fft_mkaux_t2t3_3d),
                SYM(notorious_fft_mkaux_t4_3d)};
    auto mkd = {SYM(notorious_fft_mkaux_dft), SYM(notorious_fft_mkaux_realdft), SYM(notorious_fft_mkaux_t2t3),
                SYM(notorious_fft_mkaux_t4)};
    int fam = 0;
    for (auto m1 : mk1) {
        EXPECT(m1 != nullptr);
        if (!m1) continue;
        EXPECT(m1(0) == nullptr);
        EXPECT(m1(-5) == nullptr);
        if (fam > 0) EXPECT(m1(12) == nullptr);  // powers of two only
        fam++;
    }
    for (auto m2 : mk2) {
        EXPECT(m2 != nullptr);
        if (!m2) continue;
        EXPECT(m2(0, 4) == nullptr);
        EXPECT(m2(4, -1) == nullptr);
    }
    for (auto m3 : mk3) {
        EXPECT(m3 != nullptr);
        if (m3) EXPECT(m3(4, 4, 0) == nullptr);
    }
    for (auto md : mkd) {
        EXPECT(md != nullptr);
        if (!md) continue;
        EXPECT(md(0, nullptr) == nullptr);
        int dims[2] = {4, -1};
        EXPECT(md(2, dims) == nullptr);
    }
    auto fr = SYM(notorious_fft_free_aux);
    EXPECT(fr != nullptr);
    if (fr) fr(nullptr);
    // transforms with a NULL aux are no-ops
    notorious_fft_real xr[64], yr[64];
    notorious_fft_cmpl xc[32], yc[32];
    memset(xc, 0, sizeof xc);
    for (int i = 0; i < 64; i++) xr[i] = (notorious_fft_real)i, yr[i] = (notorious_fft_real)-1;
    memset(yc, 0xff, sizeof yc);
    notorious_fft_cmpl yc0[32];
    memcpy(yc0, yc, sizeof yc);
    auto dft = SYM(notorious_fft_dft), inv = SYM(notorious_fft_invdft);
    auto rd = SYM(notorious_fft_realdft);
    auto ird = SYM(notorious_fft_invrealdft);
    if (dft) dft(xc, yc, nullptr);
    if (inv) inv(xc, yc, nullptr);
    if (rd) rd(xr, yc, nullptr);
    EXPECT(memcmp(yc, yc0, sizeof yc) == 0);
    if (ird) ird(xc, yr, nullptr);
    auto trig = {SYM(notorious_fft_dct2), SYM(notorious_fft_dst2), SYM(notorious_fft_dct3), SYM(notorious_fft_dst3),
                 SYM(notorious_fft_dct4), SYM(notorious_fft_dst4)};
    for (auto t : trig) {
        EXPECT(t != nullptr);
        if (t) t(xr, yr, nullptr);
    }
    for (int i = 0; i < 64; i++) EXPECT(yr[i] == (notorious_fft_real)-1);
    // a valid constructor: an aux where a device is visible, NULL where none is
    auto m1 = SYM(notorious_fft_mkaux_dft_1d);
    if (m1 && fr) {
        notorious_fft_aux *a = m1(16);
        if (a) {
            EXPECT(a->N == 16);  // int N first, as the reference's aux
            fr(a);
        }
    }
#undef SYM
#undef EXPECT
    dlclose(h);
    return failures;
}
