// sanitize_args.cpp -- drives the host-side argument and error paths of the
// C ABIs (extended API of libnotorious_fft_hip.so, and the reference's
// 27-function API of both drop-in libraries) in a build whose host code is
// instrumented with AddressSanitizer / UndefinedBehaviorSanitizer
// (`make -C notorious-fft_amd/csrc sanitize`).  It needs no GPU: every call
// either rejects its arguments before any device work or fails cleanly where
// no device is visible.  Exit status 0 when every expectation holds; the
// sanitizers abort on the first memory or UB error.
//
// Test infrastructure (run by tests/test_sanitize.py); not part of the
// library.
#include <stdio.h>
#include <string.h>

#include "../../include/notorious_fft_hip.h"

static int failures = 0;
#define EXPECT(cond)                                                                                  \
    do {                                                                                              \
        if (!(cond)) {                                                                                \
            failures++;                                                                               \
            fprintf(stderr, "FAIL line %d: %s (last error: %s)\n", __LINE__, #cond, nfft_last_error()); \
        }                                                                                             \
    } while (0)

static void core_paths()
{
    nfft_plan *p = nullptr;
    const int d1[1] = {1024}, d0[1] = {0}, dneg[1] = {-3}, d12[1] = {12};
    const int dbig[2] = {1 << 16, 1 << 16};
    EXPECT(nfft_plan_create(nullptr, NFFT_KIND_DFT, NFFT_F32, 1, d1, 1, 0) == NFFT_E_INVALID);
    EXPECT(nfft_plan_create(&p, 7, NFFT_F32, 1, d1, 1, 0) == NFFT_E_INVALID && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, 5, 1, d1, 1, 0) == NFFT_E_INVALID && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F32, 0, d1, 1, 0) == NFFT_E_INVALID && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F32, 17, d1, 1, 0) == NFFT_E_INVALID && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F32, 1, nullptr, 1, 0) == NFFT_E_INVALID && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F32, 1, d1, 0, 0) == NFFT_E_INVALID && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F32, 1, d1, 1, 3) == NFFT_E_INVALID && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F64, 1, d0, 1, 0) == NFFT_E_SIZE && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F64, 1, dneg, 1, 0) == NFFT_E_SIZE && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F64, 2, dbig, 1, 0) == NFFT_E_UNSUPPORTED && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_REALDFT, NFFT_F32, 1, d12, 1, 0) == NFFT_E_SIZE && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_T2T3, NFFT_F32, 1, d12, 1, 0) == NFFT_E_SIZE && !p);
    EXPECT(nfft_plan_create(&p, NFFT_KIND_T4, NFFT_F64, 1, d12, 1, 0) == NFFT_E_SIZE && !p);
    EXPECT(strlen(nfft_last_error()) > 0);
    // a valid request: a plan where a device is visible, a clean error where none is
    const int st = nfft_plan_create(&p, NFFT_KIND_DFT, NFFT_F32, 1, d1, 4, 0);
    EXPECT((st == NFFT_OK) == (p != nullptr));
    if (p) EXPECT(nfft_plan_destroy(p) == NFFT_OK);
    char buf[16];
    EXPECT(nfft_execute(nullptr, NFFT_OP_DFT, buf, buf, nullptr) == NFFT_E_INVALID);
    EXPECT(nfft_execute_noalloc(nullptr, NFFT_OP_DFT, buf, buf, nullptr) == NFFT_E_INVALID);
    EXPECT(nfft_vjp(nullptr, NFFT_OP_DFT, buf, buf, nullptr) == NFFT_E_INVALID);
    double loss[1];
    EXPECT(nfft_dft_weighted_power(nullptr, buf, buf, buf, loss, nullptr) == NFFT_E_INVALID);
    EXPECT(nfft_plan_destroy(nullptr) == NFFT_OK);
    EXPECT(nfft_plan_check(nullptr, nullptr) == NFFT_E_INVALID);
    EXPECT(nfft_plan_workspace_bytes(nullptr) == 0);
    EXPECT(strcmp(nfft_plan_describe(nullptr), "") == 0);
    nfft_ragged *r = nullptr;
    const int sz[3] = {16, 0, 100}, szbig[2] = {16, 65537}, szok[2] = {16, 100};
    EXPECT(nfft_ragged_create(nullptr, NFFT_F32, 2, szok, 0) == NFFT_E_INVALID);
    EXPECT(nfft_ragged_create(&r, NFFT_F32, 0, szok, 0) == NFFT_E_INVALID && !r);
    EXPECT(nfft_ragged_create(&r, 9, 2, szok, 0) == NFFT_E_INVALID && !r);
    EXPECT(nfft_ragged_create(&r, NFFT_F32, 2, nullptr, 0) == NFFT_E_INVALID && !r);
    EXPECT(nfft_ragged_create(&r, NFFT_F64, 3, sz, 0) == NFFT_E_SIZE && !r);
    EXPECT(nfft_ragged_create(&r, NFFT_F64, 2, szbig, 0) == NFFT_E_UNSUPPORTED && !r);
    EXPECT(nfft_ragged_execute(nullptr, 0, buf, buf, nullptr) == NFFT_E_INVALID);
    EXPECT(nfft_ragged_check(nullptr, nullptr) == NFFT_E_INVALID);
    EXPECT(nfft_ragged_destroy(nullptr) == NFFT_OK);
    EXPECT(nfft_ragged_total(nullptr) == 0);
}

// The reference's API through each drop-in library (sanitize_dropin.cpp,
// compiled once per precision against the matching header types).
int dropin_paths_f64(const char *path);
int dropin_paths_f32(const char *path);

int main(int argc, char **argv)
{
    core_paths();
    if (argc > 2) {
        failures += dropin_paths_f64(argv[1]);
        failures += dropin_paths_f32(argv[2]);
    }
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("host argument-path sanitizer checks passed\n");
    return 0;
}
