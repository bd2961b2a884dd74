// nfft_plan.cpp -- planner and executor behind the C ABI.
//
// Replaces the reference's planner (src/06_plan.h:28-246: power-of-two vs
// Bluestein, one slab of tables per plan) and its aux tree
// (src/07_legacy.h:1389-1700) with device-resident plans: tables are
// generated in FP64 on the GPU and cached per device, scratch lives in one
// HBM allocation per plan, and every transform is a short sequence of kernel
// launches on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/notorious_fft_hip.h"
#include "nfft_internal.h"
#include "nfft_plan.h"

namespace nfft {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string t_err;

namespace {
std::mutex g_alloc_mu;
std::string g_fail_env;
long long g_fail_at = 0, g_alloc_count = 0;
}  // namespace

static thread_local long long t_alloc_failures = 0;
long long alloc_failures() { return t_alloc_failures; }

hipError_t dev_malloc_raw(void **p, size_t bytes)
{
    {
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        const char *v = getenv("NFFT_FAIL_ALLOC");
        const std::string cur = v ? v : "";
        if (cur != g_fail_env) {  // (re)armed: count from here; "k" or "k:tag"
            g_fail_env = cur;
            g_fail_at = cur.empty() ? 0 : atoll(cur.c_str());
            g_alloc_count = 0;
        }
        if (g_fail_at > 0 && ++g_alloc_count == g_fail_at) {
            *p = nullptr;
            t_alloc_failures++;
            return hipErrorOutOfMemory;
        }
    }
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) t_alloc_failures++;
    return e;
}
// Set for the duration of an allocation-free execute: engines then use only
// the scratch they are handed (no plan-owned counter blocks or slots).
static thread_local int t_noalloc = 0;
void set_error(const std::string &msg) { t_err = msg; }
void clear_error() { t_err.clear(); }
const char *last_error() { return t_err.c_str(); }

static bool hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

static bool is_pow2(long long n) { return n > 0 && (n & (n - 1)) == 0; }
static int ilog2ll(long long n)
{
    int l = 0;
    while ((1LL << l) < n) l++;
    return l;
}

// ---------------------------------------------------------------------------
// kernel timer
// ---------------------------------------------------------------------------
namespace {
struct ProfRec {
    std::string family;
    int prec, L;
    double bytes;
    int slot;
};
struct Prof {
    std::mutex mu;
    bool on = false;
    std::vector<hipEvent_t> ev;  // 2 per slot
    int used = 0;
    bool overflow = false;
    std::vector<ProfRec> recs;
};
Prof g_prof;
}  // namespace

bool prof_on() { return g_prof.on; }

int prof_begin(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_prof.mu);
    if (!g_prof.on) return -1;
    if (2 * (g_prof.used + 1) > (int)g_prof.ev.size()) {
        g_prof.overflow = true;
        return -1;
    }
    const int slot = g_prof.used++;
    if (hipEventRecord(g_prof.ev[2 * slot], s) != hipSuccess) return -1;
    return slot;
}

void prof_end(int slot, hipStream_t s, const char *family, int prec, int L, double bytes)
{
    std::lock_guard<std::mutex> lk(g_prof.mu);
    if (slot < 0 || 2 * slot + 1 >= (int)g_prof.ev.size()) return;
    (void)hipEventRecord(g_prof.ev[2 * slot + 1], s);
    g_prof.recs.push_back(ProfRec{family, prec, L, bytes, slot});
}

// ---------------------------------------------------------------------------
// per-device table cache
// ---------------------------------------------------------------------------
namespace {
struct Cache {
    std::mutex mu;
    std::map<std::string, void *> tabs;
    hipStream_t st = nullptr;
};
Cache g_cache[64];

Cache &cur_cache()
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    return g_cache[dev & 63];
}
}  // namespace

hipStream_t setup_stream()
{
    Cache &c = cur_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    if (!c.st) {
        if (hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking) != hipSuccess) c.st = nullptr;
    }
    return c.st;
}

template <typename Gen> static void *cache_get(const std::string &key, size_t bytes, Gen gen)
{
    hipStream_t st = setup_stream();
    if (!st) {
        set_error("cannot create the setup stream");
        return nullptr;
    }
    Cache &c = cur_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.tabs.find(key);
    if (it != c.tabs.end()) return it->second;
    void *p = nullptr;
    if (!hip_ok(dev_malloc(&p, bytes < 16 ? 16 : bytes), "dev_malloc(table)")) return nullptr;
    if (!hip_ok(gen(p, st), "table generation") || !hip_ok(hipStreamSynchronize(st), "table generation")) {
        (void)hipFree(p);
        return nullptr;
    }
    c.tabs[key] = p;
    return p;
}

struct StageInfo {
    int S, R0, twn;
    int ns(int s) const { return s == 0 ? 1 : R0 << (4 * (s - 1)); }
    int twoff(int s) const
    {
        int o = 0;
        for (int t = 1; t < s; t++) o += ns(t) * 4;
        return o;
    }
};
static StageInfo stage_info(int L)
{
    StageInfo si;
    const int lg = ilog2ll(L);
    si.S = lg <= 4 ? 1 : (lg + 3) / 4;
    si.R0 = L >> (4 * (si.S - 1));
    si.twn = si.S <= 1 ? 0 : si.twoff(si.S);
    return si;
}

// Stage twiddles of an L-point line: FP64 base powers [w, w^2, w^4, w^8] of
// w = e^{-2 pi i k/(16 Ns)} for every butterfly index k of every stage >= 1
// (the same table serves both precisions).
const void *stage_table(int /*prec*/, int L)
{
    const StageInfo si = stage_info(L);
    const std::string key = "stage64/" + std::to_string(L);
    return cache_get(key, (size_t)std::max(1, si.twn) * 16, [&](void *p, hipStream_t st) {
        for (int s = 1; s < si.S; s++) {
            const long long cnt = (long long)si.ns(s) * 4;
            const int off = si.twoff(s);
            hipError_t e = launch_gen_phase((double2 *)p + off, nullptr, cnt, 3, 0, 0, (long long)si.ns(s) * 16, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    });
}

const void *phase_table(int prec, long long count, long long a, long long c, long long den)
{
    const std::string key = "ph/" + std::to_string(prec) + "/" + std::to_string(count) + "/" + std::to_string(a) +
                            "/" + std::to_string(c) + "/" + std::to_string(den);
    return cache_get(key, (size_t)std::max(1LL, count) * cx_bytes(prec), [&](void *p, hipStream_t st) {
        return launch_gen_phase(prec == F64 ? (double2 *)p : nullptr, prec == F32 ? (float2 *)p : nullptr, count,
                                0, a, c, den, st);
    });
}

// FP32 copy of the two-level table; rows of 1024 points read at most the
// first kTwinHi hi entries (exponents below N / 16).
bool tw2f_table(long long n, Tw2f *out)
{
    const int lg = ilog2ll(n);
    const int sh = (lg + 1) / 2;
    const long long lo_n = 1LL << sh, hi_n = (n + lo_n - 1) >> sh;
    if (lo_n > kTwinLo || hi_n < kTwinHi || ((n / 1024 - 1) * 64 >> sh) >= kTwinHi) return false;
    out->hi = phase_table(F32, hi_n, lo_n, 0, n);
    out->lo = phase_table(F32, lo_n, 1, 0, n);
    out->sh = sh;
    out->mask = (int)(lo_n - 1);
    return out->hi && out->lo;
}

bool tw2_table(long long n, Tw2 *out)
{
    const int lg = ilog2ll(n);
    const int sh = (lg + 1) / 2;
    const long long lo_n = 1LL << sh, hi_n = (n + lo_n - 1) >> sh;
    const void *hi = phase_table(F64, hi_n, lo_n, 0, n);
    const void *lo = phase_table(F64, lo_n, 1, 0, n);
    if (!hi || !lo) return false;
    out->hi = (const double2 *)hi;
    out->lo = (const double2 *)lo;
    out->sh = sh;
    out->mask = (int)(lo_n - 1);
    return true;
}

const void *chirp_table(int prec, long long n, const double2 **chirp64)
{
    auto gen = [&](int pr) {
        const std::string key = "chirp/" + std::to_string(pr) + "/" + std::to_string(n);
        return cache_get(key, (size_t)n * cx_bytes(pr), [&](void *p, hipStream_t st) {
            return launch_gen_phase(pr == F64 ? (double2 *)p : nullptr, pr == F32 ? (float2 *)p : nullptr, n, 2, 0,
                                    0, 2 * n, st);
        });
    };
    if (chirp64) *chirp64 = (const double2 *)gen(F64);
    return gen(prec);
}

// ---------------------------------------------------------------------------
// 1D complex engine
// ---------------------------------------------------------------------------
static int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

// ---------------------------------------------------------------------------
// mixed radix vs Bluestein for 7-smooth lengths
// ---------------------------------------------------------------------------
// Tile width for `len` columns/rows at most `wmax` wide, balanced so that no
// tile is much narrower than the others.
static int balanced_width(int len, int wmax)
{
    wmax = std::max(1, std::min(len, wmax));
    const int tiles = (len + wmax - 1) / wmax;
    return (len + tiles - 1) / tiles;
}

// Measured on MI355X (2^28 FP32 / 2^27 FP64 elements): mixed-radix lines beat
// Bluestein for every FP64 length and for FP32 up to 512 points; the
// mixed-radix four-step wins for FP64 unless its tiles are a single column
// or row wide, and for FP32 from 10^4 points with tiles >= 4 wide.
static bool smooth_default(int prec, long long n, bool line, int wmin)
{
    if (line) return prec == F64 || n <= 512;
    return prec == F64 ? wmin >= 2 : (n >= 10000 && wmin >= 4);
}

static std::string smooth_key(int prec, long long n)
{
    return "c2c-smooth p" + std::to_string(prec) + " n" + std::to_string(n);
}

// Tuning results of this process, by record key (value: the text after the
// key), so that a length is measured once per process even when no
// NFFT_TUNE_CACHE file is named.
static std::mutex g_tune_mu;
static std::map<std::string, std::string> g_tune_mem;

static bool tune_mem_get(const std::string &key, std::string *val)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_mem.find(key);
    if (it == g_tune_mem.end()) return false;
    *val = it->second;
    return true;
}

static void tune_mem_put(const std::string &key, const std::string &val)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_mem[key] = val;
}

// Choice recorded in the NFFT_TUNE_CACHE file (the last record wins), else
// measured earlier in this process (-1: none).
static int tuned_choice(int prec, long long n)
{
    const std::string key = smooth_key(prec, n);
    int found = -1;
    const char *path = getenv("NFFT_TUNE_CACHE");
    if (path && *path) {
        if (FILE *f = fopen(path, "r")) {
            char line[256];
            while (fgets(line, sizeof line, f)) {
                if (strncmp(line, key.c_str(), key.size()) == 0 && line[key.size()] == ' ')
                    found = atoi(line + key.size() + 1);
            }
            fclose(f);
        }
    }
    std::string mem;
    if (found < 0 && tune_mem_get(key, &mem)) found = atoi(mem.c_str());
    return found;
}

// Project-written autotuner: times the mixed-radix plan and the Bluestein
// plan of length n on a synthetic batch of about 2^22 points and records the
// faster one (1: mixed radix, 0: Bluestein) in NFFT_TUNE_CACHE.
static int tune_smooth(int n, int prec)
{
    C2C cand[2];
    if (!cand[0].init(n, prec, 0) || !cand[1].init(n, prec, 1)) {
        cand[0].release();
        cand[1].release();
        return -1;
    }
    const long long lines = std::max(1LL, (1LL << 22) / n);
    const size_t cb = cx_bytes(prec);
    hipStream_t st = setup_stream();
    void *x = nullptr, *y = nullptr, *scr = nullptr;
    const long long sp = std::max(cand[0].scratch_per_line, cand[1].scratch_per_line);
    float ms[2] = {1e30f, 1e30f};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = dev_malloc(&x, (size_t)lines * n * cb) == hipSuccess && dev_malloc(&y, (size_t)lines * n * cb) == hipSuccess &&
              (sp == 0 || dev_malloc(&scr, (size_t)lines * sp * cb) == hipSuccess) && hipEventCreate(&e0) == hipSuccess &&
              hipEventCreate(&e1) == hipSuccess &&
              launch_fill_uniform(x, 2 * lines * n, 5, 0, prec, st) == hipSuccess;
    for (int c = 0; c < 2 && ok; c++) {
        for (int w = 0; w < 2 && ok; w++) ok = cand[c].run(x, y, lines, n, n, -1, 1.0, scr, lines, st) == hipSuccess;
        ok = ok && hipEventRecord(e0, st) == hipSuccess;
        for (int r = 0; r < 5 && ok; r++) ok = cand[c].run(x, y, lines, n, n, -1, 1.0, scr, lines, st) == hipSuccess;
        ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms[c], e0, e1) == hipSuccess;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void *p : {x, y, scr})
        if (p) (void)hipFree(p);
    cand[0].release();
    cand[1].release();
    if (!ok) {
        (void)hipGetLastError();
        return -1;
    }
    const int pick = ms[1] < ms[0] ? 1 : 0;
    tune_mem_put(smooth_key(prec, n), std::to_string(pick));
    const char *path = getenv("NFFT_TUNE_CACHE");
    if (path && *path) {
        if (FILE *f = fopen(path, "a")) {
            fprintf(f, "%s %d\n", smooth_key(prec, n).c_str(), pick);
            fclose(f);
        }
    }
    return pick;
}

// Four-step choice: rows of 2^lg2 points, row pass on the prefetching grid
// (pf = 1) or one workgroup per tile (pf = 0).  A record in NFFT_TUNE_CACHE
// ("c2c-four p<prec> n<N> <lg2> <pf>") overrides the default rule; the tuner
// below writes one.  The tuner's own candidates are built with t_four_force
// set, so they neither read nor write records.
struct FourChoice {
    int lg2 = -1, pf = -1;
};
static thread_local FourChoice t_four_force;

static std::string four_key(int prec, long long n)
{
    return "c2c-four p" + std::to_string(prec) + " n" + std::to_string(n);
}

static FourChoice tuned_four(int prec, long long n)
{
    FourChoice c;
    const std::string key = four_key(prec, n);
    const char *path = getenv("NFFT_TUNE_CACHE");
    if (path && *path) {
        if (FILE *f = fopen(path, "r")) {
            char line[256];
            while (fgets(line, sizeof line, f)) {
                int a = -1, b = -1;
                if (strncmp(line, key.c_str(), key.size()) == 0 && line[key.size()] == ' ' &&
                    sscanf(line + key.size() + 1, "%d %d", &a, &b) == 2) {
                    c.lg2 = a;
                    c.pf = b;
                }
            }
            fclose(f);
        }
    }
    std::string mem;
    if (c.lg2 < 0 && tune_mem_get(key, &mem) && sscanf(mem.c_str(), "%d %d", &c.lg2, &c.pf) != 2) c = FourChoice{};
    return c;
}

// Project-written autotuner for power-of-two four-steps: times every row
// length the kernels have (2^8..2^11 points, columns within their limit) with
// both row passes on a synthetic batch of 1 GiB (2^27 FP32 / 2^26 FP64
// points: four times the memory-side cache, so that the passes stream from
// HBM as large batches do; NFFT_TUNE_LG overrides the log2 of the points)
// and records the fastest in NFFT_TUNE_CACHE.  Each candidate's time is the
// best of three blocks of five runs, and the default (dflt) is kept unless a
// candidate beats it by more than 2 % (run-to-run noise on these batches is
// 2-4 %).  Returns the choice (lg2 < 0: nothing measured).
static FourChoice tune_four(int n, int prec, FourChoice dflt)
{
    FourChoice best;
    const int lgp = std::max(12, std::min(30, env_int("NFFT_TUNE_LG", prec == F32 ? 27 : 26)));
    const long long lines = std::max(1LL, (1LL << lgp) / n);
    const size_t cb = cx_bytes(prec);
    hipStream_t st = setup_stream();
    if (!st) return best;
    void *x = nullptr, *y = nullptr, *scr = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = dev_malloc(&x, (size_t)lines * n * cb) == hipSuccess && dev_malloc(&y, (size_t)lines * n * cb) == hipSuccess &&
              dev_malloc(&scr, (size_t)lines * n * cb) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
              hipEventCreate(&e1) == hipSuccess && launch_fill_uniform(x, 2 * lines * n, 5, 0, prec, st) == hipSuccess;
    float best_ms = 1e30f, dflt_ms = 1e30f;
    const int lg = ilog2ll(n);
    for (int lg2 = 8; lg2 <= 11 && ok; lg2++) {
        const long long n2 = 1LL << lg2, n1 = n / n2;
        if (n2 > rowT_max(prec) || n1 < 2 || n1 > col_max(prec) || lg2 >= lg) continue;
        for (int pf = 1; pf >= 0 && ok; pf--) {
            C2C c;
            t_four_force = FourChoice{lg2, pf};
            const bool built = c.init(n, prec, 0, 0);
            t_four_force = FourChoice{};
            if (!built || c.algo != ALG_FOUR) {
                c.release();
                continue;
            }
            float ms = 1e30f;
            for (int w = 0; w < 2 && ok; w++) ok = c.run(x, y, lines, n, n, -1, 1.0, scr, lines, st) == hipSuccess;
            for (int blk = 0; blk < 3 && ok; blk++) {
                float b = 0;
                ok = hipEventRecord(e0, st) == hipSuccess;
                for (int r = 0; r < 5 && ok; r++) ok = c.run(x, y, lines, n, n, -1, 1.0, scr, lines, st) == hipSuccess;
                ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                     hipEventElapsedTime(&b, e0, e1) == hipSuccess;
                ms = std::min(ms, b);
            }
            c.release();
            if (lg2 == dflt.lg2 && pf == dflt.pf) dflt_ms = ms;
            if (ok && ms < best_ms) {
                best_ms = ms;
                best = FourChoice{lg2, pf};
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void *p : {x, y, scr})
        if (p) (void)hipFree(p);
    if (!ok) {
        (void)hipGetLastError();
        return FourChoice{};
    }
    if (dflt_ms <= 1.02f * best_ms) best = dflt;
    if (best.lg2 >= 0) tune_mem_put(four_key(prec, n), std::to_string(best.lg2) + " " + std::to_string(best.pf));
    const char *path = getenv("NFFT_TUNE_CACHE");
    if (best.lg2 >= 0 && path && *path) {
        if (FILE *f = fopen(path, "a")) {
            fprintf(f, "%s %d %d\n", four_key(prec, n).c_str(), best.lg2, best.pf);
            fclose(f);
        }
    }
    return best;
}

// Dependency waits of the persistent four-step give up after this long
// (wall clock); a healthy wait lasts microseconds.
static const double kPersistTimeoutSec = 2.0;
// Where the persistent four-step is the default (NFFT_PERSIST=0/1 forces it
// off/on).  FP64 N = 2^15 and 2^16: its intermediate goes through one
// 16-byte agent-scope access per point, and the single launch beats the
// two-launch path with 1024-point rows (2^27 points: 2^15 1.55 -> 1.38 ms,
// 2^16 1.55 -> 1.38 ms); from 2^17 its smaller tiles lose (2^17 1.57 vs
// 1.73 ms, 2^18 1.57 vs 2.28).  FP32 gains nothing over the default paths
// (2^16 1.53 vs 1.53 ms, 2^17 1.59 vs 1.59, 2^18 1.59 vs 2.40).
static bool persist_default(int prec, long long n) { return prec == F64 && n >= (1LL << 15) && n <= (1LL << 16); }
// Intermediate the persistent kernel keeps in flight (slots), sized to stay
// in the 256 MB Infinity Cache.
static const long long kPersistSlotBytes = 128LL << 20;
// Transforms per persistent launch (bounds the counter block).
static const long long kPersistCap = 1LL << 15;

static bool persist_setup(C2C &c)
{
    int dev = 0, ncu = 0, khz = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice") ||
        !hip_ok(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "device attribute") ||
        !hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "device attribute"))
        return false;
    const size_t bytes = (size_t)persist_ctr_words(kPersistCap) * sizeof(unsigned);
    if (!hip_ok(dev_malloc(&c.pctr, bytes), "hipMalloc")) return false;
    if (!hip_ok(hipMemset(c.pctr, 0, bytes), "hipMemset")) {
        (void)hipFree(c.pctr);
        c.pctr = nullptr;
        return false;
    }
    c.pcap = kPersistCap;
    c.pgrid = std::max(1, ncu) * c.pcfg.per_cu;
    c.ptimeout = (long long)(kPersistTimeoutSec * 1000.0 * std::max(khz, 1000));
    c.pgrab = std::max(1, env_int("NFFT_PERSIST_GRAB", 1));
    c.plag = std::max(0, env_int("NFFT_PERSIST_LAG", 0));
    long long mb = env_int("NFFT_PERSIST_MB", 0);
    const long long bytes_slots = mb > 0 ? (mb << 20) : kPersistSlotBytes;
    c.pslots = std::max(1LL, bytes_slots / ((long long)c.n * (long long)cx_bytes(c.prec)));
    return true;
}

// Transforms per XCD-affine launch (bounds its counter block: 24 words per
// transform).
static const long long kXcdCap = 1LL << 12;

// The XCD-affine four-step for FP32 lengths that have an instance, where it
// is the default: N = 2^20 (2^28 points: 1.62-1.68 ms vs 1.75 ms for the two
// launches; 2^19 1.60 vs 1.60 ms and 2^21 1.80-1.82 vs 1.75 ms keep the two
// launches, and so do 2^16 / 2^17 / 2^18 at 4.2 / 3.2 / 1.95 ms vs 1.50 /
// 1.59 / 1.59 ms, where a transform has too few tiles to occupy an XCD's 32
// workgroups with two slots).  NFFT_XCD=0 turns it off, NFFT_XCD=2 uses it
// for every length that has an instance (N = 2^16 ... 2^21).
// Returns 1 when the path is set up, 0 when it does not apply, -1 on an error
// (error set).
static int xcd_setup(C2C &c)
{
    int n1 = 0, per_cu = 0;
    const int mode = env_int("NFFT_XCD", 1);
    if (c.prec != F32 || mode == 0 || (mode == 1 && c.n != (1 << 20)) || !xcd_supported(c.n, &n1, &per_cu))
        return 0;
    int dev = 0, ncu = 0, khz = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice") ||
        !hip_ok(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "device attribute") ||
        !hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "device attribute"))
        return -1;
    const size_t bytes = (size_t)xcd_ctr_words(kXcdCap) * sizeof(unsigned);
    if (!hip_ok(dev_malloc(&c.xctr, bytes), "hipMalloc")) return -1;
    if (!hip_ok(hipMemset(c.xctr, 0, bytes), "hipMemset")) {
        (void)hipFree(c.xctr);
        c.xctr = nullptr;
        return -1;
    }
    c.xcap = kXcdCap;
    c.xgrid = std::max(1, ncu) * per_cu;
    if (c.xgrid > kXcdMaxGrid || kXcdCap * std::max(XShape<512>::B, XShape<2048>::A) > kXcdMaxTiles ||
        !xcd_topology(&c.xnq, c.xqmap)) {  // not applicable on this device
        (void)hipFree(c.xctr);
        c.xctr = nullptr;
        return 0;
    }
    c.ptimeout = (long long)(kPersistTimeoutSec * 1000.0 * std::max(khz, 1000));
    c.n1 = n1;
    c.n2 = (int)(c.n / n1);
    return 1;
}

bool C2C::init(int n_, int prec_, int smooth_mode, int persist_mode)
{
    n = n_;
    prec = prec_;
    if (n < 1) {
        set_error("transform length must be >= 1");
        return false;
    }
    // Lines with their exchanges in two LDS halves (k_line_h, one line per
    // workgroup):
    //   - twice the LDS-resident length (FP32 2^15, FP64 2^14): one pass over
    //     HBM instead of the four-step's two (2^28 points: FP32 1.50 -> 0.96
    //     ms, FP64 3.04 -> 1.64 ms), unless a four-step split is asked for (the
    //     R2C row pass that absorbs the unpack, the four-step tuner);
    //   - 8192 points (FP32 0.786 -> 0.757-0.773 ms, FP64 1.637 -> 1.532 ms)
    //     and FP32 16384 (0.977 -> 0.808 ms), where half the LDS lets two or
    //     more lines share a CU.
    // These are all the lengths the form exists for.  NFFT_LINE_H=0 turns it
    // off.
    const int line_h_env = env_int("NFFT_LINE_H", 1);
    const bool line_h = n == 2LL * line_max(prec) && t_four_force.lg2 < 0 && line_h_env != 0;
    if (is_pow2(n) && (n <= line_max(prec) || line_h)) {
        algo = ALG_LINE;
        L = n;
        halves = line_h || (line_h_env != 0 && line_h_supported(prec, n)) ? 1 : 0;
        tw_a = stage_table(prec, L);
        return tw_a != nullptr;
    }
    if (n > kMaxLen) {
        set_error("transform length " + std::to_string(n) + " exceeds the library's limit of 2^30");
        return false;
    }
    if (is_pow2(n) && (long long)n > (long long)col_max(prec) * rowT_max(prec)) {
        // beyond the four-step: N = N1 x N2 over two engines (six-step:
        // transpose, N1-point lines, twiddle + transpose, N2-point lines,
        // transpose), N1 >= N2
        const int lg = ilog2ll(n);
        n2 = 1 << (lg / 2);
        n1 = n / n2;
        algo = ALG_BIG;
        sub1 = std::make_shared<C2C>();
        sub2 = std::make_shared<C2C>();
        if (!sub1->init(n1, prec) || !sub2->init(n2, prec) || !tw2_table(n, &tw2)) return false;
        scratch_per_line = 2LL * n + std::max(sub1->scratch_per_line * n2, sub2->scratch_per_line * n1);
        return true;
    }
    if (is_pow2(n)) {
        algo = ALG_FOUR;
        scratch_per_line = n;
        if (persist_mode != 0 && env_int("NFFT_PERSIST", persist_default(prec, n) ? 1 : 0) && persist_find(prec, n, env_int("NFFT_PERSIST_VAR", 0), &pcfg)) {
            n1 = pcfg.n1;
            n2 = pcfg.n2;
            if (!persist_setup(*this)) return false;
        } else {
            // Rows of 1024 points (the prefetching row grid with 16-row tiles
            // in FP32) over columns of 16..2048 points measured fastest at
            // every length from 2^14 to 2^21, C2C and R2C, both precisions
            // (2^28 points FP32 / 2^27 FP64: e.g. FP32 2^19 512x1024 1.64 ms
            // vs 1024x512 1.80; FP64 2^16 64x1024 3.04 ms vs 256x256 3.15);
            // beyond that the square split.
            const int lg = ilog2ll(n);
            int lg2 = (lg - 10 >= 4 && lg - 10 <= 11) ? 10 : lg / 2;
            lg2 = std::min(lg2, ilog2ll(rowT_max(prec)));
            // a tuned choice (NFFT_TUNE_CACHE record, or measured now with
            // NFFT_TUNE=1), then NFFT_FOUR_N2, over the rule above
            FourChoice fc = t_four_force;
            if (fc.lg2 < 0) {
                fc = tuned_four(prec, n);
                if (fc.lg2 < 0 && env_int("NFFT_TUNE", 0)) fc = tune_four(n, prec, FourChoice{lg2, 1});
            }
            const auto split_ok = [&](long long f) {
                return f > 0 && is_pow2(f) && f <= rowT_max(prec) && n / f <= col_max(prec) && n / f >= 2;
            };
            if (fc.lg2 >= 0 && fc.lg2 < 31 && split_ok(1LL << fc.lg2)) lg2 = fc.lg2;
            const int forced = env_int("NFFT_FOUR_N2", 0);
            if (split_ok(forced)) lg2 = ilog2ll(forced);
            n2 = 1 << lg2;
            n1 = n / n2;
            if (n1 > col_max(prec)) {
                set_error("power-of-two length " + std::to_string(n) + " exceeds the library's limit");
                return false;
            }
            col_lpb = env_int("NFFT_COL_LPB", 0);
            row_lpb = env_int("NFFT_ROWT_LPB", 0);
            row_pf = env_int("NFFT_ROWT_PF", fc.pf >= 0 ? fc.pf : 1) != 0 ? 1 : 0;
            // FP32 1024-point rows take the inter-pass twiddle (column pass
            // 992 -> 910 us per 2^28 points; the prefetching row grid absorbs it)
            row_twin = prec == F32 && n2 == 1024 && env_int("NFFT_ROW_TWIDDLE", 1) != 0 ? 1 : 0;
            if (row_twin && !tw2f_table(n, &tw2f)) return false;
            // batches of at least 16 transforms: the XCD-affine single launch
            // (same N1 x 1024 split, inter-pass twiddle where row_twin puts it)
            if (persist_mode != 0 && n2 == 1024) {
                const int x = xcd_setup(*this);
                if (x < 0) return false;
                if (x > 0 && (n1 != n / 1024 || n2 != 1024)) {
                    set_error("internal: XCD four-step split");
                    return false;
                }
            }
        }
        tw_a = stage_table(prec, n1);
        tw_b = stage_table(prec, n2);
        return tw_a && tw_b && tw2_table(n, &tw2);
    }
    // 7-smooth lengths: mixed radix (the reference uses Bluestein for every
    // non-power-of-two length; results agree to rounding) where it is the
    // faster of the two: a tuned choice from NFFT_TUNE_CACHE, a measurement
    // when NFFT_TUNE=1, else the measured default rule
    const int smooth_env = env_int("NFFT_SMOOTH", 1);  // 0: never, 2: whenever possible
    if (smooth_env == 2 && smooth_mode == 2) smooth_mode = 1;
    if (smooth_mode != 0 && smooth_env != 0 && !is_pow2(n)) {
        SmoothPlan p1{}, p2{};
        const int mr = smooth_radix_of(prec);
        const bool line = n <= smooth_line_max(prec) && smooth_factor(n, &p1, mr);
        const bool four = !line && (smooth_split(n, smooth_line_max(prec), &p1, &p2, mr) ||
                                    smooth_split(n, prec == F32 ? 8192 : 4096, &p1, &p2, mr));
        if (line || four) {
            const int budget = smooth_tile_points(prec);
            const int w1 = four ? balanced_width(p2.L, budget / p1.L) : 0;
            const int w2 = four ? balanced_width(p1.L, budget / p2.L) : 0;
            int use = smooth_mode == 1 ? 1 : tuned_choice(prec, n);
            if (use < 0 && env_int("NFFT_TUNE", 0)) use = tune_smooth(n, prec);
            if (use < 0) use = smooth_default(prec, n, line, std::min(w1, w2)) ? 1 : 0;
            if (use == 1) {
                sp = p1;
                if (line) {
                    algo = ALG_SMOOTH;
                    L = n;
                    sgroup = std::max(1, budget / n);
                    tw_a = phase_table(prec, n, 1, 0, n);
                    return tw_a != nullptr;
                }
                algo = ALG_SMOOTH4;
                sp2 = p2;
                n1 = p1.L;
                n2 = p2.L;
                sgroup = w1;
                sgroup2 = w2;
                tw_a = phase_table(prec, n1, 1, 0, n1);
                tw_b = phase_table(prec, n2, 1, 0, n2);
                scratch_per_line = n;
                return tw_a && tw_b && tw2_table(n, &tw2);
            }
        }
    }
    // Bluestein
    m = 1;
    while (m < 2LL * n - 1) m <<= 1;
    const double2 *chirp64 = nullptr;
    chirp = chirp_table(prec, n, &chirp64);
    if (!chirp || !chirp64) return false;
    if (m <= line_max(prec)) {
        algo = ALG_BLUE1;
        L = m;
        tw_a = stage_table(prec, m);
        if (!tw_a) return false;
    } else {
        const int lg = ilog2ll(m);
        const int lg2 = std::min(lg / 2, ilog2ll(bluerow_max(prec)));
        m2 = 1 << lg2;
        m1 = m / m2;
    }
    if (m > line_max(prec) && m1 > col_max(prec)) {
        // M beyond the three-pass form: chirp, FFT_M, x H, IFFT_M, chirp as
        // elementwise passes around an M-point engine
        if (m > kMaxLen) {
            set_error("Bluestein length " + std::to_string(n) + " exceeds the library's limit");
            return false;
        }
        algo = ALG_BLUEBIG;
        m1 = 0;
        sub1 = std::make_shared<C2C>();
        if (!sub1->init(m, prec)) return false;
        scratch_per_line = m + sub1->scratch_per_line;
    } else if (m > line_max(prec)) {
        algo = ALG_BLUE3;
        tw_a = stage_table(prec, m1);
        tw_b = stage_table(prec, m2);
        if (!tw_a || !tw_b || !tw2_table(m, &tw2)) return false;
        scratch_per_line = m;
    }
    // filter spectrum H = FFT_m(h) / m, computed in FP64 by this engine
    C2C h64;
    if (!h64.init(m, F64)) return false;
    hipStream_t st = setup_stream();
    void *hbuf = nullptr, *Hbuf = nullptr, *sbuf = nullptr;
    bool ok = hip_ok(dev_malloc(&hbuf, (size_t)m * 16), "hipMalloc") &&
              hip_ok(dev_malloc(&Hbuf, (size_t)m * 16), "hipMalloc") &&
              hip_ok(dev_malloc(&H, (size_t)m * cx_bytes(prec)), "hipMalloc");
    if (ok && h64.scratch_per_line) ok = hip_ok(dev_malloc(&sbuf, (size_t)h64.scratch_per_line * 16), "hipMalloc");
    ok = ok && hip_ok(launch_blue_filter((double2 *)hbuf, chirp64, n, m, st), "filter") &&
         hip_ok(h64.run(hbuf, Hbuf, 1, m, m, -1, 1.0, sbuf, 1, st), "filter FFT") &&
         hip_ok(launch_blue_hprep((const double2 *)Hbuf, H, prec, m, algo == ALG_BLUE3 ? m1 : 0, m2, 1.0 / m, st),
                "filter prep") &&
         hip_ok(hipStreamSynchronize(st), "filter sync");
    (void)hipFree(hbuf);
    (void)hipFree(Hbuf);
    if (sbuf) (void)hipFree(sbuf);
    h64.release();
    if (!ok) {
        if (H) (void)hipFree(H);
        H = nullptr;
    }
    return ok;
}

void C2C::release()
{
    if (H) (void)hipFree(H);
    H = nullptr;
    if (pctr) (void)hipFree(pctr);
    pctr = nullptr;
    if (xctr) (void)hipFree(xctr);
    xctr = nullptr;
    for (auto *sub : {&sub1, &sub2}) {
        if (*sub) (*sub)->release();
        sub->reset();
    }
}

bool C2C::device_fault() const
{
    for (unsigned *c : {pctr, xctr}) {
        if (!c) continue;
        unsigned w[2] = {0, 0};
        if (hipMemcpy(w, c + 1, sizeof w, hipMemcpyDeviceToHost) != hipSuccess) return true;
        if ((w[0] | w[1]) != 0) return true;
    }
    return false;
}

std::string C2C::describe() const
{
    char b[128];
    switch (algo) {
    case ALG_LINE:
        snprintf(b, sizeof b, halves ? "line %d (exchanges in two LDS halves)" : "line %d", n);
        break;
    case ALG_FOUR:
        if (pctr)
            snprintf(b, sizeof b, "four-step %dx%d persistent (tiles %d/%d, %d workgroups)", n1, n2, pcfg.lpb1,
                     pcfg.lpb2, pgrid);
        else if (xctr)
            snprintf(b, sizeof b, "four-step %dx%d XCD-affine (%d workgroups; two launches below 16 transforms)",
                     n1, n2, xgrid);
        else
            snprintf(b, sizeof b, "four-step %dx%d", n1, n2);
        break;
    case ALG_BLUE1: snprintf(b, sizeof b, "bluestein N=%d M=%d single-pass", n, m); break;
    case ALG_SMOOTH: {
        int o = snprintf(b, sizeof b, "mixed-radix %d (", n);
        for (int i = 0; i < sp.ns && o < (int)sizeof b - 8; i++)
            o += snprintf(b + o, sizeof b - o, i ? "x%d" : "%d", sp.radix[i]);
        snprintf(b + o, sizeof b - o, ")");
        break;
    }
    case ALG_SMOOTH4: snprintf(b, sizeof b, "mixed-radix four-step %dx%d", n1, n2); break;
    case ALG_BLUE3: snprintf(b, sizeof b, "bluestein N=%d M=%d (%dx%d)", n, m, m1, m2); break;
    case ALG_BIG:
        snprintf(b, sizeof b, "six-step %dx%d [%s | %s]", n1, n2, sub1 ? sub1->describe().c_str() : "?",
                 sub2 ? sub2->describe().c_str() : "?");
        break;
    case ALG_BLUEBIG:
        snprintf(b, sizeof b, "bluestein N=%d M=%d [%s]", n, m, sub1 ? sub1->describe().c_str() : "?");
        break;
    default: snprintf(b, sizeof b, "n=%d", n);
    }
    return b;
}

hipError_t C2C::run(const void *in, void *out, long long nlines, long long in_dist, long long out_dist, int dir,
                    double scale, void *scratch, long long scratch_lines, hipStream_t s) const
{
    if (nlines <= 0) return hipSuccess;
    const size_t cb = cx_bytes(prec);
    switch (algo) {
    case ALG_LINE: {
        LineArgs a{in, out, nlines, in_dist, out_dist, tw_a, scale, halves};
        return launch_line(prec, L, dir, a, s);
    }
    case ALG_BLUE1: {
        BlueArgs a{in, out, nlines, in_dist, out_dist, n, tw_a, chirp, H, dir > 0 ? 1 : 0, scale};
        return launch_blue_line(prec, m, a, s);
    }
    case ALG_SMOOTH: {
        SmoothArgs a{in, out, nlines, in_dist, out_dist, sgroup, sp, tw_a, scale};
        return launch_smooth_line(prec, dir, a, s);
    }
    case ALG_SMOOTH4: {
        if (scratch_lines < 1 || !scratch) return hipErrorInvalidValue;
        for (long long l0 = 0; l0 < nlines; l0 += scratch_lines) {
            Smooth4Args a{};
            a.in = (const char *)in + (size_t)l0 * in_dist * cb;
            a.out = (char *)out + (size_t)l0 * out_dist * cb;
            a.scratch = scratch;
            a.ntrans = std::min(scratch_lines, nlines - l0);
            a.in_dist = in_dist;
            a.out_dist = out_dist;
            a.sp1 = sp;
            a.sp2 = sp2;
            a.w1 = sgroup;
            a.w2 = sgroup2;
            a.tw1 = tw_a;
            a.tw2t = tw_b;
            a.tw2 = tw2;
            a.scale = scale;
            hipError_t e = launch_smooth4(prec, dir, a, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    case ALG_FOUR: {
        if (scratch_lines < 1 || !scratch) return hipErrorInvalidValue;
        if (pctr && !t_noalloc) {
            // one persistent launch per kPersistCap transforms; the scratch
            // holds the rotating intermediate slots
            for (long long l0 = 0; l0 < nlines; l0 += pcap) {
                const long long cnt = std::min(pcap, nlines - l0);
                PersistArgs a{};
                a.in = (const char *)in + (size_t)l0 * in_dist * cb;
                a.out = (char *)out + (size_t)l0 * out_dist * cb;
                a.ntrans = cnt;
                a.in_dist = in_dist;
                a.out_dist = out_dist;
                a.scratch = scratch;
                a.slots = (int)std::max(1LL, std::min(std::min(scratch_lines, cnt), pslots));
                a.lag = plag > 0 ? std::min(plag, a.slots - 1) : a.slots / 2;
                a.ctr = pctr;
                a.tw_a = tw_a;
                a.tw_b = tw_b;
                a.tw2 = tw2;
                a.scale = scale;
                a.timeout = ptimeout;
                a.grab = pgrab;
                const long long items = cnt * ((long long)n2 / pcfg.lpb1 + (long long)n1 / pcfg.lpb2);
                const int grid = (int)std::min<long long>(pgrid, (items + pgrab - 1) / pgrab);
                hipError_t e = launch_persist_reset(pctr, persist_ctr_words(cnt), s);
                if (e == hipSuccess) e = launch_four_persist(pcfg, dir, a, grid, s);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        }
        constexpr long long kXcdLines = (long long)kXcdQueues * kXcdSlots;
        if (xctr && !t_noalloc && nlines >= kXcdLines && scratch_lines >= (long long)xnq * kXcdSlots) {
            // one XCD-affine launch per xcap transforms; the scratch holds the
            // two intermediate slots of each XCD
            for (long long l0 = 0; l0 < nlines; l0 += xcap) {
                const long long cnt = std::min(xcap, nlines - l0);
                XcdArgs a{};
                a.in = (const char *)in + (size_t)l0 * in_dist * cb;
                a.out = (char *)out + (size_t)l0 * out_dist * cb;
                a.ntrans = cnt;
                a.in_dist = in_dist;
                a.out_dist = out_dist;
                a.scratch = scratch;
                a.ctr = xctr;
                a.tw_a = tw_a;
                a.tw_b = tw_b;
                a.twf = tw2f;
                a.tw2 = tw2;
                a.scale = scale;
                a.timeout = ptimeout;
                a.nq = xnq;
                for (int i = 0; i < kXcdIds; i++) a.qmap[i] = xqmap[i];
                hipError_t e = launch_persist_reset(xctr, xcd_ctr_words(cnt), s);
                if (e == hipSuccess) e = launch_four_xcd(n1, dir, row_twin != 0, a, xgrid, s);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        }
        for (long long l0 = 0; l0 < nlines; l0 += scratch_lines) {
            const long long cnt = std::min(scratch_lines, nlines - l0);
            ColArgs c{};
            c.in = (const char *)in + (size_t)l0 * in_dist * cb;
            c.out = scratch;
            c.ncols = n2;
            c.S = n2;
            c.in_dist = in_dist;
            c.out_dist = n;
            c.ntrans = cnt;
            c.tw = tw_a;
            c.tw2 = tw2;
            c.scale = 1.0;
            c.lpb = col_lpb;
            hipError_t e = launch_col(prec, n1, dir, row_twin ? COL_PLAIN : COL_TWIDDLE, c, s);
            if (e != hipSuccess) return e;
            RowTArgs r{scratch, (char *)out + (size_t)l0 * out_dist * cb, n1, n, out_dist, cnt, tw_b, scale, row_lpb,
                       row_pf};
            if (row_twin) r.twf = tw2f;
            e = launch_rowT(prec, n2, dir, r, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    case ALG_BIG: {
        // per chunk of transforms: scratch [A: cnt n][B: cnt n][engine scratch]
        if (scratch_lines < 1 || !scratch) return hipErrorInvalidValue;
        const long long big = (long long)n;
        for (long long l0 = 0; l0 < nlines; l0 += scratch_lines) {
            const long long cnt = std::min(scratch_lines, nlines - l0);
            char *A = (char *)scratch, *B = A + (size_t)cnt * big * cb, *E = B + (size_t)cnt * big * cb;
            const long long e_elems = (scratch_lines - cnt) * scratch_per_line + cnt * (scratch_per_line - 2 * big);
            hipError_t e = hipSuccess;
            // in (n1 x n2 per transform) -> A (n2 x n1)
            for (long long t = 0; t < cnt && e == hipSuccess; t += (in_dist == big ? cnt : 1))
                e = launch_transpose((const char *)in + (size_t)(l0 + t) * in_dist * cb, A + (size_t)t * big * cb,
                                     in_dist == big ? cnt : 1, n1, n2, (int)cb, s);
            if (e != hipSuccess) return e;
            // n1-point lines, then the twiddle w_N^{r c} with the transpose to (n1 x n2)
            const long long f1 = sub1->scratch_per_line ? e_elems / sub1->scratch_per_line : 0;
            e = sub1->run(A, A, cnt * n2, n1, n1, dir, 1.0, E, f1, s);
            if (e == hipSuccess) e = launch_tw_transpose(prec, A, B, cnt, n2, n1, tw2, dir, scale, s);
            const long long f2 = sub2->scratch_per_line ? e_elems / sub2->scratch_per_line : 0;
            if (e == hipSuccess) e = sub2->run(B, B, cnt * n1, n2, n2, dir, 1.0, E, f2, s);
            // (n1 x n2) -> out (n2 x n1): point k1 + n1 k2
            for (long long t = 0; t < cnt && e == hipSuccess; t += (out_dist == big ? cnt : 1))
                e = launch_transpose(B + (size_t)t * big * cb, (char *)out + (size_t)(l0 + t) * out_dist * cb,
                                     out_dist == big ? cnt : 1, n1, n2, (int)cb, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    case ALG_BLUEBIG: {
        // per chunk: scratch [Z: cnt m][M-point engine scratch]
        if (scratch_lines < 1 || !scratch) return hipErrorInvalidValue;
        const int conj = dir > 0 ? 1 : 0;
        for (long long l0 = 0; l0 < nlines; l0 += scratch_lines) {
            const long long cnt = std::min(scratch_lines, nlines - l0);
            char *Z = (char *)scratch, *E = Z + (size_t)cnt * m * cb;
            const long long e_elems = (scratch_lines - cnt) * scratch_per_line + cnt * (scratch_per_line - m);
            const long long fm = sub1->scratch_per_line ? e_elems / sub1->scratch_per_line : 0;
            hipError_t e = launch_blue_ew(prec, 0, (const char *)in + (size_t)l0 * in_dist * cb, Z, chirp, n, m,
                                          in_dist, cnt, conj, 1.0, s);
            if (e == hipSuccess) e = sub1->run(Z, Z, cnt, m, m, -1, 1.0, E, fm, s);
            if (e == hipSuccess) e = launch_blue_ew(prec, 1, nullptr, Z, H, n, m, m, cnt, 0, 1.0, s);
            if (e == hipSuccess) e = sub1->run(Z, Z, cnt, m, m, 1, 1.0, E, fm, s);
            if (e == hipSuccess)
                e = launch_blue_ew(prec, 2, Z, (char *)out + (size_t)l0 * out_dist * cb, chirp, n, m, out_dist, cnt,
                                   conj, scale, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    case ALG_BLUE3: {
        if (scratch_lines < 1 || !scratch) return hipErrorInvalidValue;
        for (long long l0 = 0; l0 < nlines; l0 += scratch_lines) {
            const long long cnt = std::min(scratch_lines, nlines - l0);
            ColArgs c{};
            c.in = (const char *)in + (size_t)l0 * in_dist * cb;
            c.out = scratch;
            c.ncols = m2;
            c.S = m2;
            c.in_dist = in_dist;
            c.out_dist = m;
            c.ntrans = cnt;
            c.tw = tw_a;
            c.tw2 = tw2;
            c.chirp = chirp;
            c.nv = n;
            c.conj = dir > 0 ? 1 : 0;
            c.scale = 1.0;
            hipError_t e = launch_col(prec, m1, -1, COL_BLUE_IN, c, s);
            if (e != hipSuccess) return e;
            BlueRowArgs r{scratch, m1, m, cnt, tw_b, H, tw2};
            e = launch_blue_row(prec, m2, r, s);
            if (e != hipSuccess) return e;
            ColArgs o{};
            o.in = scratch;
            o.out = (char *)out + (size_t)l0 * out_dist * cb;
            o.ncols = m2;
            o.S = m2;
            o.in_dist = m;
            o.out_dist = out_dist;
            o.ntrans = cnt;
            o.tw = tw_a;
            o.tw2 = tw2;
            o.chirp = chirp;
            o.nv = n;
            o.conj = dir > 0 ? 1 : 0;
            o.scale = scale;
            e = launch_col(prec, m1, 1, COL_BLUE_OUT, o, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    }
    return hipErrorInvalidValue;
}

hipError_t c2c_axis(const C2C &e, void *buf, long long A, long long B, int dir, double scale, void *scratch,
                    long long scratch_elems, hipStream_t s, const void *src)
{
    const int n = e.n;
    if (!src) src = buf;
    if (n == 1) {
        if (src != buf) {
            const hipError_t err =
                hipMemcpyAsync(buf, src, (size_t)(A * B) * cx_bytes(e.prec), hipMemcpyDeviceToDevice, s);
            if (err != hipSuccess) return err;
        }
        if (scale == 1.0) return hipSuccess;
        return launch_scale_real(buf, 2 * A * B, e.prec, scale, s);
    }
    if (B == 1) {
        const long long fit = e.scratch_per_line ? scratch_elems / e.scratch_per_line : 0;
        return e.run(src, buf, A, n, n, dir, scale, scratch, fit, s);
    }
    if (e.algo == ALG_LINE && n <= col_max(e.prec) && B <= 0x7fffffffLL) {
        ColArgs c{};
        c.in = src;
        c.out = buf;
        c.ncols = (int)B;
        c.S = B;
        c.in_dist = (long long)n * B;
        c.out_dist = (long long)n * B;
        c.ntrans = A;
        c.tw = e.tw_a;
        c.scale = scale;
        return launch_col(e.prec, n, dir, COL_PLAIN, c, s);
    }
    // general axis: transpose to make it contiguous, transform, transpose back
    const long long tot = A * n * B;
    if (scratch_elems < tot) return hipErrorInvalidValue;
    const size_t cb = cx_bytes(e.prec);
    hipError_t err = launch_transpose(src, scratch, A, n, B, (int)cb, s);
    if (err != hipSuccess) return err;
    const long long rest = scratch_elems - tot;
    const long long fit = e.scratch_per_line ? rest / e.scratch_per_line : 0;
    err = e.run(scratch, scratch, A * B, n, n, dir, scale, (char *)scratch + (size_t)tot * cb, fit, s);
    if (err != hipSuccess) return err;
    return launch_transpose(scratch, buf, A, B, n, (int)cb, s);
}

// ---------------------------------------------------------------------------
// real (realdft) engine along contiguous lines of N reals
// ---------------------------------------------------------------------------
bool R2C::init(int N_, int prec_)
{
    N = N_;
    prec = prec_;
    if (!is_pow2(N)) {
        set_error("realdft needs a power-of-two length");
        return false;
    }
    if (N == 1) return true;
    L = N / 2;
    rtw = phase_table(prec, L, 1, 0, N);
    if (!rtw) return false;
    // complex lines of twice the LDS-resident length in the half-LDS form (one
    // pass instead of a four-step; 2^28 reals: FP32 N = 2^16 realdft 0.88 ->
    // 0.61 ms, invrealdft 1.44 -> 0.65 ms; FP64 2^15 1.54 -> 1.06, 2.56 ->
    // 1.16 ms), FP32 16384-point lines (N = 2^15: 0.58 -> 0.55 ms, inverse
    // 0.65 -> 0.48 ms) and FP64 8192-point lines forward (N = 2^14: 1.06 ->
    // 0.91 ms; its inverse, 1.10 both ways, and FP32 8192, even, keep the
    // LDS-resident line).  NFFT_LINE_H=0 / 2 as for complex lines.
    const int line_h_env = env_int("NFFT_LINE_H", 1);
    const bool line_h = L == 2LL * line_max(prec) && line_h_env != 0 && line_h_supported(prec, L);
    if (L <= line_max(prec) || line_h) {
        single = true;
        if (line_h || (line_h_env == 2 && line_h_supported(prec, L)) ||
            (line_h_env != 0 && prec == F32 && L == 16384 && line_h_supported(prec, L)))
            halves = 3;
        else if (line_h_env != 0 && prec == F64 && L == 8192 && line_h_supported(prec, L)) halves = 1;
        tw = stage_table(prec, L);
        return tw != nullptr;
    }
    single = false;
    // the row pass of a two-launch four-step absorbs the unpack (k_rowT_r2c).
    // Its row length: shorter rows in row-pair tiles of twice the height
    // (128-byte output runs on both sides of each pair) where measured faster
    // on 2^28 reals -- FP64 256-point rows up to N = 2^20 (2^16: 1.71 ->
    // 1.53 ms, 2^20: 1.80 -> 1.71 ms), FP32 512-point rows while the columns
    // stay within 512 points (2^16: 0.97 -> 0.90 ms, 2^19: 0.95 -> 0.91 ms) --
    // and the four-step's own split otherwise.
    const bool fuse = env_int("NFFT_R2C_FUSE", 1) != 0;
    int lg2r = 0;
    if (fuse && env_int("NFFT_FOUR_N2", 0) == 0) {
        if (prec == F64 && L >= (2LL << 8) && L <= (1LL << 19)) lg2r = 8;
        if (prec == F32 && L >= (2LL << 9) && L <= (512LL << 9)) lg2r = 9;
    }
    t_four_force = lg2r ? FourChoice{lg2r, 1} : FourChoice{};
    const bool built = inner.init(L, prec, 2, fuse ? 0 : 2);
    t_four_force = FourChoice{};
    if (!built) return false;
    r2c_lpb = env_int("NFFT_R2C_LPB", lg2r && inner.n2 == (1 << lg2r) ? 2 * rowT_lpb_rt(prec, 1 << lg2r) : 0);
    fused = inner.algo == ALG_FOUR && !inner.pctr && (inner.n1 % 2) == 0 && env_int("NFFT_R2C_FUSE", 1) != 0;
    scratch_per_line = L + inner.scratch_per_line;
    return true;
}

void R2C::release() { inner.release(); }

hipError_t R2C::fwd(const void *in, void *out, long long nlines, long long in_dist, long long out_dist, double scale,
                    int dbl, void *scratch, long long scratch_elems, hipStream_t s) const
{
    if (nlines <= 0) return hipSuccess;
    if (N == 1) {
        EwArgs a{in, out, nlines, in_dist, out_dist, 0, nullptr, nullptr, scale, 1.0, 0, 1};
        return launch_ew(prec, EW_N1, a, 1, s);
    }
    if (single) {
        RtArgs a{in, out, nlines, in_dist, out_dist, tw, rtw, nullptr, scale, 1.0, dbl, halves & 1};
        return launch_r2c(prec, L, a, s);
    }
    const long long chunk = scratch_elems / scratch_per_line;
    if (chunk < 1) return hipErrorInvalidValue;
    const size_t cb = cx_bytes(prec), rb = re_bytes(prec);
    void *Z = scratch;
    void *isc = (char *)scratch + (size_t)chunk * L * cb;
    for (long long l0 = 0; l0 < nlines; l0 += chunk) {
        const long long cnt = std::min(chunk, nlines - l0);
        const void *src = (const char *)in + (size_t)l0 * in_dist * rb;
        void *dst = (char *)out + (size_t)l0 * out_dist * cb;
        hipError_t e;
        if (fused) {
            // column pass (+ inter-pass twiddle) into Z, then rows + unpack
            ColArgs c{};
            c.in = src;
            c.out = Z;
            c.ncols = inner.n2;
            c.S = inner.n2;
            c.in_dist = in_dist / 2;
            c.out_dist = L;
            c.ntrans = cnt;
            c.tw = inner.tw_a;
            c.tw2 = inner.tw2;
            c.scale = 1.0;
            c.lpb = inner.col_lpb;
            e = launch_col(prec, inner.n1, -1, COL_TWIDDLE, c, s);
            if (e != hipSuccess) return e;
            RowR2CArgs r{Z, dst, inner.n1, L, out_dist, cnt, inner.tw_b, rtw, scale, dbl, r2c_lpb};
            e = launch_rowT_r2c(prec, inner.n2, r, s);
            if (e != hipSuccess) return e;
            continue;
        }
        e = inner.run(src, Z, cnt, in_dist / 2, L, -1, 1.0, isc, chunk, s);
        if (e != hipSuccess) return e;
        EwArgs a{Z, dst, cnt, L, out_dist, L, rtw, nullptr, scale, 1.0, dbl, 0};
        e = launch_ew(prec, EW_UNPACK, a, L, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t R2C::inv(const void *in, void *out, long long nlines, long long in_dist, long long out_dist, double scale,
                    int half, void *scratch, long long scratch_elems, hipStream_t s) const
{
    if (nlines <= 0) return hipSuccess;
    if (N == 1) {
        EwArgs a{in, out, nlines, in_dist, out_dist, 0, nullptr, nullptr, scale, 1.0, 0, 2};
        return launch_ew(prec, EW_N1, a, 1, s);
    }
    if (single) {
        RtArgs a{in, out, nlines, in_dist, out_dist, tw, rtw, nullptr, scale, 1.0, half, (halves >> 1) & 1};
        return launch_c2r(prec, L, a, s);
    }
    const long long chunk = scratch_elems / scratch_per_line;
    if (chunk < 1) return hipErrorInvalidValue;
    const size_t cb = cx_bytes(prec), rb = re_bytes(prec);
    void *Z = scratch;
    void *isc = (char *)scratch + (size_t)chunk * L * cb;
    for (long long l0 = 0; l0 < nlines; l0 += chunk) {
        const long long cnt = std::min(chunk, nlines - l0);
        EwArgs a{(const char *)in + (size_t)l0 * in_dist * cb, Z, cnt, in_dist, L, L, rtw, nullptr, 1.0, 1.0, half, 0};
        hipError_t e = launch_ew(prec, EW_PACK, a, L, s);
        if (e != hipSuccess) return e;
        e = inner.run(Z, (char *)out + (size_t)l0 * out_dist * rb, cnt, L, out_dist / 2, 1, scale, isc, chunk, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// trig engine (DCT/DST 2/3 or 4) along contiguous lines of N reals
// ---------------------------------------------------------------------------
bool Trig::init(int N_, int prec_, int t4_)
{
    N = N_;
    prec = prec_;
    t4 = t4_;
    if (!is_pow2(N)) {
        set_error("trigonometric transforms need a power-of-two length");
        return false;
    }
    if (N == 1) return true;
    L = N / 2;
    if (!t4) {
        dw = phase_table(prec, L, 1, 0, 4LL * N);   // e^{-i pi k/(2N)}
        rw = phase_table(prec, L, 1, 0, N);         // e^{-2 pi i k/N}
        if (!dw || !rw) return false;
    } else {
        pre = phase_table(prec, L, 1, 0, 2LL * N);      // e^{-i pi n/N}
        post = phase_table(prec, N, 2, 1, 8LL * N);     // e^{-i pi (2n+1)/(4N)}
        if (!pre || !post) return false;
    }
    if (L <= line_max(prec)) {
        single = true;
        tw = stage_table(prec, L);
        return tw != nullptr;
    }
    single = false;
    if (!t4) {
        if (!real.init(N, prec)) return false;
        // pre buffer (L+1 complex) + post buffer (L+1 complex) + realdft scratch
        scratch_per_line = 2 * (L + 1) + real.scratch_per_line;
    } else {
        if (!cplx.init(L, prec)) return false;
        scratch_per_line = 2 * L + cplx.scratch_per_line;
    }
    return true;
}

void Trig::release()
{
    real.release();
    cplx.release();
}

hipError_t Trig::run(int op, const void *in, void *out, long long nlines, long long in_dist, long long out_dist,
                     double scale, double w0, void *scratch, long long scratch_elems, hipStream_t s) const
{
    if (nlines <= 0) return hipSuccess;
    const int top = op - NFFT_OP_DCT2;  // T_DCT2..T_DST4
    if (N == 1) {
        double f = 1.0;
        if (top == T_DCT2 || top == T_DST2) f = 2.0 * w0;
        else if (top == T_DCT3 || top == T_DST3) f = w0;
        else f = 1.41421356237309504880;
        EwArgs a{in, out, nlines, in_dist, out_dist, 0, nullptr, nullptr, f * scale, 1.0, 0, 3};
        return launch_ew(prec, EW_N1, a, 1, s);
    }
    if (single) {
        if (!t4) {
            RtArgs a{in, out, nlines, in_dist, out_dist, tw, dw, rw, scale, w0, 0};
            return launch_t23(prec, L, top, a, s);
        }
        RtArgs a{in, out, nlines, in_dist, out_dist, tw, pre, post, scale, 1.0, 0};
        return launch_t4(prec, L, top, a, s);
    }
    const long long chunk = scratch_elems / scratch_per_line;
    if (chunk < 1) return hipErrorInvalidValue;
    const size_t cb = cx_bytes(prec), rb = re_bytes(prec);
    for (long long l0 = 0; l0 < nlines; l0 += chunk) {
        const long long cnt = std::min(chunk, nlines - l0);
        const void *src = (const char *)in + (size_t)l0 * in_dist * rb;
        void *dst = (char *)out + (size_t)l0 * out_dist * rb;
        hipError_t e;
        if (!t4) {
            void *A = scratch;
            void *B = (char *)scratch + (size_t)chunk * (L + 1) * cb;
            void *rs = (char *)scratch + (size_t)chunk * 2 * (L + 1) * cb;
            const long long rsz = scratch_elems - chunk * 2 * (L + 1);
            EwArgs pa{src, A, cnt, in_dist, L + 1, L, dw, rw, 1.0, w0, 0, top};
            e = launch_ew(prec, EW_TRIG_PRE, pa, L + 1, s);
            if (e != hipSuccess) return e;
            if (top == T_DCT2 || top == T_DST2) {
                // A holds the packed reorder (N reals per line at stride 2(L+1))
                e = real.fwd(A, B, cnt, 2 * (L + 1), L + 1, 1.0, 0, rs, rsz, s);
                if (e != hipSuccess) return e;
                EwArgs qa{B, dst, cnt, L + 1, out_dist, L, dw, rw, scale, w0, 0, top};
                e = launch_ew(prec, EW_TRIG_POST, qa, L, s);
            } else {
                e = real.inv(A, B, cnt, L + 1, 2 * (L + 1), 1.0, 0, rs, rsz, s);
                if (e != hipSuccess) return e;
                EwArgs qa{B, dst, cnt, 2 * (L + 1), out_dist, L, dw, rw, scale, w0, 0, top};
                e = launch_ew(prec, EW_TRIG_POST, qa, L, s);
            }
        } else {
            void *A = scratch;
            void *B = (char *)scratch + (size_t)chunk * L * cb;
            void *cs = (char *)scratch + (size_t)chunk * 2 * L * cb;
            const long long csz = scratch_elems - chunk * 2 * L;
            const long long cfit = cplx.scratch_per_line ? csz / cplx.scratch_per_line : 0;
            EwArgs pa{src, A, cnt, in_dist, L, L, pre, post, 1.0, 1.0, 0, top};
            e = launch_ew(prec, EW_TRIG_PRE, pa, L, s);
            if (e != hipSuccess) return e;
            e = cplx.run(A, B, cnt, L, L, -1, 1.0, cs, cfit, s);
            if (e != hipSuccess) return e;
            EwArgs qa{B, dst, cnt, L, out_dist, L, pre, post, scale, 1.0, 0, top};
            e = launch_ew(prec, EW_TRIG_POST, qa, L, s);
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace nfft

// ===========================================================================
// plans
// ===========================================================================
using namespace nfft;

static long long prod(const std::vector<int> &v, size_t b, size_t e)
{
    long long p = 1;
    for (size_t i = b; i < e; i++) p *= v[i];
    return p;
}

// Whole batches up to 4 GiB of intermediate run as one launch pair: per-chunk
// launches cost a fill/drain each (measured: FP64 2^20 x 128 takes 2.27 ms in
// 64 MiB chunks, 1.68 ms in one).
static const long long kChunkBytesDefault = 4LL << 30;

// Lines of scratch an engine gets: one chunk of about NFFT_CHUNK_MB of
// intermediate (default 4 GiB, i.e. whole batches of the usual sizes).  The
// chunked passes process that many lines per launch pair; the persistent
// four-step uses at most NFFT_PERSIST_MB (default 128 MiB) of it as its
// rotating slots.
static long long chunk_lines(long long nlines, long long per_line_elems, size_t eb)
{
    if (per_line_elems <= 0) return 0;
    long long target = (long long)env_int("NFFT_CHUNK_MB", 0) << 20;
    if (target <= 0) target = kChunkBytesDefault;
    long long c = target / (per_line_elems * (long long)eb);
    c = std::max(1LL, std::min(c, nlines));
    return c;
}

// Scratch elements a complex engine needs for `lines` transforms: the
// persistent four-step only ever uses its rotating slots.
static long long engine_elems(const C2C &e, long long lines, size_t cb)
{
    if (e.pctr) return std::min(lines, e.pslots) * e.scratch_per_line;
    return chunk_lines(lines, e.scratch_per_line, cb) * e.scratch_per_line;
}

static double axis_inv_scale(int norm, int n)
{
    if (norm == NFFT_NORM_NONE) return 1.0;
    if (norm == NFFT_NORM_INVERSE) return 1.0 / n;
    return is_pow2(n) ? 1.0 : 1.0 / n;  // reference behaviour
}

bool nfft_plan::init(int kind_, int prec_, int d_, const int *dims_, long long batch_, int norm_)
{
    kind = kind_;
    prec = prec_;
    d = d_;
    norm = norm_;
    batch = batch_;
    dims.assign(dims_, dims_ + d);
    P = prod(dims, 0, d);
    const size_t cb = cx_bytes(prec), rb = re_bytes(prec);
    long long full = 0;   // elements for a full-array copy (transposes / inverse copies)
    long long eng = 0;    // engine scratch elements
    const int last = dims[d - 1];
    std::string dsc;
    if (kind == NFFT_KIND_DFT) {
        ax.resize(d);
        for (int i = 0; i < d; i++) {
            if (!ax[i].init(dims[i], prec)) return false;
            dsc += (i ? " | " : "") + ax[i].describe();
        }
        const long long lines = batch * P / last;
        eng = engine_elems(ax[d - 1], lines, cb);
        for (int i = 0; i + 1 < d; i++) {
            const C2C &e = ax[i];
            const bool direct = dims[i] == 1 || (e.algo == ALG_LINE && dims[i] <= col_max(prec));
            if (!direct) {
                full = std::max(full, batch * P);
                const long long lines_i = batch * P / dims[i];
                eng = std::max(eng, engine_elems(e, lines_i, cb));
            }
        }
        work_bytes = (size_t)(full + eng) * cb;
    } else if (kind == NFFT_KIND_REALDFT) {
        if (!real.init(last, prec)) return false;
        dsc = "realdft " + std::to_string(last) +
              (real.single ? (real.halves == 3   ? " single-pass (exchanges in two LDS halves)"
                              : real.halves == 1 ? " single-pass (forward: exchanges in two LDS halves)"
                                                 : " single-pass")
                           : " via " + real.inner.describe() + (real.fused ? " (unpack fused into the row pass)" : ""));
        const long long H = last / 2 + 1;
        const long long Pl = P / last;  // leading product
        ax.resize(d - 1);
        for (int i = 0; i + 1 < d; i++) {
            if (!ax[i].init(dims[i], prec)) return false;
            dsc += " | " + ax[i].describe();
        }
        const long long lines = batch * Pl;
        long long e_real = chunk_lines(lines, real.scratch_per_line, cb) * real.scratch_per_line;
        eng = e_real;
        long long cfull = 0;
        if (d > 1) cfull = batch * Pl * H;  // copy of the spectrum for the inverse
        for (int i = 0; i + 1 < d; i++) {
            const C2C &e = ax[i];
            const bool direct = dims[i] == 1 || (e.algo == ALG_LINE && dims[i] <= col_max(prec));
            if (!direct) {
                const long long lines_i = batch * Pl * H / dims[i];
                eng = std::max(eng, batch * Pl * H + engine_elems(e, lines_i, cb));
            }
        }
        full = cfull;
        work_bytes = (size_t)(full + eng) * cb;
    } else {
        const int t4 = kind == NFFT_KIND_T4;
        tax.resize(d);
        for (int i = 0; i < d; i++) {
            if (!tax[i].init(dims[i], prec, t4)) return false;
        }
        // the reference transforms every axis except a trailing run of unit
        // axes (src/07_legacy.h:822-861: N1 == 1 short-cut at each level)
        n_active = d;
        while (n_active > 1 && dims[n_active - 1] == 1) n_active--;
        dsc = std::string(t4 ? "dct4/dst4 " : "dct2/3 dst2/3 ") + std::to_string(P) +
              (d > 1 ? " separable" : (tax[0].single ? " single-pass" : " multi-pass"));
        long long e_max = 0;
        for (int i = 0; i < n_active; i++) {
            const long long lines_i = batch * P / dims[i];
            e_max = std::max(e_max, chunk_lines(lines_i, tax[i].scratch_per_line, rb) * tax[i].scratch_per_line);
        }
        // in reals: full-array transposes (2 copies) + the engine scratch (complex)
        full = n_active > 1 || d > 1 ? 2 * batch * P : 0;
        work_bytes = (size_t)full * rb + (size_t)e_max * cb;
    }
    desc = dsc;
    if (work_bytes) {
        if (dev_malloc(&work, work_bytes) != hipSuccess) {
            set_error("hipMalloc of " + std::to_string(work_bytes) + " workspace bytes failed");
            work = nullptr;
            return false;
        }
    }
    return true;
}

void nfft_plan::release()
{
    for (auto &e : ax) e.release();
    real.release();
    for (auto &t : tax) t.release();
    if (work) (void)hipFree(work);
    work = nullptr;
    if (wpow_part) (void)hipFree(wpow_part);
    wpow_part = nullptr;
}

// complex N-D over dims[0..nd-1] of `lines` stacked arrays, in place on buf
// except for the first (last-axis) pass which reads `in`.
static hipError_t dft_nd(const nfft_plan &p, const std::vector<C2C> &ax, const std::vector<int> &dims, int nd,
                         long long outer, const void *in, void *buf, int dir, bool forward_scale_rule,
                         double extra_scale, void *scratch, long long scratch_elems, hipStream_t s)
{
    const size_t cb = cx_bytes(p.prec);
    const long long Pn = prod(dims, 0, nd);
    auto scl = [&](int i) {
        double v = forward_scale_rule ? 1.0 : axis_inv_scale(p.norm, dims[i]);
        return v;
    };
    // last axis (contiguous), out of place
    const int lastn = dims[nd - 1];
    const long long lines = outer * Pn / lastn;
    const C2C &el = ax[nd - 1];
    hipError_t e;
    double first_scale = scl(nd - 1) * extra_scale;
    if (lastn == 1) {
        if (in != buf) {
            e = hipMemcpyAsync(buf, in, (size_t)outer * Pn * cb, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return e;
        }
        if (first_scale != 1.0) {
            e = launch_scale_real(buf, 2 * outer * Pn, p.prec, first_scale, s);
            if (e != hipSuccess) return e;
        }
    } else {
        const long long fit = el.scratch_per_line ? scratch_elems / el.scratch_per_line : 0;
        e = el.run(in, buf, lines, lastn, lastn, dir, first_scale, scratch, fit, s);
        if (e != hipSuccess) return e;
    }
    for (int i = nd - 2; i >= 0; i--) {
        const long long A = outer * prod(dims, 0, i), B = prod(dims, i + 1, nd);
        e = c2c_axis(ax[i], buf, A, B, dir, scl(i), scratch, scratch_elems, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// trig transform along axis i of a row-major (outer, dims...) real array.
static hipError_t trig_axis(const nfft_plan &p, int i, int op, const void *in, void *out, double w0, void *scr,
                            hipStream_t s)
{
    const size_t rb = re_bytes(p.prec);
    const int n = p.dims[i];
    const long long A = p.batch * prod(p.dims, 0, i), B = prod(p.dims, i + 1, p.d);
    const long long tot = A * n * B;
    void *full0 = scr, *full1 = (char *)scr + (size_t)tot * rb;
    void *eng = (char *)scr + (size_t)(p.d > 1 || p.n_active > 1 ? 2 * tot : 0) * rb;
    const long long eng_elems =
        (long long)((p.work_bytes - (size_t)(p.d > 1 || p.n_active > 1 ? 2 * tot : 0) * rb) / cx_bytes(p.prec));
    if (B == 1) return p.tax[i].run(op, in, out, A, n, n, 1.0, w0, eng, eng_elems, s);
    hipError_t e = launch_transpose(in, full0, A, n, B, (int)rb, s);
    if (e != hipSuccess) return e;
    e = p.tax[i].run(op, full0, full1, A * B, n, n, 1.0, w0, eng, eng_elems, s);
    if (e != hipSuccess) return e;
    return launch_transpose(full1, out, A, B, n, (int)rb, s);
}

int nfft_plan::execute(int op, const void *in, void *out, hipStream_t s, int vjp_mode)
{
    const size_t cb = cx_bytes(prec), rb = re_bytes(prec);
    hipError_t e = hipSuccess;
    if (kind == NFFT_KIND_DFT) {
        if (op != NFFT_OP_DFT && op != NFFT_OP_INVDFT) {
            set_error("operation does not match a DFT plan");
            return NFFT_E_INVALID;
        }
        int dir = op == NFFT_OP_DFT ? -1 : 1;
        bool fwd_rule = op == NFFT_OP_DFT;
        if (vjp_mode) {
            // vjp(dft) = DFT^H (unnormalised); vjp(invdft) = DFT with invdft's scaling
            if (op == NFFT_OP_DFT) { dir = 1; fwd_rule = true; }
            else { dir = -1; fwd_rule = false; }
        }
        const long long eng_elems = (long long)(work_bytes / cb);
        e = dft_nd(*this, ax, dims, d, batch, in, out, dir, fwd_rule, 1.0, work, eng_elems, s);
    } else if (kind == NFFT_KIND_REALDFT) {
        const int last = dims[d - 1];
        const long long H = last / 2 + 1, Pl = P / last;
        const long long lines = batch * Pl;
        const bool forward = (op == NFFT_OP_REALDFT) != (vjp_mode != 0);
        if (op != NFFT_OP_REALDFT && op != NFFT_OP_INVREALDFT) {
            set_error("operation does not match a realdft plan");
            return NFFT_E_INVALID;
        }
        if (in == out && batch > 1) {
            set_error("in-place realdft/invrealdft needs batch == 1");
            return NFFT_E_INVALID;
        }
        // scratch layout: [spectrum copy (batch*Pl*H)] [engine]
        const long long copy = d > 1 ? batch * Pl * H : 0;
        void *eng = (char *)work + (size_t)copy * cb;
        const long long eng_elems = (long long)(work_bytes / cb) - copy;
        // leading-axis complex dims of the half spectrum
        std::vector<int> cd(dims.begin(), dims.end());
        cd[d - 1] = (int)H;
        if (forward) {
            // op realdft, or the vjp of invrealdft (bins 0 < k < N/2 doubled)
            const bool vjp_inv = op == NFFT_OP_INVREALDFT;
            e = real.fwd(in, out, lines, last, H, 1.0, vjp_inv ? 1 : 0, eng, eng_elems, s);
            if (e == hipSuccess && d > 1) {
                for (int i = d - 2; i >= 0 && e == hipSuccess; i--) {
                    const long long A = batch * prod(dims, 0, i), B = prod(cd, i + 1, d);
                    const double sc = vjp_inv ? axis_inv_scale(norm, dims[i]) : 1.0;
                    e = c2c_axis(ax[i], out, A, B, -1, sc, eng, eng_elems, s);
                }
            }
        } else {
            // op invrealdft, or the vjp of realdft (C^H on leading axes, then
            // the half-weighted invrealdft)
            const bool vjp_real = op == NFFT_OP_REALDFT;
            const void *src = in;
            if (d > 1) {
                // the first leading-axis pass reads the input (left intact)
                // and writes the spectrum copy; the others work in place
                void *cp = work;
                for (int i = d - 2; i >= 0 && e == hipSuccess; i--) {
                    const long long A = batch * prod(dims, 0, i), B = prod(cd, i + 1, d);
                    const double sc = vjp_real ? 1.0 : axis_inv_scale(norm, dims[i]);
                    e = c2c_axis(ax[i], cp, A, B, 1, sc, eng, eng_elems, s, i == d - 2 ? in : nullptr);
                }
                src = cp;
            }
            if (e == hipSuccess) e = real.inv(src, out, lines, H, last, 1.0, vjp_real ? 1 : 0, eng, eng_elems, s);
        }
    } else {
        const bool t4 = kind == NFFT_KIND_T4;
        const bool ok_op = t4 ? (op == NFFT_OP_DCT4 || op == NFFT_OP_DST4)
                              : (op >= NFFT_OP_DCT2 && op <= NFFT_OP_DST3);
        if (!ok_op) {
            set_error("operation does not match this trigonometric plan");
            return NFFT_E_INVALID;
        }
        int run_op = op;
        double w0 = 1.0;
        if (vjp_mode && !t4) {
            // (REDFT10)^T = REDFT01 diag(2,1,..,1); (RODFT10)^T = RODFT01 diag(1,..,1,2)
            switch (op) {
            case NFFT_OP_DCT2: run_op = NFFT_OP_DCT3; w0 = 2.0; break;
            case NFFT_OP_DST2: run_op = NFFT_OP_DST3; w0 = 2.0; break;
            case NFFT_OP_DCT3: run_op = NFFT_OP_DCT2; w0 = 0.5; break;
            case NFFT_OP_DST3: run_op = NFFT_OP_DST2; w0 = 0.5; break;
            }
        }
        if (n_active == 1 && d == 1) {
            e = tax[0].run(run_op, in, out, batch, dims[0], dims[0], 1.0, w0, work,
                           (long long)(work_bytes / cb), s);
        } else {
            // axes n_active-1 .. 0; the first reads `in`, the rest work in place on `out`
            const void *src = in;
            for (int i = n_active - 1; i >= 0 && e == hipSuccess; i--) {
                e = trig_axis(*this, i, run_op, src, out, w0, work, s);
                src = out;
            }
        }
    }
    if (e != hipSuccess) {
        set_error(std::string("execute: ") + hipGetErrorString(e));
        return NFFT_E_HIP;
    }
    (void)rb;
    return NFFT_OK;
}

// ---------------------------------------------------------------------------
// Weighted spectral power L = sum |W Y|^2 and its gradient 2 DFT^H (W^2 Y).
// Where axis 0 has a single-pass column kernel (d >= 2), the forward
// transform of axis 0, the weighting and the adjoint of axis 0 are one fused
// pass (k_col_wpow) between the other axes' forward and adjoint passes;
// otherwise forward transform, elementwise weighting, adjoint (NFFT_FUSE=0
// forces the latter).
// ---------------------------------------------------------------------------
int nfft_plan::weighted_power(const void *x, const void *w, void *grad, double *loss, hipStream_t s)
{
    if (kind != NFFT_KIND_DFT) {
        set_error("weighted power needs a DFT plan");
        return NFFT_E_INVALID;
    }
    const size_t cb = cx_bytes(prec);
    const long long eng_elems = (long long)(work_bytes / cb);
    const long long S = P / dims[0];
    const bool fused = d >= 2 && env_int("NFFT_FUSE", 1) != 0 && ax[0].algo == ALG_LINE &&
                       col_wpow_supported(prec, dims[0]) && S <= 0x7fffffffLL;
    const long long per = fused ? col_wpow_tiles(prec, dims[0], (int)S) : wpow_mask_chunks(P);
    if (per < 1) {
        set_error("weighted power: no kernel for this shape");
        return NFFT_E_UNSUPPORTED;
    }
    if (batch * per > wpow_cap) {
        if (wpow_part) (void)hipFree(wpow_part);
        wpow_part = nullptr;
        wpow_cap = 0;
        if (dev_malloc(&wpow_part, (size_t)(batch * per) * sizeof(double)) != hipSuccess) {
            set_error("hipMalloc of the weighted-power partial sums failed");
            return NFFT_E_ALLOC;
        }
        wpow_cap = batch * per;
    }
    hipError_t e = hipSuccess;
    constexpr long long kGroup = 65535;  // transforms per weighting launch (grid.y)
    if (fused) {
        // forward along axes d-1..1 of the batch x N0 sub-arrays
        const std::vector<C2C> sub(ax.begin() + 1, ax.end());
        const std::vector<int> subd(dims.begin() + 1, dims.end());
        e = dft_nd(*this, sub, subd, d - 1, batch * dims[0], x, grad, -1, true, 1.0, work, eng_elems, s);
        for (long long t0 = 0; t0 < batch && e == hipSuccess; t0 += kGroup) {
            WpowArgs a{};
            a.buf = (char *)grad + (size_t)t0 * P * cb;
            a.w = w;
            a.partial = wpow_part + t0 * per;
            a.ncols = (int)S;
            a.dist = P;
            a.ntrans = std::min(kGroup, batch - t0);
            a.tw = ax[0].tw_a;
            e = launch_col_wpow(prec, dims[0], a, s);
        }
        // adjoint (unnormalised inverse) along axes d-1..1
        if (e == hipSuccess) e = dft_nd(*this, sub, subd, d - 1, batch * dims[0], grad, grad, 1, true, 1.0, work, eng_elems, s);
    } else {
        e = dft_nd(*this, ax, dims, d, batch, x, grad, -1, true, 1.0, work, eng_elems, s);
        for (long long t0 = 0; t0 < batch && e == hipSuccess; t0 += kGroup)
            e = launch_wpow_mask(prec, (char *)grad + (size_t)t0 * P * cb, w, wpow_part + t0 * per, P,
                                 std::min(kGroup, batch - t0), s);
        if (e == hipSuccess) e = dft_nd(*this, ax, dims, d, batch, grad, grad, 1, true, 1.0, work, eng_elems, s);
    }
    if (e == hipSuccess && loss) e = launch_wpow_reduce(wpow_part, loss, per, batch, s);
    if (e != hipSuccess) {
        set_error(std::string("weighted power: ") + hipGetErrorString(e));
        return NFFT_E_HIP;
    }
    return NFFT_OK;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (nfft_execute_noalloc): one
// array per GPU the library runs on, held until the process ends.  Host code
// uses only the device address obtained for the current device, never the
// array's name.
// ---------------------------------------------------------------------------
static constexpr size_t kStaticScratchBytes = 64u << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
void *g_scratch_ptr[kMaxDevices];
}  // namespace

// Device address of `scratch` on the current device (resolved once per
// device); nullptr with the error set on failure.
static void *static_scratch()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("no usable current device for the static scratch");
        return nullptr;
    }
    std::call_once(g_scratch_once[dev], [dev] {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[dev] = p;
    });
    if (!g_scratch_ptr[dev]) set_error("hipGetSymbolAddress of the static scratch failed");
    return g_scratch_ptr[dev];
}

// Scratch elements (complex) an execute of DFT plan p needs when it may chunk
// freely: one line of each engine that needs scratch, plus the full-array copy
// of a strided axis that is not transformed in place.
static long long dft_min_scratch(const nfft_plan &p)
{
    long long need = 0;
    const C2C &el = p.ax[p.d - 1];
    if (p.dims[p.d - 1] > 1) need = el.scratch_per_line;
    for (int i = 0; i + 1 < p.d; i++) {
        const C2C &e = p.ax[i];
        if (p.dims[i] == 1) continue;
        const long long B = prod(p.dims, i + 1, p.d);
        if (B == 1) {
            need = std::max(need, e.scratch_per_line);
        } else if (!(e.algo == ALG_LINE && p.dims[i] <= col_max(p.prec))) {
            need = std::max(need, p.batch * p.P + e.scratch_per_line);
        }
    }
    return need;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *nfft_last_error(void) { return nfft::last_error(); }

int nfft_plan_create(nfft_plan **plan, int kind, int prec, int d, const int *dims, long long batch, int norm)
{
    clear_error();
    if (!plan) {
        set_error("plan pointer is NULL");
        return NFFT_E_INVALID;
    }
    *plan = nullptr;
    if (kind < 0 || kind > 3 || (prec != NFFT_F32 && prec != NFFT_F64) || d < 1 || d > 16 || !dims || batch < 1 ||
        norm < 0 || norm > 2) {
        set_error("invalid plan arguments");
        return NFFT_E_INVALID;
    }
    long long p = 1;
    for (int i = 0; i < d; i++) {
        if (dims[i] <= 0) {
            set_error("dimensions must be positive");
            return NFFT_E_SIZE;
        }
        p *= dims[i];
        if (p > (1LL << 31)) {
            set_error("transform size exceeds 2^31 elements");
            return NFFT_E_UNSUPPORTED;
        }
    }
    // size rules of the reference's constructors (src/07_legacy.h:1405-1634)
    if (kind == NFFT_KIND_REALDFT && !is_pow2(dims[d - 1])) {
        set_error("realdft needs a power-of-two last dimension");
        return NFFT_E_SIZE;
    }
    if (kind == NFFT_KIND_T2T3 || kind == NFFT_KIND_T4) {
        for (int i = 0; i < d; i++)
            if (!is_pow2(dims[i])) {
                set_error("trigonometric transforms need power-of-two dimensions");
                return NFFT_E_SIZE;
            }
    }
    // the allocation-free entry point's scratch address, resolved at set-up
    // on this device (an execute under graph capture then only reads it)
    if (kind == NFFT_KIND_DFT && !static_scratch()) return NFFT_E_HIP;
    nfft_plan *pl = new (std::nothrow) nfft_plan();
    if (!pl) {
        set_error("out of host memory");
        return NFFT_E_ALLOC;
    }
    const long long fails0 = alloc_failures();
    if (!pl->init(kind, prec, d, dims, batch, norm)) {
        const std::string msg = last_error();
        pl->release();
        delete pl;
        set_error(msg);
        if (msg.find("limit") != std::string::npos) return NFFT_E_UNSUPPORTED;
        if (alloc_failures() > fails0 || msg.find("hipMalloc") != std::string::npos) return NFFT_E_ALLOC;
        return NFFT_E_HIP;
    }
    *plan = pl;
    return NFFT_OK;
}

int nfft_plan_destroy(nfft_plan *plan)
{
    if (!plan) return NFFT_OK;
    plan->release();
    delete plan;
    return NFFT_OK;
}

int nfft_execute(nfft_plan *plan, int op, const void *in, void *out, void *stream)
{
    clear_error();
    if (!plan || !in || !out) {
        set_error("NULL plan or buffer");
        return NFFT_E_INVALID;
    }
    return plan->execute(op, in, out, (hipStream_t)stream, 0);
}

int nfft_execute_noalloc(nfft_plan *plan, int op, const void *in, void *out, void *stream)
{
    clear_error();
    if (!plan || !in || !out) {
        set_error("NULL plan or buffer");
        return NFFT_E_INVALID;
    }
    if (plan->kind != NFFT_KIND_DFT || (op != NFFT_OP_DFT && op != NFFT_OP_INVDFT)) {
        set_error("nfft_execute_noalloc runs dft / invdft of DFT plans only");
        return NFFT_E_INVALID;
    }
    const size_t cb = cx_bytes(plan->prec);
    const long long need = dft_min_scratch(*plan);
    if ((size_t)need * cb > kStaticScratchBytes) {
        set_error("this plan needs " + std::to_string((size_t)need * cb) +
                  " bytes of scratch; the allocation-free entry point has " + std::to_string(kStaticScratchBytes));
        return NFFT_E_SIZE;
    }
    void *scr = static_scratch();
    if (!scr) return NFFT_E_HIP;
    struct Guard {
        Guard() { t_noalloc = 1; }
        ~Guard() { t_noalloc = 0; }
    } guard;
    const int dir = op == NFFT_OP_DFT ? -1 : 1;
    const hipError_t e = dft_nd(*plan, plan->ax, plan->dims, plan->d, plan->batch, in, out, dir, op == NFFT_OP_DFT,
                                1.0, scr, (long long)(kStaticScratchBytes / cb), (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error(std::string("execute_noalloc: ") + hipGetErrorString(e));
        return NFFT_E_HIP;
    }
    return NFFT_OK;
}

int nfft_dft_weighted_power(nfft_plan *plan, const void *x, const void *w, void *grad, double *loss, void *stream)
{
    clear_error();
    if (!plan || !x || !w || !grad) {
        set_error("NULL plan or buffer");
        return NFFT_E_INVALID;
    }
    return plan->weighted_power(x, w, grad, loss, (hipStream_t)stream);
}

int nfft_vjp(nfft_plan *plan, int op, const void *grad_out, void *grad_in, void *stream)
{
    clear_error();
    if (!plan || !grad_out || !grad_in) {
        set_error("NULL plan or buffer");
        return NFFT_E_INVALID;
    }
    return plan->execute(op, grad_out, grad_in, (hipStream_t)stream, 1);
}

long long nfft_plan_workspace_bytes(const nfft_plan *plan) { return plan ? (long long)plan->work_bytes : 0; }

const char *nfft_plan_describe(const nfft_plan *plan) { return plan ? plan->desc.c_str() : ""; }

int nfft_plan_check(nfft_plan *plan, void *stream)
{
    clear_error();
    if (!plan) {
        set_error("plan is NULL");
        return NFFT_E_INVALID;
    }
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
        set_error("hipStreamSynchronize failed");
        return NFFT_E_HIP;
    }
    bool bad = plan->real.inner.device_fault();
    for (const auto &e : plan->ax) bad = bad || e.device_fault();
    for (const auto &t : plan->tax) bad = bad || t.cplx.device_fault() || t.real.inner.device_fault();
    if (bad) {
        set_error("persistent four-step: a dependency wait timed out");
        return NFFT_E_HIP;
    }
    return NFFT_OK;
}

int nfft_prof_enable(int on, int capacity)
{
    clear_error();
    std::lock_guard<std::mutex> lk(g_prof.mu);
    if (on) {
        if (capacity < 1) capacity = 4096;
        while ((int)g_prof.ev.size() < 2 * capacity) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) {
                set_error("hipEventCreate failed");
                return NFFT_E_HIP;
            }
            g_prof.ev.push_back(e);
        }
        g_prof.used = 0;
        g_prof.overflow = false;
        g_prof.recs.clear();
    }
    g_prof.on = on != 0;
    return NFFT_OK;
}

int nfft_prof_count(void)
{
    std::lock_guard<std::mutex> lk(g_prof.mu);
    std::map<std::string, int> fam;
    for (auto &r : g_prof.recs) fam[r.family + "<" + (r.prec ? "f64," : "f32,") + std::to_string(r.L) + ">"] = 1;
    return (int)fam.size() + (g_prof.overflow ? 1000000 : 0);
}

int nfft_prof_get(int idx, char *name, int name_len, long long *launches, double *total_ms, double *bytes)
{
    std::lock_guard<std::mutex> lk(g_prof.mu);
    struct Agg {
        long long n = 0;
        double ms = 0, bytes = 0;
    };
    std::map<std::string, Agg> fam;
    for (auto &r : g_prof.recs) {
        float ms = 0;
        if (hipEventSynchronize(g_prof.ev[2 * r.slot + 1]) != hipSuccess ||
            hipEventElapsedTime(&ms, g_prof.ev[2 * r.slot], g_prof.ev[2 * r.slot + 1]) != hipSuccess) {
            set_error("event timing failed");
            return NFFT_E_HIP;
        }
        Agg &a = fam[r.family + "<" + (r.prec ? "f64," : "f32,") + std::to_string(r.L) + ">"];
        a.n++;
        a.ms += ms;
        a.bytes += r.bytes;
    }
    int i = 0;
    for (auto &kv : fam) {
        if (i++ != idx) continue;
        if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", kv.first.c_str());
        if (launches) *launches = kv.second.n;
        if (total_ms) *total_ms = kv.second.ms;
        if (bytes) *bytes = kv.second.bytes;
        return NFFT_OK;
    }
    set_error("no such profile entry");
    return NFFT_E_INVALID;
}

int nfft_fill_uniform(void *out, long long count, unsigned long long seed, unsigned long long start, int prec,
                      void *stream)
{
    clear_error();
    if (!out || count < 0 || (prec != NFFT_F32 && prec != NFFT_F64)) {
        set_error("invalid fill arguments");
        return NFFT_E_INVALID;
    }
    hipError_t e = launch_fill_uniform(out, count, seed, start, prec, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error(hipGetErrorString(e));
        return NFFT_E_HIP;
    }
    return NFFT_OK;
}

int nfft_fill_gaussian(void *out, long long count, unsigned long long seed, unsigned long long start, int prec,
                       void *stream)
{
    clear_error();
    if (!out || count < 0 || (prec != NFFT_F32 && prec != NFFT_F64)) {
        set_error("invalid fill arguments");
        return NFFT_E_INVALID;
    }
    hipError_t e = launch_fill_gaussian(out, count, seed, start, prec, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error(hipGetErrorString(e));
        return NFFT_E_HIP;
    }
    return NFFT_OK;
}

}  // extern "C"
