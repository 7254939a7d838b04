// d2bi_tune.cuh -- the autotuner: launch choices that cannot change a bit of
// any result, timed on the device the first time a shape is seen.
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
//
// Tuned choices (each leaves every image bit-identical, because no thread's
// arithmetic or summation order depends on it):
//   TUNE_INT_BLOCK   threads per workgroup of the FAST integrators (64, 128
//                    or 256): a lane integrates one pixel column on its own
//   TUNE_TILE_ORDER  order of the FAST voxel kernel's tiles over the
//                    workgroups (0: contiguous ranges per XCD, 1: linear, 2:
//                    rows dealt to the XCDs in turn): a wave sums one tile on
//                    its own
//   TUNE_BRANCHES    concurrent branches of a batch's graph (1..4): a
//                    condition's kernels and inputs do not depend on its branch
// A candidate replaces the default only when it runs 3 % faster, so near
// ties keep the default.  Results are kept per shape for the process and,
// with D2BI_TUNE_CACHE=<file>,
// appended to that file and read back by later processes.  D2BI_TUNE=0 turns
// tuning off (the defaults: 256 threads, rows dealt to the XCDs, four
// branches where they fit); D2BI_INT_BLOCK, D2BI_TILE_ORDER and D2BI_BATCH_BRANCHES force a
// choice.  Nothing is timed while a graph is being captured, or on a
// workspace that must not allocate: those launches use cached or default
// choices.
#pragma once

namespace d2bi {

enum { TUNE_INT_BLOCK = 1, TUNE_TILE_ORDER = 2, TUNE_BRANCHES = 3 };

typedef std::array<int64_t, 9> TuneKey;        // kind, then up to 8 shape numbers

static std::map<TuneKey, int> g_tune;
static bool g_tune_loaded = false;
static bool g_capturing = false;                // a batch graph is being captured

static bool tuning_on()
{
    const char* e = std::getenv("D2BI_TUNE");
    return !(e && e[0] == '0' && e[1] == 0);
}

static int forced_int(const char* name, int lo, int hi)
{
    const char* e = std::getenv(name);
    if (!e || !*e) return -1;
    char* end = nullptr;
    const long v = std::strtol(e, &end, 10);
    if (*end || v < lo || v > hi) return -1;
    return (int)v;
}

static void tune_load()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char* path = std::getenv("D2BI_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    TuneKey k;
    long long v[10];
    while (std::fscanf(f, "%lld %lld %lld %lld %lld %lld %lld %lld %lld %lld", &v[0], &v[1], &v[2], &v[3], &v[4],
                       &v[5], &v[6], &v[7], &v[8], &v[9]) == 10) {
        for (int i = 0; i < 9; i++) k[i] = v[i];
        g_tune[k] = (int)v[9];
    }
    std::fclose(f);
}

// cached choice for a key, or -1
static int tune_get(const TuneKey& k)
{
    tune_load();
    auto it = g_tune.find(k);
    return it == g_tune.end() ? -1 : it->second;
}

static void tune_put(const TuneKey& k, int value)
{
    g_tune[k] = value;
    const char* path = std::getenv("D2BI_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    for (int i = 0; i < 9; i++) std::fprintf(f, "%lld ", (long long)k[i]);
    std::fprintf(f, "%d\n", value);
    std::fclose(f);
}

// Device time (ms) of launch() on stream s: the last of reps runs (the
// first of two warms the caches).
template <class Launch>
static float time_launch(Launch launch, hipStream_t s, int reps = 2)
{
    float ms = -1.0f;
    for (int rep = 0; rep < reps; rep++) {
        if (hipEventRecord(g.ev[7], s) != hipSuccess) return -1.0f;
        launch();
        if (hipEventRecord(g.ev[8], s) != hipSuccess || hipEventSynchronize(g.ev[8]) != hipSuccess) return -1.0f;
        (void)hipEventElapsedTime(&ms, g.ev[7], g.ev[8]);
    }
    return ms;
}

// Choose among candidates by timing: returns the cached or forced choice,
// times the candidates when tuning is allowed (fallback dflt otherwise).
template <class Run>
static int tune_choose(const TuneKey& k, const char* force_env, const int* cand, int ncand, int dflt, bool may_time,
                       hipStream_t s, Run run)
{
    int lo = cand[0], hi = cand[0];
    for (int i = 1; i < ncand; i++) {
        lo = std::min(lo, cand[i]);
        hi = std::max(hi, cand[i]);
    }
    const int f = forced_int(force_env, lo, hi);
    for (int i = 0; i < ncand; i++)
        if (f == cand[i]) return f;                 // a forced value that is a candidate
    if (!tuning_on()) return dflt;
    const int c = tune_get(k);
    if (c >= 0) return c;
    if (!may_time || g_capturing) return dflt;
    // another candidate replaces the default only when it is 3 % faster, so
    // that near ties (within the run-to-run spread) keep the default
    float t[8];
    float tdef = 3.4e38f;
    for (int i = 0; i < ncand && i < 8; i++) {
        t[i] = time_launch([&] { run(cand[i]); }, s);
        if (t[i] < 0.0f) {
            (void)hipGetLastError();
            return dflt;
        }
        if (cand[i] == dflt) tdef = t[i];
    }
    int best = dflt;
    float tbest = 0.97f * tdef;
    for (int i = 0; i < ncand && i < 8; i++)
        if (cand[i] != dflt && t[i] < tbest) {
            tbest = t[i];
            best = cand[i];
        }
    tune_put(k, best);
    return best;
}

}  // namespace d2bi
