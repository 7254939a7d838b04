// d2bi_run.cuh -- one imaging condition: planning, binning, voxelisation (with the
// FAST variant choice), integration, the aperture sum and the field API.
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
#pragma once

namespace d2bi {

// ------------------------------------------------------------------ orchestration

struct Timer {
    hipEvent_t* ev;
    hipStream_t s;
    int record(int i)
    {
        HIPCHK(hipEventRecord(ev[i], s));
        return D2BI_OK;
    }
    double ms(int a, int b)
    {
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, ev[a], ev[b]);
        return t;
    }
};

static void fill_vox_params(const d2bi_record* r, const d2bi_plan* p, const AxisDev& ax, VoxParams* V)
{
    V->Mx = p->Mx; V->My = p->My; V->Mz = p->Mz;
    V->Ex = p->Ex; V->Ey = p->Ey; V->Ez = p->Ez;
    V->a[0] = p->a_col[0]; V->a[1] = p->a_col[4]; V->a[2] = p->a_col[8];
    V->rc2 = p->rc2; V->i2s2 = p->i2s2; V->sigma = p->sigma;
    V->use_density = r->use_density ? 1 : 0;
    V->ax = ax;
}

// Tile depth of a FAST kernel variant: 0 (a lane per voxel) and 1 (rows of
// 4 x 4 x 4 tiles) use 4 voxels, 2 (rows of 4 x 4 x 8 tiles) 8.
static int variant_tz(int variant) { return variant == 2 ? 8 : 4; }
// Voxels along x of a variant's tiles: 8 for variant 3 (rows of 8 x 4 x 4 tiles).
static int variant_tx(int variant) { return variant == 3 ? 8 : 4; }

// The order of the FAST voxel kernel's tiles over its workgroups and the
// grid it needs (8 per_xcd workgroups): orders 0 and 1 launch one workgroup
// per tile (rounded up to a multiple of 8); order 2 deals whole rows of ntx
// tiles, so each XCD gets ceil(rows / 8) rows.
static void set_tile_order(FastParams* F, int order)
{
    F->order = order;
    if (order == 2) {
        const int rows = (F->ntiles + F->ntx - 1) / F->ntx;
        F->per_xcd = ((rows + 7) / 8) * F->ntx;
    } else {
        F->per_xcd = (F->ntiles + 7) / 8;
    }
}

// Link cells in the largest unwrapped cell range of a tx x 4 x tz tile.
static int tile_cells(const d2bi_plan* p, const AxisTables& tab, int tz, int tx = TILE)
{
    int worst = 1;
    for (int a = 0; a < 3; a++) {
        const int M = a == 0 ? p->Mx : a == 1 ? p->My : p->Mz;
        const int ts = a == 2 ? tz : a == 0 ? tx : TILE;
        int mx = 0;
        for (int t0 = 0; t0 < M; t0 += ts)
            mx = std::max(mx, tab.cell[a][std::min(M, t0 + ts) - 1] - tab.cell[a][t0] + 3);
        worst *= mx;
    }
    return worst;
}

// Launch parameters of the FAST voxel kernel for a plan, tiles of tx x 4 x tz.
static int fast_params(const d2bi_plan* p, const AxisTables& tab, const VoxParams& V, FastParams* Fo, int tz,
                       int tx = TILE)
{
    FastParams& F = *Fo;
    F.v = V;
    F.tz = tz;
    F.ntx = (p->Mx + tx - 1) / tx;
    F.nty = (p->My + TILE - 1) / TILE;
    F.ntz = (p->Mz + tz - 1) / tz;
    F.ntiles = F.ntx * F.nty * F.ntz;
    set_tile_order(&F, 2);
    F.app_cs = nullptr;
    F.app_stride = 0;
    F.field_stride = 0;
    // guard band: the FP32 exponent carries <= ~2e-6 relative error (inputs
    // rounded once from FP64, three FMAs); +-1e-5 relative covers it 5x
    const double k2 = -p->i2s2 * 1.4426950408889634;
    F.k2 = (float)k2;
    F.e_rc = (float)(k2 * p->rc2);
    F.e_band = (float)(std::fabs(k2 * p->rc2) * 1.0e-5);
    F.w_scale = (float)std::exp2((double)F.e_rc);
    F.m2_min = (float)(1.0e-32 / ((double)F.w_scale * (double)F.w_scale));
    F.inv_k2 = (float)(1.0 / k2);
    F.rc_cull2 = (float)(p->rc2 * (1.0 + 1.0e-4));
    F.cellf[0] = (float)p->a_col[0]; F.cellf[1] = (float)p->a_col[4]; F.cellf[2] = (float)p->a_col[8];
    // cells thinner than 3 sigma by more than rounding (the guard band, ~2e-5
    // A, covers less): only z can be, x and y are wide by construction
    const double rc = std::sqrt(p->rc2), thin = rc * (1.0 - 1.0e-7);
    if (p->a_col[0] < thin || p->a_col[4] < thin)
        return seterr(D2BI_ERR_ARG, "link cells %.6f x %.6f narrower than 3 sigma = %.6f", p->a_col[0],
                      p->a_col[4], rc);
    F.thin_z = p->a_col[8] < thin ? 1 : 0;
    // the unwrapped cell range of a tile must fit the per-tile cell table
    const int worst = tile_cells(p, tab, tz, tx);
    if (worst > TILE_MAXC) return seterr(D2BI_ERR_CAPACITY, "voxel tile spans %d link cells (max %d)", worst, TILE_MAXC);
    return D2BI_OK;
}

// Launch parameters of a variant.  A 4 x 4 x 8 tile that spans more link
// cells than the per-tile table holds falls back to the 4 x 4 x 4 row kernel
// (*variant becomes 1), and so does an 8 x 4 x 4 tile (variant 3) that
// spans too many.
static int effective_variant(const d2bi_plan* p, const AxisTables& tab, int variant)
{
    if (variant == 3 && tile_cells(p, tab, 4, 8) > TILE_MAXC) variant = 1;
    if (variant == 2 && tile_cells(p, tab, 8) > TILE_MAXC) variant = 1;
    return variant;
}

static int variant_params(const d2bi_plan* p, const AxisTables& tab, const VoxParams& V, int* variant,
                          FastParams* F)
{
    *variant = effective_variant(p, tab, *variant);
    return fast_params(p, tab, V, F, variant_tz(*variant), variant_tx(*variant));
}

// Kernel instance for (diagnostics, density, variant); link cells thinner
// than 3 sigma along z take a row kernel (variant 1, 2 or 3) with the
// z-offset test.
template <bool DENS, int ROW, class Go>
static void pick_row(int tz, int tx, Go go)
{
    if (tx == 8) go(k_voxel_tile<DENS, false, ROW, 4, 8>);
    else if (tz == 8) go(k_voxel_tile<DENS, false, ROW, 8>);
    else go(k_voxel_tile<DENS, false, ROW, 4>);
}

template <class Go>
static void pick_tile(bool diag, bool dens, int variant, bool thin_z, Go go)
{
    if (diag) { go(k_voxel_tile<true, true, 0, 4>); return; }
    const int tz = variant_tz(variant), tx = variant_tx(variant);
    if (!thin_z && variant == 0) {
        if (dens) go(k_voxel_tile<true, false, 0, 4>);
        else go(k_voxel_tile<false, false, 0, 4>);
    } else if (thin_z) {
        if (dens) pick_row<true, 2>(tz, tx, go);
        else pick_row<false, 2>(tz, tx, go);
    } else {
        if (dens) pick_row<true, 1>(tz, tx, go);
        else pick_row<false, 1>(tz, tx, go);
    }
}

template <bool DIAG>
static void launch_tile(const FastParams& F, int variant, bool dens, const int* cell_start,
                        const double4* slot_frac, const float4* p32, const float* sn32, float4* field, int* cnt,
                        float* r2m, unsigned long long* gc, hipStream_t s)
{
    pick_tile(DIAG, dens, variant, F.thin_z != 0, [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(8 * F.per_xcd), dim3(64), 0, s, F, cell_start, slot_frac, p32, sn32,
                           field, cnt, r2m, gc);
    });
}

// Choice of the FAST voxel kernel variant (0: a lane per voxel, 1: a lane
// per row of four voxels and quarter of the candidates of a 4 x 4 x 4 tile,
// 2: a lane per row and half of the candidates of a 4 x 4 x 8 tile).  The
// variants sum the same atoms in different orders, so their fields agree to
// FP32 rounding, not bit for bit; the choice is therefore fixed, variant 2
// (the fastest on every measured workload; variant_params falls back to 1 for
// a tile whose cell range does not fit), so that a condition images the same
// in every process.  D2BI_FAST_VARIANT=0|1|2 forces a variant;
// D2BI_FAST_TUNE=1 instead times all three on the first condition of each
// voxel-grid shape and keeps the fastest (results then depend on the timing).
static std::map<std::array<int, 7>, int> g_variant_cache;

static bool tune_variants()
{
    const char* e = std::getenv("D2BI_FAST_TUNE");
    return e && e[0] == '1' && e[1] == 0;
}

static int forced_variant()
{
    const char* env = std::getenv("D2BI_FAST_VARIANT");
    if (env && env[0] >= '0' && env[0] <= '3' && env[1] == 0) return env[0] - '0';
    return -1;
}

// the variant for a shape: forced, fixed, or tuned (-1 if not tuned yet)
static int known_variant(const d2bi_plan* p, bool dens)
{
    const int f = forced_variant();
    if (f >= 0) return f;
    if (!tune_variants()) return 2;
    const std::array<int, 7> key{p->Mx, p->My, p->Mz, p->Ex, p->Ey, p->Ez, dens ? 1 : 0};
    auto it = g_variant_cache.find(key);
    return it == g_variant_cache.end() ? -1 : it->second;
}

template <class TimeFn>
static int fast_variant(const d2bi_plan* p, bool dens, bool diag, bool thin_z, TimeFn time_variant)
{
    auto& cache = g_variant_cache;
    const int f = forced_variant();
    if (f >= 0) return f;
    if (diag) return 0;
    if (!tune_variants()) return 2;
    const std::array<int, 7> key{p->Mx, p->My, p->Mz, p->Ex, p->Ey, p->Ez, dens ? 1 : 0};
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    // each variant twice, the second (warm-cache) timing counts; thin z
    // cells have only the row kernels
    const int v0 = thin_z ? 1 : 0;
    float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int rep = 0; rep < 2; rep++)
        for (int v = v0; v < 4; v++) {
            t[v] = time_variant(v);
            if (t[v] < 0) return seterr(D2BI_ERR_HIP, "variant timing failed");
        }
    int best = v0;
    for (int v = v0 + 1; v < 4; v++)
        if (t[v] < t[best]) best = v;
    cache[key] = best;
    return best;
}

// Tile orders the tuner may choose; 2 (rows dealt to the XCDs) is the
// default, the fastest of the three on c1, c2 and c3 (up to 12 % over per-XCD
// ranges, whose z slabs leave the XCDs given vacuum above and below a padded
// foil idle early).
static const int TILE_ORDERS[3] = {0, 1, 2};

// Tuner key of the FAST voxel kernel's tile order: the shape and the variant
// that runs (after the fallback of variant_params).
static TuneKey tile_order_key(const d2bi_plan* p, const AxisTables& tab, bool dens, int variant)
{
    const int v = effective_variant(p, tab, variant);
    return TuneKey{TUNE_TILE_ORDER, p->Mx, p->My, p->Mz, p->Ex, p->Ey, p->Ez, dens ? 1 : 0, v};
}

// Voxelise into the mode's field buffers of workspace w; optional count /
// r2min outputs.
static int run_voxel(const d2bi_record* r, const d2bi_plan* p, const AxisTables& tab, const AxisDev& ax,
                     int mode, bool want_diag, int64_t* n_guard, hipStream_t s, Workspace& w, Timer* T = nullptr)
{
    const int64_t nvox = (int64_t)p->Mx * p->My * p->Mz;
    VoxParams V;
    fill_vox_params(r, p, ax, &V);
    int* dcnt = nullptr;
    void* dr2 = nullptr;
    if (want_diag) {
        CHK(w.cnt.ensure(sizeof(int) * nvox));
        CHK(w.r2m.ensure(sizeof(double) * nvox));
        dcnt = w.cnt.as<int>();
        dr2 = w.r2m.p;
    }
    if (mode == D2BI_MODE_FP64) {
        CHK(w.f64_re.ensure(sizeof(double) * nvox));
        CHK(w.f64_im.ensure(sizeof(double) * nvox));
        CHK(w.f64_rho.ensure(sizeof(double) * nvox));
        dim3 grid((p->Mx + 7) / 8, (p->My + 7) / 8, (p->Mz + 3) / 4);
        if (T) CHK(T->record(2));
        hipLaunchKernelGGL(k_voxel_fp64<1>, grid, dim3(256), 0, s, V, w.cell_start.as<int>(),
                           w.slot_frac.as<double4>(), w.slot_cs64.as<double2>(), (int64_t)0, w.f64_re.as<double>(),
                           w.f64_im.as<double>(), (int64_t)0, w.f64_rho.as<double>(), dcnt, (double*)dr2);
        HIPCHK(hipGetLastError());
        if (T) CHK(T->record(3));
        return D2BI_OK;
    }
    CHK(w.f32_field.ensure(sizeof(float4) * nvox));
    FastParams F;
    CHK(fast_params(p, tab, V, &F, 4));
    CHK(w.counters.ensure(64));
    unsigned long long* gc = (unsigned long long*)(w.counters.as<char>() + 8);
    const size_t gp_bytes = sizeof(unsigned long long) * GUARD_SLOTS * GUARD_STRIDE;
    CHK(w.guard_part.ensure(gp_bytes));
    unsigned long long* gp = w.guard_part.as<unsigned long long>();
    const bool dens = r->use_density != 0;
    // launch parameters follow the variant's tile depth; order: the tiles'
    // order over the workgroups (0 per XCD, 1 linear)
    auto run_variant = [&](int v, int order = 2) -> int {
        FastParams Fv;
        CHK(variant_params(p, tab, V, &v, &Fv));
        set_tile_order(&Fv, order);
        pick_tile(want_diag, dens, v, Fv.thin_z != 0, [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(8 * Fv.per_xcd), dim3(64), 0, s, Fv, w.cell_start.as<int>(),
                               w.slot_frac.as<double4>(), w.slot_p32.as<float4>(),
                               w.slot_sn32.as<float>(), w.f32_field.as<float4>(), dcnt, (float*)dr2, gp);
        });
        return D2BI_OK;
    };
    const int variant = fast_variant(p, dens, want_diag, F.thin_z != 0, [&](int v) -> float {
        // a variant whose tiles span too many link cells is not timed
        if (tile_cells(p, tab, variant_tz(v), variant_tx(v)) > TILE_MAXC) return 3.4e38f;
        HIPCHK(hipEventRecord(g.ev[7], s));
        CHK(run_variant(v));
        HIPCHK(hipEventRecord(g.ev[8], s));
        HIPCHK(hipEventSynchronize(g.ev[8]));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, g.ev[7], g.ev[8]);
        return ms;
    });
    if (variant < 0) return variant;
    // the tile order: timed the first time a shape is seen (each launch
    // overwrites the whole field, so the timing runs leave no trace)
    int order = 2;
    if (!want_diag) {
        order = tune_choose(tile_order_key(p, tab, dens, variant), "D2BI_TILE_ORDER", TILE_ORDERS, 3, 2, true, s,
                            [&](int o) { (void)run_variant(variant, o); });
    }
    fill_i32((int*)gp, (int64_t)(gp_bytes / sizeof(int)), 0, s);   // tuning launches counted too
    if (T) CHK(T->record(2));
    CHK(run_variant(variant, order));
    HIPCHK(hipGetLastError());
    if (T) CHK(T->record(3));
    hipLaunchKernelGGL(k_sum_guard, dim3(1), dim3(256), 0, s, gp, GUARD_SLOTS, GUARD_STRIDE, gc);
    if (n_guard) {
        unsigned long long v = 0;
        HIPCHK(hipMemcpyAsync(&v, gc, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *n_guard = (int64_t)v;
    }
    return D2BI_OK;
}

// FAST integration of 8 or more precession points per pass through the
// shared step matrix (k_integrate_poly); D2BI_FAST_INTEGRATOR=stage keeps the
// stage-by-stage march for every pass.
static bool fast_integrator_poly()
{
    const char* e = std::getenv("D2BI_FAST_INTEGRATOR");
    return !(e && std::strcmp(e, "stage") == 0);
}

// Points of the next FAST integration pass when `left` remain.
static int pass_points(int left)
{
    return left >= 32 ? 32 : left >= 16 ? 16 : left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
}

// Tuner key of the FAST integrators' workgroup size for a pass of np points.
static TuneKey int_block_key(const d2bi_plan* p, int np, bool dens)
{
    return TuneKey{TUNE_INT_BLOCK, p->Mx, p->My, p->Mz, np, dens ? 1 : 0, fast_integrator_poly() ? 1 : 0, 0, 0};
}

// Whether every tuned launch choice of a FAST condition is forced, already
// timed, or not tuned at all.
static bool condition_tuned(const d2bi_plan* p, const AxisTables& tab, bool dens, int variant)
{
    if (!tuning_on()) return true;
    if (forced_int("D2BI_TILE_ORDER", 0, 2) < 0 && tune_get(tile_order_key(p, tab, dens, variant)) < 0) return false;
    if (forced_int("D2BI_INT_BLOCK", 64, 256) >= 0) return true;
    for (int p0 = 0; p0 < p->nPrec;) {
        const int np = pass_points(p->nPrec - p0);
        if (tune_get(int_block_key(p, np, dens)) < 0) return false;
        p0 += np;
    }
    return true;
}

// One pass of the FAST integrators over NP points, workgroups of `block`
// threads (a lane per pixel column; any block size gives the same image).
template <int NP>
static void launch_int_fast(const FastIntParams& P, bool dens, const void* field, double* img, int acc,
                            double inv, int block, hipStream_t s)
{
    const int64_t plane = (int64_t)P.Mx * P.My;
    const dim3 grid((unsigned)((plane + block - 1) / block)), thr(block);
    // from 8 points on, the shared step matrix costs less than it saves
    if constexpr (NP >= 8) {
        if (fast_integrator_poly()) {
            if (dens) hipLaunchKernelGGL((k_integrate_poly<NP, true>), grid, thr, 0, s, P, field, img, acc, inv);
            else hipLaunchKernelGGL((k_integrate_poly<NP, false>), grid, thr, 0, s, P, field, img, acc, inv);
            return;
        }
    }
    if (dens) hipLaunchKernelGGL((k_integrate_fast<NP, true>), grid, thr, 0, s, P, field, img, acc, inv);
    else hipLaunchKernelGGL((k_integrate_fast<NP, false>), grid, thr, 0, s, P, field, img, acc, inv);
}

static void launch_int_pass(int np, const FastIntParams& P, bool dens, const void* f, double* img, int acc,
                            double inv, int block, hipStream_t s)
{
    switch (np) {
        case 32: launch_int_fast<32>(P, dens, f, img, acc, inv, block, s); break;
        case 16: launch_int_fast<16>(P, dens, f, img, acc, inv, block, s); break;
        case 8: launch_int_fast<8>(P, dens, f, img, acc, inv, block, s); break;
        case 4: launch_int_fast<4>(P, dens, f, img, acc, inv, block, s); break;
        case 2: launch_int_fast<2>(P, dens, f, img, acc, inv, block, s); break;
        default: launch_int_fast<1>(P, dens, f, img, acc, inv, block, s); break;
    }
}

// FAST integration of all precession points, in passes of up to 32.
// dens: the field holds float4 {x, rho} (density field on), else float2 x.
// acc_in: add to d_img from the first pass on; inv: weight of one
// integration (0: 1/nPrec).  tune_ws: a workspace whose tune_img buffer may
// grow, for timing the workgroup sizes of a pass that adds to d_img (null:
// such a pass is not timed).
static int launch_integrate_fast(const d2bi_plan* p, bool dens, const void* f, double* d_img, hipStream_t s,
                                 bool acc_in = false, double inv = 0.0, Workspace* tune_ws = nullptr)
{
    if (inv == 0.0) inv = 1.0 / p->nPrec;
    for (int p0 = 0, first = acc_in ? 0 : 1; p0 < p->nPrec; first = 0) {
        const int np = pass_points(p->nPrec - p0);
        FastIntParams P;
        P.Mx = p->Mx; P.My = p->My; P.Mz = p->Mz; P.np = np;
        P.b0 = (float)p->b0; P.bg = (float)p->bg; P.dz = (float)p->dz;
        for (int i = 0; i < 32; i++) {
            P.aa[i] = i < np ? (float)p->aa[p0 + i] : 0.0f;
            P.al[i] = i < np ? (float)(2.0 * p->dz * p->aa[p0 + i]) : 0.0f;
        }
        // workgroup size: timed on a launch that writes (not adds to) an
        // image, the output itself for the first pass, else a tuning buffer
        double* timg = first ? d_img : nullptr;
        const TuneKey key = int_block_key(p, np, dens);
        if (!timg && tune_ws && !tune_ws->tune_img.arena && !g_capturing && tuning_on() && tune_get(key) < 0) {
            CHK(tune_ws->tune_img.ensure(sizeof(double) * (int64_t)P.Mx * P.My));
            timg = tune_ws->tune_img.as<double>();
        }
        static const int cand[3] = {64, 128, 256};
        const int block = tune_choose(key, "D2BI_INT_BLOCK", cand, 3, 256, timg != nullptr, s, [&](int b) {
            launch_int_pass(np, P, dens, f, timg, 0, inv, b, s);
        });
        launch_int_pass(np, P, dens, f, d_img, !first, inv, block, s);
        HIPCHK(hipGetLastError());
        p0 += np;
    }
    return D2BI_OK;
}

// accumulate: add each column's sum over precession points to d_img;
// otherwise write sum / nPrec (div = 0) or sum / div.
// The FP64 field is read from workspace w (the library's by default).
static int launch_integrate_fp64(const d2bi_plan* p, const double* d_aa, double* d_img, hipStream_t s,
                                 int accumulate = 0, double div = 0.0, Workspace* wp = nullptr,
                                 int64_t field_off = 0)
{
    Workspace& w = wp ? *wp : g.ws;
    const int64_t plane = (int64_t)p->Mx * p->My;
    IntParams P{p->Mx, p->My, p->Mz, p->nPrec, p->b0, p->bg, p->dz, d_aa, accumulate,
                div == 0.0 ? (double)p->nPrec : div};
    hipLaunchKernelGGL(k_integrate_fp64, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, P,
                       w.f64_re.as<double>() + field_off, w.f64_im.as<double>() + field_off,
                       w.f64_rho.as<double>(), d_img);
    HIPCHK(hipGetLastError());
    return D2BI_OK;
}

// Integrate the field of workspace w into d_img (Mx*My doubles).
static int run_integrate(const d2bi_plan* p, bool dens, int mode, hipStream_t s, Workspace& w, double* d_img,
                         Timer* T = nullptr)
{
    if (mode == D2BI_MODE_FP64) {
        CHK(w.aa.ensure(sizeof(double) * D2BI_MAX_PREC));
        double* d_aa = w.aa.as<double>();
        HIPCHK(hipMemcpyAsync(d_aa, p->aa, sizeof(double) * p->nPrec, hipMemcpyHostToDevice, s));
        if (T) CHK(T->record(4));
        CHK(launch_integrate_fp64(p, d_aa, d_img, s, 0, 0.0, &w));
        if (T) CHK(T->record(5));
        // the aa copy reads p->aa (caller memory): finish before returning
        HIPCHK(hipStreamSynchronize(s));
        return D2BI_OK;
    }
    if (T) CHK(T->record(4));
    CHK(launch_integrate_fast(p, dens, w.f32_field.p, d_img, s, false, 0.0, &w));
    if (T) CHK(T->record(5));
    return D2BI_OK;
}

// Whole condition: bin -> voxelise -> integrate on workspace w, the image
// written to d_img (device memory of Mx*My doubles) or, when d_img is null,
// left in w.img.
static int image_condition(const d2bi_record* r, const double* d_x, int64_t n, double precAngle, int nPrec,
                           int mode, d2bi_stats* st, d2bi_plan* plan_out, Workspace& w, double* d_img = nullptr)
{
    if (mode != D2BI_MODE_FP64 && mode != D2BI_MODE_FAST) return seterr(D2BI_ERR_ARG, "unknown mode %d", mode);
    d2bi_plan plan;
    CHK(plan_image(r, n, precAngle, nPrec, &plan));
    AxisTables tab;
    CHK(axis_tables(r, &plan, &tab));
    BinParams BP;
    CHK(bin_params(r, &plan, &BP));
    BP.fast = mode == D2BI_MODE_FAST ? 1 : 0;
    if (!d_img) {
        CHK(w.img.ensure(sizeof(double) * (int64_t)plan.Mx * plan.My));
        d_img = w.img.as<double>();
    }
    hipStream_t s = g.stream;
    Timer T{g.ev, s};
    CHK(T.record(0));
    AxisDev ax;
    CHK(upload_axis(tab, &ax, s, w));
    Binned b;
    CHK(run_binning(d_x, n, BP, &b, s, w));
    CHK(T.record(1));
    CHK(run_voxel(r, &plan, tab, ax, mode, false, nullptr, s, w, &T));
    CHK(run_integrate(&plan, r->use_density != 0, mode, s, w, d_img, &T));
    CHK(T.record(6));
    // counters (largest cell count at 0, FAST guard re-tests at 8) come back
    // with the final synchronisation
    int64_t hc[2] = {0, 0};
    if (st) HIPCHK(hipMemcpyAsync(hc, w.counters.p, 16, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        const int mpc = (int)(hc[0] & 0xffffffff);
        st->n_points += b.nr;
        st->n_guard += mode == D2BI_MODE_FAST ? hc[1] : 0;
        st->max_per_cell = std::max(st->max_per_cell, mpc);
        st->n_cells = b.ncell;
        st->ms_bin += T.ms(0, 1);
        st->ms_voxel += T.ms(2, 3);          // the voxel kernel alone
        st->ms_integrate += T.ms(4, 5);      // the integration kernels alone
        st->ms_total += T.ms(0, 6);
    }
    if (plan_out) *plan_out = plan;
    return D2BI_OK;
}

__global__ void k_div_f64(double* __restrict__ p, int64_t n, double d)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = p[i] / d;
}

// Aperture points per shared voxel pass: D2BI_APERTURE_SHARED=0 bins and
// voxelises every point on its own instead (the same image to the bit).
static bool aperture_shared()
{
    const char* e = std::getenv("D2BI_APERTURE_SHARED");
    return !(e && e[0] == '0' && e[1] == 0);
}

// One shared voxel pass over na = 2 or 4 aperture points whose phases are in
// w.app_cs (a stride of nr slots): FP64 fields to f64_re/f64_im at a stride of
// nvox, FAST to f32_field at a stride of nvox records; variant is the FAST
// row kernel (1 or 2).
static int run_voxel_apps(const d2bi_record* r, const d2bi_plan* p, const AxisTables& tab, const AxisDev& ax,
                          int mode, int na, int variant, int64_t nr, hipStream_t s, Workspace& w)
{
    const int64_t nvox = (int64_t)p->Mx * p->My * p->Mz;
    VoxParams V;
    fill_vox_params(r, p, ax, &V);
    if (mode == D2BI_MODE_FP64) {
        CHK(w.f64_re.ensure(sizeof(double) * nvox * na));
        CHK(w.f64_im.ensure(sizeof(double) * nvox * na));
        CHK(w.f64_rho.ensure(sizeof(double) * nvox));
        dim3 grid((p->Mx + 7) / 8, (p->My + 7) / 8, (p->Mz + 3) / 4);
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, V, w.cell_start.as<int>(), w.slot_frac.as<double4>(),
                               w.app_cs.as<double2>(), nr, w.f64_re.as<double>(), w.f64_im.as<double>(), nvox,
                               w.f64_rho.as<double>(), (int*)nullptr, (double*)nullptr);
        };
        if (na == 4) go(k_voxel_fp64<4>);
        else go(k_voxel_fp64<2>);
        HIPCHK(hipGetLastError());
        return D2BI_OK;
    }
    const bool dens = r->use_density != 0;
    CHK(w.f32_field.ensure((dens ? sizeof(float4) : sizeof(float2)) * nvox * na));
    FastParams F;
    CHK(variant_params(p, tab, V, &variant, &F));
    if (variant != 1 && variant != 2) return seterr(D2BI_ERR_ARG, "shared aperture pass: variant %d", variant);
    F.app_cs = w.app_cs.as<float2>();
    F.app_stride = nr;
    F.field_stride = nvox;
    const int fo = forced_int("D2BI_TILE_ORDER", 0, 2);
    set_tile_order(&F, fo >= 0 ? fo : 2);
    CHK(w.counters.ensure(64));
    unsigned long long* gc = (unsigned long long*)(w.counters.as<char>() + 8);
    const size_t gp_bytes = sizeof(unsigned long long) * GUARD_SLOTS * GUARD_STRIDE;
    CHK(w.guard_part.ensure(gp_bytes));
    unsigned long long* gp = w.guard_part.as<unsigned long long>();
    fill_i32((int*)gp, (int64_t)(gp_bytes / sizeof(int)), 0, s);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(8 * F.per_xcd), dim3(64), 0, s, F, w.cell_start.as<int>(),
                           w.slot_frac.as<double4>(), w.slot_p32.as<float4>(), w.slot_sn32.as<float>(),
                           w.f32_field.as<float4>(), (int*)nullptr, (float*)nullptr, gp);
    };
    auto pick = [&](auto dens_c, auto row_c, auto na_c) {
        constexpr bool D = decltype(dens_c)::value;
        constexpr int R = decltype(row_c)::value, N = decltype(na_c)::value;
        if (variant == 2) go(k_voxel_tile<D, false, R, 8, 4, N>);
        else go(k_voxel_tile<D, false, R, 4, 4, N>);
    };
    auto pick_na = [&](auto dens_c, auto row_c) {
        if (na == 4) pick(dens_c, row_c, std::integral_constant<int, 4>{});
        else pick(dens_c, row_c, std::integral_constant<int, 2>{});
    };
    auto pick_row_c = [&](auto dens_c) {
        if (F.thin_z) pick_na(dens_c, std::integral_constant<int, 2>{});
        else pick_na(dens_c, std::integral_constant<int, 1>{});
    };
    if (dens) pick_row_c(std::true_type{});
    else pick_row_c(std::false_type{});
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_sum_guard, dim3(1), dim3(256), 0, s, gp, GUARD_SLOTS, GUARD_STRIDE, gc);
    return D2BI_OK;
}

// computeD2BI_img with apertureAngle and nAperturePoints > 1
// (Lib_DynamicalTwoBeamImaging.f90:1775-1786, 1853-1866, 1905-1951).  Every
// aperture point has its own g-vector for g.r (findGdotR_aperture :1269-1396)
// and so its own phase field; the link cells, ranges and densities are those
// of the unshifted condition.  The atoms are binned once; each pass then
// evaluates the phases of up to four points per slot (k_app_phase) and
// voxelises them together, one neighbour walk for all of them as the
// reference's loop over the neighbour list does (:1905-1922), and each
// point's field is integrated at its nPrec deviation parameters aa(app, :).
// The column sums are added up point by point and the image is divided by
// nApp * nPrec at the end (the reference adds over (prec, app) in the other
// nesting; FP64 sums agree to rounding).  Result left in ws.img.
static int image_aperture(const d2bi_record* r, const double* d_x, int64_t n, double precAngle, int nPrec,
                          double apAngle, int nApp, int mode, d2bi_stats* st, d2bi_plan* plan_out)
{
    if (mode != D2BI_MODE_FP64 && mode != D2BI_MODE_FAST) return seterr(D2BI_ERR_ARG, "unknown mode %d", mode);
    d2bi_plan plan;
    CHK(plan_image(r, n, precAngle, nPrec, &plan));
    std::vector<double> gshift(3 * (size_t)nApp), aa((size_t)nApp * nPrec);
    CHK(aperture_setup(r, precAngle, nPrec, apAngle, nApp, gshift.data(), aa.data()));
    AxisTables tab;
    CHK(axis_tables(r, &plan, &tab));
    hipStream_t s = g.stream;
    Workspace& w = g.ws;
    AxisDev ax;
    CHK(upload_axis(tab, &ax, s, w));
    const int64_t plane = (int64_t)plan.Mx * plan.My;
    const int64_t nvox = plane * plan.Mz;
    CHK(w.img.ensure(sizeof(double) * plane));
    CHK(w.aa.ensure(sizeof(double) * D2BI_MAX_PREC));
    double* d_img = w.img.as<double>();
    const bool dens = r->use_density != 0;
    const double total = (double)nApp * nPrec;
    const bool f64 = mode == D2BI_MODE_FP64;
    // the plan of aperture point app: its g and deviation parameters
    auto point_plan = [&](int app) {
        d2bi_plan pk = plan;
        for (int i = 0; i < 3; i++) pk.gg[i] = gshift[3 * app + i];
        for (int q = 0; q < nPrec; q++) pk.aa[q] = aa[app + (size_t)nApp * q];
        return pk;
    };
    // aperture point app's field (FP64: offset into f64_re/f64_im; FAST: the
    // field record array) integrated into d_img
    auto integrate_point = [&](int app, int64_t f64_off, const void* f32) -> int {
        const d2bi_plan pk = point_plan(app);
        if (f64) {
            double* d_aa = w.aa.as<double>();
            HIPCHK(hipMemcpyAsync(d_aa, pk.aa, sizeof(double) * nPrec, hipMemcpyHostToDevice, s));
            // point 0 writes its sum (div 1), the others add theirs
            CHK(launch_integrate_fp64(&pk, d_aa, d_img, s, app > 0 ? 1 : 0, 1.0, &w, f64_off));
            // pk.aa (host) is read by the copy above: finish before it goes
            HIPCHK(hipStreamSynchronize(s));
        } else {
            CHK(launch_integrate_fast(&pk, dens, f32, d_img, s, app > 0, 1.0 / total, &w));
        }
        return D2BI_OK;
    };
    auto add_stats = [&](const Binned& b, int points) -> int {
        int64_t hc[2] = {0, 0};
        if (st) HIPCHK(hipMemcpyAsync(hc, w.counters.p, 16, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (st) {
            st->n_points += b.nr * points;
            st->n_guard += f64 ? 0 : hc[1];
            st->max_per_cell = std::max(st->max_per_cell, (int)(hc[0] & 0xffffffff));
            st->n_cells = b.ncell;
        }
        return D2BI_OK;
    };

    if (!aperture_shared() || (f64 && n > (int64_t)INT32_MAX)) {
        // every point binned and voxelised on its own
        for (int app = 0; app < nApp; app++) {
            d2bi_plan pk = point_plan(app);
            BinParams BP;
            CHK(bin_params(r, &pk, &BP));
            BP.fast = f64 ? 0 : 1;
            Binned b;
            CHK(run_binning(d_x, n, BP, &b, s, w));
            CHK(run_voxel(r, &pk, tab, ax, mode, false, nullptr, s, w));
            CHK(integrate_point(app, 0, w.f32_field.p));
            CHK(add_stats(b, 1));
        }
    } else {
        // one binning, at point 0's g (its phases are the binning's own)
        d2bi_plan p0 = point_plan(0);
        BinParams BP;
        CHK(bin_params(r, &p0, &BP));
        BP.fast = f64 ? 0 : 1;
        BP.keep_j = f64 ? 1 : 0;
        Binned b;
        CHK(run_binning(d_x, n, BP, &b, s, w));
        const int64_t nr = b.nr;
        // shared passes of four or two points: FP64, or the FAST row kernels
        // of rows of four voxels (the default variant 2, or a forced 1 or 2)
        int variant = forced_variant();
        if (variant < 0) variant = 2;
        variant = effective_variant(&plan, tab, variant);
        const int na_max = (f64 || variant == 1 || variant == 2) ? 4 : 1;
        if (na_max > 1) CHK(w.app_cs.ensure((f64 ? sizeof(double2) : sizeof(float2)) * std::max<int64_t>(nr, 1) * na_max));
        const int nblk = (int)std::max<int64_t>(1, (n + BIN_TILE - 1) / BIN_TILE);
        for (int app = 0; app < nApp;) {
            const int left = nApp - app;
            const int na = na_max >= 4 && left >= 4 ? 4 : (na_max >= 2 && left >= 2 ? 2 : 1);
            if (nr > 0 && (app > 0 || na > 1)) {
                AppG G;
                std::memset(&G, 0, sizeof G);
                for (int a = 0; a < na; a++)
                    for (int i = 0; i < 3; i++) G.g[a][i] = gshift[3 * (app + a) + i];
                const dim3 grid((unsigned)((nr + 255) / 256));
                auto go = [&](auto kern) {
                    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, nr, w.cell_ids.as<int>(), w.pt_j.as<int>(), d_x,
                                       BP, w.blk_off.as<int>(), nblk, G, nr,
                                       na > 1 ? w.app_cs.as<double2>() : w.slot_cs64.as<double2>(),
                                       w.app_cs.as<float2>(), w.slot_p32.as<float4>(), w.slot_sn32.as<float>());
                };
                if (f64) {
                    if (na == 4) go(k_app_phase<true, 4>);
                    else if (na == 2) go(k_app_phase<true, 2>);
                    else go(k_app_phase<true, 1>);
                } else {
                    if (na == 4) go(k_app_phase<false, 4>);
                    else if (na == 2) go(k_app_phase<false, 2>);
                    else go(k_app_phase<false, 1>);
                }
                HIPCHK(hipGetLastError());
            }
            if (na == 1) {
                const d2bi_plan pk = point_plan(app);
                CHK(run_voxel(r, &pk, tab, ax, mode, false, nullptr, s, w));
                CHK(integrate_point(app, 0, w.f32_field.p));
            } else {
                CHK(run_voxel_apps(r, &plan, tab, ax, mode, na, variant, nr, s, w));
                const size_t rec = dens ? sizeof(float4) : sizeof(float2);
                for (int a = 0; a < na; a++)
                    CHK(integrate_point(app + a, nvox * a, w.f32_field.as<char>() + rec * nvox * a));
            }
            CHK(add_stats(b, na));
            app += na;
        }
    }
    if (f64) {
        hipLaunchKernelGGL(k_div_f64, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, d_img, plane, total);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }
    if (plan_out) *plan_out = plan;
    return D2BI_OK;
}

static int upload_atoms_tmp(const double* x, int64_t ld, int64_t n, double** d_x)
{
    if (n < 0 || (n > 0 && (!x || ld < 3))) return seterr(D2BI_ERR_ARG, "bad atom array");
    CHK(g.ws.x.ensure(sizeof(double) * 3 * std::max<int64_t>(n, 1)));
    void* hp;
    CHK(pinned(sizeof(double) * 3 * std::max<int64_t>(n, 1), &hp));
    double* h = (double*)hp;
    for (int64_t j = 0; j < n; j++) {
        h[3 * j] = x[ld * j]; h[3 * j + 1] = x[ld * j + 1]; h[3 * j + 2] = x[ld * j + 2];
    }
    if (n > 0) HIPCHK(hipMemcpyAsync(g.ws.x.p, h, sizeof(double) * 3 * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    *d_x = g.ws.x.as<double>();
    return D2BI_OK;
}


}  // namespace d2bi
