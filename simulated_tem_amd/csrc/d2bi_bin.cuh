// d2bi_bin.cuh -- link-cell binning of the replica atoms: findGdotR and
// LinkCell3D add as a stable counting sort (classify, scan, scatter, place,
// sort, then gather or, FAST imaging, emit).
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
#pragma once

namespace d2bi {

// ------------------------------------------------------------------ (a) binning

static constexpr int BIN_BLOCK = 256;
static constexpr int BIN_ROUNDS = 4;
static constexpr int BIN_TILE = BIN_BLOCK * BIN_ROUNDS;
static constexpr int MAX_REPL = 32;

struct BinParams {
    double RU[9];               // column-major
    double delta[3];
    double lc_delta[3];
    double ia[9];               // link-cell inverse (column-major)
    double gg[3];
    double Ruvw[3 * MAX_REPL];
    int nuvw;
    int Ex, Ey, Ez;
    int64_t n;                  // atoms
    int generic;                // 1: LinkCell3D add1 (no transform, no filter)
    int fast;                   // FAST imaging: FP32 records instead of g.r and the FP64 slot arrays
    int keep_j;                 // FP64 imaging: also keep each point's atom number (aperture phases)
    double cell_a[3];           // diagonal of the link-cell matrix (FAST-mode records)
    // replica pre-classification: c ~= M x + off[nn] approximates the
    // link-cell coordinates of replica nn (products folded, so a few
    // roundings away from the reference's order); a replica is decided
    // from it unless a coordinate lies within eps of a box face
    int prefilter;
    double M[9];                // ia . RU (column-major)
    double off[3 * MAX_REPL];   // ia . (RU . Ruvw[nn] + delta - lc_delta)
    double mabs, oabs;          // magnitude bounds: per unit |x|, and the rest
};

// Link-cell index of a reduced coordinate: mod(floor(y) + 16*N, N)
// (whichCell, Lib_SimpleSupercells.f90:371-373).  For points further than
// 16 boxes away the reference's expression goes negative; here the floor is
// clamped and the modulo kept non-negative, which changes nothing for
// points within range and keeps every index inside the cell array.
__device__ inline int cell_of(double fl, int n)
{
    const double c = fmin(fmax(fl, -1.0e9), 1.0e9);
    const int r = ((int)c + 16 * n) % n;
    return r < 0 ? r + n : r;
}

// Transform atom j for replica nn: yy (rotated lab position), zz (link-cell
// frame), and whether it lies in the link-cell box.  Operation order of
// findGdotR (Lib_DynamicalTwoBeamImaging.f90:1225-1240) and pointInSupercell0
// (Lib_SimpleSupercells.f90:1077-1084).
__device__ inline bool bin_point(const BinParams& P, const double* x3, int nn, double yy[3],
                                 double cyy[3])
{
    if (P.generic) {
        yy[0] = x3[0]; yy[1] = x3[1]; yy[2] = x3[2];
        for (int i = 0; i < 3; i++)
            cyy[i] = (P.ia[i] * yy[0] + P.ia[3 + i] * yy[1]) + P.ia[6 + i] * yy[2];
        return true;
    }
    double y0[3], zz[3];
    for (int i = 0; i < 3; i++) y0[i] = x3[i] + P.Ruvw[3 * nn + i];
    for (int i = 0; i < 3; i++) yy[i] = (P.RU[i] * y0[0] + P.RU[3 + i] * y0[1]) + P.RU[6 + i] * y0[2];
    for (int i = 0; i < 3; i++) {
        yy[i] = yy[i] + P.delta[i];
        zz[i] = yy[i] - P.lc_delta[i];
    }
    for (int i = 0; i < 3; i++)
        cyy[i] = (P.ia[i] * zz[0] + P.ia[3 + i] * zz[1]) + P.ia[6 + i] * zz[2];
    // |floor| beyond 2^30 cannot be inside the box; clamping keeps int conversion defined
    int ix = (int)fmin(fmax(floor(cyy[0]), -1.0e9), 1.0e9);
    int iy = (int)fmin(fmax(floor(cyy[1]), -1.0e9), 1.0e9);
    int iz = (int)fmin(fmax(floor(cyy[2]), -1.0e9), 1.0e9);
    bool is = (ix == (ix + 256 * P.Ex) % P.Ex);
    is = is && (iy == (iy + 256 * P.Ey) % P.Ey);
    is = is && (iz == (iz + 256 * P.Ez) % P.Ez);
    return is;
}

// Replica bits set in any lane of the wave (wave-uniform).
__device__ __forceinline__ uint32_t wave_or(uint32_t m)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m |= (uint32_t)__shfl_xor((int)m, o, 64);
    return __builtin_amdgcn_readfirstlane(m);
}

// Pass 1: replica predicate bits per atom and per-(replica, block) counts.
__global__ __launch_bounds__(BIN_BLOCK) void k_bin_classify(const double* __restrict__ x, BinParams P,
                                                            uint32_t* __restrict__ mask,
                                                            int* __restrict__ blk_cnt, int nblk)
{
    __shared__ int cnt[MAX_REPL];
    if (threadIdx.x < MAX_REPL) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int r = 0; r < BIN_ROUNDS; r++) {
        int64_t j = (int64_t)blockIdx.x * BIN_TILE + r * BIN_BLOCK + threadIdx.x;
        uint32_t m = 0;
        if (j < P.n) {
            double x3[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
            if (P.prefilter) {
                double cb[3];
                for (int i = 0; i < 3; i++) cb[i] = P.M[i] * x3[0] + P.M[3 + i] * x3[1] + P.M[6 + i] * x3[2];
                // |approximate - exact| is a few units of roundoff (1e-16) of
                // the magnitudes involved; eps is 1e-10 of their bound
                const double eps =
                    1.0e-10 * (1.0 + P.mabs * (fabs(x3[0]) + fabs(x3[1]) + fabs(x3[2])) + P.oabs);
                // c in [0, E) per axis is |c - E/2| < E/2: certainly outside
                // when |c - E/2| >= E/2 + eps, certainly inside when <= E/2 - eps
                const double h[3] = {0.5 * P.Ex, 0.5 * P.Ey, 0.5 * P.Ez};
                double ch[3], hout[3], hin[3];
                for (int i = 0; i < 3; i++) {
                    ch[i] = cb[i] - h[i];
                    hout[i] = h[i] + eps;
                    hin[i] = h[i] - eps;
                }
                for (int nn = 0; nn < P.nuvw; nn++) {
                    bool out = false, sure = true;
                    for (int i = 0; i < 3; i++) {
                        const double u = fabs(ch[i] + P.off[3 * nn + i]);
                        out = out || u >= hout[i];
                        sure = sure && u <= hin[i];
                    }
                    if (out) continue;                       // outside on some axis for certain
                    if (sure) { m |= 1u << nn; continue; }   // inside on every axis for certain
                    double yy[3], cyy[3];
                    if (bin_point(P, x3, nn, yy, cyy)) m |= 1u << nn;
                }
            } else {
                for (int nn = 0; nn < P.nuvw; nn++) {
                    double yy[3], cyy[3];
                    if (bin_point(P, x3, nn, yy, cyy)) m |= 1u << nn;
                }
            }
            mask[j] = m;
        }
        // only the replicas some lane of the wave lies in
        for (uint32_t wm = wave_or(m); wm; wm &= wm - 1) {
            const int nn = __builtin_ctz(wm);
            const unsigned long long b = __ballot((m >> nn) & 1u);
            if (lane == 0) atomicAdd(&cnt[nn], (int)__popcll(b));
        }
    }
    __syncthreads();
    if (threadIdx.x < P.nuvw) blk_cnt[(int64_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// Pass 2: point number ii in (replica, atom) order, link cell and a
// per-cell arrival slot from an atomic counter; FP64 imaging and link-cell
// queries also the offset in cell and g.r, FAST imaging the atom number
// (its record is emitted in slot order after the sort).
__global__ __launch_bounds__(BIN_BLOCK) void k_bin_scatter(const double* __restrict__ x, BinParams P,
                                                           const uint32_t* __restrict__ mask,
                                                           const int* __restrict__ blk_off, int nblk,
                                                           int* __restrict__ pt_cell,
                                                           int* __restrict__ pt_slot,
                                                           double* __restrict__ pt_frac,
                                                           double* __restrict__ pt_gdotr,
                                                           int* __restrict__ pt_j,
                                                           int* __restrict__ cell_cnt)
{
    __shared__ int run[MAX_REPL];
    __shared__ int wcnt[BIN_BLOCK / 64][MAX_REPL];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < P.nuvw) run[threadIdx.x] = blk_off[(int64_t)threadIdx.x * nblk + blockIdx.x];
    __syncthreads();
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < BIN_ROUNDS; r++) {
        int64_t j = (int64_t)blockIdx.x * BIN_TILE + r * BIN_BLOCK + threadIdx.x;
        uint32_t m = (j < P.n) ? mask[j] : 0u;
        // the wave's count per replica: zero, then the replicas some lane lies in
        const uint32_t wm0 = wave_or(m);
        if (lane < P.nuvw) wcnt[wave][lane] = 0;
        for (uint32_t wm = wm0; wm; wm &= wm - 1) {
            const int nn = __builtin_ctz(wm);
            const unsigned long long b = __ballot((m >> nn) & 1u);
            if (lane == 0) wcnt[wave][nn] = (int)__popcll(b);
        }
        __syncthreads();
        if (m) {
            double x3[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
            for (uint32_t wm = wm0; wm; wm &= wm - 1) {
                const int nn = __builtin_ctz(wm);
                unsigned long long b = __ballot((m >> nn) & 1u);   // uniform among lanes with m != 0
                if (!((m >> nn) & 1u)) continue;
                int off = run[nn];
                for (int w = 0; w < wave; w++) off += wcnt[w][nn];
                int ii = off + (int)__popcll(b & lt);
                double yy[3], cyy[3];
                bin_point(P, x3, nn, yy, cyy);
                double fl[3] = {floor(cyy[0]), floor(cyy[1]), floor(cyy[2])};
                // a replica point passed pointInSupercell, so 0 <= floor < E
                // and mod(floor + 16E, E) is the floor itself
                const int cx = P.generic ? cell_of(fl[0], P.Ex) : (int)fl[0];
                const int cy = P.generic ? cell_of(fl[1], P.Ey) : (int)fl[1];
                const int cz = P.generic ? cell_of(fl[2], P.Ez) : (int)fl[2];
                int cell = cx + P.Ex * (cy + P.Ey * cz);
                pt_cell[ii] = cell;
                if (P.fast | P.keep_j) pt_j[ii] = (int)j;
                if (!P.fast) {
                    pt_frac[3 * (int64_t)ii + 0] = cyy[0] - fl[0];
                    pt_frac[3 * (int64_t)ii + 1] = cyy[1] - fl[1];
                    pt_frac[3 * (int64_t)ii + 2] = cyy[2] - fl[2];
                    pt_gdotr[ii] = (P.gg[0] * yy[0] + P.gg[1] * yy[1]) + P.gg[2] * yy[2];
                }
                // arrival slot: lanes of this wave landing in one cell share
                // one atomic (the lowest such lane adds the group's size);
                // the grouping is ALU-only, so each lane waits for one atomic
                // return.  Order inside a cell is restored by ii later.
                unsigned long long rem = __ballot(1), grp = 0;
                int leader = lane;
                while (rem) {
                    const int l = __ffsll((long long)rem) - 1;
                    const int lc = __shfl(cell, l, 64);
                    const unsigned long long gm = __ballot(cell == lc);
                    if (cell == lc) { leader = l; grp = gm; }
                    rem &= ~gm;
                }
                int base = 0;
                if (lane == leader) base = atomicAdd(&cell_cnt[cell], (int)__popcll(grp));
                base = __shfl(base, leader, 64);
                pt_slot[ii] = base + (int)__popcll(grp & lt);
            }
        }
        __syncthreads();
        if (threadIdx.x < P.nuvw) {
            int s = 0;
            for (int w = 0; w < BIN_BLOCK / 64; w++) s += wcnt[w][threadIdx.x];
            run[threadIdx.x] += s;
        }
        __syncthreads();
    }
}

// Pass 3: place each point at its cell's start + arrival slot.
__global__ void k_bin_place(int64_t nr, const int* __restrict__ pt_cell, const int* __restrict__ pt_slot,
                            const int* __restrict__ cell_start, int* __restrict__ cell_ids)
{
    int64_t ii = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= nr) return;
    cell_ids[cell_start[pt_cell[ii]] + pt_slot[ii]] = (int)ii;
}

// Pass 4a: restore insertion order inside each cell (sort by ii, which is
// the reference's add order) and the largest cell count (one atomic per
// wave).  A block takes 256 consecutive cells: their points are one
// contiguous range, loaded into LDS with coalesced reads when it fits,
// sorted there segment by segment (one thread per cell: cells of up to
// SORT_K points are ranked in registers, fully unrolled; larger ones are
// insertion-sorted), and written back coalesced.  A range that does not fit
// is sorted in place in memory.
static constexpr int SORT_K = 16;
static constexpr int SORT_LDS = 4096;                // ints of LDS per block

// n >= 1.  Every read stays inside the cell: the index is clamped before
// the select, so a load the compiler issues unconditionally is still in
// bounds.
template <typename Ids>
__device__ __forceinline__ void sort_cell(Ids ids, int s0, int n)
{
    if (n <= SORT_K) {
        int v[SORT_K];
#pragma unroll
        for (int k = 0; k < SORT_K; k++) {
            const int x = ids[s0 + min(k, n - 1)];
            v[k] = k < n ? x : 0x7fffffff;
        }
#pragma unroll
        for (int i = 0; i < SORT_K; i++) {
            int r = 0;
#pragma unroll
            for (int j = 0; j < SORT_K; j++) r += v[j] < v[i] ? 1 : 0;
            if (i < n) ids[s0 + r] = v[i];
        }
    } else {
        for (int i = s0 + 1; i < s0 + n; i++) {
            const int v = ids[i];
            int k = i - 1;
            while (k >= s0 && ids[k] > v) { ids[k + 1] = ids[k]; k--; }
            ids[k + 1] = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_bin_sort(int ncell, const int* __restrict__ cell_start,
                                                  int* __restrict__ cell_ids, int* __restrict__ blk_max)
{
    __shared__ int wmax[4];
    __shared__ int sid[SORT_LDS];
    const int c0 = blockIdx.x * blockDim.x;
    const int c = c0 + threadIdx.x;
    const int cl = min(c0 + (int)blockDim.x, ncell);
    const int R0 = cell_start[c0], R1 = cell_start[cl];
    const bool in_lds = R1 - R0 <= SORT_LDS;
    int m = 0, s0 = 0, n = 0;
    if (c < ncell) {
        s0 = cell_start[c];
        n = cell_start[c + 1] - s0;
        m = n;
    }
    if (in_lds) {
        for (int k = R0 + (int)threadIdx.x; k < R1; k += blockDim.x) sid[k - R0] = cell_ids[k];
        __syncthreads();
        if (n > 1) sort_cell(sid, s0 - R0, n);
        __syncthreads();
        for (int k = R0 + (int)threadIdx.x; k < R1; k += blockDim.x) cell_ids[k] = sid[k - R0];
    } else if (n > 1) {
        sort_cell(cell_ids, s0, n);
    }
    // the block's largest cell count, stored (no atomics on one address)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) blk_max[blockIdx.x] = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
}

// Largest of n per-block counts into *out (one block).
__global__ __launch_bounds__(256) void k_max_i32(const int* __restrict__ v, int n, int* __restrict__ out)
{
    __shared__ int wmax[4];
    int m = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = max(m, v[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) *out = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
}

// Sum of the FAST guard counters into *out (one block).
__global__ __launch_bounds__(256) void k_sum_guard(const unsigned long long* __restrict__ part, int n, int stride,
                                                   unsigned long long* __restrict__ out)
{
    __shared__ unsigned long long ws[4];
    unsigned long long v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[(int64_t)i * stride];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// Pass 4b: slot-ordered point data, one thread per slot: 1-based id, FP64
// offset in cell, FP64 (cos, sin) of g.r, and the FAST record (offset from
// the cell origin in A for cuboidal cells, cos g.r; sin g.r beside it), all
// rounded from FP64.
__global__ __launch_bounds__(256) void k_bin_gather(int64_t nr, const int* __restrict__ cell_ids,
                                                    const double* __restrict__ pt_frac,
                                                    const double* __restrict__ pt_gdotr, double a0, double a1,
                                                    double a2, int* __restrict__ slot_id,
                                                    double4* __restrict__ slot_frac, double2* __restrict__ slot_cs64,
                                                    float4* __restrict__ slot_p32, float* __restrict__ slot_sn32)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nr) return;
    const int ii = cell_ids[s];
    slot_id[s] = ii + 1;
    const double f0 = pt_frac[3 * (int64_t)ii], f1 = pt_frac[3 * (int64_t)ii + 1], f2 = pt_frac[3 * (int64_t)ii + 2];
    slot_frac[s] = make_double4(f0, f1, f2, 0.0);
    double sn, cs;
    sincos(pt_gdotr[ii], &sn, &cs);
    slot_cs64[s] = make_double2(cs, sn);
    slot_p32[s] = make_float4((float)(a0 * f0), (float)(a1 * f1), (float)(a2 * f2), (float)cs);
    slot_sn32[s] = (float)sn;
}

// Pass 4b, FAST imaging: each slot's point evaluated again from its atom
// and replica (the transform of k_bin_scatter, so the same cell offsets to
// the bit) and written in slot order: the FAST record (offset from the cell
// origin in A and cos g.r; sin g.r beside it), rounded once from FP64, and
// the FP64 offset in cell for the guard re-test.  The replica of point ii
// is the last one whose first point number is <= ii (numbers run in
// (replica, atom) order; rep_first is padded to 32 entries with INT_MAX).
__global__ __launch_bounds__(256) void k_bin_emit_fast(int64_t nr, const int* __restrict__ cell_ids,
                                                       const int* __restrict__ pt_j, const double* __restrict__ x,
                                                       BinParams P, const int* __restrict__ blk_off, int nblk,
                                                       float4* __restrict__ slot_p32, float* __restrict__ slot_sn32,
                                                       double4* __restrict__ slot_frac)
{
    __shared__ int rep_first[MAX_REPL];
    if (threadIdx.x < MAX_REPL)
        rep_first[threadIdx.x] = threadIdx.x < P.nuvw ? blk_off[(int64_t)threadIdx.x * nblk] : 0x7fffffff;
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nr) return;
    const int ii = cell_ids[s];
    const int64_t j = pt_j[ii];
    int nn = 0;                                       // rep_first[0] = 0 <= ii
#pragma unroll
    for (int step = MAX_REPL / 2; step > 0; step >>= 1)
        if (rep_first[nn + step] <= ii) nn += step;
    const double x3[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
    double yy[3], cyy[3];
    bin_point(P, x3, nn, yy, cyy);
    const double f0 = cyy[0] - floor(cyy[0]), f1 = cyy[1] - floor(cyy[1]), f2 = cyy[2] - floor(cyy[2]);
    const double gd = (P.gg[0] * yy[0] + P.gg[1] * yy[1]) + P.gg[2] * yy[2];
    double sn, cs;
    sincos(gd, &sn, &cs);
    slot_p32[s] = make_float4((float)(P.cell_a[0] * f0), (float)(P.cell_a[1] * f1), (float)(P.cell_a[2] * f2),
                              (float)cs);
    slot_sn32[s] = (float)sn;
    slot_frac[s] = make_double4(f0, f1, f2, 0.0);
}

// g-vectors of up to four aperture points (findGdotR_aperture,
// Lib_DynamicalTwoBeamImaging.f90:1269-1396: the points of a condition differ
// only in g, so they share its binned points and take their own g.r)
struct AppG {
    double g[4][3];
};

// cos and sin of g_a.r for every slot and the NA aperture points of G, the
// slot's point evaluated again from its atom and replica as k_bin_emit_fast
// does (g.r with k_bin_scatter's expression, so the phases are those a
// binning at g_a would give).  F64: double2 (cos, sin) at out64[a stride +
// slot].  FAST: NA = 1 rewrites the FP32 phase of the slot's record in place
// (slot_p32.w, slot_sn32); NA > 1 writes float2 (cos, sin) at out32[a stride
// + slot].
template <bool F64, int NA>
__global__ __launch_bounds__(256) void k_app_phase(int64_t nr, const int* __restrict__ cell_ids,
                                                   const int* __restrict__ pt_j, const double* __restrict__ x,
                                                   BinParams P, const int* __restrict__ blk_off, int nblk, AppG G,
                                                   int64_t stride, double2* __restrict__ out64,
                                                   float2* __restrict__ out32, float4* __restrict__ slot_p32,
                                                   float* __restrict__ slot_sn32)
{
    __shared__ int rep_first[MAX_REPL];
    if (threadIdx.x < MAX_REPL)
        rep_first[threadIdx.x] = threadIdx.x < P.nuvw ? blk_off[(int64_t)threadIdx.x * nblk] : 0x7fffffff;
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nr) return;
    const int ii = cell_ids[s];
    const int64_t j = pt_j[ii];
    int nn = 0;
#pragma unroll
    for (int step = MAX_REPL / 2; step > 0; step >>= 1)
        if (rep_first[nn + step] <= ii) nn += step;
    const double x3[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
    double yy[3], cyy[3];
    bin_point(P, x3, nn, yy, cyy);
#pragma unroll
    for (int a = 0; a < NA; a++) {
        const double gd = (G.g[a][0] * yy[0] + G.g[a][1] * yy[1]) + G.g[a][2] * yy[2];
        double sn, cs;
        sincos(gd, &sn, &cs);
        if (F64) {
            out64[a * stride + s] = make_double2(cs, sn);
        } else if (NA == 1) {
            reinterpret_cast<float*>(slot_p32 + s)[3] = (float)cs;
            slot_sn32[s] = (float)sn;
        } else {
            out32[a * stride + s] = make_float2((float)cs, (float)sn);
        }
    }
}

struct Binned {
    int64_t nr = 0;
    int ncell = 0;
    int max_per_cell = 0;
};

// Pass 4 (sort within cells, then gather or emit by slot), on workspace w;
// d_x, blk_off and nblk are the condition's atoms and per-(replica, block)
// point offsets (FAST imaging).
static void launch_bin_finish(int ncell, int64_t nr, const BinParams& P, int* mpc, hipStream_t s, Workspace& w,
                              const double* d_x, const int* blk_off, int nblk)
{
    const int nb = (ncell + 255) / 256;         // blk_max holds >= nb ints (sized with cell_start)
    hipLaunchKernelGGL(k_bin_sort, dim3(nb), dim3(256), 0, s, ncell, w.cell_start.as<int>(), w.cell_ids.as<int>(),
                       w.blk_max.as<int>());
    hipLaunchKernelGGL(k_max_i32, dim3(1), dim3(256), 0, s, w.blk_max.as<int>(), nb, mpc);
    if (nr > 0 && P.fast)
        hipLaunchKernelGGL(k_bin_emit_fast, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, nr,
                           w.cell_ids.as<int>(), w.pt_j.as<int>(), d_x, P, blk_off, nblk,
                           w.slot_p32.as<float4>(), w.slot_sn32.as<float>(), w.slot_frac.as<double4>());
    else if (nr > 0)
        hipLaunchKernelGGL(k_bin_gather, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, nr,
                           w.cell_ids.as<int>(), w.pt_frac.as<double>(), w.pt_gdotr.as<double>(), P.cell_a[0],
                           P.cell_a[1], P.cell_a[2], w.slot_id.as<int>(), w.slot_frac.as<double4>(),
                           w.slot_cs64.as<double2>(), w.slot_p32.as<float4>(), w.slot_sn32.as<float>());
}

// Run binning for either a condition (generic = 0) or plain points, on
// workspace w.
static int run_binning(const double* d_x, int64_t n, BinParams& P, Binned* out, hipStream_t s, Workspace& w)
{
    const int64_t ncell64 = (int64_t)P.Ex * P.Ey * P.Ez;
    if (ncell64 > (int64_t)INT32_MAX / 2) return seterr(D2BI_ERR_CAPACITY, "too many link cells");
    const int ncell = (int)ncell64;
    const int nblk = (int)std::max<int64_t>(1, (n + BIN_TILE - 1) / BIN_TILE);
    if ((P.fast || P.keep_j) && n > (int64_t)INT32_MAX) return seterr(D2BI_ERR_CAPACITY, "more than 2^31 - 1 atoms");
    P.n = n;
    CHK(w.mask.ensure(sizeof(uint32_t) * std::max<int64_t>(n, 1)));
    CHK(w.blk_cnt.ensure(sizeof(int) * ((int64_t)P.nuvw * nblk + 1)));
    CHK(w.blk_off.ensure(sizeof(int) * ((int64_t)P.nuvw * nblk + 1)));
    CHK(w.cell_cnt.ensure(sizeof(int) * (ncell + 1)));
    CHK(w.cell_start.ensure(sizeof(int) * (ncell + 1)));
    CHK(w.blk_max.ensure(sizeof(int) * ((ncell + 255) / 256 + 1)));
    CHK(w.counters.ensure(64));
    fill_i32(w.cell_cnt.as<int>(), ncell + 1, 0, s);
    fill_i32(w.counters.as<int>(), 16, 0, s);
    if (n > 0) {
        hipLaunchKernelGGL(k_bin_classify, dim3(nblk), dim3(BIN_BLOCK), 0, s, d_x, P,
                           w.mask.as<uint32_t>(), w.blk_cnt.as<int>(), nblk);
    } else {
        fill_i32(w.blk_cnt.as<int>(), (int64_t)P.nuvw * nblk + 1, 0, s);
    }
    HIPCHK(hipGetLastError());
    const int64_t nbc = (int64_t)P.nuvw * nblk;
    CHK(scan_exclusive(w.blk_cnt.as<int>(), w.blk_off.as<int>(), nbc, s, w));
    int nr_host = 0;
    HIPCHK(hipMemcpyAsync(&nr_host, w.blk_off.as<int>() + nbc, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (nr_host < 0) return seterr(D2BI_ERR_CAPACITY, "point count overflow");
    const int64_t nr = nr_host;
    const int64_t nrc = std::max<int64_t>(nr, 1);
    CHK(w.pt_cell.ensure(sizeof(int) * nrc));
    CHK(w.pt_slot.ensure(sizeof(int) * nrc));
    CHK(w.cell_ids.ensure(sizeof(int) * nrc));
    CHK(w.slot_p32.ensure(sizeof(float4) * nrc));
    CHK(w.slot_sn32.ensure(sizeof(float) * nrc));
    CHK(w.slot_frac.ensure(sizeof(double4) * nrc));
    if (P.fast || P.keep_j) CHK(w.pt_j.ensure(sizeof(int) * nrc));
    if (!P.fast) {
        CHK(w.pt_frac.ensure(sizeof(double) * 3 * nrc));
        CHK(w.pt_gdotr.ensure(sizeof(double) * nrc));
        CHK(w.slot_id.ensure(sizeof(int) * nrc));
        CHK(w.slot_cs64.ensure(sizeof(double2) * nrc));
    }
    if (n > 0) {
        hipLaunchKernelGGL(k_bin_scatter, dim3(nblk), dim3(BIN_BLOCK), 0, s, d_x, P,
                           w.mask.as<uint32_t>(), w.blk_off.as<int>(), nblk, w.pt_cell.as<int>(),
                           w.pt_slot.as<int>(), w.pt_frac.as<double>(), w.pt_gdotr.as<double>(),
                           w.pt_j.as<int>(), w.cell_cnt.as<int>());
        HIPCHK(hipGetLastError());
    }
    CHK(scan_exclusive(w.cell_cnt.as<int>(), w.cell_start.as<int>(), ncell, s, w));
    if (nr > 0) {
        hipLaunchKernelGGL(k_bin_place, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, nr,
                           w.pt_cell.as<int>(), w.pt_slot.as<int>(), w.cell_start.as<int>(),
                           w.cell_ids.as<int>());
        HIPCHK(hipGetLastError());
    }
    launch_bin_finish(ncell, nr, P, w.counters.as<int>(), s, w, d_x, w.blk_off.as<int>(), nblk);
    HIPCHK(hipGetLastError());
    out->nr = nr;
    out->ncell = ncell;
    return D2BI_OK;
}

// Fill the binning parameters of a condition from the plan.
static int bin_params(const d2bi_record* r, const d2bi_plan* p, BinParams* P)
{
    std::memset(P, 0, sizeof *P);
    if (r->nuvw < 1 || r->nuvw > MAX_REPL)
        return seterr(D2BI_ERR_ARG, "nuvw = %d not supported (1..%d)", r->nuvw, MAX_REPL);
    std::memcpy(P->RU, p->RU, sizeof P->RU);
    std::memcpy(P->delta, r->delta, sizeof P->delta);
    std::memcpy(P->lc_delta, p->lc_delta, sizeof P->lc_delta);
    std::memcpy(P->ia, p->ia_col, sizeof P->ia);
    std::memcpy(P->gg, p->gg, sizeof P->gg);
    for (int nn = 0; nn < r->nuvw; nn++) {
        const int* u = &r->uvw[3 * nn];
        for (int i = 0; i < 3; i++) {       // getSuperVector (Lib_SimpleSupercells.f90:904-918)
            double s1 = r->super_a[i] * r->super_n[0];
            double s2 = r->super_a[3 + i] * r->super_n[1];
            double s3 = r->super_a[6 + i] * r->super_n[2];
            P->Ruvw[3 * nn + i] = (u[0] * s1 + u[1] * s2) + u[2] * s3;
        }
    }
    P->nuvw = r->nuvw;
    P->Ex = p->Ex; P->Ey = p->Ey; P->Ez = p->Ez;
    P->generic = 0;
    P->cell_a[0] = p->a_col[0]; P->cell_a[1] = p->a_col[4]; P->cell_a[2] = p->a_col[8];
    // folded transform of the pre-classification (any rounding order: it
    // only decides replicas that are further than eps from a box face)
    double ia_max = 0.0, ru_max = 0.0, t_max = 0.0;
    for (int i = 0; i < 9; i++) {
        ia_max = std::max(ia_max, std::fabs(P->ia[i]));
        ru_max = std::max(ru_max, std::fabs(P->RU[i]));
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += P->ia[3 * k + i] * P->RU[3 * j + k];
            P->M[3 * j + i] = v;
        }
    for (int nn = 0; nn < P->nuvw; nn++) {
        double z[3];
        for (int k = 0; k < 3; k++) {
            double v = P->delta[k] - P->lc_delta[k];
            double t = std::fabs(P->delta[k]) + std::fabs(P->lc_delta[k]);
            for (int j = 0; j < 3; j++) {
                v += P->RU[3 * j + k] * P->Ruvw[3 * nn + j];
                t += std::fabs(P->RU[3 * j + k] * P->Ruvw[3 * nn + j]) + std::fabs(P->Ruvw[3 * nn + j]);
            }
            z[k] = v;
            t_max = std::max(t_max, t);
        }
        for (int i = 0; i < 3; i++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += P->ia[3 * k + i] * z[k];
            P->off[3 * nn + i] = v;
        }
    }
    // bounds on the magnitude of every intermediate of both evaluations:
    // |x| terms scale with 3 max|ia| max(1, max|RU|), the rest with
    // 3 max|ia| (|RU Ruvw| + |Ruvw| + |delta| + |lc_delta|)
    P->mabs = 3.0 * ia_max * std::max(1.0, ru_max) * std::max(1.0, 3.0 * ru_max);
    P->oabs = 3.0 * ia_max * t_max;
    // box faces at 0 and E <= 2^20, so that E + eps and the exact path's
    // integer test are exact
    P->prefilter = (P->Ex <= (1 << 20) && P->Ey <= (1 << 20) && P->Ez <= (1 << 20)) ? 1 : 0;
    return D2BI_OK;
}

}  // namespace d2bi
