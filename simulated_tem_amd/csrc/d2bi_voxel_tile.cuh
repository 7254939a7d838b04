// d2bi_voxel_tile.cuh -- the FAST voxel kernel k_voxel_tile: LDS-staged voxel
// tiles, the row variants and the FP64 guard re-test.
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
#pragma once

namespace d2bi {

// ------------------------------------------------------------------ (b) voxelise, FAST

static constexpr int TILE_CAP = 160;                 // staged candidates per pass (lane per voxel)
static constexpr int GUARD_SLOTS = 1024;             // guard counters (power of two) ...
static constexpr int GUARD_STRIDE = 16;              // ... 16 x 8 B = 128 B apart
static constexpr int TILE_MAXC = 128;                // cells in a tile's unwrapped range (<= 5^3)
static constexpr int ROW_BLOCK = 32;                 // row kernel: steps per guard test and Kahan merge

// Inclusive prefix sum over the 64 lanes of a wave in six DPP steps: within
// each row of 16 lanes (row_shr 1, 2, 4, 8 with zero fill), then row 0's
// total into row 1 and row 2's into row 3 (row_bcast:15), then rows 0-1's
// into rows 2-3 (row_bcast:31).  All lanes must be active.
__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
    return v;
}

struct FastParams {
    VoxParams v;
    int ntx, nty, ntz;          // tiles per axis
    int ntiles, per_xcd;        // tiles; workgroups per XCD (the grid is 8 per_xcd)
    float e_rc, e_band;         // k2 (3 sigma)^2 and the guard half-width around it
    float rc_cull2;             // (3 sigma)^2 widened for the tile-box cull
    float k2;                   // -log2(e)/(2 sigma^2)
    float w_scale;              // 2^e_rc
    float m2_min;               // (1e-16)^2 / w_scale^2
    float inv_k2;               // 1 / k2
    float cellf[3];             // link-cell sides in FP32
    int thin_z;                 // link cells thinner than 3 sigma along z
    int tz;                     // tile depth in voxels (4, or 8 for the 4 x 4 x 8 row kernel)
    int order;                  // tiles over workgroups: 0 contiguous ranges per XCD, 1 linear, 2 rows dealt to XCDs
    // aperture points sharing one pass (kernels with NA > 1): point a's
    // (cos, sin) of g.r at app_cs[a app_stride + slot], its field at
    // field[a field_stride + voxel]
    const float2* app_cs;
    int64_t app_stride, field_stride;
};

// One wave per voxel tile: 4x4x4 voxels, or 4x4x8 for the row kernel with
// TZ = 8 (x, y, z; x fastest).  The tile's candidates are the points of
// every link cell in the unwrapped range [cell(first voxel)-1,
// cell(last voxel)+1] on each axis; an unwrapped cell u stands for cell
// mod(u, E) displaced by u - mod(u, E) cells, which is exactly the periodic
// image neighbourList1b sees through its mod() (Lib_LinkCell3D.f90:1567-1575).
// Cells, then points, farther than 3 sigma from the tile's voxel box are
// dropped; the rest are compacted into LDS (ballot + popcount) and every
// voxel of the tile tests them all.
//
// Staging.  The kept non-empty cells are listed once per tile with their
// origin in the tile frame and the offset from candidate index to point
// slot.  Candidates are then taken 64 at a time; the cell of candidate t is
// the number of kept cells starting at or before t, less one, found from a
// 64-bit mask of the cells starting inside the round (a flag per lane, a
// ballot and a popcount).
//
// Arithmetic.  Positions are taken relative to the tile centre c (FP32,
// rounded once from FP64 values).  For a candidate p and voxel q the Gaussian
// exponent e = k2*|p - q|^2 (k2 = -log2(e)/(2 sigma^2), so the weight is 2^e)
// is evaluated as
//     e = (-2 k2 p).q + (k2 |p|^2 + k2 |q|^2)
// from a per-candidate record (-2 k2 p, k2 |p|^2): three FMAs and an add.
// Membership is e >= k2 (3 sigma)^2; where |e - k2 (3 sigma)^2| is within
// the guard band the pair is re-tested in FP64 with the expression of
// neighbourList1b (and the neighbour offset must be -1, 0 or +1, as in its
// 27-cell loop), so the set of summed atoms is exactly the FP64 set.
// cos/sin of g.r were evaluated in FP64.
//
// ROW = false: lane = voxel; S = sum w (cos, sin) is accumulated in blocks of
// 32 candidates whose sums are merged with Kahan compensation.
//
// ROW = true: the wave is the tile's voxel rows along x (one per (y, z): 16
// for TZ = 4, 32 for TZ = 8) times NQ candidate parts (4 or 2); lane (row, s)
// walks candidates s, s+NQ, s+2NQ, ... for the TX voxels of its row (TX = 4,
// or 8 for 8 x 4 x 4 tiles).  The Gaussian factorises along x:
//     2^(e_i - e_rc) = X_i * 2^t,   X_i = 2^(k2 ((p_x - q_x,i)^2 + p_y^2 + p_z^2)),
//     t = -2 k2 (p_y q_y + p_z q_z) + k2 (q_y^2 + q_z^2) - e_rc   (e_rc = k2 (3 sigma)^2),
// and the TX x positions q_x,i are the tile's, common to every row, so
// X_0..X_TX-1 (which carry the candidate's own y-z term) are evaluated once
// per candidate when it is staged and a lane spends two FMAs and one
// v_exp_f32 per candidate on TX voxels.  Membership e_i >= e_rc
// is u_i = X_i 2^t - 1 >= 0; a block in which some |u_i| lies within the
// guard band is walked again and the pairs of the band are re-tested in FP64,
// correcting the sums where the FP64 test disagrees.  The parts' sums are
// added in a fixed order through LDS at the end, after which lane (row, s)
// holds voxels s NV .. s NV + NV - 1 of its row (NV = TX / NQ).
//
// Thin cells.  Link cells are at least 3 sigma wide along x and y (Ex <=
// Dx/deltaM + 2 with deltaM >= 3 sigma / min(a_x, a_y)), so along those axes
// every point within 3 sigma of a voxel lies in a neighbouring cell and the
// range test alone decides membership (the guard band covers rounding of the
// cell sides).  Along z the cell height Mz a_z / Ez can be smaller (a_z <
// min(a_x, a_y) with deltaM dividing Mz, or Ez = 3 for thin foils); then a
// point two cells away can lie within 3 sigma, and neighbourList1b does not
// see it.  ROW = 2 applies that rule: a candidate whose cell is not within one
// cell of the row's along z gets weight 0 for the row.  The diagnostic kernel
// tests the z offset of every pair when F.thin_z is set.
//
// Aperture points (NA = 2 or 4, row kernels with rows of four voxels): the NA
// points of a condition's aperture sum share its binned points, weights and
// membership and differ only in g.r (Lib_DynamicalTwoBeamImaging.f90:
// 1905-1922 sums them in one neighbour loop).  Each staged candidate carries
// NA (cos, sin) pairs, every step adds the pair's weight into NA sums, and a
// guard correction is applied to each.  Point a's sums take the same terms in
// the same order as an NA = 1 pass with that point's phases, so its field is
// that pass's to the bit.
template <bool DENSITY, bool DIAG, int ROW, int TZ, int TX = 4, int NA = 1>
__global__ __launch_bounds__(64) void k_voxel_tile(FastParams F, const int* __restrict__ cell_start,
                                                   const double4* __restrict__ slot_frac,
                                                   const float4* __restrict__ p32,
                                                   const float* __restrict__ sn32,
                                                   float4* __restrict__ field,
                                                   int* __restrict__ out_cnt,
                                                   float* __restrict__ out_r2m,
                                                   unsigned long long* __restrict__ guard_count)
{
    static_assert(!(DIAG && ROW), "the diagnostic kernel tests every pair itself");
    static_assert(TZ == 4 || (ROW && TZ == 8), "tiles of 4 x 4 x 4, or 4 x 4 x 8 for the row kernel");
    static_assert(TX == 4 || (ROW && TX == 8 && TZ == 4), "rows of 8 voxels: 8 x 4 x 4 tiles of the row kernel");
    static_assert(NA == 1 || ((NA == 2 || NA == 4) && ROW && TX == 4), "aperture points: 2 or 4, row kernels of 4");
    // ROW: candidate parts per row (4 for 16 rows, 2 for 32 rows), voxels a
    // lane finishes (4 / NQ)
    constexpr int NQ = ROW ? (TZ == 8 ? 2 : 4) : 1;
    constexpr int NV = ROW ? TX / NQ : 1;
    constexpr int XR = TX / 4;                          // float4 records of a candidate's X factors
    // XCD-aware order: consecutive workgroups go round-robin to the 8 XCDs;
    // order 0 gives each XCD a contiguous range of tiles so neighbouring
    // tiles, which read the same link cells, share one L2 (the ranges are
    // slabs along z); order 1 is linear; order 2 deals rows of tiles along x
    // to the XCDs in turn (row r to XCD r mod 8), so that every XCD holds
    // rows at every depth and a row's tiles share one L2.  A tile's sums do
    // not depend on the order.
    const int bid = blockIdx.x;
    int L;
    if (F.order == 0) L = (bid & 7) * F.per_xcd + (bid >> 3);
    else if (F.order == 1) L = bid;
    else {
        const int idx = bid >> 3, row = (idx / F.ntx) * 8 + (bid & 7);
        L = row * F.ntx + (idx - (idx / F.ntx) * F.ntx);
    }
    if (L >= F.ntiles) return;
    const VoxParams& P = F.v;
    const int tx = L % F.ntx, ty = (L / F.ntx) % F.nty, tz = L / (F.ntx * F.nty);
    const int lane = threadIdx.x;
    const int x0 = tx * TX, y0 = ty * TILE, z0 = tz * TZ;
    const int x1 = min(P.Mx, x0 + TX) - 1, y1 = min(P.My, y0 + TILE) - 1, z1 = min(P.Mz, z0 + TZ) - 1;
    // the lane's row of voxels along x (y fastest) and, lane per voxel, its x
    const int rrow = ROW ? lane / NQ : (lane >> 2);
    const int ix = x0 + (lane & 3), iy = y0 + (rrow & 3), iz = z0 + (rrow >> 2);
    const bool row_act = iy <= y1 && iz <= z1;
    const bool act = ix <= x1 && row_act;
    const int jx = min(ix, x1), jy = min(iy, y1), jz = min(iz, z1);
    const QL* const qlz_tab = TZ == 8 ? P.ax.ql[3] : P.ax.ql[2];   // z relative to the tile's first cell
    const QL* const qlx_tab = TX == 8 ? P.ax.ql[4] : P.ax.ql[0];   // x relative to the tile's first cell

    // origin cell of the tile; the voxel's FP32 position relative to it and
    // its cell offset (host tables)
    const int c0x = P.ax.cell[0][x0], c0y = P.ax.cell[1][y0], c0z = P.ax.cell[2][z0];
    const QL qlx = qlx_tab[jx], qly = P.ax.ql[1][jy], qlz = qlz_tab[jz];
    const int lcx = qlx.lc, lcy = qly.lc, lcz = qlz.lc;
    const float qx = qlx.q, qy = qly.q, qz = qlz.q;
    // the tile's voxel box: positions do not decrease with the voxel index
    // (link cells are monotonic along each axis), so its corners are the
    // first and last voxels
    const QL ex0 = qlx_tab[x0], ex1 = qlx_tab[x1];
    const QL ey0 = P.ax.ql[1][y0], ey1 = P.ax.ql[1][y1];
    const QL ez0 = qlz_tab[z0], ez1 = qlz_tab[z1];
    const float bx0 = ex0.q, bx1 = ex1.q, by0 = ey0.q, by1 = ey1.q, bz0 = ez0.q, bz1 = ez1.q;
    // tile-centred frame
    const float ccx = 0.5f * (bx0 + bx1), ccy = 0.5f * (by0 + by1), ccz = 0.5f * (bz0 + bz1);
    const float hx = bx1 - ccx, hy = by1 - ccy, hz = bz1 - ccz;
    const float qcx = qx - ccx, qcy = qy - ccy, qcz = qz - ccz;
    const float k2 = F.k2;
    const float kq2 = k2 * fmaf(qcx, qcx, fmaf(qcy, qcy, qcz * qcz));
    // ROW: the tile's four x positions (clamped at the grid's edge), the
    // lane's candidate part and the y-z part of its exponent
    float qxt[TX];
#pragma unroll
    for (int i = 0; i < TX; i++) qxt[i] = qlx_tab[min(x0 + i, x1)].q - ccx;
    const int squart = lane & (NQ - 1);
    const float kyz = fmaf(k2, fmaf(qcy, qcy, qcz * qcz), -F.e_rc);

    // unwrapped cell range (cells are monotonic, so ux0 + dx lies in [-1, E])
    const int ux0 = c0x - 1, uy0 = c0y - 1, uz0 = c0z - 1;
    const int ncx = ex1.lc + 3, ncy = ey1.lc + 3, ncz = ez1.lc + 3;
    const int ncell = ncx * ncy * ncz;                 // <= TILE_MAXC (checked on the host)

    // staged candidates: ROW = false, (-2 k2 p, k2 |p|^2) tile-centred, then
    // (cos, sin) of g.r, slot, cell offsets; ROW = true, (X_0..X_3), then
    // (-2 k2 p_y, -2 k2 p_z, sin, cos), then (sin, slot, cell offsets, 0),
    // NA > 1 then (cos, sin) of the aperture points two to a record, and four
    // zero records after the last candidate of a pass
    constexpr int NPH = NA > 1 ? NA / 2 : 0;
    constexpr int REC = ROW ? XR + 2 + NPH : 2;
    // ROW: 96 staged candidates per pass keep the wave's LDS at 7.6 KB, so
    // its 88 VGPRs, not LDS, bound occupancy (20 waves per CU; 128 allowed 17
    // and ran 15 % slower on c1)
    constexpr int CAP = ROW ? 96 : TILE_CAP;
    constexpr int NREC = ROW ? CAP + 4 : CAP;
    __shared__ float4 crec[TILE_MAXC];         // kept cells: origin in the tile frame, slot - index
    __shared__ int cupk[TILE_MAXC];            // kept cells: dx | dy << 8 | dz << 16
    __shared__ int sflag[64];                  // cells starting at a lane's candidate this round
    __shared__ float4 sA[REC * NREC];
    int* const cstart = reinterpret_cast<int*>(sA);    // kept cells' first candidate (before staging)
    static_assert(sizeof(int) * TILE_MAXC <= sizeof(float4) * REC * NREC, "cstart overlay");
    static_assert(!ROW || 3 * 64 * TX * sizeof(float) <= sizeof(float4) * REC * NREC, "row sums overlay");

    // set bits of a wave mask below this lane (v_mbcnt)
    auto below = [](unsigned long long m) -> int {
        return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    };

    // kept cells: in range of the tile box and not empty, in cell order
    int K = 0, T = 0;
    {
        // (c + 0.5) / n lies at least 1/(2n) >= 1/256 from an integer, far
        // beyond the error of v_rcp_f32, so the truncations below are exact
        const float inx = __builtin_amdgcn_rcpf((float)ncx), inxy = __builtin_amdgcn_rcpf((float)(ncx * ncy));
        for (int cb = 0; cb < ncell; cb += 64) {
            const int c = cb + lane;
            int cnt = 0, b = 0, upk = 0;
            float lx = 0.0f, ly = 0.0f, lz = 0.0f;
            if (c < ncell) {
                // c = dx + ncx (dy + ncy dz); c < 128 so the float quotients are exact
                const int dz = (int)(((float)c + 0.5f) * inxy);
                const int r = c - dz * (ncx * ncy);
                const int dy = (int)(((float)r + 0.5f) * inx);
                const int dx = r - dy * ncx;
                lx = F.cellf[0] * (float)(dx - 1) - ccx;
                ly = F.cellf[1] * (float)(dy - 1) - ccy;
                lz = F.cellf[2] * (float)(dz - 1) - ccz;
                const float gx = fmaxf(0.0f, fmaxf(lx - hx, -hx - (lx + F.cellf[0])));
                const float gy = fmaxf(0.0f, fmaxf(ly - hy, -hy - (ly + F.cellf[1])));
                const float gz = fmaxf(0.0f, fmaxf(lz - hz, -hz - (lz + F.cellf[2])));
                if (fmaf(gx, gx, fmaf(gy, gy, gz * gz)) <= F.rc_cull2) {
                    int rx = ux0 + dx, ry = uy0 + dy, rz = uz0 + dz;
                    rx = rx < 0 ? rx + P.Ex : (rx >= P.Ex ? rx - P.Ex : rx);
                    ry = ry < 0 ? ry + P.Ey : (ry >= P.Ey ? ry - P.Ey : ry);
                    rz = rz < 0 ? rz + P.Ez : (rz >= P.Ez ? rz - P.Ez : rz);
                    const int rc = rx + P.Ex * (ry + P.Ey * rz);
                    b = cell_start[rc];
                    cnt = cell_start[rc + 1] - b;
                    upk = dx | (dy << 8) | (dz << 16);
                }
            }
            const int incl = wave_incl_scan(cnt);
            const unsigned long long m = __ballot(cnt > 0);
            if (cnt > 0) {
                const int k = K + below(m);
                const int st = T + incl - cnt;
                cstart[k] = st;
                crec[k] = make_float4(lx, ly, lz, __int_as_float(b - st));
                cupk[k] = upk;
            }
            K += (int)__popcll(m);
            T += __shfl(incl, 63, 64);
        }
    }
    __syncthreads();
    T = __builtin_amdgcn_readfirstlane(T);
    K = __builtin_amdgcn_readfirstlane(K);
    // starts of kept cells lane and lane + 64 (K <= 128)
    const int s_lo = lane < K ? cstart[lane] : 0x7fffffff;
    const int s_hi = lane + 64 < K ? cstart[lane + 64] : 0x7fffffff;
    __syncthreads();

    // exact FP64 test of neighbourList1b (Lib_LinkCell3D.f90:1565-1596) for
    // a staged candidate (packed unwrapped cell offsets u, point slot) and
    // the voxel of x index vx (cell offset lx) in the lane's row
    auto exact_at = [&](int lx, int vx, int u, int slot) -> bool {
        const int kx = (u & 255) - 1 - lx, ky = ((u >> 8) & 255) - 1 - lcy, kz = (u >> 16) - 1 - lcz;
        if (kx < -1 || kx > 1 || ky < -1 || ky > 1 || kz < -1 || kz > 1) return false;
        const double4 f = slot_frac[slot];                              // FP64 offset in cell
        const double zix = P.ax.zi[0][vx], ziy = P.ax.zi[1][jy], ziz = P.ax.zi[2][jz];
        const double d1 = P.a[0] * (f.x + (kx - zix));
        const double d2 = P.a[1] * (f.y + (ky - ziy));
        const double d3 = P.a[2] * (f.z + (kz - ziz));
        return ((d1 * d1 + d2 * d2) + d3 * d3) <= P.rc2;
    };
    auto kahan_add = [](float& s, float& c, float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    };
    float sre = 0.0f, sim = 0.0f, cre = 0.0f, cim = 0.0f;
    float emax = -3.4028234663852886e38f;
    int cnt = 0;
    unsigned nguard = 0;
    // ROW: sums per aperture point (with their Kahan compensations) and
    // largest u of the row's voxels over this lane's part of the candidates
    float rre[NA][TX], rim[NA][TX], kre[NA][TX], kim[NA][TX], umax[TX];
#pragma unroll
    for (int i = 0; i < TX; i++) {
#pragma unroll
        for (int a = 0; a < NA; a++) {
            rre[a][i] = 0.0f; rim[a][i] = 0.0f; kre[a][i] = 0.0f; kim[a][i] = 0.0f;
        }
        umax[i] = -1.0f;
    }

    int t_next = 0, kbase = 0;
    while (t_next < T) {
        // stage: compact candidates within 3 sigma of the tile box
        int n = 0;
        while (t_next < T && n <= CAP - 64) {
            // cell of candidate t_next + lane
            sflag[lane] = 0;
            __syncthreads();
            if (s_lo >= t_next && s_lo < t_next + 64) sflag[s_lo - t_next] = 1;
            if (s_hi >= t_next && s_hi < t_next + 64) sflag[s_hi - t_next] = 1;
            __syncthreads();
            const unsigned long long sm = __ballot(sflag[lane] != 0);
            const int k = kbase + below(sm) + ((sm >> lane) & 1ull ? 1 : 0) - 1;
            kbase += (int)__popcll(sm);
            const int t = t_next + lane;
            bool keep = false;
            float4 r0, r1, r2, r0b, rph[NPH > 0 ? NPH : 1];
            if (t < T) {
                const float4 cr = crec[k];
                const int slot = t + __float_as_int(cr.w);
                const float4 pc = p32[slot];
                const float px = pc.x + cr.x, py = pc.y + cr.y, pz = pc.z + cr.z;
                const float ex = fmaxf(0.0f, fabsf(px) - hx);
                const float ey = fmaxf(0.0f, fabsf(py) - hy);
                const float ez = fmaxf(0.0f, fabsf(pz) - hz);
                keep = fmaf(ex, ex, fmaf(ey, ey, ez * ez)) <= F.rc_cull2;
                if (keep) {
                    const float m2k = -2.0f * k2;
                    if (ROW) {
                        float xv[TX];
                        const float kyzp = k2 * fmaf(py, py, pz * pz);
#pragma unroll
                        for (int i = 0; i < TX; i++) {
                            const float d = px - qxt[i];
                            xv[i] = __builtin_amdgcn_exp2f(fmaf(k2, d * d, kyzp));
                        }
                        r0 = make_float4(xv[0], xv[1], xv[2], xv[3]);
                        if (TX == 8) r0b = make_float4(xv[4 % TX], xv[5 % TX], xv[6 % TX], xv[7 % TX]);
                        r1 = make_float4(m2k * py, m2k * pz, NA == 1 ? sn32[slot] : 0.0f, pc.w);
                        r2 = make_float4(NA == 1 ? sn32[slot] : 0.0f, __int_as_float(slot),
                                         __int_as_float(cupk[k]), 0.0f);
#pragma unroll
                        for (int h = 0; h < NPH; h++) {
                            const float2 c0 = F.app_cs[(2 * h) * F.app_stride + slot];
                            const float2 c1 = F.app_cs[(2 * h + 1) * F.app_stride + slot];
                            rph[h] = make_float4(c0.x, c0.y, c1.x, c1.y);
                        }
                    } else {
                        r0 = make_float4(m2k * px, m2k * py, m2k * pz, k2 * fmaf(px, px, fmaf(py, py, pz * pz)));
                        r1 = make_float4(pc.w, sn32[slot], __int_as_float(slot), __int_as_float(cupk[k]));
                    }
                }
            }
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const int p = n + below(m);
                sA[REC * p] = r0;
                if (ROW && TX == 8) sA[REC * p + 1] = r0b;
                sA[REC * p + (ROW ? XR : 1)] = r1;
                if (ROW) sA[REC * p + XR + 1] = r2;
#pragma unroll
                for (int h = 0; h < NPH; h++) sA[REC * p + XR + 2 + h] = rph[h];
            }
            n += (int)__popcll(m);
            t_next += 64;
        }
        n = __builtin_amdgcn_readfirstlane(n);
        if (ROW && lane < 4) {
            // zero records after the last candidate: X = 0, never a member
            // and never in the guard band (u = -1)
            const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int k = 0; k < REC; k++) sA[REC * (n + lane) + k] = z4;
        }
        __syncthreads();
        if (DIAG) {
            // every voxel against every staged candidate, per-pair exact
            // membership with the counts and nearest ranges of the diagnostics
            for (int jb = 0; jb < n; jb += 16) {
                const int je = min(jb + 16, n);
                float pr = 0.0f, pi = 0.0f;
                for (int j = jb; j < je; j++) {
                    const float4 A = sA[REC * j];
                    const float4 C = sA[REC * j + 1];
                    const float e = fmaf(A.x, qcx, fmaf(A.y, qcy, fmaf(A.z, qcz, A.w + kq2)));
                    bool in = act && e >= F.e_rc;
                    if (act && fabsf(e - F.e_rc) <= F.e_band) {
                        nguard++;
                        in = exact_at(lcx, jx, __float_as_int(C.w), __float_as_int(C.z));
                    }
                    if (F.thin_z) in = in && (unsigned)((__float_as_int(C.w) >> 16) - lcz) <= 2u;
                    const float w = in ? __builtin_amdgcn_exp2f(e) : 0.0f;
                    pr = fmaf(w, C.x, pr);
                    pi = fmaf(w, C.y, pi);
                    emax = in ? fmaxf(emax, e) : emax;
                    cnt += in ? 1 : 0;
                }
                kahan_add(sre, cre, pr);
                kahan_add(sim, cim, pi);
            }
        } else if (!ROW) {
            // Branch-free pass over blocks of 32 candidates in the frame
            // t = e - k2 (3 sigma)^2 (member iff t >= 0, weight 2^t scaled by
            // 2^-e_rc), tracking the smallest |t|.  A block in which some
            // lane came within the guard band is summed again with the exact
            // FP64 test for those pairs (uniform branch, rare).
            const float kq2r = kq2 - F.e_rc;
            for (int jb = 0; jb < n; jb += 32) {
                const int je = min(jb + 32, n);
                float pr = 0.0f, pi = 0.0f, dmin = 3.4028234663852886e38f;
#pragma unroll 4
                for (int j = jb; j < je; j++) {
                    const float4 A = sA[REC * j];
                    const float4 C = sA[REC * j + 1];
                    const float t = fmaf(A.x, qcx, fmaf(A.y, qcy, fmaf(A.z, qcz, A.w + kq2r)));
                    const float w = t >= 0.0f ? __builtin_amdgcn_exp2f(t) : 0.0f;
                    pr = fmaf(w, C.x, pr);
                    pi = fmaf(w, C.y, pi);
                    dmin = fminf(dmin, fabsf(t));
                    if (DENSITY) emax = fmaxf(emax, t);
                }
                if (__ballot(dmin <= F.e_band)) {
                    pr = 0.0f; pi = 0.0f;
                    for (int j = jb; j < je; j++) {
                        const float4 A = sA[REC * j];
                        const float4 C = sA[REC * j + 1];
                        const float t = fmaf(A.x, qcx, fmaf(A.y, qcy, fmaf(A.z, qcz, A.w + kq2r)));
                        bool in = t >= 0.0f;
                        if (fabsf(t) <= F.e_band) {
                            nguard++;
                            in = exact_at(lcx, jx, __float_as_int(C.w), __float_as_int(C.z));
                        }
                        const float w = in ? __builtin_amdgcn_exp2f(t) : 0.0f;
                        pr = fmaf(w, C.x, pr);
                        pi = fmaf(w, C.y, pi);
                    }
                }
                kahan_add(sre, cre, pr);
                kahan_add(sim, cim, pi);
            }
        } else {
            // ROW: steps of NQ candidates (one per part), blocks of ROW_BLOCK
            // steps; the guard band in u is the exponent's band (|u| ~ ln2 |t|
            // near the threshold, so it is 1.44 times as wide).  The member
            // weight is X E = u + 1, taken as w = s (u + 1) with the step
            // s = clamp(2^40 u) in [0, 1]: 1 for u >= 2^-40, 0 for u <= 0, and
            // fractional only inside the guard band, whose pairs the block's
            // re-walk settles.  Each block's sums are added to the running
            // sums with Kahan compensation, and so are the re-walk's
            // corrections (the block sums are not live across the re-walk,
            // which sets the register peak).
            const int nst = (n + NQ - 1) / NQ;
            auto step = [](float u) { return __builtin_amdgcn_fmed3f(u * 1099511627776.0f, 0.0f, 1.0f); };
            // the X factors of staged candidate j
            auto load_x = [&](int j, float* xv) {
#pragma unroll
                for (int c = 0; c < XR; c++) {
                    const float4 X = sA[REC * j + c];
                    xv[4 * c] = X.x; xv[4 * c + 1] = X.y; xv[4 * c + 2] = X.z; xv[4 * c + 3] = X.w;
                }
            };
            // the (cos, sin) pairs of staged candidate j (B = its record XR,
            // sn its sine when NA = 1)
            auto load_ph = [&](int j, const float4& B, float sn, float* pc, float* ps) {
                if (NA == 1) {
                    pc[0] = B.w;
                    ps[0] = sn;
                } else {
#pragma unroll
                    for (int h = 0; h < NPH; h++) {
                        const float4 Ph = sA[REC * j + XR + 2 + h];
                        pc[2 * h] = Ph.x; ps[2 * h] = Ph.y; pc[2 * h + 1] = Ph.z; ps[2 * h + 1] = Ph.w;
                    }
                }
            };
            for (int kb = 0; kb < nst; kb += ROW_BLOCK) {
                const int ke = min(kb + ROW_BLOCK, nst);
                float dmin = 3.4028234663852886e38f;
                float bre[NA][TX], bim[NA][TX];
#pragma unroll
                for (int i = 0; i < TX; i++)
#pragma unroll
                    for (int a = 0; a < NA; a++) { bre[a][i] = 0.0f; bim[a][i] = 0.0f; }
#pragma unroll 2
                for (int kk = kb; kk < ke; kk++) {
                    const int j = NQ * kk + squart;
                    float xv[TX];
                    load_x(j, xv);
                    const float4 B = sA[REC * j + XR];
                    float E = __builtin_amdgcn_exp2f(fmaf(B.x, qcy, fmaf(B.y, qcz, kyz)));
                    const float sn = B.z;
                    if (ROW == 2) {
                        const float4 C = sA[REC * j + XR + 1];
                        E = (unsigned)((__float_as_int(C.z) >> 16) - lcz) <= 2u ? E : 0.0f;
                    }
                    float pc[NA], ps[NA];
                    load_ph(j, B, sn, pc, ps);
#pragma unroll
                    for (int i = 0; i < TX; i++) {
                        const float u = fmaf(xv[i], E, -1.0f);
                        const float s = step(u);
                        const float w = fmaf(s, u, s);
#pragma unroll
                        for (int a = 0; a < NA; a++) {
                            bre[a][i] = fmaf(w, pc[a], bre[a][i]);
                            bim[a][i] = fmaf(w, ps[a], bim[a][i]);
                        }
                        dmin = fminf(dmin, fabsf(u));
                        if (DENSITY) umax[i] = fmaxf(umax[i], u);
                    }
                }
#pragma unroll
                for (int i = 0; i < TX; i++)
#pragma unroll
                    for (int a = 0; a < NA; a++) {
                        kahan_add(rre[a][i], kre[a][i], bre[a][i]);
                        kahan_add(rim[a][i], kim[a][i], bim[a][i]);
                    }
                if (__ballot(dmin <= F.e_band)) {
                    for (int kk = kb; kk < ke; kk++) {
                        const int j = NQ * kk + squart;
                        float xv[TX];
                        load_x(j, xv);
                        const float4 B = sA[REC * j + XR];
                        const float4 C = sA[REC * j + XR + 1];
                        float E = __builtin_amdgcn_exp2f(fmaf(B.x, qcy, fmaf(B.y, qcz, kyz)));
                        if (ROW == 2) E = (unsigned)((__float_as_int(C.z) >> 16) - lcz) <= 2u ? E : 0.0f;
#pragma unroll
                        for (int i = 0; i < TX; i++) {
                            const float u = fmaf(xv[i], E, -1.0f);
                            if (fabsf(u) <= F.e_band) {
                                const int vx = min(x0 + i, x1);
                                // one re-test per aperture point it settles
                                if (row_act && x0 + i <= x1) nguard += NA;
                                const bool in = exact_at(qlx_tab[vx].lc, vx, __float_as_int(C.z),
                                                         __float_as_int(C.y));
                                // replace the pass's s (u + 1) by the exact test's weight
                                const float s = step(u);
                                const float d = (in ? u + 1.0f : 0.0f) - fmaf(s, u, s);
                                float pc[NA], ps[NA];
                                load_ph(j, B, C.x, pc, ps);
#pragma unroll
                                for (int a = 0; a < NA; a++) {
                                    kahan_add(rre[a][i], kre[a][i], d * pc[a]);
                                    kahan_add(rim[a][i], kim[a][i], d * ps[a]);
                                }
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
    }

    // the lane's voxels: ROW, voxels s NV .. s NV + NV - 1 of its row, from
    // the parts' sums added in part order through LDS; otherwise its own
    float vre[NA][NV], vim[NA][NV], vem[NV];
    if (ROW) {
        float2* red = reinterpret_cast<float2*>(sA);                  // [64 lanes][TX voxels]
        float* redm = reinterpret_cast<float*>(sA) + 2 * 64 * TX;     // [64 lanes][TX voxels]
        const int rb = lane & ~(NQ - 1);
#pragma unroll
        for (int ap = 0; ap < NA; ap++) {
            if (ap > 0) __syncthreads();                              // the previous point's reads
#pragma unroll
            for (int i = 0; i < TX; i++) {
                red[TX * lane + i] = make_float2(rre[ap][i], rim[ap][i]);
                if (DENSITY && ap == 0) redm[TX * lane + i] = umax[i];
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < NV; t++) {
                const int i = squart * NV + t;
                float a = 0.0f, b = 0.0f, um = -1.0f;
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const float2 v = red[TX * (rb + q) + i];
                    a += v.x;
                    b += v.y;
                    if (DENSITY && ap == 0) um = fmaxf(um, redm[TX * (rb + q) + i]);
                }
                vre[ap][t] = a;
                vim[ap][t] = b;
                // largest exponent in the t frame: 2^t = u + 1 (u = -1: no candidate)
                if (ap == 0)
                    vem[t] = (DENSITY && um > -1.0f) ? __builtin_amdgcn_logf(um + 1.0f) : -3.4028234663852886e38f;
            }
        }
    } else {
        vre[0][0] = sre;
        vim[0][0] = sim;
        vem[0] = emax;
    }

#pragma unroll
    for (int t = 0; t < NV; t++) {
        const int vx = ROW ? x0 + squart * NV + t : ix;
        if (!(ROW ? vx <= x1 && row_act : act)) continue;
        const int64_t o = (int64_t)vx + (int64_t)P.Mx * (iy + (int64_t)P.My * iz);
        const float emax_ = vem[t];
        // nearest squared range from the largest exponent (huge(1.0) if none).
        // Without DIAG the largest exponent is over every candidate, members
        // or not: it differs from the members' only when no candidate is
        // nearer than about 3 sigma, where the density is 0 either way.
        float r2m;
        if (DIAG) r2m = emax_ > -3.4028234663852886e38f ? emax_ / k2 : 3.4028234663852886e38f;
        else r2m = emax_ > -3.4028234663852886e38f ? (emax_ + F.e_rc) * F.inv_k2 : 3.4028234663852886e38f;
        float rh = 1.0f;
        if (DENSITY && P.use_density) {
            // convertScaledRangeToRho (Lib_DynamicalTwoBeamImaging.f90:1973-1994) in FP32
            const float sg = (float)P.sigma;
            if (r2m < sg * sg) rh = 1.0f;
            else if (r2m > 4 * sg * sg) rh = 0.0f;
            else {
                const float xx = sqrtf(r2m) / sg - 1.0f;
                rh = 1.0f - xx * xx * xx * fmaf(6.0f * xx, xx, fmaf(-15.0f, xx, 10.0f));
            }
        }
#pragma unroll
        for (int ap = 0; ap < NA; ap++) {
            const float sre_ = vre[ap][t], sim_ = vim[ap][t];
            // x = S / max(|S|, 1e-16).  Without DIAG the sums carry the factor
            // 2^-e_rc of the t frame (weights 2^t = 2^(e - e_rc)); it cancels in
            // S/|S| and moves the threshold.
            const float m2 = fmaf(sre_, sre_, sim_ * sim_);
            const float ww = m2 > (DIAG ? 1.0e-32f : F.m2_min) ? __builtin_amdgcn_rsqf(m2)
                                                              : (DIAG ? 1.0e16f : 1.0e16f * F.w_scale);
            const int64_t oa = o + ap * F.field_stride;
            if (DENSITY) field[oa] = make_float4(sre_ * ww, sim_ * ww, rh, 0.0f);
            else reinterpret_cast<float2*>(field)[oa] = make_float2(sre_ * ww, sim_ * ww);
        }
        if (DIAG) {
            out_cnt[o] = cnt;
            out_r2m[o] = r2m;
        }
    }
    // spread over GUARD_SLOTS lines: many waves adding into one address
    // serialise.  ROW lanes counted only their active voxels.
    if ((ROW || act) && nguard)
        atomicAdd(guard_count + GUARD_STRIDE * (blockIdx.x & (GUARD_SLOTS - 1)), (unsigned long long)nguard);
}

}  // namespace d2bi
