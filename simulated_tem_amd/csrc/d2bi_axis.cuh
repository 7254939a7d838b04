// d2bi_axis.cuh -- per-axis voxel tables (link cell, offset in the cell, FP32
// position of each voxel coordinate), uploaded once per condition.
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
#pragma once

namespace d2bi {

// ------------------------------------------------------------------ voxel axis tables

// The voxel and link cells are cuboidal (findTransformedSupercell gives a
// diagonal D, Lib_DiffractionConditions.f90:1074-1077), so a voxel's link
// cell and in-cell offset factorise per axis: for axis x,
//   xx = a_vox(1,1)*ix - lc_delta(1),  zi = ia(1,1)*xx,
//   cell = mod(floor(zi) + 16*Ex, Ex), zi = zi - cell
// (neighbourList1b, Lib_LinkCell3D.f90:1549-1565).  The zero off-diagonal
// terms of the reference's 3x3 products add +-0 and change no bit.
static constexpr int TILE = 4;                       // FAST voxel tiles: 4 voxels along x and y (z: 4 or 8)

// FP32 voxel position along one axis relative to the link cell of its
// FAST tile's first voxel, and that cell offset
struct QL {
    float q;
    int lc;
};

// Voxel positions relative to the first link cell of their tile: table k
// holds axis QL_AXIS[k] for tiles of QL_TILE[k] voxels along it.
static constexpr int NQL = 5;
static constexpr int QL_AXIS[NQL] = {0, 1, 2, 2, 0};
static constexpr int QL_TILE[NQL] = {4, 4, 4, 8, 8};

struct AxisTables {
    std::vector<int> cell[3];
    std::vector<double> zi[3];
    std::vector<QL> ql[NQL];       // x, y, z for 4-voxel tiles; z, x for 8-voxel tiles
    std::vector<int> start[3];     // first voxel of each link cell (size E+1)
    int lo[3], hi[3];              // active link-cell range
};

static bool diagonal(const double* m)
{
    return m[1] == 0.0 && m[2] == 0.0 && m[3] == 0.0 && m[5] == 0.0 && m[6] == 0.0 && m[7] == 0.0;
}

static int axis_tables(const d2bi_record* r, const d2bi_plan* p, AxisTables* t)
{
    if (!diagonal(r->gk_a) || !diagonal(p->a_col) || !diagonal(p->ia_col))
        return seterr(D2BI_ERR_ARG, "voxel/link cells must be cuboidal (diagonal cell matrices)");
    const int M[3] = {p->Mx, p->My, p->Mz};
    const int E[3] = {p->Ex, p->Ey, p->Ez};
    for (int a = 0; a < 3; a++) {
        t->cell[a].resize(M[a]);
        t->zi[a].resize(M[a]);
        t->start[a].assign(E[a] + 1, M[a]);
        const double av = r->gk_a[4 * a], ia = p->ia_col[4 * a];
        for (int i = 0; i < M[a]; i++) {
            double xx = av * i - p->lc_delta[a];
            double zi = ia * xx;
            int c = ((int)floor(zi) + 16 * E[a]) % E[a];
            t->cell[a][i] = c;
            t->zi[a][i] = zi - c;
        }
        for (int i = 1; i < M[a]; i++)
            if (t->cell[a][i] < t->cell[a][i - 1])
                return seterr(D2BI_ERR_ARG, "voxel column wraps the link-cell box (axis %d)", a);
        const double side = p->a_col[4 * a];
        for (int k = 0; k < NQL; k++) {
            if (QL_AXIS[k] != a) continue;
            const int ts = QL_TILE[k];
            t->ql[k].resize(M[a]);
            for (int i = 0; i < M[a]; i++) {
                const int lc = t->cell[a][i] - t->cell[a][(i / ts) * ts];
                t->ql[k][i] = QL{(float)(side * (t->zi[a][i] + lc)), lc};
            }
        }
        for (int i = M[a] - 1; i >= 0; i--) t->start[a][t->cell[a][i]] = i;
        // cells with no voxels take the start of the next occupied cell
        for (int c = E[a] - 1; c >= 0; c--)
            if (t->start[a][c] > t->start[a][c + 1]) t->start[a][c] = t->start[a][c + 1];
        t->lo[a] = t->cell[a][0];
        t->hi[a] = t->cell[a][M[a] - 1];
    }
    return D2BI_OK;
}

// Upload tables: ints [cellx|celly|cellz], doubles [zix|ziy|ziz], ints [startx|starty|startz]
struct AxisDev {
    const int* cell[3];
    const double* zi[3];
    const int* start[3];
    const QL* ql[NQL];
};

// Copy the axis tables into workspace w.  t must stay alive until the
// stream has executed the copies.
static int upload_axis(const AxisTables& t, AxisDev* d, hipStream_t s, Workspace& w)
{
    size_t nc = t.cell[0].size() + t.cell[1].size() + t.cell[2].size();
    size_t ns = t.start[0].size() + t.start[1].size() + t.start[2].size();
    size_t nq = 0;
    for (int k = 0; k < NQL; k++) nq += t.ql[k].size();
    CHK(w.ax_cell.ensure(sizeof(int) * nc));
    CHK(w.ax_zi.ensure(sizeof(double) * nc));
    CHK(w.ax_start.ensure(sizeof(int) * ns));
    CHK(w.ax_ql.ensure(sizeof(QL) * nq));
    if (w.ax_cell.arena) {
        // a workspace that must not allocate: copy from the tables
        // themselves, not through the (growable) pinned staging buffer
        size_t oc = 0, os = 0, oq = 0;
        for (int k = 0; k < NQL; k++) {
            d->ql[k] = w.ax_ql.as<QL>() + oq;
            HIPCHK(hipMemcpyAsync((void*)d->ql[k], t.ql[k].data(), sizeof(QL) * t.ql[k].size(),
                                  hipMemcpyHostToDevice, s));
            oq += t.ql[k].size();
        }
        for (int a = 0; a < 3; a++) {
            d->cell[a] = w.ax_cell.as<int>() + oc;
            d->zi[a] = w.ax_zi.as<double>() + oc;
            d->start[a] = w.ax_start.as<int>() + os;
            HIPCHK(hipMemcpyAsync((void*)d->cell[a], t.cell[a].data(), sizeof(int) * t.cell[a].size(),
                                  hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync((void*)d->zi[a], t.zi[a].data(), sizeof(double) * t.zi[a].size(),
                                  hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync((void*)d->start[a], t.start[a].data(), sizeof(int) * t.start[a].size(),
                                  hipMemcpyHostToDevice, s));
            oc += t.cell[a].size();
            os += t.start[a].size();
        }
        return D2BI_OK;
    }
    size_t bytes = sizeof(int) * nc + sizeof(double) * nc + sizeof(int) * ns + sizeof(QL) * nq;
    void* hp;
    CHK(pinned(bytes, &hp));
    char* h = (char*)hp;
    int* hc = (int*)h;
    double* hz = (double*)(h + sizeof(int) * nc);
    int* hs = (int*)(h + sizeof(int) * nc + sizeof(double) * nc);
    QL* hq = (QL*)(h + sizeof(int) * nc + sizeof(double) * nc + sizeof(int) * ns);
    size_t oc = 0, os = 0, oq = 0;
    for (int k = 0; k < NQL; k++) {
        std::memcpy(hq + oq, t.ql[k].data(), sizeof(QL) * t.ql[k].size());
        d->ql[k] = w.ax_ql.as<QL>() + oq;
        oq += t.ql[k].size();
    }
    for (int a = 0; a < 3; a++) {
        std::memcpy(hc + oc, t.cell[a].data(), sizeof(int) * t.cell[a].size());
        std::memcpy(hz + oc, t.zi[a].data(), sizeof(double) * t.zi[a].size());
        std::memcpy(hs + os, t.start[a].data(), sizeof(int) * t.start[a].size());
        d->cell[a] = w.ax_cell.as<int>() + oc;
        d->zi[a] = w.ax_zi.as<double>() + oc;
        d->start[a] = w.ax_start.as<int>() + os;
        oc += t.cell[a].size();
        os += t.start[a].size();
    }
    HIPCHK(hipMemcpyAsync(w.ax_cell.p, hc, sizeof(int) * nc, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(w.ax_zi.p, hz, sizeof(double) * nc, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(w.ax_start.p, hs, sizeof(int) * ns, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(w.ax_ql.p, hq, sizeof(QL) * nq, hipMemcpyHostToDevice, s));
    // the pinned buffer is reused by later calls: they wait for this event
    HIPCHK(hipEventRecord(g.ev[9], s));
    g.pinned_busy = true;
    return D2BI_OK;
}

struct VoxParams {
    int Mx, My, Mz;
    int Ex, Ey, Ez;
    double a[3];                // link-cell sides (diagonal of a_col)
    double rc2, i2s2, sigma;
    int use_density;
    AxisDev ax;
};

// convertScaledRangeToRho (Lib_DynamicalTwoBeamImaging.f90:1973-1994)
__device__ inline double rho_of(double r2_min, double sig)
{
    if (r2_min < sig * sig) return 1.0;
    if (r2_min > 4 * sig * sig) return 0.0;
    double xx = sqrt(r2_min) / sig - 1.0;
    return 1 - ((xx * xx) * xx) * ((10 - 15 * xx) + (6 * xx) * xx);
}

}  // namespace d2bi
