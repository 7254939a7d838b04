// d2bi_blur_query.cuh -- the separable Gaussian blur and the generic link-cell
// neighbour query.
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
#pragma once

namespace d2bi {

// ------------------------------------------------------------------ (d) blur

struct BlurParams {
    int nx, ny, nk;
};

static constexpr int BLUR_TILE = 256;
static constexpr int BLUR_MAXK = 256;

// x pass of gaussianBlurImage (d2bi.f90:1577-1599): row segments staged in
// LDS with their halo; the kernel weights come from the host (exp in FP64).
__global__ __launch_bounds__(BLUR_TILE) void k_blur_x(BlurParams B, const double* __restrict__ fin,
                                                      const double* __restrict__ kern,
                                                      double* __restrict__ tmp, double* __restrict__ wsx)
{
    __shared__ double row[BLUR_TILE + 2 * BLUR_MAXK];
    const int iy = blockIdx.y;
    const int x0 = blockIdx.x * BLUR_TILE;
    const int lo = x0 - B.nk, n = BLUR_TILE + 2 * B.nk;
    for (int t = threadIdx.x; t < n; t += BLUR_TILE) {
        const int gx = lo + t;
        row[t] = (gx >= 0 && gx < B.nx) ? fin[(int64_t)gx + (int64_t)B.nx * iy] : 0.0;
    }
    __syncthreads();
    const int ix = x0 + threadIdx.x;
    if (ix >= B.nx) return;
    const int j0 = max(0, ix - B.nk), j1 = min(B.nx - 1, ix + B.nk);
    double wf = 0.0, ws = 0.0;
    for (int jj = j0; jj <= j1; jj++) {
        const double ww = kern[abs(jj - ix)];
        wf = wf + ww * row[jj - lo];
        ws = ws + ww;
    }
    tmp[(int64_t)ix + (int64_t)B.nx * iy] = wf;
    if (iy == 0) wsx[ix] = 1 / ws;
}

// y pass (d2bi.f90:1605-1619): out = sum_y k * tmp * (1/W_x) / W_y.
__global__ __launch_bounds__(256) void k_blur_y(BlurParams B, const double* __restrict__ tmp,
                                                const double* __restrict__ kern,
                                                const double* __restrict__ wsx, double* __restrict__ out)
{
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ix >= B.nx || iy >= B.ny) return;
    const int j0 = max(0, iy - B.nk), j1 = min(B.ny - 1, iy + B.nk);
    double wf = 0.0, ws = 0.0;
    for (int jj = j0; jj <= j1; jj++) {
        const double ww = kern[abs(jj - iy)];
        wf = wf + ww * tmp[(int64_t)ix + (int64_t)B.nx * jj];
        ws = ws + ww;
    }
    out[(int64_t)ix + (int64_t)B.nx * iy] = (wf * wsx[ix]) / ws;
}

// ------------------------------------------------------------------ generic link-cell query

struct QueryParams {
    double a[9], ia[9];
    int Nx, Ny, Nz;
    double rc2;
    int maxNeigh;
};

// neighbourList1b (Lib_LinkCell3D.f90:1515-1605) for arbitrary query points
// and a general (not necessarily cuboidal) cell.
__global__ void k_lc_query(QueryParams Q, int64_t nq, const double* __restrict__ q,
                           const int* __restrict__ cell_start, const int* __restrict__ slot_id,
                           const double4* __restrict__ frac, int* __restrict__ nout,
                           int* __restrict__ idout, double* __restrict__ r2out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const double x[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    double zi[3];
    for (int k = 0; k < 3; k++) zi[k] = (Q.ia[k] * x[0] + Q.ia[3 + k] * x[1]) + Q.ia[6 + k] * x[2];
    const int ix = cell_of(floor(zi[0]), Q.Nx);
    const int iy = cell_of(floor(zi[1]), Q.Ny);
    const int iz = cell_of(floor(zi[2]), Q.Nz);
    zi[0] = zi[0] - ix; zi[1] = zi[1] - iy; zi[2] = zi[2] - iz;
    int n = 0;
    for (int kz = -1; kz <= 1; kz++) {
        const int jz = (iz + Q.Nz + kz) % Q.Nz;
        const double dzi3 = kz - zi[2];
        for (int ky = -1; ky <= 1; ky++) {
            const int jy = (iy + Q.Ny + ky) % Q.Ny;
            const double dzi2 = ky - zi[1];
            for (int kx = -1; kx <= 1; kx++) {
                const int jx = (ix + Q.Nx + kx) % Q.Nx;
                const double dzi1 = kx - zi[0];
                const int c = jx + Q.Nx * (jy + Q.Ny * jz);
                for (int s = cell_start[c]; s < cell_start[c + 1]; s++) {
                    const double4 f = frac[s];
                    const double dzj1 = f.x + dzi1, dzj2 = f.y + dzi2, dzj3 = f.z + dzi3;
                    double d[3];
                    for (int k = 0; k < 3; k++) d[k] = (Q.a[k] * dzj1 + Q.a[3 + k] * dzj2) + Q.a[6 + k] * dzj3;
                    const double dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
                    if (dd <= Q.rc2) {
                        if (n < Q.maxNeigh) {
                            idout[i * Q.maxNeigh + n] = slot_id[s];
                            r2out[i * Q.maxNeigh + n] = dd;
                        }
                        n++;
                    }
                }
            }
        }
    }
    nout[i] = n;
}

}  // namespace d2bi
