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
// LDS with their halo, and the kernel weights (exp in FP64 on the host) in
// LDS.  Every lane runs the taps t = -nk..nk, skipping those outside the row,
// so each output adds its terms in the reference's order (jj ascending).  The
// edge weight sums W_x, W_y depend only on the position and come from the host.
__global__ __launch_bounds__(BLUR_TILE) void k_blur_x(BlurParams B, const double* __restrict__ fin,
                                                      const double* __restrict__ kern,
                                                      double* __restrict__ tmp)
{
    __shared__ double row[BLUR_TILE + 2 * BLUR_MAXK];
    __shared__ double ks[BLUR_MAXK + 1];
    const int iy = blockIdx.y;
    const int x0 = blockIdx.x * BLUR_TILE;
    const int lo = x0 - B.nk, n = BLUR_TILE + 2 * B.nk;
    for (int t = threadIdx.x; t < n; t += BLUR_TILE) {
        const int gx = lo + t;
        row[t] = (gx >= 0 && gx < B.nx) ? fin[(int64_t)gx + (int64_t)B.nx * iy] : 0.0;
    }
    for (int t = threadIdx.x; t <= B.nk; t += BLUR_TILE) ks[t] = kern[t];
    __syncthreads();
    const int ix = x0 + threadIdx.x;
    if (ix >= B.nx) return;
    const double* r = row + (ix - lo);                  // r[t] = f(ix + t)
    double wf = 0.0;
    if (lo >= 0 && x0 + BLUR_TILE - 1 + B.nk < B.nx) {  // every tap inside the row (uniform)
        for (int t = -B.nk; t < 0; t++) wf = wf + ks[-t] * r[t];
        for (int t = 0; t <= B.nk; t++) wf = wf + ks[t] * r[t];
    } else {
        for (int t = -B.nk; t <= B.nk; t++) {
            const int jj = ix + t;
            if (jj >= 0 && jj < B.nx) wf = wf + ks[t < 0 ? -t : t] * r[t];
        }
    }
    tmp[(int64_t)ix + (int64_t)B.nx * iy] = wf;
}

// y pass (d2bi.f90:1605-1619): out = sum_y k * tmp * (1/W_x) / W_y.  A lane
// finishes BLUR_RY consecutive rows of its column and reads each input row of
// their window once (the rows of a wave are the same, so the tap tests are
// uniform); each output still adds its terms with jj ascending, in the
// reference's order.
static constexpr int BLUR_RY = 8;

__global__ __launch_bounds__(256) void k_blur_y(BlurParams B, const double* __restrict__ tmp,
                                                const double* __restrict__ kern,
                                                const double* __restrict__ wx_inv,
                                                const double* __restrict__ wy, double* __restrict__ out)
{
    __shared__ double ks[BLUR_MAXK + 1];
    for (int t = threadIdx.x; t <= B.nk; t += 256) ks[t] = kern[t];
    __syncthreads();
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    // the wave's first row, held in a scalar register
    const int y0 = __builtin_amdgcn_readfirstlane((blockIdx.y * 4 + (threadIdx.x >> 6)) * BLUR_RY);
    if (ix >= B.nx || y0 >= B.ny) return;
    double wf[BLUR_RY];
#pragma unroll
    for (int k = 0; k < BLUR_RY; k++) wf[k] = 0.0;
    const int r0 = max(0, y0 - B.nk), r1 = min(B.ny - 1, y0 + BLUR_RY - 1 + B.nk);
#pragma unroll 4
    for (int r = r0; r <= r1; r++) {
        const double v = tmp[(int64_t)ix + (int64_t)B.nx * r];
#pragma unroll
        for (int k = 0; k < BLUR_RY; k++) {
            const int t = r - (y0 + k);
            if (t >= -B.nk && t <= B.nk) wf[k] = wf[k] + ks[t < 0 ? -t : t] * v;
        }
    }
    const double wxi = wx_inv[ix];
#pragma unroll
    for (int k = 0; k < BLUR_RY; k++)
        if (y0 + k < B.ny) out[(int64_t)ix + (int64_t)B.nx * (y0 + k)] = (wf[k] * wxi) / wy[y0 + k];
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
