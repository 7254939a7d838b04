// d2bi_fp64.cuh -- the FP64 path: voxelisation over the 27 neighbour cells and
// the RK4 column march in the reference's arithmetic.
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
#pragma once

namespace d2bi {

// ------------------------------------------------------------------ (b) voxelise, FP64

// One thread per voxel; the 27 neighbour cells in (kz,ky,kx) order and the
// points of each cell in insertion order, exactly the neighbour list order
// of neighbourList1b, summed as the reference sums it (:1877-1892).  NA > 1:
// the aperture branch's shared neighbour loop (:1905-1922), one weight per
// member and one sum per aperture point a, whose phases are cs64[a cs_stride
// + slot] and whose fields go to out_re/out_im[a out_stride + voxel]; each
// point's sum takes the same terms in the same order as an NA = 1 pass over
// that point's phases.
template <int NA>
__global__ __launch_bounds__(256) void k_voxel_fp64(VoxParams P, const int* __restrict__ cell_start,
                                                    const double4* __restrict__ frac,
                                                    const double2* __restrict__ cs64, int64_t cs_stride,
                                                    double* __restrict__ out_re,
                                                    double* __restrict__ out_im, int64_t out_stride,
                                                    double* __restrict__ out_rho,
                                                    int* __restrict__ out_cnt,
                                                    double* __restrict__ out_r2m)
{
    const int ix = blockIdx.x * 8 + (threadIdx.x & 7);
    const int iy = blockIdx.y * 8 + ((threadIdx.x >> 3) & 7);
    const int iz = blockIdx.z * 4 + (threadIdx.x >> 6);
    if (ix >= P.Mx || iy >= P.My || iz >= P.Mz) return;
    const int cx = P.ax.cell[0][ix], cy = P.ax.cell[1][iy], cz = P.ax.cell[2][iz];
    const double zx = P.ax.zi[0][ix], zy = P.ax.zi[1][iy], zz = P.ax.zi[2][iz];
    double sre[NA], sim[NA];
#pragma unroll
    for (int a = 0; a < NA; a++) { sre[a] = 0.0; sim[a] = 0.0; }
    double r2min = 3.4028234663852886e38;
    int n = 0;
    for (int kz = -1; kz <= 1; kz++) {
        const int jz = (cz + P.Ez + kz) % P.Ez;
        const double dzi3 = kz - zz;
        for (int ky = -1; ky <= 1; ky++) {
            const int jy = (cy + P.Ey + ky) % P.Ey;
            const double dzi2 = ky - zy;
            for (int kx = -1; kx <= 1; kx++) {
                const int jx = (cx + P.Ex + kx) % P.Ex;
                const double dzi1 = kx - zx;
                const int c = jx + P.Ex * (jy + P.Ey * jz);
                const int s1 = cell_start[c + 1];
                for (int s = cell_start[c]; s < s1; s++) {
                    const double4 f = frac[s];
                    const double d1 = P.a[0] * (f.x + dzi1);
                    const double d2 = P.a[1] * (f.y + dzi2);
                    const double d3 = P.a[2] * (f.z + dzi3);
                    const double dd = (d1 * d1 + d2 * d2) + d3 * d3;
                    if (dd <= P.rc2) {
                        n++;
                        r2min = r2min < dd ? r2min : dd;
                        const double e = exp(-dd * P.i2s2);
#pragma unroll
                        for (int a = 0; a < NA; a++) {
                            const double2 cs = cs64[a * cs_stride + s];
                            sre[a] = sre[a] + e * cs.x;
                            sim[a] = sim[a] + e * cs.y;
                        }
                    }
                }
            }
        }
    }
    const int64_t o = (int64_t)ix + (int64_t)P.Mx * (iy + (int64_t)P.My * iz);
#pragma unroll
    for (int a = 0; a < NA; a++) {
        double ww = hypot(sre[a], sim[a]);
        ww = 1 / (1.0e-16 > ww ? 1.0e-16 : ww);
        out_re[a * out_stride + o] = sre[a] * ww;
        out_im[a * out_stride + o] = sim[a] * ww;
    }
    out_rho[o] = P.use_density ? rho_of(r2min, P.sigma) : 1.0;
    if (out_cnt) out_cnt[o] = n;
    if (out_r2m) out_r2m[o] = r2min;
}

// ------------------------------------------------------------------ (c) integrate, FP64

struct IntParams {
    int Mx, My, Mz, nPrec;
    double b0, bg, dz;
    const double* aa;           // device array of 2 pi s_g per precession point
    int accumulate;             // add the column sum to img instead of writing
    double div;                 // written value: sum / div
};

// dynamicalTwoBeamIntegrationColumn_phaserho (Lib_DynamicalTwoBeamImaging.f90:2316-2480):
// one thread per pixel column, per-double-step renormalisation, precession
// average (:1897-1900, 1951).
__global__ __launch_bounds__(256) void k_integrate_fp64(IntParams P, const double* __restrict__ fre,
                                                        const double* __restrict__ fim,
                                                        const double* __restrict__ frho,
                                                        double* __restrict__ img)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)P.Mx * P.My;
    if (col >= plane) return;
    const int nz = P.Mz;
    const double b0 = P.b0, bg = P.bg, dz = P.dz;
    double acc = 0.0;
    for (int prec = 0; prec < P.nPrec; prec++) {
        const double a = P.aa[prec];
        double phi0r = 1.0, phi0i = 0.0, phigr = 0.0, phigi = 0.0;
        double dphi0r, dphi0i, dphigr, dphigi;
        double phi0r1, phi0i1, phigr1, phigi1, phi0r2, phi0i2, phigr2, phigi2;
        // coefficients at sample k (sample Mz is sample 0, :1894-1895)
        auto coef = [&](int k, double& c0, double& ca, double& cgr, double& cgi) {
            const int64_t o = col + plane * (k == nz ? 0 : k);
            const double rh = frho[o];
            c0 = b0 * rh;
            ca = b0 * rh + a;
            cgr = (bg * rh) * fre[o];
            cgi = (bg * rh) * fim[o];
        };
        double c0a, caa, cgra, cgia;
        if (nz >= 2) coef(0, c0a, caa, cgra, cgia);
        for (int ii = 0; ii <= nz - 2; ii += 2) {
            double c0b, cab, cgrb, cgib, c0c, cac, cgrc, cgic;
            coef(ii + 1, c0b, cab, cgrb, cgib);
            coef(ii + 2, c0c, cac, cgrc, cgic);
#define DERIV(c0, ca, cgr, cgi)                                      \
    dphi0r = ((-(c0 * phi0i)) - cgi * phigr) - cgr * phigi;          \
    dphi0i = ((c0 * phi0r) + cgr * phigr) - cgi * phigi;             \
    dphigr = ((cgi * phi0r) - cgr * phi0i) - ca * phigi;             \
    dphigi = ((cgr * phi0r) + cgi * phi0i) + ca * phigr;
            DERIV(c0a, caa, cgra, cgia)
            phi0r2 = phi0r; phi0i2 = phi0i; phigr2 = phigr; phigi2 = phigi;
            phi0r1 = phi0r + (dphi0r * dz) / 3; phi0i1 = phi0i + (dphi0i * dz) / 3;
            phigr1 = phigr + (dphigr * dz) / 3; phigi1 = phigi + (dphigi * dz) / 3;
            phi0r = phi0r + dphi0r * dz; phi0i = phi0i + dphi0i * dz;
            phigr = phigr + dphigr * dz; phigi = phigi + dphigi * dz;
            DERIV(c0b, cab, cgrb, cgib)
            phi0r = phi0r2 + dphi0r * dz; phi0i = phi0i2 + dphi0i * dz;
            phigr = phigr2 + dphigr * dz; phigi = phigi2 + dphigi * dz;
            phi0r1 = phi0r1 + ((dphi0r * 2) * dz) / 3; phi0i1 = phi0i1 + ((dphi0i * 2) * dz) / 3;
            phigr1 = phigr1 + ((dphigr * 2) * dz) / 3; phigi1 = phigi1 + ((dphigi * 2) * dz) / 3;
            DERIV(c0b, cab, cgrb, cgib)
            phi0r = phi0r2 + (dphi0r * 2) * dz; phi0i = phi0i2 + (dphi0i * 2) * dz;
            phigr = phigr2 + (dphigr * 2) * dz; phigi = phigi2 + (dphigi * 2) * dz;
            phi0r1 = phi0r1 + ((dphi0r * 2) * dz) / 3; phi0i1 = phi0i1 + ((dphi0i * 2) * dz) / 3;
            phigr1 = phigr1 + ((dphigr * 2) * dz) / 3; phigi1 = phigi1 + ((dphigi * 2) * dz) / 3;
            DERIV(c0c, cac, cgrc, cgic)
#undef DERIV
            phi0r = phi0r1 + (dphi0r * dz) / 3; phi0i = phi0i1 + (dphi0i * dz) / 3;
            phigr = phigr1 + (dphigr * dz) / 3; phigi = phigi1 + (dphigi * dz) / 3;
            double phimag = ((phi0r * phi0r + phi0i * phi0i) + phigr * phigr) + phigi * phigi;
            phimag = 1 / sqrt(phimag);
            phi0r = phi0r * phimag; phi0i = phi0i * phimag;
            phigr = phigr * phimag; phigi = phigi * phimag;
            c0a = c0c; caa = cac; cgra = cgrc; cgia = cgic;
        }
        acc = acc + (phigr * phigr + phigi * phigi);
    }
    img[col] = P.accumulate ? img[col] + acc : acc / P.div;
}

}  // namespace d2bi
