/*
 * d2bi_oracle.c -- FP64 CPU restatement of d2bi's imaging hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only as the checker.  The product library never links or calls it.
 *
 * Every function restates one routine of the reference (mason-daniel/
 * simulated_tem, Fortran) with the same operation order, evaluated in IEEE
 * double with FMA contraction disabled (-ffp-contract=off), so that floor()
 * membership decisions and r^2 <= (3 sigma)^2 tests are reproducible bit for
 * bit.  The reference is compiled by gfortran -O2 -march=native, which may
 * contract to FMA; the oracle fixes the un-contracted order instead, and the
 * GPU FP64 mode follows the oracle.
 *
 * Pinning (what checks each function; tests/test_oracle_kat.py and
 * tests/test_oracle_pins.py):
 *   orc_wavelength/velocity   KAT testLib_RelativisticElectrons.f90:27-29
 *   orc_det3/orc_inverse3     closed form on diagonal / random matrices (M*N = I)
 *   orc_lc_* (add, neighbours) KAT testLib_LinkCell3D.f90:271-290 (ids, order,
 *                              dx, counts) + O(N^2) brute force on random boxes
 *   orc_estimate_max_count    KAT testLib_LinkCell3D.f90 header (6 per cell)
 *   orc_integrate_column      closed form intensityAtDepthZ
 *                              (Lib_DiffractionConditions.f90:1209-1258),
 *                              norm conservation, empty column -> 0
 *   voxel_sum (x, rho)        perfect bcc crystal with h+k+l even -> one
 *                              constant unit phase factor on every voxel;
 *                              smoothstep values of rho
 *   orc_find_gdotr            brute force: every image of every atom over
 *                              5x5x5 supercell replicas, rotated, offset and
 *                              kept inside the link-cell box, is the oracle's
 *                              point set (cell coordinates and g.r)
 *   orc_compute_img           golden image Data/test/...ng1.00.dat (approx);
 *                              perfect crystal -> every column equals the
 *                              x = 1 column, within RK4 error (2e-6) of
 *                              intensityAtDepthZ
 *   orc_blur                  constant image invariance, delta response
 *                              = normalised separable Gaussian, symmetry
 *   orc_plan                  golden .dat header (|g|, scale, image space)
 *   orc_sunflower             radii sqrt((2k-1)/(2n-1)) and golden-angle steps
 *                              (closed form), all points inside the unit disk
 *   orc_aperture_setup        nApp = 1 and apertureAngle = 0 reduce to orc_plan's
 *                              aa; |gshift - g| = modg * |point|, shift in x, y only
 *   orc_compute_img_aperture  nApp = 1, and apertureAngle = 0 for any nApp,
 *                              reproduce orc_compute_img; g.r per point differs
 *                              from the unshifted one by dg.y exactly
 *                              (the 25 tabulated point sets are the fixture
 *                              tests/golden/aperture_points.json)
 */
#include <math.h>
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "../include/d2bi_gpu.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* Lib_DynamicalTwoBeamImaging.f90:55 and Lib_DiffractionConditions.f90:22 */
static const double PI_IMG = 3.141592653589790;
static const double PI_DC = 3.141592653589790;
/* Lib_RelativisticElectrons.f90:9-13 */
static const double ME = 0.056856301036;
static const double EQ = 1.0;
static const double CL = 2997.92458;
static const double HP = 4.135667696;
#define HBAR (HP / (2 * 3.14159265390))

#define M3(m, i, j) ((m)[3 * ((j) - 1) + ((i) - 1)]) /* 1-based Fortran indexing */

/* ------------------------------------------------------------------ scalars */

/* Lib_RelativisticElectrons.f90:24-36 */
double orc_wavelength(double V)
{
    if (V <= 0) return 0.0;
    double lambda = 2 * ME * EQ * V * (1 + EQ * V / (2 * ME * CL * CL));
    return HP / sqrt(lambda);
}

/* Lib_RelativisticElectrons.f90:40-50 */
double orc_velocity(double V)
{
    double u = 1 + EQ * V / (ME * CL * CL);
    return CL * sqrt(1 - 1 / (u * u));
}

/* Lib_SimpleSupercells.f90:1238-1255: (1/3) * sum of nine cofactor terms */
double orc_det3(const double* M)
{
    double dd[9];
    dd[0] = M3(M,1,1) * (M3(M,2,2) * M3(M,3,3) - M3(M,2,3) * M3(M,3,2));
    dd[1] = M3(M,1,2) * (M3(M,2,3) * M3(M,3,1) - M3(M,2,1) * M3(M,3,3));
    dd[2] = M3(M,1,3) * (M3(M,2,1) * M3(M,3,2) - M3(M,2,2) * M3(M,3,1));
    dd[3] = M3(M,2,1) * (M3(M,3,2) * M3(M,1,3) - M3(M,3,3) * M3(M,1,2));
    dd[4] = M3(M,2,2) * (M3(M,3,3) * M3(M,1,1) - M3(M,3,1) * M3(M,1,3));
    dd[5] = M3(M,2,3) * (M3(M,3,1) * M3(M,1,2) - M3(M,3,2) * M3(M,1,1));
    dd[6] = M3(M,3,1) * (M3(M,1,2) * M3(M,2,3) - M3(M,1,3) * M3(M,2,2));
    dd[7] = M3(M,3,2) * (M3(M,1,3) * M3(M,2,1) - M3(M,1,1) * M3(M,2,3));
    dd[8] = M3(M,3,3) * (M3(M,1,1) * M3(M,2,2) - M3(M,1,2) * M3(M,2,1));
    double s = 0.0;
    for (int i = 0; i < 9; i++) s = s + dd[i];
    return (1.0 / 3.0) * s;
}

/* Lib_SimpleSupercells.f90:1205-1236 */
void orc_inverse3(const double* M, double* N)
{
    double idd = orc_det3(M);
    if (fabs(idd) < 2.2250738585072014e-308) {
        for (int i = 0; i < 9; i++) N[i] = (i % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    idd = 1.0 / idd;
    M3(N,1,1) = (M3(M,2,2) * M3(M,3,3) - M3(M,2,3) * M3(M,3,2)) * idd;
    M3(N,2,1) = (M3(M,2,3) * M3(M,3,1) - M3(M,2,1) * M3(M,3,3)) * idd;
    M3(N,3,1) = (M3(M,2,1) * M3(M,3,2) - M3(M,2,2) * M3(M,3,1)) * idd;
    M3(N,1,2) = (M3(M,1,3) * M3(M,3,2) - M3(M,1,2) * M3(M,3,3)) * idd;
    M3(N,2,2) = (M3(M,1,1) * M3(M,3,3) - M3(M,1,3) * M3(M,3,1)) * idd;
    M3(N,3,2) = (M3(M,1,2) * M3(M,3,1) - M3(M,1,1) * M3(M,3,2)) * idd;
    M3(N,1,3) = (M3(M,1,2) * M3(M,2,3) - M3(M,1,3) * M3(M,2,2)) * idd;
    M3(N,2,3) = (M3(M,1,3) * M3(M,2,1) - M3(M,1,1) * M3(M,2,3)) * idd;
    M3(N,3,3) = (M3(M,1,1) * M3(M,2,2) - M3(M,1,2) * M3(M,2,1)) * idd;
}

/* Fortran matmul(A,B) for 3x3: C(i,j) = sum_k A(i,k) B(k,j), k ascending */
static void matmul3(const double* A, const double* B, double* C)
{
    double T[9];
    for (int j = 1; j <= 3; j++)
        for (int i = 1; i <= 3; i++)
            M3(T,i,j) = (M3(A,i,1) * M3(B,1,j) + M3(A,i,2) * M3(B,2,j)) + M3(A,i,3) * M3(B,3,j);
    memcpy(C, T, sizeof T);
}

/* norm2 of a 3-vector.  gfortran evaluates NORM2 with a scaled sum of
 * squares; for the one-non-zero-component columns of the cuboidal voxel
 * cells both forms return |x| exactly.  The scaled form is restated. */
static double norm2_3(const double* v)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < 3; i++) {
        if (v[i] != 0.0) {
            double ax = fabs(v[i]);
            if (scale < ax) { ssq = 1.0 + ssq * (scale / ax) * (scale / ax); scale = ax; }
            else            { ssq = ssq + (ax / scale) * (ax / scale); }
        }
    }
    return scale * sqrt(ssq);
}

/* getReciprocalLatticeVectors (Lib_DiffractionConditions.f90:1630-1638) */
void orc_recip(const double* A, double* B)
{
    double At[9];
    for (int i = 1; i <= 3; i++) for (int j = 1; j <= 3; j++) M3(At,i,j) = M3(A,j,i);
    orc_inverse3(At, B);
    for (int i = 0; i < 9; i++) B[i] = 2 * PI_DC * B[i];
}

/* reflection1 (Lib_DiffractionConditions.f90:1655-1667) */
static void reflection(const double* B, const double* hkl, double* g)
{
    for (int i = 1; i <= 3; i++)
        g[i - 1] = (M3(B,i,1) * hkl[0] + M3(B,i,2) * hkl[1]) + M3(B,i,3) * hkl[2];
}

/* energy_deviation_parameter (Lib_DiffractionConditions.f90:1265-1299) */
static double energy_deviation_parameter(double v, const double* khat, const double* g)
{
    const double HBAR2ON2M = HBAR * HBAR / (2 * ME);
    double cst = 2 * ME * v / HBAR;
    double eps = (khat[0] * g[0] + khat[1] * g[1]) + khat[2] * g[2];
    eps = cst * eps;
    eps = ((eps + g[0] * g[0]) + g[1] * g[1]) + g[2] * g[2];
    return eps * HBAR2ON2M;
}

/* energyToLengthDeviationParameter (Lib_DiffractionConditions.f90:1194-1204) */
static double e2l_deviation(double v, double eps) { return -eps / (2 * PI_DC * HBAR * v); }

/* getDeviationParameter(this,g) (Lib_DynamicalTwoBeamImaging.f90:268-282) */
double orc_deviation_parameter(double E_keV, const double* g)
{
    static const double khat[3] = {0.0, 0.0, 1.0};
    double modk = 2 * PI_IMG / orc_wavelength(E_keV * 1000);
    double vv = HBAR * modk / ME;
    double eps = energy_deviation_parameter(vv, khat, g);
    return e2l_deviation(vv, eps);
}

/* getAperture (Lib_DynamicalTwoBeamImaging.f90:1137-1156) */
static void get_aperture(const d2bi_record* r, const double* tilt, double* g)
{
    double B[9], RU[9], gc[3];
    orc_recip(r->a_cell, B);
    reflection(B, r->g, gc);
    matmul3(r->R, r->U, RU);
    if (tilt) matmul3(tilt, RU, RU);
    for (int i = 1; i <= 3; i++)
        g[i - 1] = (M3(RU,i,1) * gc[0] + M3(RU,i,2) * gc[1]) + M3(RU,i,3) * gc[2];
}

/* estimateMaxCountPerCell1 (Lib_LinkCell3D.f90:320-349) */
int orc_estimate_max_count(double nCells, int nTotal)
{
    double rho = nTotal / nCells;
    double pp = exp(-rho), psum = pp;
    for (int kk = 1; kk <= nTotal; kk++) {
        if (psum >= 0.99) return kk;
        pp = pp * rho / kk;
        psum = psum + pp;
    }
    return nTotal;
}

/* ------------------------------------------------------------------ plan */

/* computeD2BI_img prologue (Lib_DynamicalTwoBeamImaging.f90:1635-1836) with
 * nProcs = 1 (factoriseParallel -> Cx = Cy = 1, Lib_FactoriseParallel.f90:104)
 * and nAperturePoints = 1 (aperture offset (0,0,0): evenlySpaced.h n = 1). */
int orc_plan(const d2bi_record* r, int64_t nAtoms, double precAngle, int nPrec, d2bi_plan* p)
{
    memset(p, 0, sizeof *p);
    if (nPrec < 1 || nPrec > D2BI_MAX_PREC) return -1;
    /* findImageParameters :450-482 */
    double B[9], gc[3];
    orc_recip(r->a_cell, B);
    reflection(B, r->g, gc);
    p->modg = norm2_3(gc);
    for (int i = 0; i < 3; i++) p->xmax[i] = norm2_3(&r->gk_a[3 * i]) * r->gk_n[i];
    p->Mx = r->gk_n[0]; p->My = r->gk_n[1]; p->Mz = r->gk_n[2];
    p->aperpx = sqrt(p->xmax[0] * p->xmax[1] / (double)(p->Mx * p->My));
    p->nPrec = nPrec;
    double costheta = 1.0, sintheta = 0.0;
    if (nPrec > 1) { costheta = cos(precAngle); sintheta = sin(precAngle); }
    /* :1676-1681 */
    p->sigma = r->a0 / 2;
    p->i2s2 = 1 / (2 * p->sigma * p->sigma);
    double l1 = norm2_3(&r->gk_a[0]), l2 = norm2_3(&r->gk_a[3]);
    p->deltaM = (int)ceil(3 * p->sigma / (l1 < l2 ? l1 : l2));
    p->nMax = orc_estimate_max_count((double)p->Mx * p->My * p->Mz, (int)nAtoms) * p->deltaM * p->deltaM;
    /* :1714-1760, Cx = Cy = 1; real(Dx)/deltaM is a default-real (FP32) expression */
    int Dx = (int)ceilf((float)p->Mx / 1.0f), Dy = (int)ceilf((float)p->My / 1.0f);
    int ex = (int)floorf((float)Dx / (float)p->deltaM + 2.0f);
    int ey = (int)floorf((float)Dy / (float)p->deltaM + 2.0f);
    int ez = (int)floorf((float)p->Mz / (float)p->deltaM);
    p->Ex = ex > 3 ? ex : 3; p->Ey = ey > 3 ? ey : 3; p->Ez = ez > 3 ? ez : 3;
    for (int i = 0; i < 3; i++) {
        p->a_col[i]     = ((double)(Dx + 2 * p->deltaM) * r->gk_a[i]) / p->Ex;
        p->a_col[3 + i] = ((double)(Dy + 2 * p->deltaM) * r->gk_a[3 + i]) / p->Ey;
        p->a_col[6 + i] = ((double)p->Mz * r->gk_a[6 + i]) / p->Ez;
    }
    orc_inverse3(p->a_col, p->ia_col);
    /* :1852  lc3d_delta = cellToRealSpace(gksuper,0,0,0) - cellToRealSpace(super,1,1,0) */
    for (int i = 0; i < 3; i++) {
        double g0 = (r->gk_a[i] * 0 + r->gk_a[3 + i] * 0) + r->gk_a[6 + i] * 0;
        double s1 = (p->a_col[i] * 1 + p->a_col[3 + i] * 1) + p->a_col[6 + i] * 0;
        p->lc_delta[i] = g0 - s1;
    }
    matmul3(r->R, r->U, p->RU);
    get_aperture(r, NULL, p->gg);
    /* :1795-1798 */
    p->dz = p->xmax[2] / p->Mz;
    p->bg = PI_IMG / r->xig;
    p->b0 = PI_IMG / r->xi0;
    /* :1805-1836 */
    for (int prec = 1; prec <= nPrec; prec++) {
        double gg[3];
        if (prec == 1) {
            get_aperture(r, NULL, gg);
        } else {
            double phi = (prec - 1) * 2 * PI_IMG / (nPrec - 1);
            double ux = sin(phi), uy = cos(phi), T[9];
            M3(T,1,1) = costheta + ux * ux * (1 - costheta); M3(T,1,2) = ux * uy * (1 - costheta);             M3(T,1,3) = uy * sintheta;
            M3(T,2,1) = ux * uy * (1 - costheta);             M3(T,2,2) = costheta + uy * uy * (1 - costheta); M3(T,2,3) = -ux * sintheta;
            M3(T,3,1) = -uy * sintheta;                       M3(T,3,2) = ux * sintheta;                       M3(T,3,3) = costheta;
            get_aperture(r, T, gg);
        }
        double g2[3] = {gg[0] + 0.0, gg[1] + 0.0, gg[2] + 0.0};
        double sg = orc_deviation_parameter(r->E, g2);
        p->aa[prec - 1] = 2 * PI_IMG * sg;
    }
    p->rc2 = (3 * p->sigma) * (3 * p->sigma);
    return 0;
}

/* ------------------------------------------------------------------ link cells */

/* type LinkCell3D (Lib_LinkCell3D.f90:48-63): per cell an id list and
 * fractional offsets in insertion order; storage doubles on overflow
 * (doubleCell :388), fatal beyond 16x the initial capacity (:716-723). */
typedef struct {
    int nindx, nmax;
    int* indx;
    double* x;     /* 3*nmax */
} orc_cell;

typedef struct {
    int Nx, Ny, Nz;
    double a[9], ia[9];
    int nMax0, nMax;
    orc_cell* c;
} orc_lc;

orc_lc* orc_lc_new(const double* a, int Nx, int Ny, int Nz, int nMax)
{
    orc_lc* lc = (orc_lc*)calloc(1, sizeof *lc);
    lc->Nx = Nx; lc->Ny = Ny; lc->Nz = Nz;
    memcpy(lc->a, a, 9 * sizeof(double));
    orc_inverse3(lc->a, lc->ia);
    lc->nMax0 = nMax; lc->nMax = nMax;
    int64_t nc = (int64_t)Nx * Ny * Nz;
    lc->c = (orc_cell*)calloc(nc, sizeof(orc_cell));
    for (int64_t i = 0; i < nc; i++) {
        lc->c[i].nmax = nMax;
        lc->c[i].indx = (int*)malloc(sizeof(int) * (nMax > 0 ? nMax : 1));
        lc->c[i].x = (double*)malloc(sizeof(double) * 3 * (nMax > 0 ? nMax : 1));
    }
    return lc;
}

void orc_lc_free(orc_lc* lc)
{
    if (!lc) return;
    int64_t nc = (int64_t)lc->Nx * lc->Ny * lc->Nz;
    for (int64_t i = 0; i < nc; i++) { free(lc->c[i].indx); free(lc->c[i].x); }
    free(lc->c);
    free(lc);
}

/* clear0 (Lib_LinkCell3D.f90:576-588) */
void orc_lc_clear(orc_lc* lc)
{
    int64_t nc = (int64_t)lc->Nx * lc->Ny * lc->Nz;
    for (int64_t i = 0; i < nc; i++) lc->c[i].nindx = 0;
    lc->nMax = 0;
}

static inline int64_t cidx(const orc_lc* lc, int ix, int iy, int iz)
{
    return (int64_t)ix + (int64_t)lc->Nx * (iy + (int64_t)lc->Ny * iz);
}

/* realSpaceToCell0 (Lib_SimpleSupercells.f90:348-355) */
static inline void real_to_cell(const double* ia, const double* y, double* x)
{
    for (int i = 1; i <= 3; i++)
        x[i - 1] = (M3(ia,i,1) * y[0] + M3(ia,i,2) * y[1]) + M3(ia,i,3) * y[2];
}

/* add0 (Lib_LinkCell3D.f90:691-748); returns 0, or -1 on the fatal overflow */
int orc_lc_add(orc_lc* lc, int id, const double* x)
{
    double yy[3], zz[3];
    real_to_cell(lc->ia, x, yy);
    for (int i = 0; i < 3; i++) zz[i] = yy[i] - floor(yy[i]);
    /* whichCell(cellspace=.true.) (Lib_SimpleSupercells.f90:358-384) */
    int ix = ((int)floor(yy[0]) + 16 * lc->Nx) % lc->Nx;
    int iy = ((int)floor(yy[1]) + 16 * lc->Ny) % lc->Ny;
    int iz = ((int)floor(yy[2]) + 16 * lc->Nz) % lc->Nz;
    orc_cell* c = &lc->c[cidx(lc, ix, iy, iz)];
    int nn = c->nindx + 1;
    if (nn > c->nmax) {
        if (c->nmax == 16 * lc->nMax0) return -1;
        int newmax = c->nmax * 2 > 2 ? c->nmax * 2 : 2;
        c->indx = (int*)realloc(c->indx, sizeof(int) * newmax);
        c->x = (double*)realloc(c->x, sizeof(double) * 3 * newmax);
        c->nmax = newmax;
    }
    c->nindx = nn;
    c->indx[nn - 1] = id;
    c->x[3 * (nn - 1) + 0] = zz[0];
    c->x[3 * (nn - 1) + 1] = zz[1];
    c->x[3 * (nn - 1) + 2] = zz[2];
    if (nn > lc->nMax) lc->nMax = nn;
    return 0;
}

/* neighbourList1b (Lib_LinkCell3D.f90:1515-1605), with the separation
 * vector also returned as in neighbourList0a (:1214-1287) when dx != NULL. */
int orc_lc_neighbours(const orc_lc* lc, const double* x, double rc,
                      int* id, double* r2, double* dx)
{
    const int Nx = lc->Nx, Ny = lc->Ny, Nz = lc->Nz;
    const double* a = lc->a;
    double zi[3];
    real_to_cell(lc->ia, x, zi);
    int ix = ((int)floor(zi[0]) + 16 * Nx) % Nx;
    int iy = ((int)floor(zi[1]) + 16 * Ny) % Ny;
    int iz = ((int)floor(zi[2]) + 16 * Nz) % Nz;
    zi[0] = zi[0] - ix; zi[1] = zi[1] - iy; zi[2] = zi[2] - iz;
    const double rc2 = rc * rc;
    int n = 0;
    for (int kz = -1; kz <= 1; kz++) {
        int jz = (iz + Nz + kz) % Nz;
        double dzi3 = kz - zi[2];
        for (int ky = -1; ky <= 1; ky++) {
            int jy = (iy + Ny + ky) % Ny;
            double dzi2 = ky - zi[1];
            for (int kx = -1; kx <= 1; kx++) {
                int jx = (ix + Nx + kx) % Nx;
                double dzi1 = kx - zi[0];
                const orc_cell* c = &lc->c[cidx(lc, jx, jy, jz)];
                for (int jk = 0; jk < c->nindx; jk++) {
                    double dzj1 = c->x[3 * jk + 0] + dzi1;
                    double dzj2 = c->x[3 * jk + 1] + dzi2;
                    double dzj3 = c->x[3 * jk + 2] + dzi3;
                    double d[3];
                    for (int i = 1; i <= 3; i++)
                        d[i - 1] = (M3(a,i,1) * dzj1 + M3(a,i,2) * dzj2) + M3(a,i,3) * dzj3;
                    double dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
                    if (dd <= rc2) {
                        id[n] = c->indx[jk];
                        r2[n] = dd;
                        if (dx) { dx[3 * n] = d[0]; dx[3 * n + 1] = d[1]; dx[3 * n + 2] = d[2]; }
                        n++;
                    }
                }
            }
        }
    }
    return n;
}

/* Accessors so tests can read the cell lists. */
int orc_lc_count(const orc_lc* lc, int ix, int iy, int iz) { return lc->c[cidx(lc, ix, iy, iz)].nindx; }
int orc_lc_storage(const orc_lc* lc, int ix, int iy, int iz) { return lc->c[cidx(lc, ix, iy, iz)].nmax; }
int orc_lc_nmax(const orc_lc* lc) { return lc->nMax; }
int orc_lc_id(const orc_lc* lc, int ix, int iy, int iz, int k) { return lc->c[cidx(lc, ix, iy, iz)].indx[k]; }
void orc_lc_frac(const orc_lc* lc, int ix, int iy, int iz, int k, double* f)
{
    const double* s = &lc->c[cidx(lc, ix, iy, iz)].x[3 * k];
    f[0] = s[0]; f[1] = s[1]; f[2] = s[2];
}

/* Flatten the cell lists in cell order c = cx + Ex*(cy + Ey*cz). */
int64_t orc_lc_flatten(const orc_lc* lc, int32_t* cell_start, int32_t* ids, double* frac)
{
    int64_t nc = (int64_t)lc->Nx * lc->Ny * lc->Nz, k = 0;
    for (int64_t c = 0; c < nc; c++) {
        if (cell_start) cell_start[c] = (int32_t)k;
        for (int j = 0; j < lc->c[c].nindx; j++, k++) {
            if (ids) ids[k] = lc->c[c].indx[j];
            if (frac) memcpy(&frac[3 * k], &lc->c[c].x[3 * j], 3 * sizeof(double));
        }
    }
    if (cell_start) cell_start[nc] = (int32_t)k;
    return k;
}

/* pointInSupercell0 (Lib_SimpleSupercells.f90:1063-1102) */
static inline int point_in_box(const double* ia, int Nx, int Ny, int Nz, const double* x)
{
    int ix = (int)floor((M3(ia,1,1) * x[0] + M3(ia,1,2) * x[1]) + M3(ia,1,3) * x[2]);
    int is = (ix == (ix + 256 * Nx) % Nx);
    int iy = (int)floor((M3(ia,2,1) * x[0] + M3(ia,2,2) * x[1]) + M3(ia,2,3) * x[2]);
    is = is && (iy == (iy + 256 * Ny) % Ny);
    int iz = (int)floor((M3(ia,3,1) * x[0] + M3(ia,3,2) * x[1]) + M3(ia,3,3) * x[2]);
    is = is && (iz == (iz + 256 * Nz) % Nz);
    return is;
}

/* findGdotR (Lib_DynamicalTwoBeamImaging.f90:1160-1265).  x is (ld,nAtoms)
 * column-major.  Fills lc (cleared first) and gdotr[0..Nr-1]; returns Nr,
 * or -1 on link-cell overflow, -2 if gdotr_cap is too small. */
int64_t orc_find_gdotr(const d2bi_record* r, const d2bi_plan* p, const double* x, int64_t ld,
                       int64_t nAtoms, orc_lc* lc, double* gdotr, int64_t gdotr_cap)
{
    orc_lc_clear(lc);
    const double* RU = p->RU;
    const double* gg = p->gg;
    int64_t ii = 0;
    for (int nn = 0; nn < r->nuvw; nn++) {
        double Ruvw[3];
        const int* u = &r->uvw[3 * nn];
        for (int i = 0; i < 3; i++) {
            double s1 = r->super_a[i] * r->super_n[0];     /* getSuperVector :904-918 */
            double s2 = r->super_a[3 + i] * r->super_n[1];
            double s3 = r->super_a[6 + i] * r->super_n[2];
            Ruvw[i] = (u[0] * s1 + u[1] * s2) + u[2] * s3;
        }
        for (int64_t jj = 0; jj < nAtoms; jj++) {
            const double* xa = &x[ld * jj];
            double y0[3], yy[3], zz[3];
            for (int i = 0; i < 3; i++) y0[i] = xa[i] + Ruvw[i];
            for (int i = 1; i <= 3; i++)
                yy[i - 1] = (M3(RU,i,1) * y0[0] + M3(RU,i,2) * y0[1]) + M3(RU,i,3) * y0[2];
            for (int i = 0; i < 3; i++) { yy[i] = yy[i] + r->delta[i]; zz[i] = yy[i] - p->lc_delta[i]; }
            if (point_in_box(lc->ia, lc->Nx, lc->Ny, lc->Nz, zz)) {
                ii++;
                if (ii > gdotr_cap) return -2;
                if (orc_lc_add(lc, (int)ii, zz)) return -1;
                gdotr[ii - 1] = (gg[0] * yy[0] + gg[1] * yy[1]) + gg[2] * yy[2];
            }
        }
    }
    return ii;
}

/* convertScaledRangeToRho (Lib_DynamicalTwoBeamImaging.f90:1973-1994) */
double orc_rho(double r2_min, double sig)
{
    if (r2_min < sig * sig) return 1.0;
    if (r2_min > 4 * sig * sig) return 0.0;
    double xx = sqrt(r2_min) / sig - 1.0;
    return 1 - ((xx * xx) * xx) * ((10 - 15 * xx) + (6 * xx) * xx);
}

/* dynamicalTwoBeamIntegrationColumn_phaserho (Lib_DynamicalTwoBeamImaging.f90:2316-2480).
 * xr/xi/rho hold nz+1 samples (0..nz). */
double orc_integrate_column(double a, double b0, double bg, double dz,
                            const double* xr, const double* xi, const double* rho, int nz)
{
    double phi0r = 1.0, phi0i = 0.0, phigr = 0.0, phigi = 0.0;
    double dphi0r, dphi0i, dphigr, dphigi;
    double phi0r1, phi0i1, phigr1, phigi1, phi0r2, phi0i2, phigr2, phigi2;
#define COEF(k, c0, ca, cgr, cgi) \
    double c0 = b0 * rho[k], ca = b0 * rho[k] + a, cgr = (bg * rho[k]) * xr[k], cgi = (bg * rho[k]) * xi[k];
#define DERIV(c0, ca, cgr, cgi)                                           \
    dphi0r = ((-(c0 * phi0i)) - cgi * phigr) - cgr * phigi;               \
    dphi0i = ((c0 * phi0r) + cgr * phigr) - cgi * phigi;                  \
    dphigr = ((cgi * phi0r) - cgr * phi0i) - ca * phigi;                  \
    dphigi = ((cgr * phi0r) + cgi * phi0i) + ca * phigr;
    for (int ii = 0; ii <= nz - 2; ii += 2) {
        COEF(ii, c0a, caa, cgra, cgia)
        COEF(ii + 1, c0b, cab, cgrb, cgib)
        COEF(ii + 2, c0c, cac, cgrc, cgic)
        /* step 1 */
        DERIV(c0a, caa, cgra, cgia)
        phi0r2 = phi0r; phi0i2 = phi0i; phigr2 = phigr; phigi2 = phigi;
        phi0r1 = phi0r + (dphi0r * dz) / 3; phi0i1 = phi0i + (dphi0i * dz) / 3;
        phigr1 = phigr + (dphigr * dz) / 3; phigi1 = phigi + (dphigi * dz) / 3;
        phi0r = phi0r + dphi0r * dz; phi0i = phi0i + dphi0i * dz;
        phigr = phigr + dphigr * dz; phigi = phigi + dphigi * dz;
        /* step 2 */
        DERIV(c0b, cab, cgrb, cgib)
        phi0r = phi0r2 + dphi0r * dz; phi0i = phi0i2 + dphi0i * dz;
        phigr = phigr2 + dphigr * dz; phigi = phigi2 + dphigi * dz;
        phi0r1 = phi0r1 + ((dphi0r * 2) * dz) / 3; phi0i1 = phi0i1 + ((dphi0i * 2) * dz) / 3;
        phigr1 = phigr1 + ((dphigr * 2) * dz) / 3; phigi1 = phigi1 + ((dphigi * 2) * dz) / 3;
        /* step 3 */
        DERIV(c0b, cab, cgrb, cgib)
        phi0r = phi0r2 + (dphi0r * 2) * dz; phi0i = phi0i2 + (dphi0i * 2) * dz;
        phigr = phigr2 + (dphigr * 2) * dz; phigi = phigi2 + (dphigi * 2) * dz;
        phi0r1 = phi0r1 + ((dphi0r * 2) * dz) / 3; phi0i1 = phi0i1 + ((dphi0i * 2) * dz) / 3;
        phigr1 = phigr1 + ((dphigr * 2) * dz) / 3; phigi1 = phigi1 + ((dphigi * 2) * dz) / 3;
        /* step 4 */
        DERIV(c0c, cac, cgrc, cgic)
        phi0r = phi0r1 + (dphi0r * dz) / 3; phi0i = phi0i1 + (dphi0i * dz) / 3;
        phigr = phigr1 + (dphigr * dz) / 3; phigi = phigi1 + (dphigi * dz) / 3;
        double phimag = ((phi0r * phi0r + phi0i * phi0i) + phigr * phigr) + phigi * phigi;
        phimag = 1 / sqrt(phimag);
        phi0r = phi0r * phimag; phi0i = phi0i * phimag;
        phigr = phigr * phimag; phigi = phigi * phimag;
        (void)c0c; (void)cac; (void)cgrc; (void)cgic; (void)c0a; (void)caa; (void)cgra; (void)cgia;
    }
#undef COEF
#undef DERIV
    return phigr * phigr + phigi * phigi;
}

/* intensityAtDepthZ (Lib_DiffractionConditions.f90:1209-1258): closed form
 * for a perfect crystal; used only to check orc_integrate_column. */
double orc_intensity_at_depth(double xi0, double xig, double sg, double z)
{
    double b0 = PI_DC / xi0, bg = PI_DC / xig, aa = 2 * PI_DC * sg;
    double dd = sqrt(aa * aa / 4 + bg * bg);
    double l1 = (b0 + aa / 2) + dd, l2 = (b0 + aa / 2) - dd;
    double complex phig = cexp(I * (l1 * z)) - cexp(I * (l2 * z));
    phig = phig * bg / (2 * dd);
    return creal(phig * conj(phig));
}

/* One voxel: lines :1877-1892.  Returns the neighbour count. */
static int voxel_sum(const orc_lc* lc, const d2bi_plan* p, const d2bi_record* r,
                     const double* gdotr, int ix, int iy, int iz,
                     int* id, double* r2, double* xre, double* xim, double* rho, double* r2out)
{
    double xx[3];
    for (int i = 0; i < 3; i++) {
        double c = (r->gk_a[i] * ix + r->gk_a[3 + i] * iy) + r->gk_a[6 + i] * iz;
        xx[i] = c - p->lc_delta[i];
    }
    int nn = orc_lc_neighbours(lc, xx, 3 * p->sigma, id, r2, NULL);
    double r2_min = 3.4028234663852886e38;     /* huge(1.0): default-real huge */
    double complex s = 0.0;
    for (int kk = 0; kk < nn; kk++) {
        r2_min = r2_min < r2[kk] ? r2_min : r2[kk];
        s = s + cexp(CMPLX(-r2[kk] * p->i2s2, gdotr[id[kk] - 1]));
    }
    if (rho) *rho = orc_rho(r2_min, p->sigma);
    double ww = cabs(s);
    ww = 1 / (1.0e-16 > ww ? 1.0e-16 : ww);
    *xre = creal(s) * ww;
    *xim = cimag(s) * ww;
    if (r2out) *r2out = r2_min;
    return nn;
}

typedef struct {
    orc_lc* lc;
    double* gdotr;
    int64_t nPoints;
    d2bi_plan plan;
} orc_state;

/* Build the link cells and g.r for one condition. */
orc_state* orc_prepare(const d2bi_record* r, const double* x, int64_t ld, int64_t nAtoms,
                       double precAngle, int nPrec, int* err)
{
    orc_state* s = (orc_state*)calloc(1, sizeof *s);
    *err = 0;
    if (orc_plan(r, nAtoms, precAngle, nPrec, &s->plan)) { *err = -1; free(s); return NULL; }
    const d2bi_plan* p = &s->plan;
    s->lc = orc_lc_new(p->a_col, p->Ex, p->Ey, p->Ez, p->nMax);
    int64_t cap = (int64_t)r->nuvw * nAtoms + 1;
    s->gdotr = (double*)malloc(sizeof(double) * cap);
    s->nPoints = orc_find_gdotr(r, p, x, ld, nAtoms, s->lc, s->gdotr, cap);
    if (s->nPoints < 0) { *err = (int)s->nPoints - 10; }
    return s;
}

void orc_state_free(orc_state* s)
{
    if (!s) return;
    orc_lc_free(s->lc);
    free(s->gdotr);
    free(s);
}

int64_t orc_state_npoints(const orc_state* s) { return s->nPoints; }
const d2bi_plan* orc_state_plan(const orc_state* s) { return &s->plan; }
orc_lc* orc_state_lc(orc_state* s) { return s->lc; }
const double* orc_state_gdotr(const orc_state* s) { return s->gdotr; }

/* Voxel field over a window of columns [ix0,ix1) x [iy0,iy1), all z.
 * Layout: index (ix-ix0) + wx*((iy-iy0) + wy*iz). */
void orc_field(const orc_state* s, const d2bi_record* r, int ix0, int ix1, int iy0, int iy1,
               double* xre, double* xim, double* rho, int32_t* count, double* r2min, int nthreads)
{
    const d2bi_plan* p = &s->plan;
    const int wx = ix1 - ix0, wy = iy1 - iy0, Mz = p->Mz;
    int cap = 27 * (p->nMax > 0 ? p->nMax : 1) * 16 + 64;
    (void)nthreads;
#ifdef _OPENMP
    #pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
#endif
    {
        int* id = (int*)malloc(sizeof(int) * cap);
        double* r2 = (double*)malloc(sizeof(double) * cap);
#ifdef _OPENMP
        #pragma omp for schedule(dynamic, 1)
#endif
        for (int iy = iy0; iy < iy1; iy++)
            for (int ix = ix0; ix < ix1; ix++)
                for (int iz = 0; iz < Mz; iz++) {
                    int64_t o = (int64_t)(ix - ix0) + (int64_t)wx * ((iy - iy0) + (int64_t)wy * iz);
                    double a, b, rh = 1.0, rm;
                    int n = voxel_sum(s->lc, p, r, s->gdotr, ix, iy, iz, id, r2, &a, &b,
                                      r->use_density ? &rh : NULL, &rm);
                    if (xre) xre[o] = a;
                    if (xim) xim[o] = b;
                    if (rho) rho[o] = rh;
                    if (count) count[o] = n;
                    if (r2min) r2min[o] = rm;
                }
        free(id);
        free(r2);
    }
}

/* computeD2BI_img pixel loop (:1870-1951) over a window of columns; img has
 * (ix1-ix0)*(iy1-iy0) entries, x fastest.  nthreads > 1 splits rows across
 * OpenMP threads, the way MPI ranks split patches; each pixel's arithmetic
 * is unchanged. */
void orc_image_window(const orc_state* s, const d2bi_record* r, int ix0, int ix1, int iy0, int iy1,
                      double* img, int nthreads)
{
    const d2bi_plan* p = &s->plan;
    const int Mz = p->Mz, wx = ix1 - ix0;
    int cap = 27 * (p->nMax > 0 ? p->nMax : 1) * 16 + 64;
    (void)nthreads;
#ifdef _OPENMP
    #pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
#endif
    {
        int* id = (int*)malloc(sizeof(int) * cap);
        double* r2 = (double*)malloc(sizeof(double) * cap);
        double* xr = (double*)malloc(sizeof(double) * (Mz + 1));
        double* xi = (double*)malloc(sizeof(double) * (Mz + 1));
        double* rh = (double*)malloc(sizeof(double) * (Mz + 1));
#ifdef _OPENMP
        #pragma omp for schedule(dynamic, 1)
#endif
        for (int iy = iy0; iy < iy1; iy++)
            for (int ix = ix0; ix < ix1; ix++) {
                for (int iz = 0; iz <= Mz; iz++) rh[iz] = 1.0;   /* rho = 1 (:1792) */
                for (int iz = 0; iz < Mz; iz++)
                    voxel_sum(s->lc, p, r, s->gdotr, ix, iy, iz, id, r2, &xr[iz], &xi[iz],
                              r->use_density ? &rh[iz] : NULL, NULL);
                xr[Mz] = xr[0]; xi[Mz] = xi[0]; rh[Mz] = rh[0];
                double acc = 0.0;
                for (int prec = 0; prec < p->nPrec; prec++)
                    acc = acc + orc_integrate_column(p->aa[prec], p->b0, p->bg, p->dz, xr, xi, rh, Mz);
                img[(int64_t)(ix - ix0) + (int64_t)wx * (iy - iy0)] = acc / (1 * p->nPrec);
            }
        free(id); free(r2); free(xr); free(xi); free(rh);
    }
}

/* computeD2BI_img for the whole image; returns 0 or a negative code. */
int orc_compute_img(const d2bi_record* r, const double* x, int64_t ld, int64_t nAtoms,
                    double precAngle, int nPrec, double* img, int nthreads)
{
    int err;
    orc_state* s = orc_prepare(r, x, ld, nAtoms, precAngle, nPrec, &err);
    if (!s) return err;
    if (err) { orc_state_free(s); return err; }
    orc_image_window(s, r, 0, s->plan.Mx, 0, s->plan.My, img, nthreads);
    orc_state_free(s);
    return 0;
}

/* ------------------------------------------------------------------ aperture */

/* sunflower(n, alpha=0) with radius(k, n, b)
 * (Lib_EvenlySpacedPointsInCircle.f90:438-478), that module's
 * PI = 3.141592654; the point set evenlySpacedPointsInCircle returns for
 * n > 25 (:93-97).  xy is (2, n). */
void orc_sunflower(int n, double* xy)
{
    const double pi = 3.141592654;
    const double phi = (sqrt(5.0) + 1.0) / 2;
    const double alpha = 0.0;
    const int bb = (int)lround(sqrt(alpha * alpha * n));
    for (int kk = 1; kk <= n; kk++) {
        double rr;
        if (kk > n - bb) rr = 1;
        else rr = sqrt((double)(2 * kk - 1) / (2 * n - bb - 1));
        double theta = 2 * pi * kk / (phi * phi);
        xy[2 * (kk - 1)] = rr * cos(theta);
        xy[2 * (kk - 1) + 1] = rr * sin(theta);
    }
}

/* computeD2BI_img aperture sampling (Lib_DynamicalTwoBeamImaging.f90:1655-1667,
 * 1805-1836) and findGdotR_aperture's shifted g-vectors (:1336-1344).
 * xy: the nApp unit-disk points (2, nApp).  gshift (3, nApp); aa (nApp, nPrec),
 * app fastest. */
int orc_aperture_setup(const d2bi_record* r, double precAngle, int nPrec, double apertureAngle, int nApp,
                       const double* xy, double* gshift, double* aa)
{
    if (nPrec < 1 || nPrec > D2BI_MAX_PREC || nApp < 1) return -1;
    double modk = 2 * PI_IMG / orc_wavelength(r->E * 1000);
    double modg = apertureAngle * modk;
    double gg0[3];
    get_aperture(r, NULL, gg0);
    for (int app = 0; app < nApp; app++) {
        double dgx = modg * xy[2 * app], dgy = modg * xy[2 * app + 1];
        gshift[3 * app] = gg0[0] + dgx;
        gshift[3 * app + 1] = gg0[1] + dgy;
        gshift[3 * app + 2] = gg0[2];
    }
    double costheta = 1.0, sintheta = 0.0;
    if (nPrec > 1) { costheta = cos(precAngle); sintheta = sin(precAngle); }
    for (int prec = 1; prec <= nPrec; prec++) {
        double gg[3];
        if (prec == 1) {
            get_aperture(r, NULL, gg);
        } else {
            double phi = (prec - 1) * 2 * PI_IMG / (nPrec - 1);
            double ux = sin(phi), uy = cos(phi), T[9];
            M3(T,1,1) = costheta + ux * ux * (1 - costheta); M3(T,1,2) = ux * uy * (1 - costheta);             M3(T,1,3) = uy * sintheta;
            M3(T,2,1) = ux * uy * (1 - costheta);             M3(T,2,2) = costheta + uy * uy * (1 - costheta); M3(T,2,3) = -ux * sintheta;
            M3(T,3,1) = -uy * sintheta;                       M3(T,3,2) = ux * sintheta;                       M3(T,3,3) = costheta;
            get_aperture(r, T, gg);
        }
        for (int app = 0; app < nApp; app++) {
            double dg[3] = {modg * xy[2 * app], modg * xy[2 * app + 1], modg * 0.0};
            double g2[3] = {gg[0] + dg[0], gg[1] + dg[1], gg[2] + dg[2]};
            aa[app + nApp * (prec - 1)] = 2 * PI_IMG * orc_deviation_parameter(r->E, g2);
        }
    }
    return 0;
}

/* findGdotR_aperture (Lib_DynamicalTwoBeamImaging.f90:1269-1396): as
 * orc_find_gdotr, with gdotr(ii, kk) = gshift(:,kk).yy for every aperture
 * point; gdotr is (cap, nApp) column-major. */
int64_t orc_find_gdotr_aperture(const d2bi_record* r, const d2bi_plan* p, const double* x, int64_t ld,
                                int64_t nAtoms, orc_lc* lc, const double* gshift, int nApp, double* gdotr,
                                int64_t cap)
{
    orc_lc_clear(lc);
    const double* RU = p->RU;
    int64_t ii = 0;
    for (int nn = 0; nn < r->nuvw; nn++) {
        double Ruvw[3];
        const int* u = &r->uvw[3 * nn];
        for (int i = 0; i < 3; i++) {
            double s1 = r->super_a[i] * r->super_n[0];
            double s2 = r->super_a[3 + i] * r->super_n[1];
            double s3 = r->super_a[6 + i] * r->super_n[2];
            Ruvw[i] = (u[0] * s1 + u[1] * s2) + u[2] * s3;
        }
        for (int64_t jj = 0; jj < nAtoms; jj++) {
            const double* xa = &x[ld * jj];
            double y0[3], yy[3], zz[3];
            for (int i = 0; i < 3; i++) y0[i] = xa[i] + Ruvw[i];
            for (int i = 1; i <= 3; i++)
                yy[i - 1] = (M3(RU,i,1) * y0[0] + M3(RU,i,2) * y0[1]) + M3(RU,i,3) * y0[2];
            for (int i = 0; i < 3; i++) { yy[i] = yy[i] + r->delta[i]; zz[i] = yy[i] - p->lc_delta[i]; }
            if (point_in_box(lc->ia, lc->Nx, lc->Ny, lc->Nz, zz)) {
                ii++;
                if (ii > cap) return -2;
                if (orc_lc_add(lc, (int)ii, zz)) return -1;
                for (int kk = 0; kk < nApp; kk++) {
                    const double* g = &gshift[3 * kk];
                    gdotr[(ii - 1) + cap * kk] = (g[0] * yy[0] + g[1] * yy[1]) + g[2] * yy[2];
                }
            }
        }
    }
    return ii;
}

/* computeD2BI_img's nApp > 1 branch (:1905-1934) over a window of columns:
 * per voxel one neighbour list, r2_min and rho, a phase sum per aperture
 * point; per column the integrations in (prec outer, app inner) order,
 * divided by nApp*nPrec after the sum (:1951).  gdotr as above. */
void orc_image_aperture_window(const d2bi_record* r, const d2bi_plan* p, const orc_lc* lc, const double* gdotr,
                               int64_t cap, int nApp, const double* aa, int ix0, int ix1, int iy0, int iy1,
                               double* img, int nthreads)
{
    const int Mz = p->Mz, wx = ix1 - ix0;
    int ncap = 27 * (p->nMax > 0 ? p->nMax : 1) * 16 + 64;
    (void)nthreads;
#ifdef _OPENMP
    #pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
#endif
    {
        int* id = (int*)malloc(sizeof(int) * ncap);
        double* r2 = (double*)malloc(sizeof(double) * ncap);
        double* xr = (double*)malloc(sizeof(double) * (Mz + 1) * nApp);
        double* xi = (double*)malloc(sizeof(double) * (Mz + 1) * nApp);
        double* rh = (double*)malloc(sizeof(double) * (Mz + 1));
        double complex* s = (double complex*)malloc(sizeof(double complex) * nApp);
#ifdef _OPENMP
        #pragma omp for schedule(dynamic, 1)
#endif
        for (int iy = iy0; iy < iy1; iy++)
            for (int ix = ix0; ix < ix1; ix++) {
                for (int iz = 0; iz <= Mz; iz++) rh[iz] = 1.0;
                for (int iz = 0; iz < Mz; iz++) {
                    double xx[3];
                    for (int i = 0; i < 3; i++) {
                        double c = (r->gk_a[i] * ix + r->gk_a[3 + i] * iy) + r->gk_a[6 + i] * iz;
                        xx[i] = c - p->lc_delta[i];
                    }
                    int nn = orc_lc_neighbours(lc, xx, 3 * p->sigma, id, r2, NULL);
                    double r2_min = 3.4028234663852886e38;
                    for (int app = 0; app < nApp; app++) s[app] = 0.0;
                    for (int kk = 0; kk < nn; kk++) {
                        r2_min = r2_min < r2[kk] ? r2_min : r2[kk];
                        for (int app = 0; app < nApp; app++)
                            s[app] = s[app] + cexp(CMPLX(-r2[kk] * p->i2s2, gdotr[(id[kk] - 1) + cap * app]));
                    }
                    if (r->use_density) rh[iz] = orc_rho(r2_min, p->sigma);
                    for (int app = 0; app < nApp; app++) {
                        double ww = cabs(s[app]);
                        ww = 1 / (1.0e-16 > ww ? 1.0e-16 : ww);
                        xr[iz + (Mz + 1) * app] = creal(s[app]) * ww;
                        xi[iz + (Mz + 1) * app] = cimag(s[app]) * ww;
                    }
                }
                for (int app = 0; app < nApp; app++) {
                    xr[Mz + (Mz + 1) * app] = xr[(Mz + 1) * app];
                    xi[Mz + (Mz + 1) * app] = xi[(Mz + 1) * app];
                }
                rh[Mz] = rh[0];
                double acc = 0.0;
                for (int prec = 0; prec < p->nPrec; prec++)
                    for (int app = 0; app < nApp; app++)
                        acc = acc + orc_integrate_column(aa[app + nApp * prec], p->b0, p->bg, p->dz,
                                                         xr + (Mz + 1) * app, xi + (Mz + 1) * app, rh, Mz);
                img[(int64_t)(ix - ix0) + (int64_t)wx * (iy - iy0)] = acc / (nApp * p->nPrec);
            }
        free(id); free(r2); free(xr); free(xi); free(rh); free(s);
    }
}

/* Whole aperture-sampled image (computeD2BI_img with apertureAngle,
 * nAperturePoints = nApp); xy the unit-disk points.  Returns 0 or < 0. */
int orc_compute_img_aperture(const d2bi_record* r, const double* x, int64_t ld, int64_t nAtoms,
                             double precAngle, int nPrec, double apertureAngle, int nApp, const double* xy,
                             double* img, int nthreads)
{
    d2bi_plan plan;
    if (orc_plan(r, nAtoms, precAngle, nPrec, &plan)) return -1;
    double* gshift = (double*)malloc(sizeof(double) * 3 * nApp);
    double* aa = (double*)malloc(sizeof(double) * nApp * nPrec);
    if (orc_aperture_setup(r, precAngle, nPrec, apertureAngle, nApp, xy, gshift, aa)) {
        free(gshift); free(aa);
        return -1;
    }
    orc_lc* lc = orc_lc_new(plan.a_col, plan.Ex, plan.Ey, plan.Ez, plan.nMax);
    int64_t cap = (int64_t)r->nuvw * nAtoms + 1;
    double* gdotr = (double*)malloc(sizeof(double) * cap * nApp);
    int64_t np = orc_find_gdotr_aperture(r, &plan, x, ld, nAtoms, lc, gshift, nApp, gdotr, cap);
    int rc = 0;
    if (np < 0) rc = (int)np - 10;
    else orc_image_aperture_window(r, &plan, lc, gdotr, cap, nApp, aa, 0, plan.Mx, 0, plan.My, img, nthreads);
    orc_lc_free(lc);
    free(gdotr); free(gshift); free(aa);
    return rc;
}

/* gaussianBlurImage (d2bi.f90:1543-1624) */
void orc_blur(const double* fin, int Nx, int Ny, double t, double* fout)
{
    int Nk = (int)ceil(t * 5);
    if (Nk < 5) Nk = 5;
    double* kernel = (double*)malloc(sizeof(double) * (Nk + 1));
    double i2s2 = 1 / (2 * t * t);
    kernel[0] = 1.0;
    for (int ix = 1; ix <= Nk; ix++) kernel[ix] = exp((double)(-ix * ix) * i2s2);
    double* wstripe = (double*)malloc(sizeof(double) * Nx);
    double* fstripe = (double*)malloc(sizeof(double) * Ny);
#define FIN(i, j) fin[(int64_t)(i) + (int64_t)Nx * (j)]
#define FOUT(i, j) fout[(int64_t)(i) + (int64_t)Nx * (j)]
    for (int ix = 0; ix < Nx; ix++) {
        double wf = 0.0, ws = 0.0;
        int lo = ix - Nk > 0 ? ix - Nk : 0, hi = ix + Nk < Nx - 1 ? ix + Nk : Nx - 1;
        for (int jj = lo; jj <= hi; jj++) {
            double ww = kernel[abs(jj - ix)];
            wf = wf + ww * FIN(jj, 0);
            ws = ws + ww;
        }
        FOUT(ix, 0) = wf;
        wstripe[ix] = 1 / ws;
    }
    for (int iy = 1; iy < Ny; iy++)
        for (int ix = 0; ix < Nx; ix++) {
            double wf = 0.0;
            int lo = ix - Nk > 0 ? ix - Nk : 0, hi = ix + Nk < Nx - 1 ? ix + Nk : Nx - 1;
            for (int jj = lo; jj <= hi; jj++) wf = wf + kernel[abs(jj - ix)] * FIN(jj, iy);
            FOUT(ix, iy) = wf;
        }
    for (int ix = 0; ix < Nx; ix++) {
        for (int iy = 0; iy < Ny; iy++) fstripe[iy] = FOUT(ix, iy);
        for (int iy = 0; iy < Ny; iy++) {
            double wf = 0.0, ws = 0.0;
            int lo = iy - Nk > 0 ? iy - Nk : 0, hi = iy + Nk < Ny - 1 ? iy + Nk : Ny - 1;
            for (int jj = lo; jj <= hi; jj++) {
                double ww = kernel[abs(jj - iy)];
                wf = wf + ww * fstripe[jj];
                ws = ws + ww;
            }
            FOUT(ix, iy) = (wf * wstripe[ix]) / ws;
        }
    }
#undef FIN
#undef FOUT
    free(kernel); free(wstripe); free(fstripe);
}
