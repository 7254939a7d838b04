// d2bi_host.cpp -- host-side setup of the imaging library.
//
// The geometry a condition needs before any kernel runs: the
// computeD2BI_img prologue (link-cell geometry, s_g per precession point),
// the foil tilt for a requested s_g, the previously-determined-tilt path,
// relativistic electron constants, and a seeded synthetic crystal generator.
// All of it is FP64 host arithmetic evaluated in the reference's operation
// order (the library is compiled with -ffp-contract=off), because the link
// cell geometry and inverse matrices computed here fix every floor()
// membership decision the kernels make.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/d2bi_gpu.h"
#include "d2bi_internal.h"

namespace d2bi {

// Lib_DynamicalTwoBeamImaging.f90:55, Lib_DiffractionConditions.f90:22
static const double PI = 3.141592653589790;
// Lib_RelativisticElectrons.f90:9-13
static const double ME = 0.056856301036;
static const double EQ = 1.0;
static const double CL = 2997.92458;
static const double HP = 4.135667696;
static const double HBAR = HP / (2 * 3.14159265390);
// huge(1.0) of a default real, as a double
static const double HUGE_R4 = 3.4028234663852886e38;

static inline double& m3(double* m, int i, int j) { return m[3 * (j - 1) + (i - 1)]; }
static inline double m3(const double* m, int i, int j) { return m[3 * (j - 1) + (i - 1)]; }

// Lib_SimpleSupercells.f90:1238-1255
double det3(const double* M)
{
    double dd[9];
    dd[0] = m3(M,1,1) * (m3(M,2,2) * m3(M,3,3) - m3(M,2,3) * m3(M,3,2));
    dd[1] = m3(M,1,2) * (m3(M,2,3) * m3(M,3,1) - m3(M,2,1) * m3(M,3,3));
    dd[2] = m3(M,1,3) * (m3(M,2,1) * m3(M,3,2) - m3(M,2,2) * m3(M,3,1));
    dd[3] = m3(M,2,1) * (m3(M,3,2) * m3(M,1,3) - m3(M,3,3) * m3(M,1,2));
    dd[4] = m3(M,2,2) * (m3(M,3,3) * m3(M,1,1) - m3(M,3,1) * m3(M,1,3));
    dd[5] = m3(M,2,3) * (m3(M,3,1) * m3(M,1,2) - m3(M,3,2) * m3(M,1,1));
    dd[6] = m3(M,3,1) * (m3(M,1,2) * m3(M,2,3) - m3(M,1,3) * m3(M,2,2));
    dd[7] = m3(M,3,2) * (m3(M,1,3) * m3(M,2,1) - m3(M,1,1) * m3(M,2,3));
    dd[8] = m3(M,3,3) * (m3(M,1,1) * m3(M,2,2) - m3(M,1,2) * m3(M,2,1));
    double s = 0.0;
    for (int i = 0; i < 9; i++) s = s + dd[i];
    return (1.0 / 3.0) * s;
}

// Lib_SimpleSupercells.f90:1205-1236
void inverse3(const double* M, double* N)
{
    double idd = det3(M);
    if (std::fabs(idd) < 2.2250738585072014e-308) {
        for (int i = 0; i < 9; i++) N[i] = (i % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    idd = 1.0 / idd;
    m3(N,1,1) = (m3(M,2,2) * m3(M,3,3) - m3(M,2,3) * m3(M,3,2)) * idd;
    m3(N,2,1) = (m3(M,2,3) * m3(M,3,1) - m3(M,2,1) * m3(M,3,3)) * idd;
    m3(N,3,1) = (m3(M,2,1) * m3(M,3,2) - m3(M,2,2) * m3(M,3,1)) * idd;
    m3(N,1,2) = (m3(M,1,3) * m3(M,3,2) - m3(M,1,2) * m3(M,3,3)) * idd;
    m3(N,2,2) = (m3(M,1,1) * m3(M,3,3) - m3(M,1,3) * m3(M,3,1)) * idd;
    m3(N,3,2) = (m3(M,1,2) * m3(M,3,1) - m3(M,1,1) * m3(M,3,2)) * idd;
    m3(N,1,3) = (m3(M,1,2) * m3(M,2,3) - m3(M,1,3) * m3(M,2,2)) * idd;
    m3(N,2,3) = (m3(M,1,3) * m3(M,2,1) - m3(M,1,1) * m3(M,2,3)) * idd;
    m3(N,3,3) = (m3(M,1,1) * m3(M,2,2) - m3(M,1,2) * m3(M,2,1)) * idd;
}

// Fortran matmul for 3x3 (k ascending); C may alias A or B.
void matmul3(const double* A, const double* B, double* Cm)
{
    double T[9];
    for (int j = 1; j <= 3; j++)
        for (int i = 1; i <= 3; i++)
            m3(T,i,j) = (m3(A,i,1) * m3(B,1,j) + m3(A,i,2) * m3(B,2,j)) + m3(A,i,3) * m3(B,3,j);
    std::memcpy(Cm, T, sizeof T);
}

// NORM2 as a scaled sum of squares.
double norm2_3(const double* v)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < 3; i++) {
        if (v[i] != 0.0) {
            double ax = std::fabs(v[i]);
            if (scale < ax) { ssq = 1.0 + ssq * (scale / ax) * (scale / ax); scale = ax; }
            else            { ssq = ssq + (ax / scale) * (ax / scale); }
        }
    }
    return scale * std::sqrt(ssq);
}

// getReciprocalLatticeVectors (Lib_DiffractionConditions.f90:1630-1638)
static void recip(const double* A, double* B)
{
    double At[9];
    for (int i = 1; i <= 3; i++) for (int j = 1; j <= 3; j++) m3(At,i,j) = m3(A,j,i);
    inverse3(At, B);
    for (int i = 0; i < 9; i++) B[i] = 2 * PI * B[i];
}

// reflection (Lib_DiffractionConditions.f90:1643-1667)
static void reflection(const double* B, const double* hkl, double* g)
{
    for (int i = 1; i <= 3; i++)
        g[i - 1] = (m3(B,i,1) * hkl[0] + m3(B,i,2) * hkl[1]) + m3(B,i,3) * hkl[2];
}

// Lib_RelativisticElectrons.f90:24-36
double wavelength(double V)
{
    if (V <= 0) return 0.0;
    double lambda = 2 * ME * EQ * V * (1 + EQ * V / (2 * ME * CL * CL));
    return HP / std::sqrt(lambda);
}

// Lib_RelativisticElectrons.f90:40-50
double velocity(double V)
{
    double u = 1 + EQ * V / (ME * CL * CL);
    return CL * std::sqrt(1 - 1 / (u * u));
}

// energy_deviation_parameter (Lib_DiffractionConditions.f90:1265-1299)
static double energy_deviation(double v, const double* khat, const double* g)
{
    const double HBAR2ON2M = HBAR * HBAR / (2 * ME);
    double cst = 2 * ME * v / HBAR;
    double eps = (khat[0] * g[0] + khat[1] * g[1]) + khat[2] * g[2];
    eps = cst * eps;
    eps = ((eps + g[0] * g[0]) + g[1] * g[1]) + g[2] * g[2];
    return eps * HBAR2ON2M;
}

// energyToLengthDeviationParameter (Lib_DiffractionConditions.f90:1194-1204)
static double e2l(double v, double eps) { return -eps / (2 * PI * HBAR * v); }

// getDeviationParameter(this,g) (Lib_DynamicalTwoBeamImaging.f90:268-282)
double deviation_parameter(double E_keV, const double* g)
{
    static const double khat[3] = {0.0, 0.0, 1.0};
    double modk = 2 * PI / wavelength(E_keV * 1000);
    double vv = HBAR * modk / ME;
    return e2l(vv, energy_deviation(vv, khat, g));
}

// getAperture (Lib_DynamicalTwoBeamImaging.f90:1137-1156)
static void aperture(const d2bi_record* r, const double* tilt, double* g)
{
    double B[9], RU[9], gc[3];
    recip(r->a_cell, B);
    reflection(B, r->g, gc);
    matmul3(r->R, r->U, RU);
    if (tilt) matmul3(tilt, RU, RU);
    for (int i = 1; i <= 3; i++)
        g[i - 1] = (m3(RU,i,1) * gc[0] + m3(RU,i,2) * gc[1]) + m3(RU,i,3) * gc[2];
}

// estimateMaxCountPerCell1 (Lib_LinkCell3D.f90:320-349)
int estimate_max_count(double nCells, int nTotal)
{
    double rho = nTotal / nCells;
    double pp = std::exp(-rho), psum = pp;
    for (int kk = 1; kk <= nTotal; kk++) {
        if (psum >= 0.99) return kk;
        pp = pp * rho / kk;
        psum = psum + pp;
    }
    return nTotal;
}

// computeD2BI_img prologue (Lib_DynamicalTwoBeamImaging.f90:1635-1836) for
// one rank: factoriseParallel(Mx,My,1) gives one patch (Lib_FactoriseParallel.f90:104),
// nAperturePoints = 1 gives a zero aperture offset.
int plan_image(const d2bi_record* r, int64_t nAtoms, double precAngle, int nPrec, d2bi_plan* p)
{
    std::memset(p, 0, sizeof *p);
    if (nPrec < 1 || nPrec > D2BI_MAX_PREC) return seterr(D2BI_ERR_ARG, "nPrec out of range");
    if (nAtoms < 0 || nAtoms > INT32_MAX) return seterr(D2BI_ERR_ARG, "nAtoms out of range");
    for (int i = 0; i < 3; i++)
        if (r->gk_n[i] < 1) return seterr(D2BI_ERR_ARG, "voxel counts must be positive");
    double B[9], gc[3];
    recip(r->a_cell, B);
    reflection(B, r->g, gc);
    p->modg = norm2_3(gc);
    for (int i = 0; i < 3; i++) p->xmax[i] = norm2_3(&r->gk_a[3 * i]) * r->gk_n[i];
    p->Mx = r->gk_n[0]; p->My = r->gk_n[1]; p->Mz = r->gk_n[2];
    p->aperpx = std::sqrt(p->xmax[0] * p->xmax[1] / (double)(p->Mx * p->My));
    p->nPrec = nPrec;
    double costheta = 1.0, sintheta = 0.0;
    if (nPrec > 1) { costheta = std::cos(precAngle); sintheta = std::sin(precAngle); }
    p->sigma = r->a0 / 2;
    p->i2s2 = 1 / (2 * p->sigma * p->sigma);
    double l1 = norm2_3(&r->gk_a[0]), l2 = norm2_3(&r->gk_a[3]);
    p->deltaM = (int)std::ceil(3 * p->sigma / (l1 < l2 ? l1 : l2));
    p->nMax = estimate_max_count((double)p->Mx * p->My * p->Mz, (int)nAtoms) * p->deltaM * p->deltaM;
    // real(Dx)/deltaM + 2 is a default-real (FP32) expression in the reference
    int Dx = (int)std::ceil((float)p->Mx / 1.0f), Dy = (int)std::ceil((float)p->My / 1.0f);
    int ex = (int)std::floor((float)Dx / (float)p->deltaM + 2.0f);
    int ey = (int)std::floor((float)Dy / (float)p->deltaM + 2.0f);
    int ez = (int)std::floor((float)p->Mz / (float)p->deltaM);
    p->Ex = std::max(3, ex); p->Ey = std::max(3, ey); p->Ez = std::max(3, ez);
    for (int i = 0; i < 3; i++) {
        p->a_col[i]     = ((double)(Dx + 2 * p->deltaM) * r->gk_a[i]) / p->Ex;
        p->a_col[3 + i] = ((double)(Dy + 2 * p->deltaM) * r->gk_a[3 + i]) / p->Ey;
        p->a_col[6 + i] = ((double)p->Mz * r->gk_a[6 + i]) / p->Ez;
    }
    inverse3(p->a_col, p->ia_col);
    for (int i = 0; i < 3; i++) {
        double g0 = (r->gk_a[i] * 0 + r->gk_a[3 + i] * 0) + r->gk_a[6 + i] * 0;
        double s1 = (p->a_col[i] * 1 + p->a_col[3 + i] * 1) + p->a_col[6 + i] * 0;
        p->lc_delta[i] = g0 - s1;
    }
    matmul3(r->R, r->U, p->RU);
    aperture(r, nullptr, p->gg);
    p->dz = p->xmax[2] / p->Mz;
    p->bg = PI / r->xig;
    p->b0 = PI / r->xi0;
    for (int prec = 1; prec <= nPrec; prec++) {
        double gg[3];
        if (prec == 1) {
            aperture(r, nullptr, gg);
        } else {
            // Rodrigues rotation by the precession angle about (sin phi, cos phi, 0) (:1814-1820)
            double phi = (prec - 1) * 2 * PI / (nPrec - 1);
            double ux = std::sin(phi), uy = std::cos(phi), T[9];
            m3(T,1,1) = costheta + ux * ux * (1 - costheta); m3(T,1,2) = ux * uy * (1 - costheta);             m3(T,1,3) = uy * sintheta;
            m3(T,2,1) = ux * uy * (1 - costheta);             m3(T,2,2) = costheta + uy * uy * (1 - costheta); m3(T,2,3) = -ux * sintheta;
            m3(T,3,1) = -uy * sintheta;                       m3(T,3,2) = ux * sintheta;                       m3(T,3,3) = costheta;
            aperture(r, T, gg);
        }
        double g2[3] = {gg[0] + 0.0, gg[1] + 0.0, gg[2] + 0.0};
        p->aa[prec - 1] = 2 * PI * deviation_parameter(r->E, g2);
    }
    p->rc2 = (3 * p->sigma) * (3 * p->sigma);
    return D2BI_OK;
}

// cross_product (Lib_SimpleSupercells.f90:1194-1202)
static void cross3(const double* x, const double* y, double* z)
{
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
}

// rotateToStandardFrame0 (Lib_DiffractionConditions.f90:65-108)
void rotate_to_standard_frame(const double* z0, const double* A, const double* hkl, double* U)
{
    double BB[9], gg[3], zz[3], xx[3], yy[3];
    recip(A, BB);
    reflection(BB, hkl, gg);
    for (int i = 1; i <= 3; i++) zz[i - 1] = (m3(A,i,1) * z0[0] + m3(A,i,2) * z0[1]) + m3(A,i,3) * z0[2];
    double nz = norm2_3(zz);
    for (int i = 0; i < 3; i++) zz[i] = zz[i] / nz;
    double gz = (gg[0] * zz[0] + gg[1] * zz[1]) + gg[2] * zz[2];
    for (int i = 0; i < 3; i++) xx[i] = gg[i] - gz * zz[i];
    double nx = norm2_3(xx);
    for (int i = 0; i < 3; i++) xx[i] = xx[i] / nx;
    cross3(zz, xx, yy);
    for (int j = 1; j <= 3; j++) { m3(U,1,j) = xx[j - 1]; m3(U,2,j) = yy[j - 1]; m3(U,3,j) = zz[j - 1]; }
}

// findTransformedSupercell (Lib_DiffractionConditions.f90:996-1102): the
// cuboid D = diag(d), the offset delta centring the rotated box on it, and
// always the 27 replicas uvw in (-1..1)^3 with x fastest.
static void transformed_supercell(const double* A_super, const double* RU, const double* d,
                                  double* D, double* delta, int32_t* uvw, int32_t* nuvw)
{
    for (int i = 0; i < 9; i++) D[i] = 0.0;
    D[0] = d[0]; D[4] = d[1]; D[8] = d[2];
    double xx[3] = {0.5, 0.5, 0.5}, yy[3];
    for (int i = 1; i <= 3; i++) yy[i - 1] = (m3(A_super,i,1) * xx[0] + m3(A_super,i,2) * xx[1]) + m3(A_super,i,3) * xx[2];
    for (int i = 1; i <= 3; i++) xx[i - 1] = (m3(RU,i,1) * yy[0] + m3(RU,i,2) * yy[1]) + m3(RU,i,3) * yy[2];
    for (int i = 0; i < 3; i++) delta[i] = d[i] / 2 - xx[i];
    int n = 0;
    for (int iz = -1; iz <= 1; iz++)
        for (int iy = -1; iy <= 1; iy++)
            for (int ix = -1; ix <= 1; ix++) {
                uvw[3 * n] = ix; uvw[3 * n + 1] = iy; uvw[3 * n + 2] = iz;
                n++;
            }
    *nuvw = n;
}

static int nint(double v) { return (int)std::lround(v); }

// Common tail of setSupercellTilt0/2: replicas, delta and the voxel cell
// (Lib_DynamicalTwoBeamImaging.f90:772-786, 1070-1104).
static void finish_tilt(d2bi_record* r, int mxy, int mzz, const double* d)
{
    double RU[9], A_super[9], D[9];
    matmul3(r->R, r->U, RU);
    for (int j = 0; j < 3; j++)                      // getSuperA (Lib_SimpleSupercells.f90:851-860)
        for (int i = 0; i < 3; i++) A_super[3 * j + i] = r->super_a[3 * j + i] * r->super_n[j];
    transformed_supercell(A_super, RU, d, D, r->delta, r->uvw, &r->nuvw);
    int nx = std::max(mxy, nint(mxy * D[0] / r->a0));
    int ny = std::max(mxy, nint(mxy * D[4] / r->a0));
    int nz = std::max(mzz, nint(mzz * D[8] / r->a0));
    D[0] = D[0] / nx; D[4] = D[4] / ny; D[8] = D[8] / nz;
    std::memcpy(r->gk_a, D, sizeof D);
    r->gk_n[0] = nx; r->gk_n[1] = ny; r->gk_n[2] = nz;
    double gg[3];
    aperture(r, nullptr, gg);
    r->sg = deviation_parameter(r->E, gg);
}

static void rot_y(double theta, double* R)
{
    // reshape((/cos,0,-sin, 0,1,0, sin,0,cos/),(/3,3/)) (:659)
    double c = std::cos(theta), s = std::sin(theta);
    R[0] = c;  R[1] = 0.0; R[2] = -s;
    R[3] = 0.0; R[4] = 1.0; R[5] = 0.0;
    R[6] = s;  R[7] = 0.0; R[8] = c;
}

// setSupercellTilt0 (Lib_DynamicalTwoBeamImaging.f90:567-788)
int setup_tilt_sg(d2bi_record* r, const double* k, double sg, int mxy, int mzz, const double* d)
{
    double hkl[3] = {(double)nint(r->g[0]), (double)nint(r->g[1]), (double)nint(r->g[2])};
    rotate_to_standard_frame(k, r->a_cell, hkl, r->U);
    double a_cell0[9], b_cell[9], a_cell[9], gg[3];
    matmul3(r->U, r->a_cell, a_cell0);
    recip(a_cell0, b_cell);
    reflection(b_cell, r->g, gg);
    double modg = norm2_3(gg);
    static const double khat[3] = {0.0, 0.0, 1.0};
    double modk = 2 * PI / wavelength(r->E * 1000);
    double vv = HBAR * modk / ME;
    double eps = energy_deviation(vv, khat, gg);
    static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (std::isnan(sg)) {
        r->sg = e2l(vv, eps);
        std::memcpy(r->R, I3, sizeof I3);
    } else {
        double aa = 2 * PI * sg + HBAR * modg * modg / (2 * ME * vv);
        double cc = gg[0] * gg[0] + gg[2] * gg[2];
        double bb = cc - aa * aa;
        if (bb < 0) {
            // "s_g/|g| too large": keep zero tilt (:659-662)
            r->sg = e2l(vv, eps);
            std::memcpy(r->R, I3, sizeof I3);
        } else {
            bb = std::sqrt(bb);
            cc = 1 / cc;
            double th[4];
            th[0] = std::acos(-(gg[2] * aa + gg[0] * bb) * cc);
            th[1] = std::acos(-(gg[2] * aa - gg[0] * bb) * cc);
            th[2] = -th[0];
            th[3] = -th[1];
            double sgi[4];
            for (int i = 0; i < 4; i++) {
                rot_y(th[i], r->R);
                matmul3(r->R, a_cell0, a_cell);
                recip(a_cell, b_cell);
                reflection(b_cell, r->g, gg);
                sgi[i] = e2l(vv, energy_deviation(vv, khat, gg));
                if (std::fabs(th[i]) * 2 > PI) sgi[i] = HUGE_R4;
            }
            int best = 0;                       // minloc: first minimum
            double bv = std::fabs(sg - sgi[0]);
            for (int i = 1; i < 4; i++)
                if (std::fabs(sg - sgi[i]) < bv) { bv = std::fabs(sg - sgi[i]); best = i; }
            rot_y(th[best], r->R);
        }
        matmul3(r->R, a_cell0, a_cell);
        recip(a_cell, b_cell);
        reflection(b_cell, r->g, gg);
        r->sg = e2l(vv, energy_deviation(vv, khat, gg));
    }
    finish_tilt(r, mxy, mzz, d);
    return D2BI_OK;
}

// setSupercellTilt2 (Lib_DynamicalTwoBeamImaging.f90:1048-1106)
int setup_tilt_ur(d2bi_record* r, const double* U, const double* R, int mxy, int mzz, const double* d)
{
    std::memcpy(r->U, U, 9 * sizeof(double));
    std::memcpy(r->R, R, 9 * sizeof(double));
    finish_tilt(r, mxy, mzz, d);
    return D2BI_OK;
}

// ------------------------------------------------------------------ synthetic crystals

namespace {
struct Rng {                       // splitmix64 stream
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * PI * u2);
    }
};
}  // namespace

int synth_bcc(const d2bi_synth_params* p, double* x, int64_t cap, int64_t* nAtoms, double* box)
{
    if (!p || !nAtoms || p->nx < 1 || p->ny < 1 || p->nz < 1 || p->a0 <= 0)
        return seterr(D2BI_ERR_ARG, "bad synthetic crystal parameters");
    const double L[3] = {p->a0 * p->nx, p->a0 * p->ny, p->a0 * p->nz};
    if (box) { box[0] = L[0]; box[1] = L[1]; box[2] = L[2]; }
    Rng rng(p->seed ? p->seed : 1);
    std::vector<double> cen(3 * std::max(p->n_centres, 0)), amp(std::max(p->n_centres, 0));
    for (int c = 0; c < p->n_centres; c++) {
        for (int i = 0; i < 3; i++) cen[3 * c + i] = rng.uniform() * L[i];
        amp[c] = p->A_min + (p->A_max - p->A_min) * rng.uniform();
    }
    std::vector<double> vc(3 * std::max(p->n_voids, 0)), vr(std::max(p->n_voids, 0));
    for (int v = 0; v < p->n_voids; v++) {
        for (int i = 0; i < 3; i++) vc[3 * v + i] = rng.uniform() * L[i];
        vr[v] = p->r_min + (p->r_max - p->r_min) * rng.uniform();
    }
    const int64_t nsites = 2LL * p->nx * p->ny * p->nz;
    const uint64_t noise_seed = rng.next();
    auto mimg = [&](double d, int i) { return d - L[i] * std::nearbyint(d / L[i]); };
    // site s: cell (cx,cy,cz) = s/2 with x fastest, basis b = s%2
    std::vector<uint8_t> keep(nsites);
    std::vector<double> pos(3 * nsites);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](unsigned t) {
        for (int64_t s = t; s < nsites; s += nth) {
            int64_t cell = s >> 1;
            int b = (int)(s & 1);
            int cx = (int)(cell % p->nx), cy = (int)((cell / p->nx) % p->ny), cz = (int)(cell / ((int64_t)p->nx * p->ny));
            double r0[3] = {(cx + 0.5 * b) * p->a0, (cy + 0.5 * b) * p->a0, (cz + 0.5 * b) * p->a0};
            bool in_void = false;
            for (int v = 0; v < p->n_voids && !in_void; v++) {
                double d0 = mimg(r0[0] - vc[3 * v], 0), d1 = mimg(r0[1] - vc[3 * v + 1], 1),
                       d2 = mimg(r0[2] - vc[3 * v + 2], 2);
                in_void = d0 * d0 + d1 * d1 + d2 * d2 < vr[v] * vr[v];
            }
            keep[s] = !in_void;
            double u[3] = {0.0, 0.0, 0.0};
            for (int c = 0; c < p->n_centres; c++) {
                double d[3] = {mimg(r0[0] - cen[3 * c], 0), mimg(r0[1] - cen[3 * c + 1], 1),
                               mimg(r0[2] - cen[3 * c + 2], 2)};
                double rr = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (rr < 1e-9) continue;
                double m = amp[c] / (rr * rr);                   // |u| = A/r^2
                if (m > p->u_cap) m = p->u_cap;
                for (int i = 0; i < 3; i++) u[i] += m * d[i] / rr;
            }
            Rng nr(noise_seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(s + 1)));
            for (int i = 0; i < 3; i++) {
                double v = r0[i] + u[i] + p->noise * nr.normal();
                v = v - L[i] * std::floor(v / L[i]);            // wrap into the periodic box
                pos[3 * s + i] = v + p->offset[i];
            }
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; t++) th.emplace_back(work, t);
    for (auto& t : th) t.join();
    int64_t n = 0;
    for (int64_t s = 0; s < nsites; s++) {
        if (!keep[s]) continue;
        if (x) {
            if (n >= cap) return seterr(D2BI_ERR_CAPACITY, "synthetic crystal larger than cap");
            x[3 * n] = pos[3 * s]; x[3 * n + 1] = pos[3 * s + 1]; x[3 * n + 2] = pos[3 * s + 2];
        }
        n++;
    }
    *nAtoms = n;
    return D2BI_OK;
}

}  // namespace d2bi

// ------------------------------------------------------------------ C-ABI (host part)

extern "C" {

int d2bi_plan_image(const d2bi_record* rec, int64_t nAtoms, double precAngle, int32_t nPrec, d2bi_plan* plan)
{
    if (!rec || !plan) return d2bi::seterr(D2BI_ERR_ARG, "null argument");
    return d2bi::plan_image(rec, nAtoms, precAngle, nPrec, plan);
}

int d2bi_setup_tilt_sg(d2bi_record* rec, const double k[3], double sg_target, int32_t mxy, int32_t mzz,
                       const double d[3])
{
    if (!rec || !k || !d || mxy < 1 || mzz < 1) return d2bi::seterr(D2BI_ERR_ARG, "bad argument");
    if (!(d[0] > 0 && d[1] > 0 && d[2] > 0)) return d2bi::seterr(D2BI_ERR_ARG, "voxel cuboid sides must be > 0");
    return d2bi::setup_tilt_sg(rec, k, sg_target, mxy, mzz, d);
}

int d2bi_setup_tilt_ur(d2bi_record* rec, const double U[9], const double R[9], int32_t mxy, int32_t mzz,
                       const double d[3])
{
    if (!rec || !U || !R || !d || mxy < 1 || mzz < 1) return d2bi::seterr(D2BI_ERR_ARG, "bad argument");
    if (!(d[0] > 0 && d[1] > 0 && d[2] > 0)) return d2bi::seterr(D2BI_ERR_ARG, "voxel cuboid sides must be > 0");
    return d2bi::setup_tilt_ur(rec, U, R, mxy, mzz, d);
}

int d2bi_rotate_to_standard_frame(const double k[3], const double a_cell[9], const double hkl[3], double U[9])
{
    if (!k || !a_cell || !hkl || !U) return d2bi::seterr(D2BI_ERR_ARG, "null argument");
    d2bi::rotate_to_standard_frame(k, a_cell, hkl, U);
    return D2BI_OK;
}

double d2bi_wavelength(double V) { return d2bi::wavelength(V); }
double d2bi_velocity(double V) { return d2bi::velocity(V); }
double d2bi_deviation_parameter(double E_keV, const double g_lab[3]) { return d2bi::deviation_parameter(E_keV, g_lab); }

int d2bi_synth_bcc(const d2bi_synth_params* p, double* x, int64_t cap, int64_t* nAtoms, double box[3])
{
    return d2bi::synth_bcc(p, x, cap, nAtoms, box);
}

}  // extern "C"
