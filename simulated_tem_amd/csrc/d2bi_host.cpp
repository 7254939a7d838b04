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
#include <cstdio>
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

#define HCHK(x)                      \
    do {                             \
        const int rc_ = (x);         \
        if (rc_ < 0) return rc_;     \
    } while (0)

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

// The precession-tilted g-vector of point prec of nPrec (computeD2BI_img
// :1805-1822): prec = 1 is the untilted getAperture; otherwise the Rodrigues
// rotation by the precession angle about (sin phi, cos phi, 0),
// phi = (prec - 1) 2 pi / (nPrec - 1).
static void precession_g(const d2bi_record* r, double costheta, double sintheta, int prec, int nPrec, double* gg)
{
    if (prec == 1) {
        aperture(r, nullptr, gg);
        return;
    }
    double phi = (prec - 1) * 2 * PI / (nPrec - 1);
    double ux = std::sin(phi), uy = std::cos(phi), T[9];
    m3(T,1,1) = costheta + ux * ux * (1 - costheta); m3(T,1,2) = ux * uy * (1 - costheta);             m3(T,1,3) = uy * sintheta;
    m3(T,2,1) = ux * uy * (1 - costheta);             m3(T,2,2) = costheta + uy * uy * (1 - costheta); m3(T,2,3) = -ux * sintheta;
    m3(T,3,1) = -uy * sintheta;                       m3(T,3,2) = ux * sintheta;                       m3(T,3,3) = costheta;
    aperture(r, T, gg);
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
        precession_g(r, costheta, sintheta, prec, nPrec, gg);
        double g2[3] = {gg[0] + 0.0, gg[1] + 0.0, gg[2] + 0.0};
        p->aa[prec - 1] = 2 * PI * deviation_parameter(r->E, g2);
    }
    p->rc2 = (3 * p->sigma) * (3 * p->sigma);
    return D2BI_OK;
}

// The arrangements of n = 1..25 points in the unit disk that
// evenlySpacedPointsInCircle returns from its table (src/evenlySpaced.h,
// xKnown): the same values, written as the default-real (FP32) literals they
// are there; set n holds its n points (x, y) and starts at n (n - 1).
static const int APERTURE_KNOWN = 25;
static const float APERTURE_XY[APERTURE_KNOWN * (APERTURE_KNOWN + 1)] = {
    /* n = 1 */ 0.0000f, 0.0000f,
    /* n = 2 */ 0.0033f, 0.4139f, 0.0013f, -0.4234f,
    /* n = 3 */ 0.2856f, -0.4698f, 0.2730f, 0.4802f, -0.5436f, -0.0069f,
    /* n = 4 */ -0.4189f, 0.4201f, 0.4250f, -0.4263f, -0.4162f, -0.4287f, 0.4276f, 0.4157f,
    /* n = 5 */ 0.0000f, 0.6230f, 0.5925f, 0.1925f, -0.5925f, 0.1925f, 0.3662f, -0.5040f, -0.3662f, -0.5040f,
    /* n = 6 */ 0.0962f, 0.6590f, -0.0001f, -0.0001f, -0.4775f, -0.4790f, 0.6566f, 0.1037f, -0.5890f, 0.2967f, 0.2995f, -0.5873f,
    /* n = 7 */ 0.3594f, 0.6018f, -0.6860f, -0.0086f, 0.3376f, -0.6123f, -0.3477f, 0.5971f, -0.3484f, -0.5998f, 0.0086f, -0.0080f, 0.7051f, -0.0060f,
    /* n = 8 */ 0.6889f, -0.1686f, 0.0038f, 0.7079f, -0.3111f, -0.6400f, 0.0016f, 0.0005f, 0.5710f, 0.4348f, 0.3032f, -0.6378f, -0.5613f, 0.4424f, -0.6888f, -0.1576f,
    /* n = 9 */ 0.6646f, 0.2658f, 0.2842f, 0.6581f, -0.6610f, 0.2900f, -0.6678f, -0.2635f, -0.2833f, -0.6631f, 0.6514f, -0.2895f, -0.0023f, 0.0005f, -0.2673f, 0.6766f, 0.2629f, -0.6723f,
    /* n = 10 */ 0.2373f, -0.6995f, -0.1931f, 0.7037f, 0.6264f, -0.2817f, -0.6888f, -0.2767f, 0.1319f, 0.2148f, -0.1001f, -0.2027f, -0.6039f, 0.2969f, -0.3247f, -0.6950f, 0.3586f, 0.6971f, 0.7013f, 0.2768f,
    /* n = 11 */ 0.3003f, 0.6941f, 0.6673f, -0.2784f, 0.7005f, 0.2817f, 0.1181f, -0.2787f, 0.1747f, 0.2016f, -0.2704f, 0.6578f, 0.3067f, -0.7240f, -0.2731f, 0.0124f, -0.7039f, 0.2876f, -0.2797f, -0.6767f, -0.6963f, -0.2989f,
    /* n = 12 */ -0.3180f, -0.6993f, 0.7739f, 0.1049f, -0.4182f, 0.6592f, -0.7286f, 0.2559f, 0.6023f, -0.4005f, 0.0807f, 0.7225f, -0.6525f, -0.2809f, -0.2132f, 0.1957f, 0.2107f, -0.7415f, 0.2788f, 0.1516f, 0.5652f, 0.5639f, -0.0172f, -0.2730f,
    /* n = 13 */ -0.1471f, -0.7076f, 0.7065f, -0.3687f, 0.3496f, -0.7076f, -0.2945f, -0.0995f, 0.0604f, 0.3004f, 0.5327f, 0.5633f, 0.7069f, 0.1266f, 0.1402f, 0.7875f, -0.7937f, -0.0975f, -0.5862f, -0.5237f, 0.2236f, -0.2110f, -0.3006f, 0.7006f, -0.6406f, 0.3619f,
    /* n = 14 */ -0.7275f, -0.1520f, -0.5471f, -0.5864f, 0.7013f, -0.3425f, -0.2605f, 0.1259f, -0.3260f, 0.6751f, 0.1584f, 0.3358f, -0.1551f, -0.3397f, 0.2597f, -0.1274f, 0.5492f, 0.5868f, -0.1374f, -0.7951f, 0.7228f, 0.1550f, 0.3247f, -0.6801f, 0.1308f, 0.7898f, -0.7022f, 0.3449f,
    /* n = 15 */ -0.1422f, -0.2654f, 0.5549f, 0.5720f, 0.3199f, -0.1904f, -0.3215f, 0.1717f, 0.5268f, -0.6078f, -0.6725f, -0.3334f, -0.7810f, 0.1009f, 0.7698f, -0.2567f, -0.2621f, 0.7040f, 0.1703f, 0.7755f, 0.7312f, 0.1747f, 0.1118f, -0.7450f, -0.6463f, 0.5109f, 0.1538f, 0.2725f, -0.3652f, -0.7051f,
    /* n = 16 */ -0.1425f, -0.8263f, 0.3192f, 0.7491f, -0.3469f, -0.1454f, -0.0851f, -0.4567f, -0.5411f, 0.5850f, -0.7588f, -0.2732f, 0.6504f, -0.4274f, 0.7421f, 0.0413f, 0.6535f, 0.4786f, 0.2439f, 0.3046f, -0.2068f, 0.2518f, -0.7447f, 0.1788f, -0.1177f, 0.7505f, -0.5076f, -0.6124f, 0.2879f, -0.7184f, 0.2160f, -0.1549f,
    /* n = 17 */ 0.3161f, -0.6961f, 0.7604f, 0.0421f, 0.3137f, -0.2210f, -0.1790f, 0.8009f, 0.0068f, 0.0087f, -0.5505f, 0.5547f, -0.3949f, 0.0686f, -0.1523f, -0.3575f, -0.7933f, 0.1922f, -0.0899f, 0.3987f, 0.6943f, -0.4170f, 0.3265f, 0.2362f, -0.7301f, -0.2495f, 0.6556f, 0.4892f, -0.5217f, -0.6202f, 0.2699f, 0.7318f, -0.1202f, -0.7983f,
    /* n = 18 */ -0.7911f, -0.2109f, 0.0114f, 0.0026f, -0.5840f, -0.5665f, 0.4079f, -0.0799f, 0.5603f, -0.5419f, -0.2173f, -0.7520f, 0.2176f, 0.7831f, -0.5727f, 0.5910f, -0.2718f, 0.2916f, 0.0768f, -0.3942f, 0.7977f, -0.1884f, 0.2155f, -0.7916f, 0.2083f, 0.3514f, 0.5645f, 0.5914f, -0.1812f, 0.7497f, 0.7394f, 0.2412f, -0.7427f, 0.1961f, -0.3600f, -0.1834f,
    /* n = 19 */ 0.5823f, 0.5849f, -0.2985f, -0.3013f, 0.5655f, -0.5519f, 0.4129f, -0.0942f, -0.5733f, -0.5940f, -0.5599f, 0.5592f, -0.1124f, 0.4167f, -0.7626f, -0.2152f, -0.4058f, 0.1159f, 0.2056f, 0.7598f, 0.7620f, 0.2131f, 0.7966f, -0.2076f, -0.2079f, 0.7953f, -0.7934f, 0.2078f, 0.3019f, 0.3115f, 0.0020f, 0.0068f, 0.1144f, -0.4037f, 0.2255f, -0.7911f, -0.1946f, -0.7611f,
    /* n = 20 */ 0.4424f, 0.7036f, -0.2951f, 0.7814f, -0.0001f, 0.0001f, 0.5143f, -0.6032f, 0.4078f, -0.1321f, -0.8061f, 0.2028f, -0.2948f, -0.2994f, 0.6825f, 0.4634f, 0.3019f, 0.3235f, 0.7812f, 0.1268f, 0.0898f, 0.7776f, -0.5925f, 0.5199f, 0.7769f, -0.2838f, -0.5867f, -0.5830f, 0.1010f, -0.4216f, -0.2134f, -0.7668f, -0.1463f, 0.4002f, -0.7656f, -0.2059f, -0.4318f, 0.1006f, 0.1718f, -0.8091f,
    /* n = 21 */ 0.8165f, -0.1460f, 0.5352f, 0.6043f, -0.8184f, -0.1478f, 0.7606f, 0.2509f, -0.4306f, -0.0769f, 0.0159f, -0.7999f, -0.3625f, -0.7562f, 0.3833f, -0.7482f, -0.6344f, -0.4840f, 0.6463f, -0.4836f, -0.2084f, 0.7760f, 0.0104f, 0.0054f, -0.5717f, 0.6030f, 0.3575f, 0.2481f, 0.0662f, 0.4520f, 0.2036f, -0.4038f, -0.7528f, 0.2409f, -0.1885f, -0.4090f, -0.2975f, 0.3129f, 0.4449f, -0.1051f, 0.1974f, 0.8049f,
    /* n = 22 */ 0.0085f, 0.4460f, -0.0187f, -0.7993f, 0.0108f, 0.8365f, -0.6532f, 0.5300f, -0.2132f, -0.3979f, 0.1917f, -0.4111f, 0.4432f, -0.1274f, -0.6452f, -0.4921f, -0.3593f, 0.2846f, 0.8167f, -0.2173f, 0.3603f, 0.7114f, -0.3781f, -0.7477f, 0.3468f, -0.7635f, 0.6209f, -0.5145f, -0.7900f, 0.2032f, -0.3486f, 0.7302f, 0.0012f, -0.0004f, 0.3639f, 0.2640f, -0.8332f, -0.1632f, 0.6752f, 0.5002f, -0.4575f, -0.0914f, 0.7897f, 0.1464f,
    /* n = 23 */ -0.4765f, -0.6921f, -0.6688f, 0.4623f, 0.7500f, -0.3959f, -0.3329f, -0.3293f, -0.1610f, -0.7879f, 0.5511f, 0.5885f, -0.2581f, 0.3898f, 0.2623f, 0.8028f, -0.0002f, -0.0008f, -0.8257f, 0.1656f, -0.4415f, 0.7195f, 0.4960f, -0.6411f, 0.0725f, -0.4485f, -0.6994f, -0.4664f, 0.1960f, -0.8234f, -0.7959f, -0.1691f, 0.4021f, -0.2222f, -0.4644f, 0.0538f, 0.4364f, 0.1729f, -0.0950f, 0.7977f, 0.7973f, -0.0551f, 0.7866f, 0.3060f, 0.1419f, 0.4365f,
    /* n = 24 */ 0.6013f, -0.5800f, 0.8014f, 0.0627f, 0.3947f, 0.1186f, -0.1415f, 0.4891f, 0.2353f, -0.5040f, 0.4815f, -0.2330f, -0.2643f, 0.8082f, -0.3289f, -0.1665f, 0.4505f, 0.7163f, 0.6824f, 0.4334f, -0.5605f, 0.5859f, -0.4159f, 0.2031f, -0.7863f, 0.2494f, 0.0984f, -0.1833f, -0.5930f, -0.5287f, 0.3545f, -0.7979f, 0.0269f, -0.8269f, -0.3280f, -0.7932f, 0.0994f, 0.8117f, 0.8023f, -0.3052f, -0.0135f, 0.1498f, -0.7804f, -0.1618f, -0.1447f, -0.4981f, 0.2362f, 0.4391f,
    /* n = 25 */ 0.8145f, 0.1926f, 0.4560f, 0.6887f, 0.6773f, -0.4332f, 0.1368f, -0.8009f, 0.4587f, -0.6934f, 0.1329f, 0.4714f, 0.4648f, 0.3014f, 0.2515f, -0.3455f, -0.8148f, -0.1139f, 0.8440f, -0.1374f, -0.4449f, -0.2826f, -0.7389f, -0.4359f, 0.1303f, 0.8328f, -0.2030f, -0.8254f, 0.7119f, 0.4914f, -0.2229f, 0.7635f, -0.5702f, 0.5641f, 0.1746f, 0.0914f, -0.4918f, -0.6601f, -0.1979f, 0.3184f, -0.1157f, -0.1161f, -0.1193f, -0.4867f, -0.4816f, 0.0866f, 0.5036f, -0.0519f, -0.8046f, 0.2541f,
};

// evenlySpacedPointsInCircle(x) without a file
// (Lib_EvenlySpacedPointsInCircle.f90:29-100): n <= 25 points from the table,
// more from the sunflower placement with alpha = 0 (:438-463) and radius
// (:465-478), in that module's PI = 3.141592654.  xy is (2, n), x first.
int aperture_points(int n, double* xy)
{
    if (n < 1) return seterr(D2BI_ERR_ARG, "aperture points must be >= 1");
    if (n <= APERTURE_KNOWN) {
        const float* t = APERTURE_XY + n * (n - 1);
        for (int k = 0; k < 2 * n; k++) xy[k] = (double)t[k];
        return D2BI_OK;
    }
    const double pi = 3.141592654;
    const double phi = (std::sqrt(5.0) + 1.0) / 2;
    const double alpha = 0.0;
    const int bb = (int)std::lround(std::sqrt(alpha * alpha * n));
    for (int kk = 1; kk <= n; kk++) {
        double rr = 1;
        if (kk <= n - bb) rr = std::sqrt((double)(2 * kk - 1) / (2 * n - bb - 1));
        const double theta = 2 * pi * kk / (phi * phi);
        xy[2 * (kk - 1)] = rr * std::cos(theta);
        xy[2 * (kk - 1) + 1] = rr * std::sin(theta);
    }
    return D2BI_OK;
}

// computeD2BI_img with apertureAngle and nAperturePoints
// (Lib_DynamicalTwoBeamImaging.f90:1655-1667, 1805-1836) and
// findGdotR_aperture (:1336-1344): the aperture points scaled to
// modg = apertureAngle 2 pi / lambda, the shifted g-vector of each point for
// g.r (the untilted g plus the point's x and y), and
// aa(app, prec) = 2 pi s_g(g_prec + dg_app), app fastest.
int aperture_setup(const d2bi_record* r, double precAngle, int nPrec, double apertureAngle, int nApp,
                   double* gshift, double* aa)
{
    if (nPrec < 1 || nPrec > D2BI_MAX_PREC) return seterr(D2BI_ERR_ARG, "nPrec out of range");
    if (nApp < 1 || nApp > D2BI_MAX_APERTURE) return seterr(D2BI_ERR_ARG, "nAperture out of range");
    const double modk = 2 * PI / wavelength(r->E * 1000);
    const double modg = apertureAngle * modk;
    std::vector<double> xy(2 * nApp);
    HCHK(aperture_points(nApp, xy.data()));
    double gg0[3];
    aperture(r, nullptr, gg0);
    for (int app = 0; app < nApp; app++) {
        const double dg[3] = {modg * xy[2 * app], modg * xy[2 * app + 1], modg * 0.0};
        gshift[3 * app] = gg0[0] + dg[0];
        gshift[3 * app + 1] = gg0[1] + dg[1];
        gshift[3 * app + 2] = gg0[2];
    }
    double costheta = 1.0, sintheta = 0.0;
    if (nPrec > 1) { costheta = std::cos(precAngle); sintheta = std::sin(precAngle); }
    for (int prec = 1; prec <= nPrec; prec++) {
        double gg[3];
        precession_g(r, costheta, sintheta, prec, nPrec, gg);
        for (int app = 0; app < nApp; app++) {
            const double dg[3] = {modg * xy[2 * app], modg * xy[2 * app + 1], modg * 0.0};
            double g2[3] = {gg[0] + dg[0], gg[1] + dg[1], gg[2] + dg[2]};
            aa[app + nApp * (prec - 1)] = 2 * PI * deviation_parameter(r->E, g2);
        }
    }
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

// ------------------------------------------------------------------ (g, ng) tilt search
//
// setSupercellTilt1a / setSupercellTilt1 without atom positions
// (Lib_DynamicalTwoBeamImaging.f90:795-1045): find the crystal rotation U and
// the stage tilt R that put the (ng g) reflection at the Bragg condition with
// the weakest excitation of every other reflection (findBestRotationMatrix,
// Lib_DiffractionConditions.f90:129-357), then tilt about y to the dark-field
// minimum of |phi_g(L)|^2 (tiltToDarkField, :751-927).  A fractional ng
// interpolates the tilts of floor(ng) and ceiling(ng) by quaternion slerp.

// Lib_DiffractionConditions.f90:24-29
static const int NTHETA = 2000;
static const double RHO_MAX = 10.0;
static const double THETA_MAX = 5.0;                       // degrees
static const double DEGTORAD = 3.141592654 / 180;
// Lib_Lattices.f90:65-69, 100-102
enum { LATT_SC = 1, LATT_BCC = 2, LATT_FCC = 3, LATT_HCP = 4 };
enum { SYM_CUBIC = 1, SYM_HEX = 2 };

// Fortran NINT: nearest integer, halves away from zero
static int nint_away(double v) { return (int)std::lround(v); }

// Fortran MOD for integers (sign of the dividend), as C's %.
static int fmod_i(int a, int b) { return a % b; }

// permittedReflections (Lib_Lattices.f90:678-735): [hkl] with
// h^2 + k^2 + l^2 <= rho^2 (hcp: its own metric) allowed by the lattice,
// l outer, k, h inner.
static void permitted_reflections(int latt, double rho, std::vector<int>& hkl)
{
    const int mm = (int)std::ceil(rho + 1.0e-12);
    hkl.clear();
    for (int lp = -mm; lp <= mm; lp++)
        for (int kp = -mm; kp <= mm; kp++)
            for (int hp = -mm; hp <= mm; hp++) {
                int rr = lp * lp + kp * kp + hp * hp;
                bool ok;
                switch (latt) {
                    case LATT_FCC: {
                        if (rr > rho * rho + 1.0e-6) continue;
                        int hx = fmod_i(hp, 2), kx = fmod_i(kp, 2), lx = fmod_i(lp, 2);
                        ok = (hx * hx + kx * kx + lx * lx) - (hx * kx + kx * lx + lx * hx) == 0;
                        break;
                    }
                    case LATT_BCC:
                        if (rr > rho * rho + 1.0e-6) continue;
                        ok = fmod_i(hp + kp + lp, 2) == 0;
                        break;
                    case LATT_HCP:
                        // (hp-kp)^2/3 is an integer division in the reference
                        rr = (int)std::floor((double)((hp + kp) * (hp + kp) + (hp - kp) * (hp - kp) / 3) +
                                             lp * lp * 0.375);
                        if (rr > rho * rho + 1.0e-6) continue;
                        ok = !((fmod_i(hp + 2 * kp, 3) == 0) && (fmod_i(std::abs(lp), 2) == 1));
                        break;
                    default:
                        if (rr > rho * rho + 1.0e-6) continue;
                        ok = true;
                }
                if (ok) { hkl.push_back(hp); hkl.push_back(kp); hkl.push_back(lp); }
            }
}

// find_gvectors_permittedReflections (Lib_DiffractionConditions.f90:1727-1789):
// g-vectors of the permitted reflections; [000] is swapped to the first
// place and [hkl] to the second.  The second swap exchanges the index
// triples of places 1 and ihkl (not 2 and ihkl) in the reference; the
// extinction distances are looked up from these triples, so the same
// exchange is made here.
static void permitted_gvectors(int latt, const double* B, const int* hkl, double rho, std::vector<double>& g,
                               std::vector<int>& hp)
{
    permitted_reflections(latt, rho, hp);
    const int n = (int)hp.size() / 3;
    g.assign(3 * (size_t)n, 0.0);
    int i000 = 0, ihkl = 0;                 // 1-based, 0 = absent
    for (int ii = 1; ii <= n; ii++) {
        const int* h = &hp[3 * (ii - 1)];
        if (h[0] * h[0] + h[1] * h[1] + h[2] * h[2] == 0) i000 = ii;
        int d0 = h[0] - hkl[0], d1 = h[1] - hkl[1], d2 = h[2] - hkl[2];
        if (d0 * d0 + d1 * d1 + d2 * d2 == 0) ihkl = ii;
        double hd[3] = {(double)h[0], (double)h[1], (double)h[2]};
        reflection(B, hd, &g[3 * (ii - 1)]);
    }
    auto swap_g = [&](int a, int b) { for (int k = 0; k < 3; k++) std::swap(g[3 * (a - 1) + k], g[3 * (b - 1) + k]); };
    auto swap_h = [&](int a, int b) { for (int k = 0; k < 3; k++) std::swap(hp[3 * (a - 1) + k], hp[3 * (b - 1) + k]); };
    if (i000 != 0) { swap_g(1, i000); swap_h(1, i000); }
    if (ihkl != 0) { swap_g(2, ihkl); swap_h(1, ihkl); }
}

// symmetry (Lib_ReadExtinctionDistances.f90:111-170): standard form of a
// reflection under the lattice symmetry.
static void standard_reflection(const int* in, int symm, int* out)
{
    int hh = std::abs(in[0]), kk = std::abs(in[1]), ll = std::abs(in[2]);
    if (symm == SYM_CUBIC) {
        if (hh >= std::max(kk, ll)) {
            if (kk >= ll) { out[0] = hh; out[1] = kk; out[2] = ll; }
            else          { out[0] = hh; out[1] = ll; out[2] = kk; }
        } else if (kk >= std::max(hh, ll)) {
            if (hh >= ll) { out[0] = kk; out[1] = hh; out[2] = ll; }
            else          { out[0] = kk; out[1] = ll; out[2] = hh; }
        } else {
            if (hh >= kk) { out[0] = ll; out[1] = hh; out[2] = kk; }
            else          { out[0] = ll; out[1] = kk; out[2] = hh; }
        }
    } else if (symm == SYM_HEX) {
        if (in[0] * in[1] >= 0) {
            if (hh >= kk) { out[0] = hh; out[1] = kk; out[2] = ll; }
            else          { out[0] = kk; out[1] = hh; out[2] = ll; }
        } else {
            if (hh >= kk) { out[0] = hh; out[1] = -kk; out[2] = ll; }
            else          { out[0] = kk; out[1] = -hh; out[2] = ll; }
        }
    } else {
        out[0] = in[0]; out[1] = in[1]; out[2] = in[2];
    }
}

// readExtinctionDistance (Lib_ReadExtinctionDistances.f90:29-109) for a list
// of reflections: two header lines, then up to 1000 records
// "element h k l T V Re(U_g) Im(U_g) xi_g"; a reflection matches a record
// when their standard forms agree.  Missing reflections keep huge(1.0)
// (their beams are too weak to matter).  Returns whether all were found.
static int read_xi_table(const char* filename, const std::vector<int>& hkl, int symm, std::vector<double>& xi,
                         bool* all_found)
{
    const size_t n = hkl.size() / 3;
    xi.assign(n, HUGE_R4);
    std::vector<char> found(n, 0);
    std::vector<int> hs(3 * n);
    for (size_t j = 0; j < n; j++) standard_reflection(&hkl[3 * j], symm, &hs[3 * j]);
    FILE* f = std::fopen(filename, "r");
    if (!f) return seterr(D2BI_ERR_ARG, "extinction distance file \"%s\" not found", filename);
    char line[512];
    int ok = 1;
    for (int i = 0; i < 2 && ok; i++) ok = std::fgets(line, sizeof line, f) != nullptr;
    if (!ok) { std::fclose(f); return seterr(D2BI_ERR_ARG, "cannot read the header of \"%s\"", filename); }
    size_t nfound = 0;
    for (int rec = 0; rec < 1000 && nfound < n; rec++) {
        if (!std::fgets(line, sizeof line, f)) break;
        char name[64];
        int id[3];
        double rd[5];
        if (std::sscanf(line, "%63s %d %d %d %lf %lf %lf %lf %lf", name, &id[0], &id[1], &id[2], &rd[0], &rd[1],
                        &rd[2], &rd[3], &rd[4]) != 9)
            break;                              // end of the table / another format
        int is[3];
        standard_reflection(id, symm, is);
        for (size_t j = 0; j < n; j++)
            if (is[0] == hs[3 * j] && is[1] == hs[3 * j + 1] && is[2] == hs[3 * j + 2]) {
                if (!found[j]) nfound++;
                found[j] = 1;
                xi[j] = rd[4];
            }
    }
    std::fclose(f);
    *all_found = nfound == n;
    return D2BI_OK;
}

// intensityAtDepthZ (Lib_DiffractionConditions.f90:1209-1258): |phi_g(z)|^2
// in a perfect crystal from the two Bloch-wave eigenvalues.
static double intensity_at_depth(double xi0, double xig, double sg, double z)
{
    const double b0 = PI / xi0, bg = PI / xig, aa = 2 * PI * sg;
    const double dd = std::sqrt(aa * aa / 4 + bg * bg);
    const double l1 = (b0 + aa / 2) + dd, l2 = (b0 + aa / 2) - dd;
    double re = std::cos(l1 * z) - std::cos(l2 * z);
    double im = std::sin(l1 * z) - std::sin(l2 * z);
    const double f = bg / (2 * dd);
    re = re * f;
    im = im * f;
    return re * re - im * (-im);
}

// findRotationsForBraggCondition2 (Lib_DiffractionConditions.f90:1382-1518):
// azimuths phi (rotation about y after theta about x) putting g at the
// Bragg condition.  limit: 0 = |phi - phi0| <= THETA_MAX, 1 = twice that,
// 2 = no limit.
static int bragg_azimuths(double v, const double* g, double cosq, double sinq, double phi0, int limit, double* phi)
{
    int nphi = 0;
    const double xx = g[1] * sinq + g[2] * cosq;
    double aa = (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2];
    aa = -((aa * HBAR) / ((2 * ME) * v));
    double uu = g[0] * g[0] + xx * xx;
    double vv = uu - aa * aa;
    if (vv < 0) return 0;
    uu = 1 / uu;
    vv = (std::sqrt(vv) * g[0]) * uu;
    uu = (aa * xx) * uu;
    const double lim = limit == 0 ? THETA_MAX * DEGTORAD : 2 * THETA_MAX * DEGTORAD;
    for (int branch = 0; branch < 2; branch++) {
        const double c = branch == 0 ? uu + vv : uu - vv;
        if (std::fabs(c) <= 1) {
            double pp = std::acos(c);
            const int jj = nint_away((pp - phi0) / (2 * PI));
            pp = pp - 2 * jj * PI;
            if (limit == 2 || std::fabs(pp - phi0) <= lim) {
                phi[nphi] = pp;
                phi[nphi + 1] = -pp;
                nphi += 2;
            }
        }
    }
    return nphi;
}

// findAllEulerAngles (Lib_DiffractionConditions.f90:1565-1623): incident
// directions khat = (-sin phi, sin theta cos phi, cos theta cos phi) at
// the Bragg condition for theta on a grid of NTHETA steps.
static void all_euler_angles(double v, const double* g, double theta0, double phi0, double theta_max, int limit,
                             std::vector<double>& khat)
{
    khat.clear();
    const double theta_min = std::max(0.0, theta0 - theta_max / 2);
    for (int ii = 1; ii <= NTHETA; ii++) {
        const double theta = theta_min + theta_max * ii / NTHETA;
        const double cost = std::cos(theta), sint = std::sin(theta);
        double phi[4];
        const int nphi = bragg_azimuths(v, g, cost, sint, phi0, limit, phi);
        for (int jj = 0; jj < nphi; jj++) {
            const double cosp = std::cos(phi[jj]), sinp = std::sin(phi[jj]);
            khat.push_back(-sinp);
            khat.push_back(sint * cosp);
            khat.push_back(cost * cosp);
        }
    }
}

// findBestRotationMatrix (Lib_DiffractionConditions.f90:129-357) without the
// near-Bragg list: U from rotateToStandardFrame, R the tilt whose incident
// direction minimises the summed two-beam intensity of every other
// permitted reflection (scoreForIncidentVector1/2, :1835-1889).
static int best_rotation(double v, const double* z0, const double* A_cell, int latt, const int* hkl,
                         const char* xifile, double foil, double* U, double* R)
{
    double BB[9], gg[3], ghkl0[3];
    recip(A_cell, BB);
    const double hd[3] = {(double)hkl[0], (double)hkl[1], (double)hkl[2]};
    reflection(BB, hd, gg);
    rotate_to_standard_frame(z0, A_cell, hd, U);
    for (int i = 1; i <= 3; i++) ghkl0[i - 1] = (m3(U,i,1) * gg[0] + m3(U,i,2) * gg[1]) + m3(U,i,3) * gg[2];
    std::vector<double> pg;
    std::vector<int> ph;
    permitted_gvectors(latt, BB, hkl, RHO_MAX, pg, ph);
    const int nn = (int)ph.size() / 3;
    std::vector<double> xi;
    bool all_found = false;
    HCHK(read_xi_table(xifile, ph, latt == LATT_HCP ? SYM_HEX : SYM_CUBIC, xi, &all_found));
    std::vector<double> khat;
    all_euler_angles(v, ghkl0, 0.0, 0.0, THETA_MAX * DEGTORAD, 0, khat);
    if (khat.empty()) all_euler_angles(v, ghkl0, 0.0, 0.0, 2 * THETA_MAX * DEGTORAD, 0, khat);
    if (khat.empty()) all_euler_angles(v, ghkl0, 0.0, 0.0, THETA_MAX * DEGTORAD, 2, khat);
    const int nk = (int)khat.size() / 3;
    if (nk == 0) return seterr(D2BI_ERR_ARG, "no tilt brings [%d %d %d] to the Bragg condition", hkl[0], hkl[1], hkl[2]);
    // scores in parallel over incident directions; each is the reference's
    // serial sum over reflections 3..nn
    std::vector<double> score(nk);
    const int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    auto work = [&](int t) {
        for (int ii = t; ii < nk; ii += nth) {
            const double* k = &khat[3 * ii];
            double s = 0;
            for (int j = 3; j <= nn; j++) {
                const double eps = energy_deviation(v, k, &pg[3 * (j - 1)]);
                s = s + intensity_at_depth(xi[0], xi[j - 1], e2l(v, eps), foil);
            }
            score[ii] = s;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; t++) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    double best = HUGE_R4;
    int ib = -1;
    for (int ii = 0; ii < nk; ii++)
        if (score[ii] < best) { best = score[ii]; ib = ii; }
    if (ib < 0) return seterr(D2BI_ERR_ARG, "tilt search found no finite score");
    const double* k = &khat[3 * ib];
    const double phi = -std::asin(k[0]);
    const double theta = std::acos(k[2] / std::cos(phi));
    const double ct = std::cos(theta), st = std::sin(theta), cp = std::cos(phi), sp = std::sin(phi);
    m3(R,1,1) = cp;   m3(R,1,2) = st * sp; m3(R,1,3) = ct * sp;
    m3(R,2,1) = 0.0;  m3(R,2,2) = ct;      m3(R,2,3) = -st;
    m3(R,3,1) = -sp;  m3(R,3,2) = st * cp; m3(R,3,3) = ct * cp;
    if (m3(R,3,3) < 0)
        for (int i = 1; i <= 3; i++) { m3(R,i,3) = -m3(R,i,3); m3(R,i,2) = -m3(R,i,2); }
    return D2BI_OK;
}

// GoldenSection (Lib_GoldenSection.f90): bracket, then golden-section
// minimisation with the reference's tolerances and step limits.
namespace {
struct Golden {
    double x1, x2, x3, y1, y2, y3, xtol, ytol;
    bool bounded = false;
    int step = 0;
    static constexpr double TOL = 1.0e-6, MINSTEP = 50, BOUNDSTEP = 10;
    double gold() const { return (1 + std::sqrt(5.0)) / 2; }
    Golden(double a, double b, double ya, double yb)
    {
        if (a < b) { x1 = a; x3 = b; y1 = ya; y3 = yb; }
        else       { x3 = b; x1 = a; y3 = yb; y1 = ya; }
        x2 = 0.0;
        y2 = HUGE_R4;
        xtol = (x3 - x1) * TOL;
        ytol = std::max(TOL, std::fabs(y3 - y1)) * TOL;
    }
    double next() const
    {
        const double G = gold();
        if (!bounded) {
            if (step == 0) return x1 + (x3 - x1) * (G - 1);
            if (y1 < y3) return x1 - (x3 - x1) / G;
            return x3 + (x3 - x1) / G;
        }
        return x1 + (x3 - x2);
    }
    void check_bounds(double xt, double yt)
    {
        if (step == 1) {
            if (yt < std::min(y1, y3)) { bounded = true; x2 = xt; y2 = yt; }
            else if ((yt > y1) && (y1 > y3)) { bounded = false; x1 = xt; y1 = yt; }
            else if ((yt > y3) && (y3 > y1)) { bounded = false; x3 = xt; y3 = yt; }
            else if ((y1 > yt) && (yt > y3)) { bounded = false; x1 = xt; y1 = yt; }
            else if ((y3 > yt) && (yt > y1)) { bounded = false; x3 = xt; y3 = yt; }
            return;
        }
        if (y1 < y3) {
            bounded = !(yt < y1);
            x2 = x1; x1 = xt; y2 = y1; y1 = yt;
        } else {
            bounded = !(yt < y3);
            x2 = x3; x3 = xt; y2 = y3; y3 = yt;
        }
    }
    // returns converged; *within = false when the step limit is reached
    bool minimise(double& xt, double& yt, bool* within)
    {
        step++;
        if (!bounded) {
            check_bounds(xt, yt);
            *within = bounded || (step <= BOUNDSTEP);
            return false;
        }
        *within = step <= MINSTEP;
        if (!*within) return false;
        if (yt < y2) {
            if (xt < x2) { x3 = x2; y3 = y2; x2 = xt; y2 = yt; }
            else         { x1 = x2; y1 = y2; x2 = xt; y2 = yt; }
        } else {
            if (xt > x2) { x3 = xt; y3 = yt; }
            else         { x1 = xt; y1 = yt; }
        }
        const bool is = (std::fabs(x3 - x1) < xtol) || (std::fabs(y2 - std::max(y1, y3)) < ytol);
        if (is) { xt = x2; yt = y2; }
        return is;
    }
};
}  // namespace

// tiltToDarkField (Lib_DiffractionConditions.f90:751-927): tilt about y by
// the angle in +-THETA_MAX/100 degrees that minimises the perfect-crystal
// |phi_g(L)|^2, if that is below its untilted value; R = Ry R.
static void tilt_to_dark_field(double v, const double* A_cell, const double* hkl, double xi0, double xig,
                               double foil, const double* U, double* R)
{
    (void)xi0;
    double BB[9], kk[3], RU[9], gg[3];
    recip(A_cell, BB);
    reflection(BB, hkl, kk);
    matmul3(R, U, RU);
    for (int i = 1; i <= 3; i++) gg[i - 1] = (m3(RU,i,1) * kk[0] + m3(RU,i,2) * kk[1]) + m3(RU,i,3) * kk[2];
    const double bg = 1 / xig;
    static const double khat[3] = {0.0, 0.0, 1.0};
    const double sg = e2l(v, energy_deviation(v, khat, gg));
    auto phiL2 = [&](double s, double L) {
        const double t = bg * std::sin((PI * std::sqrt(s * s + bg * bg)) * L);
        return (t * t) / (s * s + bg * bg);
    };
    const double y0 = phiL2(sg, foil);
    auto tilted = [&](double x, double& cost, double& sint) {
        cost = std::cos(x); sint = std::sin(x);
        const double Lp = foil / cost;
        const double sgp = (sint * gg[0] - cost * gg[2]) / (2 * PI);
        return phiL2(sgp, Lp);
    };
    double cost, sint;
    const double x1 = -((THETA_MAX * 0.01) * DEGTORAD), x3 = (THETA_MAX * 0.01) * DEGTORAD;
    const double y1 = tilted(x1, cost, sint), y3 = tilted(x3, cost, sint);
    Golden gold(x1, x3, y1, y3);
    double x2, y2;
    for (;;) {
        x2 = gold.next();
        y2 = tilted(x2, cost, sint);
        bool within;
        const bool conv = gold.minimise(x2, y2, &within);
        if (conv || !within) break;
    }
    if (y2 < y0) { cost = std::cos(x2); sint = std::sin(x2); }
    else         { cost = 1.0; sint = 0.0; }
    double RR[9] = {0};
    m3(RR,1,1) = cost;  m3(RR,1,3) = sint;
    m3(RR,2,2) = 1.0;
    m3(RR,3,1) = -sint; m3(RR,3,3) = cost;
    matmul3(RR, R, R);
}

// Quaternions (Lib_Quaternions.f90): q = (w, x, y, z).
namespace {
struct Quat { double w, x, y, z; };
}

// Eigen-decomposition of a symmetric 4x4 matrix by cyclic Jacobi rotations;
// returns the unit eigenvector of the largest eigenvalue.  This stands in
// for the LAPACK DSYEV call of RotMatToQuaternion.
static void sym4_top_eigenvector(const double K[4][4], double v[4])
{
    double A[4][4], V[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) { A[i][j] = K[i][j]; V[i][j] = i == j ? 1.0 : 0.0; }
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < 4; p++)
            for (int q = p + 1; q < 4; q++) off += A[p][q] * A[p][q];
        if (off < 1.0e-300) break;
        for (int p = 0; p < 4; p++)
            for (int q = p + 1; q < 4; q++) {
                if (A[p][q] == 0.0) continue;
                const double th = (A[q][q] - A[p][p]) / (2 * A[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 4; k++) {               // A <- A J
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; k++) {               // A <- J^T A
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; k++) {               // V <- V J
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int top = 0;
    for (int i = 1; i < 4; i++)
        if (A[i][i] > A[top][top]) top = i;
    double n = 0.0;
    for (int k = 0; k < 4; k++) n += V[k][top] * V[k][top];
    n = 1 / std::sqrt(n);
    for (int k = 0; k < 4; k++) v[k] = V[k][top] * n;
}

// RotMatToQuaternion (Lib_Quaternions.f90:246-288): the eigenvector of the
// largest eigenvalue of the symmetric matrix built from R, with the sign
// fixed so its last component is not positive.
static Quat quat_from_rot(const double* R)
{
    double K[4][4] = {{0}};
    K[0][0] = m3(R,1,1) - m3(R,2,2) - m3(R,3,3);
    K[1][0] = m3(R,2,1) + m3(R,1,2);
    K[2][0] = m3(R,3,1) + m3(R,1,3);
    K[3][0] = m3(R,2,3) - m3(R,3,2);
    K[1][1] = m3(R,2,2) - m3(R,1,1) - m3(R,3,3);
    K[2][1] = m3(R,3,2) + m3(R,2,3);
    K[3][1] = m3(R,3,1) - m3(R,1,3);
    K[2][2] = m3(R,3,3) - m3(R,1,1) - m3(R,2,2);
    K[3][2] = m3(R,1,2) - m3(R,2,1);
    K[3][3] = m3(R,1,1) + m3(R,2,2) + m3(R,3,3);
    for (int i = 0; i < 4; i++)                          // lower triangle is what is given
        for (int j = i + 1; j < 4; j++) K[i][j] = K[j][i];
    double v[4];
    sym4_top_eigenvector(K, v);
    if (v[3] > 0) for (int k = 0; k < 4; k++) v[k] = -v[k];
    return Quat{-v[3], v[0], v[1], v[2]};
}

// quaternionToRotMat (Lib_Quaternions.f90:290-316)
static void rot_from_quat(const Quat& q, double* R)
{
    double ss = (((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w) / 2;
    m3(R,1,1) = ss - (q.y * q.y + q.z * q.z);
    m3(R,2,1) = (q.x * q.y + q.z * q.w);
    m3(R,3,1) = (q.x * q.z - q.y * q.w);
    m3(R,1,2) = (q.x * q.y - q.z * q.w);
    m3(R,2,2) = ss - (q.x * q.x + q.z * q.z);
    m3(R,3,2) = (q.y * q.z + q.x * q.w);
    m3(R,1,3) = (q.x * q.z + q.y * q.w);
    m3(R,2,3) = (q.y * q.z - q.x * q.w);
    m3(R,3,3) = ss - (q.x * q.x + q.y * q.y);
    ss = 1 / ss;
    for (int i = 0; i < 9; i++) R[i] = R[i] * ss;
}

// slerp (Lib_Quaternions.f90:526-564), with midpoint (:504-524) and
// normalise (:203-219)
static Quat slerp(const Quat& a, const Quat& b, double x)
{
    double x1 = ((a.w * b.w + a.x * b.x) + a.y * b.y) + a.z * b.z;
    Quat r;
    if (x1 >= 1.0) return a;
    if (x1 <= -1.0) {
        r = Quat{a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z};
        double dd = ((r.w * r.w + r.x * r.x) + r.y * r.y) + r.z * r.z;
        if (dd > 0) {
            dd = 1 / std::sqrt(dd);
            r.w *= dd; r.x *= dd; r.y *= dd; r.z *= dd;
        }
        return r;
    }
    double x2 = 2 * x1 * x1 - 1;
    const double omega = std::acos(x2);
    x2 = 1 / std::sin(omega);
    x1 = std::sin((1 - x) * omega) * x2;
    x2 = std::sin(x * omega) * x2;
    r.w = x1 * a.w + x2 * b.w;
    r.x = x1 * a.x + x2 * b.x;
    r.y = x1 * a.y + x2 * b.y;
    r.z = x1 * a.z + x2 * b.z;
    double tt = ((r.w * r.w + r.x * r.x) + r.y * r.y) + r.z * r.z;
    if (tt == 0) return Quat{1, 0, 0, 0};
    tt = 1 / std::sqrt(tt);
    r.w *= tt; r.x *= tt; r.y *= tt; r.z *= tt;
    return r;
}

// setSupercellTilt1 without atom positions (Lib_DynamicalTwoBeamImaging.f90:795-939)
// up to the tilt: U, R for the (ng g) condition; d[2] becomes the depth
// after the tilt.
static int tilt_for_ng(d2bi_record* r, const double* k, int ng, int latt, const char* xifile, double* d)
{
    const int hkl[3] = {nint_away(ng * r->g[0]), nint_away(ng * r->g[1]), nint_away(ng * r->g[2])};
    const double foil = d[2];
    const double vv = velocity(r->E * 1000);
    HCHK(best_rotation(vv, k, r->a_cell, latt, hkl, xifile, std::fabs(foil), r->U, r->R));
    double z = std::fabs(foil / m3(r->R,3,3));
    const double g0[3] = {(double)nint_away(r->g[0]), (double)nint_away(r->g[1]), (double)nint_away(r->g[2])};
    tilt_to_dark_field(velocity(r->E * 1000), r->a_cell, g0, r->xi0, r->xig, z, r->U, r->R);
    d[2] = std::fabs(d[2] / m3(r->R,3,3));
    return D2BI_OK;
}

// setSupercellTilt1a without atom positions (Lib_DynamicalTwoBeamImaging.f90:943-1045)
int setup_tilt_ng(d2bi_record* r, const double* k, double ng, int latt, const char* xifile, int mxy, int mzz,
                  double* d)
{
    if (!(ng >= 1.0) || ng > 1.0e6) return seterr(D2BI_ERR_ARG, "ng must be >= 1");
    if (latt < LATT_SC || latt > LATT_HCP) return seterr(D2BI_ERR_ARG, "unknown lattice code %d", latt);
    if (!xifile) return seterr(D2BI_ERR_ARG, "the (g,ng) tilt needs an extinction distance file");
    const double dd[3] = {d[0], d[1], d[2]};
    const int nf = (int)std::floor(ng);
    double dw[3] = {dd[0], dd[1], dd[2]};
    HCHK(tilt_for_ng(r, k, nf, latt, xifile, dw));
    const Quat q1 = quat_from_rot(r->R);
    Quat q2 = q1;
    if (ng - nf > 1.0e-4) {
        double dw2[3] = {dd[0], dd[1], dd[2]};
        HCHK(tilt_for_ng(r, k, (int)std::ceil(ng), latt, xifile, dw2));
        q2 = quat_from_rot(r->R);
    }
    rot_from_quat(slerp(q1, q2, ng - nf), r->R);
    d[0] = dd[0];
    d[1] = dd[1];
    d[2] = std::fabs(dd[2] / m3(r->R,3,3));
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

int d2bi_setup_tilt_ng(d2bi_record* rec, const double k[3], double ng, int32_t lattice, const char* xifile,
                       int32_t mxy, int32_t mzz, double d[3])
{
    if (!rec || !k || !d || !xifile || mxy < 1 || mzz < 1) return d2bi::seterr(D2BI_ERR_ARG, "bad argument");
    if (!(d[0] > 0 && d[1] > 0 && d[2] > 0)) return d2bi::seterr(D2BI_ERR_ARG, "voxel cuboid sides must be > 0");
    return d2bi::setup_tilt_ng(rec, k, ng, lattice, xifile, mxy, mzz, d);
}

int d2bi_rotate_to_standard_frame(const double k[3], const double a_cell[9], const double hkl[3], double U[9])
{
    if (!k || !a_cell || !hkl || !U) return d2bi::seterr(D2BI_ERR_ARG, "null argument");
    d2bi::rotate_to_standard_frame(k, a_cell, hkl, U);
    return D2BI_OK;
}

int d2bi_aperture_points(int32_t n, double* xy)
{
    if (!xy) return d2bi::seterr(D2BI_ERR_ARG, "null argument");
    if (n > D2BI_MAX_APERTURE) return d2bi::seterr(D2BI_ERR_ARG, "nAperture out of range");
    return d2bi::aperture_points(n, xy);
}

int d2bi_aperture_setup(const d2bi_record* rec, double precAngle, int32_t nPrec, double apertureAngle,
                        int32_t nAperture, double* gshift, double* aa)
{
    if (!rec || !gshift || !aa) return d2bi::seterr(D2BI_ERR_ARG, "null argument");
    return d2bi::aperture_setup(rec, precAngle, nPrec, apertureAngle, nAperture, gshift, aa);
}

double d2bi_wavelength(double V) { return d2bi::wavelength(V); }
double d2bi_velocity(double V) { return d2bi::velocity(V); }
double d2bi_deviation_parameter(double E_keV, const double g_lab[3]) { return d2bi::deviation_parameter(E_keV, g_lab); }

int d2bi_synth_bcc(const d2bi_synth_params* p, double* x, int64_t cap, int64_t* nAtoms, double box[3])
{
    return d2bi::synth_bcc(p, x, cap, nAtoms, box);
}

}  // extern "C"
