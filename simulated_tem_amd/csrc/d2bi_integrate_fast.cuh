// d2bi_integrate_fast.cuh -- FAST column integration: the stage-by-stage RK4
// march and the shared step matrix for many precession points.
//
// One part of the single translation unit d2bi_gpu.hip, which includes the
// parts in order so that kernels and the device helpers they call compile
// together; each part uses the definitions of the parts before it.
#pragma once

namespace d2bi {

// ------------------------------------------------------------------ (c) integrate, FAST

struct FastIntParams {
    int Mx, My, Mz, np;         // np: precession points in this pass
    float b0, bg, dz;
    float aa[32];               // 2 pi s_g per precession point
    float al[32];               // 2 dz aa (k_integrate_poly)
};

// One lane per pixel column, NP precession points marched together from
// one read of the column (the field is shared by all precession points:
// precession changes s_g only).  FP32 classic RK4 with h = 2 dz; the
// per-step renormalisation of the reference is a positive scalar on a
// linear map, so it is applied once at the end: I = |phi_g|^2/|phi|^2.
// The mean-potential term c0 = b0 rho stays in, although in the exact ODE it
// is a global phase: RK4's truncation error depends on the whole matrix, and
// over 800 double steps leaving c0 out moves |phi_g|^2 by ~2e-4.
template <int NP, bool DENS>
__global__ __launch_bounds__(256) void k_integrate_fast(FastIntParams P, const void* __restrict__ field,
                                                        double* __restrict__ img, int accumulate,
                                                        double inv_total)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)P.Mx * P.My;
    if (col >= plane) return;
    const int nz = P.Mz;
    const float dz = P.dz, dz3 = P.dz / 3.0f, dz23 = 2.0f * P.dz / 3.0f, dz2 = 2.0f * P.dz;
    float p0r[NP], p0i[NP], pgr[NP], pgi[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) { p0r[p] = 1.0f; p0i[p] = 0.0f; pgr[p] = 0.0f; pgi[p] = 0.0f; }
    // field: float4 {x, rho} with the density field, else float2 x (rho = 1);
    // sample Mz is sample 0.  Samples are loaded one double-step ahead.
    auto load = [&](int k) -> float4 {
        const int64_t o = col + plane * (k == nz ? 0 : k);
        // each sample is read once: non-temporal loads keep the stream from
        // displacing cached data
        typedef float ntf2 __attribute__((ext_vector_type(2)));
        typedef float ntf4 __attribute__((ext_vector_type(4)));
        if (DENS) {
            const ntf4 v = __builtin_nontemporal_load(static_cast<const ntf4*>(field) + o);
            return make_float4(v.x, v.y, v.z, v.w);
        }
        const ntf2 f = __builtin_nontemporal_load(static_cast<const ntf2*>(field) + o);
        return make_float4(f.x, f.y, 1.0f, 0.0f);
    };
    auto coef = [&](const float4& f, float& c0, float& cgr, float& cgi) {
        if (DENS) {
            c0 = P.b0 * f.z;
            const float bgr = P.bg * f.z;
            cgr = bgr * f.x;
            cgi = bgr * f.y;
        } else {
            c0 = P.b0;
            cgr = P.bg * f.x;
            cgi = P.bg * f.y;
        }
    };
    // samples of the next PD double steps are in flight: ring[2u] and
    // ring[2u + 1] hold samples ii + 1, ii + 2 of the double step at ii in
    // slot u (ii = i0 + 2u), reloaded with those of ii + 2 PD once used.  A
    // column is a chain of dependent loads, so the depth, not the bandwidth,
    // bounds a wave's time when few points are marched per sample.
    constexpr int PD = NP <= 4 ? 4 : 1;
    float c0a = 0, cgra = 0, cgia = 0;
    float4 ring[2 * PD];
#pragma unroll
    for (int k = 0; k < 2 * PD; k++) ring[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (nz >= 2) {
        coef(load(0), c0a, cgra, cgia);
#pragma unroll
        for (int k = 0; k < 2 * PD; k++)
            if (k + 1 <= nz) ring[k] = load(k + 1);
    }
    for (int i0 = 0; i0 <= nz - 2; i0 += 2 * PD) {
#pragma unroll
    for (int u = 0; u < PD; u++) {
        const int ii = i0 + 2 * u;
        if (ii > nz - 2) break;
        float c0b, cgrb, cgib, c0c, cgrc, cgic;
        coef(ring[2 * u], c0b, cgrb, cgib);
        coef(ring[2 * u + 1], c0c, cgrc, cgic);
        if (ii + 2 * PD + 2 <= nz) {
            ring[2 * u] = load(ii + 2 * PD + 1);
            ring[2 * u + 1] = load(ii + 2 * PD + 2);
        }
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const float a = P.aa[p];
            float d0r, d0i, dgr, dgi;
#define FDERIV(c0, cgr, cgi)                                                   \
    {                                                                          \
        const float ca_ = c0 + a;                                              \
        d0r = fmaf(-c0, p0i[p], fmaf(-cgi, pgr[p], -cgr * pgi[p]));           \
        d0i = fmaf(c0, p0r[p], fmaf(cgr, pgr[p], -cgi * pgi[p]));             \
        dgr = fmaf(cgi, p0r[p], fmaf(-cgr, p0i[p], -ca_ * pgi[p]));           \
        dgi = fmaf(cgr, p0r[p], fmaf(cgi, p0i[p], ca_ * pgr[p]));             \
    }
            const float s0r = p0r[p], s0i = p0i[p], sgr = pgr[p], sgi = pgi[p];
            FDERIV(c0a, cgra, cgia)
            float a0r = fmaf(d0r, dz3, s0r), a0i = fmaf(d0i, dz3, s0i);
            float agr = fmaf(dgr, dz3, sgr), agi = fmaf(dgi, dz3, sgi);
            p0r[p] = fmaf(d0r, dz, s0r); p0i[p] = fmaf(d0i, dz, s0i);
            pgr[p] = fmaf(dgr, dz, sgr); pgi[p] = fmaf(dgi, dz, sgi);
            FDERIV(c0b, cgrb, cgib)
            p0r[p] = fmaf(d0r, dz, s0r); p0i[p] = fmaf(d0i, dz, s0i);
            pgr[p] = fmaf(dgr, dz, sgr); pgi[p] = fmaf(dgi, dz, sgi);
            a0r = fmaf(d0r, dz23, a0r); a0i = fmaf(d0i, dz23, a0i);
            agr = fmaf(dgr, dz23, agr); agi = fmaf(dgi, dz23, agi);
            FDERIV(c0b, cgrb, cgib)
            p0r[p] = fmaf(d0r, dz2, s0r); p0i[p] = fmaf(d0i, dz2, s0i);
            pgr[p] = fmaf(dgr, dz2, sgr); pgi[p] = fmaf(dgi, dz2, sgi);
            a0r = fmaf(d0r, dz23, a0r); a0i = fmaf(d0i, dz23, a0i);
            agr = fmaf(dgr, dz23, agr); agi = fmaf(dgi, dz23, agi);
            FDERIV(c0c, cgrc, cgic)
#undef FDERIV
            p0r[p] = fmaf(d0r, dz3, a0r); p0i[p] = fmaf(d0i, dz3, a0i);
            pgr[p] = fmaf(dgr, dz3, agr); pgi[p] = fmaf(dgi, dz3, agi);
        }
        c0a = c0c; cgra = cgrc; cgia = cgic;
    }
    }
    double acc = 0.0;
#pragma unroll
    for (int p = 0; p < NP; p++) {
        if (p >= P.np) break;
        const float ng = fmaf(pgr[p], pgr[p], pgi[p] * pgi[p]);
        const float n0 = fmaf(p0r[p], p0r[p], p0i[p] * p0i[p]);
        acc += (double)ng / (double)(n0 + ng);
    }
    acc *= inv_total;
    img[col] = accumulate ? img[col] + acc : acc;
}

// Many precession points.  The system is phi' = i H(a) phi with
// H(a) = [[c0, cg], [conj(cg), c0 + a]] (the derivative above), so one RK4
// double step is phi <- M phi with
//     M = I + Q1/3 + (2/3) K2 + K3/3 + K4/6,
//     Q1 = dz A_a,  T1 = I + Q1,  K2 = dz A_b T1,  T2 = I + K2,
//     K3 = 2 dz A_b T2,  T3 = I + K3,  K4 = 2 dz A_c T3,   A_x = i H_x(a),
// the same map as the stage-by-stage march (A_a, A_b, A_c at the double
// step's three samples).  a enters only through the (2,2) element, once per
// stage, so the entries of M are polynomials in alpha = 2 dz a of degrees 2,
// 3, 3 and 4 whose coefficients depend on the column's field samples only.
// They are formed once per double step and shared: a precession point then
// costs a Horner evaluation at its alpha and one matrix-vector product
// (~40 FMAs) instead of four derivative evaluations and the stage updates
// (~80).  Results differ from k_integrate_fast by rounding only.
struct cf {
    float r, i;
};
template <int D>
struct CPoly {
    cf c[D + 1];    // coefficients of alpha^0 .. alpha^D
};

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B <= E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// One column (P1, P2) of Q = s A_x P for a polynomial matrix P, with the
// stage's scaled coefficients c = s c0, g = s cg and ax = s / (2 dz):
//     Q1 = i (c P1 + g P2),   Q2 = i (conj(g) P1 + c P2 + ax alpha P2).
// Terms are chained from the first present one, so structural zeros cost
// nothing.
template <int D1, int D2, int E1 = (D1 > D2 ? D1 : D2), int E2 = (D1 > D2 + 1 ? D1 : D2 + 1)>
__device__ __forceinline__ void rk_col(float c, float gr, float gi, float ax, const CPoly<D1>& P1, const CPoly<D2>& P2,
                                       CPoly<E1>& Q1, CPoly<E2>& Q2)
{
    static_for<0, E1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        float ur, ui;
        if constexpr (k <= D2) {
            ur = fmaf(gr, P2.c[k].r, -gi * P2.c[k].i);
            ui = fmaf(gr, P2.c[k].i, gi * P2.c[k].r);
            if constexpr (k <= D1) {
                ur = fmaf(c, P1.c[k].r, ur);
                ui = fmaf(c, P1.c[k].i, ui);
            }
        } else {
            ur = c * P1.c[k].r;
            ui = c * P1.c[k].i;
        }
        Q1.c[k] = cf{-ui, ur};
    });
    static_for<0, E2>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr bool ta = k <= D1, tb = k <= D2, tc = k >= 1 && k - 1 <= D2;
        float ur = 0.0f, ui = 0.0f;
        if constexpr (ta) {
            ur = fmaf(gr, P1.c[k].r, gi * P1.c[k].i);
            ui = fmaf(gr, P1.c[k].i, -gi * P1.c[k].r);
        }
        if constexpr (tb) {
            if constexpr (ta) {
                ur = fmaf(c, P2.c[k].r, ur);
                ui = fmaf(c, P2.c[k].i, ui);
            } else {
                ur = c * P2.c[k].r;
                ui = c * P2.c[k].i;
            }
        }
        if constexpr (tc) {
            if constexpr (ta || tb) {
                ur = fmaf(ax, P2.c[k - 1].r, ur);
                ui = fmaf(ax, P2.c[k - 1].i, ui);
            } else {
                ur = ax * P2.c[k - 1].r;
                ui = ax * P2.c[k - 1].i;
            }
        }
        Q2.c[k] = cf{-ui, ur};
    });
}

// M += s P (degree of P <= degree of M)
template <int DM, int DP>
__device__ __forceinline__ void poly_axpy(CPoly<DM>& M, float s, const CPoly<DP>& P)
{
    static_assert(DP <= DM, "degree");
#pragma unroll
    for (int k = 0; k <= DP; k++) {
        M.c[k].r = fmaf(s, P.c[k].r, M.c[k].r);
        M.c[k].i = fmaf(s, P.c[k].i, M.c[k].i);
    }
}

template <int D>
__device__ __forceinline__ cf horner(const CPoly<D>& M, float al)
{
    cf v = M.c[D];
#pragma unroll
    for (int k = D - 1; k >= 0; k--) {
        v.r = fmaf(v.r, al, M.c[k].r);
        v.i = fmaf(v.i, al, M.c[k].i);
    }
    return v;
}

template <int NP, bool DENS>
__global__ __launch_bounds__(256) void k_integrate_poly(FastIntParams P, const void* __restrict__ field,
                                                        double* __restrict__ img, int accumulate,
                                                        double inv_total)
{
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)P.Mx * P.My;
    if (col >= plane) return;
    const int nz = P.Mz;
    const float dz = P.dz, dz2 = 2.0f * P.dz, dz6 = P.dz / 3.0f;
    float p0r[NP], p0i[NP], pgr[NP], pgi[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) { p0r[p] = 1.0f; p0i[p] = 0.0f; pgr[p] = 0.0f; pgi[p] = 0.0f; }
    auto load = [&](int k) -> float4 {
        const int64_t o = col + plane * (k == nz ? 0 : k);
        if (DENS) return static_cast<const float4*>(field)[o];
        const float2 f = static_cast<const float2*>(field)[o];
        return make_float4(f.x, f.y, 1.0f, 0.0f);
    };
    auto coef = [&](const float4& f, float& c0, float& cgr, float& cgi) {
        if (DENS) {
            c0 = P.b0 * f.z;
            const float bgr = P.bg * f.z;
            cgr = bgr * f.x;
            cgi = bgr * f.y;
        } else {
            c0 = P.b0;
            cgr = P.bg * f.x;
            cgi = P.bg * f.y;
        }
    };
    float c0a = 0, cgra = 0, cgia = 0;
    float4 fb = make_float4(0.0f, 0.0f, 0.0f, 0.0f), fc = fb;
    if (nz >= 2) {
        coef(load(0), c0a, cgra, cgia);
        fb = load(1);
        fc = load(2);
    }
    for (int ii = 0; ii <= nz - 2; ii += 2) {
        float c0b, cgrb, cgib, c0c, cgrc, cgic;
        coef(fb, c0b, cgrb, cgib);
        coef(fc, c0c, cgrc, cgic);
        if (ii + 4 <= nz) {
            fb = load(ii + 3);
            fc = load(ii + 4);
        }
        // T1 = I + dz A_a
        const float ca = dz * c0a, gar = dz * cgra, gai = dz * cgia;
        CPoly<0> T11{{cf{1.0f, ca}}}, T12{{cf{-gai, gar}}}, T21{{cf{gai, gar}}};
        CPoly<1> T22{{cf{1.0f, ca}, cf{0.0f, 0.5f}}};
        // K2 = dz A_b T1, T2 = I + K2
        CPoly<0> K2_11; CPoly<1> K2_21, K2_12; CPoly<2> K2_22;
        rk_col(dz * c0b, dz * cgrb, dz * cgib, 0.5f, T11, T21, K2_11, K2_21);
        rk_col(dz * c0b, dz * cgrb, dz * cgib, 0.5f, T12, T22, K2_12, K2_22);
        CPoly<0> U11 = K2_11; CPoly<1> U21 = K2_21, U12 = K2_12; CPoly<2> U22 = K2_22;
        U11.c[0].r += 1.0f;
        U22.c[0].r += 1.0f;
        // K3 = 2 dz A_b T2, T3 = I + K3
        CPoly<1> K3_11; CPoly<2> K3_21, K3_12; CPoly<3> K3_22;
        rk_col(dz2 * c0b, dz2 * cgrb, dz2 * cgib, 1.0f, U11, U21, K3_11, K3_21);
        rk_col(dz2 * c0b, dz2 * cgrb, dz2 * cgib, 1.0f, U12, U22, K3_12, K3_22);
        CPoly<1> V11 = K3_11; CPoly<2> V21 = K3_21, V12 = K3_12; CPoly<3> V22 = K3_22;
        V11.c[0].r += 1.0f;
        V22.c[0].r += 1.0f;
        // K4/6 = (dz/3) A_c T3: the last term of M, which it starts
        CPoly<2> M11; CPoly<3> M21, M12; CPoly<4> M22;
        rk_col(dz6 * c0c, dz6 * cgrc, dz6 * cgic, 1.0f / 6.0f, V11, V21, M11, M21);
        rk_col(dz6 * c0c, dz6 * cgrc, dz6 * cgic, 1.0f / 6.0f, V12, V22, M12, M22);
        // + K3/3 + (2/3) K2 + Q1/3 + I
        poly_axpy(M11, 1.0f / 3.0f, K3_11); poly_axpy(M21, 1.0f / 3.0f, K3_21);
        poly_axpy(M12, 1.0f / 3.0f, K3_12); poly_axpy(M22, 1.0f / 3.0f, K3_22);
        poly_axpy(M11, 2.0f / 3.0f, K2_11); poly_axpy(M21, 2.0f / 3.0f, K2_21);
        poly_axpy(M12, 2.0f / 3.0f, K2_12); poly_axpy(M22, 2.0f / 3.0f, K2_22);
        M11.c[0].r += 1.0f;
        M11.c[0].i = fmaf(ca, 1.0f / 3.0f, M11.c[0].i);
        M12.c[0].r = fmaf(-gai, 1.0f / 3.0f, M12.c[0].r);
        M12.c[0].i = fmaf(gar, 1.0f / 3.0f, M12.c[0].i);
        M21.c[0].r = fmaf(gai, 1.0f / 3.0f, M21.c[0].r);
        M21.c[0].i = fmaf(gar, 1.0f / 3.0f, M21.c[0].i);
        M22.c[0].r += 1.0f;
        M22.c[0].i = fmaf(ca, 1.0f / 3.0f, M22.c[0].i);
        M22.c[1].i += 1.0f / 6.0f;
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const float al = P.al[p];
            const cf m11 = horner(M11, al), m12 = horner(M12, al), m21 = horner(M21, al), m22 = horner(M22, al);
            const float y0r = p0r[p], y0i = p0i[p], ygr = pgr[p], ygi = pgi[p];
            p0r[p] = fmaf(m11.r, y0r, fmaf(-m11.i, y0i, fmaf(m12.r, ygr, -m12.i * ygi)));
            p0i[p] = fmaf(m11.r, y0i, fmaf(m11.i, y0r, fmaf(m12.r, ygi, m12.i * ygr)));
            pgr[p] = fmaf(m21.r, y0r, fmaf(-m21.i, y0i, fmaf(m22.r, ygr, -m22.i * ygi)));
            pgi[p] = fmaf(m21.r, y0i, fmaf(m21.i, y0r, fmaf(m22.r, ygi, m22.i * ygr)));
        }
        c0a = c0c; cgra = cgrc; cgia = cgic;
    }
    double acc = 0.0;
#pragma unroll
    for (int p = 0; p < NP; p++) {
        if (p >= P.np) break;
        const float ng = fmaf(pgr[p], pgr[p], pgi[p] * pgi[p]);
        const float n0 = fmaf(p0r[p], p0r[p], p0i[p] * p0i[p]);
        acc += (double)ng / (double)(n0 + ng);
    }
    acc *= inv_total;
    img[col] = accumulate ? img[col] + acc : acc;
}

}  // namespace d2bi
