"""The shared RK4 step matrix of the FAST integrator (k_integrate_poly).

The Howie-Whelan system of Lib_DynamicalTwoBeamImaging.f90:2316-2480 is
phi' = i H(a) phi with H(a) = [[c0, cg], [conj(cg), c0 + a]] per field
sample; one RK4 double step (h = 2 dz, samples z, z + dz, z + 2 dz) is a 2x2
complex matrix M whose entries are polynomials in alpha = 2 dz a.  The kernel
builds M from the stage products

    T1 = I + dz A_a,  K2 = dz A_b T1,  T2 = I + K2,  K3 = 2 dz A_b T2,
    T3 = I + K3,  K4/6 = (dz/3) A_c T3,
    M = I + (T1 - I)/3 + (2/3) K2 + K3/3 + K4/6,   A_x = i H_x(a),

column by column with the (2,2) element's a as a shift of coefficients.
This test performs that construction in float64 and checks it against the
stage-by-stage march (the reference's order of RK4 stages) on random field
samples and deviations, and checks the entry degrees (2, 3, 3, 4) and the
one fixed coefficient, alpha^4 / 24 in the (2,2) entry.
"""
import numpy as np


def stage_march(y, sa, sb, sc, a, dz):
    def A(s):
        return 1j * np.array([[s[0], s[1]], [np.conj(s[1]), s[0] + a]])
    k1 = A(sa) @ y
    k2 = A(sb) @ (y + dz * k1)
    k3 = A(sb) @ (y + dz * k2)
    k4 = A(sc) @ (y + 2 * dz * k3)
    return y + (2 * dz / 6) * (k1 + 2 * k2 + 2 * k3 + k4)


def stage_column(c, g, ax, P1, P2):
    """Q1 = i (c P1 + g P2), Q2 = i (conj(g) P1 + c P2 + ax alpha P2):
    one column of s A_x P with c = s c0, g = s cg, ax = s / (2 dz);
    polynomials as coefficient arrays in alpha."""
    Q1 = np.zeros(max(len(P1), len(P2)), complex)
    Q2 = np.zeros(max(len(P1), len(P2) + 1), complex)
    Q1[:len(P1)] += c * P1
    Q1[:len(P2)] += g * P2
    Q2[:len(P1)] += np.conj(g) * P1
    Q2[:len(P2)] += c * P2
    Q2[1:len(P2) + 1] += ax * P2
    return 1j * Q1, 1j * Q2


def step_matrix(sa, sb, sc, dz):
    ca, ga = dz * sa[0], dz * sa[1]
    T11, T12 = np.array([1 + 1j * ca]), np.array([1j * ga])
    T21, T22 = np.array([1j * np.conj(ga)]), np.array([1 + 1j * ca, 0.5j])
    K2_11, K2_21 = stage_column(dz * sb[0], dz * sb[1], 0.5, T11, T21)
    K2_12, K2_22 = stage_column(dz * sb[0], dz * sb[1], 0.5, T12, T22)
    U11, U22 = K2_11.copy(), K2_22.copy()
    U11[0] += 1
    U22[0] += 1
    K3_11, K3_21 = stage_column(2 * dz * sb[0], 2 * dz * sb[1], 1.0, U11, K2_21)
    K3_12, K3_22 = stage_column(2 * dz * sb[0], 2 * dz * sb[1], 1.0, K2_12, U22)
    V11, V22 = K3_11.copy(), K3_22.copy()
    V11[0] += 1
    V22[0] += 1
    M11, M21 = stage_column(dz / 3 * sc[0], dz / 3 * sc[1], 1 / 6, V11, K3_21)
    M12, M22 = stage_column(dz / 3 * sc[0], dz / 3 * sc[1], 1 / 6, K3_12, V22)
    for M, K3, K2 in ((M11, K3_11, K2_11), (M21, K3_21, K2_21), (M12, K3_12, K2_12), (M22, K3_22, K2_22)):
        M[:len(K3)] += K3 / 3
        M[:len(K2)] += 2 * K2 / 3
    M11[0] += 1 + 1j * ca / 3
    M12[0] += 1j * ga / 3
    M21[0] += 1j * np.conj(ga) / 3
    M22[0] += 1 + 1j * ca / 3
    M22[1] += 1j / 6
    return M11, M12, M21, M22


def test_step_matrix_is_the_stage_march():
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(300):
        dz = rng.uniform(0.3, 1.2)
        s = [(rng.normal() * 0.05, complex(*rng.normal(size=2)) * 0.03) for _ in range(3)]
        a = rng.normal() * 0.2
        y = rng.normal(size=2) + 1j * rng.normal(size=2)
        M = step_matrix(*s, dz)
        assert [len(m) - 1 for m in M] == [2, 3, 3, 4]
        assert np.isclose(M[3][4], 1 / 24, rtol=0, atol=1e-15)
        ev = [np.polyval(m[::-1], 2 * dz * a) for m in M]
        got = np.array([ev[0] * y[0] + ev[1] * y[1], ev[2] * y[0] + ev[3] * y[1]])
        worst = max(worst, float(np.max(np.abs(got - stage_march(y, *s, a, dz)))))
    assert worst < 1e-14, worst


def test_step_matrix_limits():
    """Zero field: M is the scalar RK4 polynomial of i a h on the g beam and
    the identity on the direct beam; unitary H keeps |phi| to RK4's O(h^5)."""
    dz = 0.8
    zero = (0.0, 0.0j)
    M = step_matrix(zero, zero, zero, dz)
    al = 0.3
    z = 1j * al
    assert np.allclose([np.polyval(M[0][::-1], al)], [1.0], atol=1e-16)
    assert np.allclose([np.polyval(M[1][::-1], al), np.polyval(M[2][::-1], al)], [0, 0], atol=1e-16)
    assert np.isclose(np.polyval(M[3][::-1], al), 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24, atol=1e-16)
    s = (0.03, 0.01 + 0.02j)
    M = step_matrix(s, s, s, dz)
    ev = np.array([[np.polyval(m[::-1], al) for m in M[:2]], [np.polyval(m[::-1], al) for m in M[2:]]])
    sv = np.linalg.svd(ev, compute_uv=False)
    assert np.all(np.abs(sv - 1) < 1e-4)
