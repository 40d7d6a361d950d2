"""Pins of the oracle's per-particle formulas against closed forms, numpy and invariants.

Each test checks a restated function of the reference against something that does not share its
code: the closed-form M4 spline and its normalisation, finite differences, numpy's eigh / inv /
det, brute-force bisection of the Johnson-Cook return, and invariances of the Jaumann rate.
"""
import math

import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(1234)


# ---- cubic spline (src/kernels.cuh:26-69) -------------------------------------------------------
def m4(q):
    """Monaghan M4 spline, 3-D normalisation 1/(pi h^3), written from the textbook form."""
    if q < 1:
        return 1 - 1.5 * q * q + 0.75 * q ** 3
    if q < 2:
        return 0.25 * (2 - q) ** 3
    return 0.0


@pytest.mark.parametrize("single", [False, True])
def test_cubic_spline_closed_form(single):
    h = 1.5e-3
    pi = np.array([0.01, -0.02, 0.003])
    for _ in range(200):
        d = RNG.normal(size=3)
        d *= RNG.uniform(0.01, 2.2) * h / np.linalg.norm(d)
        w = O.cubic_spline(pi, pi - d, h, single)
        r = np.linalg.norm(d)
        q = r / h
        if r * r >= 4 * h * h or r * r < 1e-8:  # support and the absolute 1e-8 cut (Q11)
            assert np.all(w == 0.0)
            continue
        if single:  # Q9: the radius is sqrtf of the float-rounded r^2, correctly rounded
            q = float(np.sqrt(np.float32(r * r))) / h
        expect = m4(q) / (math.pi * h ** 3)
        assert w[0] == pytest.approx(expect, rel=1e-12, abs=1e-300)
        # gradient by central differences of the closed form
        eps = 1e-7 * h
        g = np.zeros(3)
        for k in range(3):
            e = np.zeros(3)
            e[k] = eps
            g[k] = (m4(np.linalg.norm(d + e) / h) - m4(np.linalg.norm(d - e) / h)) / (2 * eps) / (math.pi * h ** 3)
        np.testing.assert_allclose(w[1:], g, rtol=1e-5 if single else 1e-6, atol=1e-9 * np.abs(g).max() + 1e-300)


def test_cubic_spline_special_cases():
    h = 1e-3
    p = np.array([0.1, 0.2, 0.3])
    assert np.all(O.cubic_spline(p, p, h, False) == 0.0)                      # r^2 < 1e-8
    assert np.all(O.cubic_spline(p, p + [2 * h, 0, 0], h, False) == 0.0)      # r = 2h
    w = O.cubic_spline(p, p + [0.0, 0.0, 1.5e-4], h, False)
    assert w[0] > 0 and w[3] > 0                                              # grad along +z of x_i - x_j
    # antisymmetry of the gradient under i <-> j
    a = O.cubic_spline(p, p + [3e-4, -2e-4, 1e-4], h, False)
    b = O.cubic_spline(p + [3e-4, -2e-4, 1e-4], p, h, False)
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[1:], -b[1:])


def test_cubic_spline_partition_of_unity_and_wdeltap():
    """Lattice sum of V W -> 1 in the bulk; W(dx, 1.5 dx) is the wdeltap of the set-ups."""
    dx = 1e-3
    h = 1.5 * dx
    g = np.arange(-4, 5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    s = sum(O.cubic_spline(np.zeros(3), p, h, False)[0] for p in pts) * dx ** 3
    s += m4(0.0) / (math.pi * h ** 3) * dx ** 3  # the i = j term the kernel returns as zero
    assert s == pytest.approx(1.0, abs=5e-3)
    wd = O.cubic_spline(np.zeros(3), np.array([dx, 0, 0]), h, False)[0]
    q = 1 / 1.5
    assert wd == pytest.approx((1 / (math.pi * h ** 3)) * (1 - 1.5 * q * q * (1 - 0.5 * q)), rel=1e-14)


def test_lapl_pse_closed_form():
    h = 2e-3
    for _ in range(50):
        pi = RNG.normal(size=3) * 1e-2
        pj = pi + RNG.normal(size=3) * 2e-3
        r2 = float(np.sum((pi - pj) ** 2))
        expect = 4.0 / (h ** 5 * math.pi ** 1.5) * math.exp(-r2 / h ** 2)
        assert O.lapl_pse(pi, pj, h) == pytest.approx(expect, rel=1e-13)


# ---- Johnson-Cook (src/plasticity.cuh:28-93) -----------------------------------------------------
JC = O.Joco(0.01098, 0.01092, 0.014, 0.93, 1.1, 1878.0, 300.0, 1e-6, 1.0)


def test_sigma_yield_closed_form():
    assert O.sigma_yield(JC, 0.0, 0.0, 300.0) == pytest.approx(JC.A, rel=1e-15)
    for eps, epsd, t in [(0.1, 1e-7, 400.0), (0.5, 5e-6, 900.0), (1.2, 2e-3, 1500.0)]:
        ed = epsd / JC.eps_dot_ref
        rate = 1 + JC.C * math.log(ed) if ed > 1 else (1 + ed) ** JC.C
        th = (t - JC.Tref) / (JC.Tmelt - JC.Tref)
        expect = (JC.A + JC.B * eps ** JC.n) * rate * (1 - th ** JC.m)
        assert O.sigma_yield(JC, eps, epsd, t) == pytest.approx(expect, rel=1e-14)
    # below the reference rate the rate factor is (1 + eps_dot*)^C, not 1
    assert O.sigma_yield(JC, 0.0, 0.5e-6, 300.0) == pytest.approx(JC.A * 1.5 ** JC.C, rel=1e-14)


def _bisect_root(norm, ep, t, dt, mu):
    lo, hi = 0.0, norm / (2 * mu)
    assert O.opt_fun(JC, lo, norm, ep, t, dt, mu) > 0 > O.opt_fun(JC, hi, norm, ep, t, dt, mu)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if O.opt_fun(JC, mid, norm, ep, t, dt, mu) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_opt_fun_closed_form():
    dt, mu = 1e-5, 0.4074
    for dl, norm, ep, t in [(0.0, 0.02, 0.0, 300.0), (1e-3, 0.03, 0.2, 600.0), (-1e-4, 0.01, 0.1, 500.0)]:
        dlp = max(dl, 0.0)
        eq = ep + math.sqrt(2 / 3) * dlp
        expect = norm - 2 * mu * dl - math.sqrt(2 / 3) * O.sigma_yield(JC, eq, math.sqrt(2 / 3) * dlp / dt, t)
        assert O.opt_fun(JC, dl, norm, ep, t, dt, mu) == pytest.approx(expect, rel=1e-14, abs=1e-18)


@pytest.mark.parametrize("dt", [4e-4, 1e-5])
def test_secant_root_matches_bisection(dt):
    """Where the first secant step exceeds the absolute tolerance the iteration runs to the root
    of opt_fun; where it does not, the reference stops after one step (absolute tol, Q11-like)."""
    mu = 0.4074
    for norm, ep, t in [(0.03, 0.2, 600.0), (0.015, 0.05, 1000.0)]:
        dl, failed, iters = O.solve_secant(JC, 1e-8, 1e-6, norm, ep, t, dt, mu)
        assert not failed and 0 < iters < 100
        assert dl == pytest.approx(_bisect_root(norm, ep, t, dt, mu), rel=1e-9)
    # first step below tol: returned after iteration 0 although the root lies far away
    for norm, ep, t in [(0.02, 0.0, 300.0), (0.0105, 0.1, 500.0)]:
        dl, failed, iters = O.solve_secant(JC, 1e-8, 1e-6, norm, ep, t, dt, mu)
        assert not failed and iters == 0
        assert abs(dl - 1e-8) < 1e-6
        assert _bisect_root(norm, ep, t, dt, mu) > 2e-6


# ---- eigen decomposition (src/eigen_solver.cu:212-302) -----------------------------------------
def test_eigen_against_numpy():
    for _ in range(300):
        a = RNG.normal(size=(3, 3)) * 10 ** RNG.uniform(-4, 1)
        a = a + a.T
        ev, e1, e2, e3 = O.solve_eigen(a[0, 0], a[0, 1], a[0, 2], a[1, 1], a[1, 2], a[2, 2])
        ref = np.linalg.eigvalsh(a)
        np.testing.assert_allclose(ev, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
        V = np.stack([e1, e2, e3], 1)
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-9)
        assert np.linalg.det(V) == pytest.approx(1.0, abs=1e-9)  # right handed
        np.testing.assert_allclose(a @ V, V * ev, atol=1e-8 * np.abs(ref).max())


def test_eigen_special_cases():
    ev, e1, e2, e3 = O.solve_eigen(0, 0, 0, 0, 0, 0)
    assert np.all(ev == 0) and np.array_equal(e1, [1, 0, 0]) and np.array_equal(e3, [0, 0, 1])
    ev, *_ = O.solve_eigen(1e-9, 0, 0, 2e-9, 0, 0)        # below the 1e-8 cut-off
    assert np.all(ev == 0)
    ev, e1, e2, e3 = O.solve_eigen(3.0, 0, 0, -1.0, 0, 2.0)  # diagonal: unsorted
    np.testing.assert_array_equal(ev, [3.0, -1.0, 2.0])
    ev, *_ = O.solve_eigen(2.0, 1.0, 0.0, 2.0, 0.0, 3.0)     # repeated eigenvalue 3
    np.testing.assert_allclose(np.sort(ev), [1.0, 3.0, 3.0], atol=1e-12)


def test_artificial_stress_back_rotation():
    """The rotation matrix is filled column by column with (e1.x, e2.x, e3.x), (e1.y, ...), so its
    ROWS are the eigenvectors: with V = [e1 e2 e3] the product rot * D * rot^T is V^T D V (not the
    similarity transform V D V^T a textbook would write), D = diag(-eps lambda_k / rho^2) for
    lambda_k > 0, else 0.  The eigenpairs themselves are pinned against numpy above."""
    eps, rho, p = 0.3, 4.43, -0.002
    for _ in range(100):
        a = RNG.normal(size=(3, 3)) * 1e-2
        S = a + a.T
        R = O.art_stress(eps, rho, p, S)
        B = S - p * np.eye(3)
        lam, e1, e2, e3 = O.solve_eigen(B[0, 0], B[0, 1], B[0, 2], B[1, 1], B[1, 2], B[2, 2])
        np.testing.assert_allclose(np.sort(lam), np.linalg.eigvalsh(B), atol=1e-12)
        V = np.stack([e1, e2, e3], 1)
        d = np.where(lam > 0, -eps * lam / rho ** 2, 0.0)
        expect = V.T @ np.diag(d) @ V
        scale = max(np.abs(d).max(), 1e-30)
        np.testing.assert_allclose(R, expect, rtol=0, atol=1e-12 * scale)
        # it is symmetric and its trace is the sum of the clipped eigenvalues
        np.testing.assert_allclose(R, R.T, atol=1e-14 * scale)
        assert np.trace(R) == pytest.approx(d.sum(), abs=1e-12 * scale)


# ---- GLM matrix semantics -----------------------------------------------------------------------
def test_glm_inverse_determinant_product():
    for _ in range(100):
        M = RNG.normal(size=(3, 3))
        det, inv = O.mat3_det_inv(M)
        assert det == pytest.approx(np.linalg.det(M), rel=1e-12)
        np.testing.assert_allclose(inv, np.linalg.inv(M), rtol=1e-10, atol=1e-12)
        N = RNG.normal(size=(3, 3))
        np.testing.assert_allclose(O.mat3_mul(M, N), M @ N, rtol=1e-14, atol=1e-14)


# ---- Jaumann rate (src/actions_gpu.cu:134-190) ---------------------------------------------------
def test_jaumann_closed_form_and_rotation_invariance():
    G = 0.4074
    for _ in range(50):
        L = RNG.normal(size=(3, 3))          # L[a, b] = dv_a / dx_b
        a = RNG.normal(size=(3, 3)) * 1e-2
        S = a + a.T
        vd9 = L.reshape(9)                   # v_der[a][b] = L[a, b] in glm storage order 3*a+b
        out = O.rowmajor_from_glm(O.jaumann(G, vd9, O.glm_from_rowmajor(S)))
        # omega is filled column by column with column 0 = (0, (vy_x - vx_y)/2, (vz_x - vx_z)/2),
        # which in row/column notation is the spin W = (L - L^T)/2; the rate is then
        # 2G dev(D) + W S + S W^T with D = (L + L^T)/2
        D = 0.5 * (L + L.T)
        W = 0.5 * (L - L.T)
        expect = 2 * G * (D - np.trace(D) / 3 * np.eye(3)) + (W @ S) + (S @ W.T)
        np.testing.assert_allclose(out, expect, atol=1e-13)
    # a pure spin leaves the norm of S unchanged to first order: S : dS/dt = 0
    L = np.array([[0.0, -0.3, 0.2], [0.3, 0.0, -0.1], [-0.2, 0.1, 0.0]])
    a = RNG.normal(size=(3, 3))
    S = a + a.T
    out = O.rowmajor_from_glm(O.jaumann(G, L.reshape(9), O.glm_from_rowmajor(S)))
    assert abs(np.sum(S * out)) < 1e-12 * np.abs(S).sum() * np.abs(out).sum()
