"""Pins the FP64 CPU oracle (oracle/rf_oracle.c) before it is trusted.

Three independent kinds of check:
  * the reference's own golden fixtures (tests/reference/100_*/*.ref, written
    by the Fortran at F16.8): byte-exact text after formatting;
  * closed forms of the mathematics (Gauss-Legendre nodes, Legendre
    orthogonality, single-fiber velocity, identity diagonal blocks);
  * brute-force adaptive quadrature of the moment integrals and of the
    Stokeslet/doublet kernel, equivariance, matrix-free vs dense operator.
A plausible mistake (dropped term, wrong sign or index, transposed operand)
fails at least one of them.
"""
import numpy as np
import pytest
from scipy import integrate

import oracle
from conftest import QUANTITIES, VARIANTS, load_ref

EPS = 0.01


@pytest.fixture(scope="module")
def sims(layout100):
    pos, orient = layout100
    out = {}
    for v in VARIANTS:
        out[v] = oracle.simulate(v.startswith("analytic"), v.endswith("direct"), EPS, 0.1, 2, pos, orient,
                                 M=5, noqi=8, restart=10, max_iters=100, tol=1e-5)
    return out


@pytest.mark.parametrize("variant", VARIANTS)
def test_fixtures_byte_exact(sims, variant):
    """fortran/advect_fibers.f90:529-539 + coeffs_and_vels.f90:55-72,633-642 write
    '(*(F16.8))' one value per line; the oracle formats to the same text."""
    import os
    from conftest import GOLDEN
    out = sims[variant]
    for step in (0, 1):
        for name, key in QUANTITIES:
            path = os.path.join(GOLDEN, "reference_100", "100_" + variant, "%d_%s.ref" % (step, name))
            ref_lines = open(path).read().splitlines()
            mine = ["%16.8f" % v for v in out[key][step]]
            assert len(ref_lines) == len(mine)
            bad = [i for i, (a, b) in enumerate(zip(ref_lines, mine)) if a != b]
            assert not bad, (variant, step, name, bad[:5])


def test_gmres_iteration_counts(sims):
    # The Fortran prints 10 and 9 iterations for the run100_iter.in runs.
    assert list(sims["analytic_gmres"]["iters"]) == [10, 9]
    assert list(sims["numeric_gmres"]["iters"]) == [10, 9]


def test_quadrature_closed_form():
    for noqi in (1, 3, 8):
        pv, wv = oracle.quadrature(noqi)
        x0, w0 = np.polynomial.legendre.leggauss(3)
        edges = np.linspace(-1, 1, noqi + 1)
        exp_p = np.concatenate([(a + b) / 2 + (b - a) / 2 * x0 for a, b in zip(edges[:-1], edges[1:])])
        exp_w = np.concatenate([(b - a) / 2 * w0 for a, b in zip(edges[:-1], edges[1:])])
        np.testing.assert_allclose(pv, exp_p, atol=1e-15)
        np.testing.assert_allclose(wv, exp_w, atol=1e-15)
        for k in range(6):  # exact up to degree 5
            exact = (1 - (-1) ** (k + 1)) / (k + 1)
            assert abs(np.sum(wv * pv ** k) - exact) < 1e-14


def test_legendre_closed_form_and_orthogonality():
    pv, wv = oracle.quadrature(8)
    P = oracle.legendre(8, pv)
    for k in range(1, 9):
        c = np.zeros(k + 1)
        c[k] = 1
        np.testing.assert_allclose(P[k - 1], np.polynomial.legendre.legval(pv, c), atol=1e-13)
    P0 = np.vstack([np.ones_like(pv), P])
    for k in range(0, 6):
        for l in range(0, 6 - k):
            s = np.sum(wv * P0[k] * P0[l])
            assert abs(s - (2.0 / (2 * k + 1) if k == l else 0.0)) < 1e-14
    with pytest.raises(ValueError):
        oracle.legendre(9, pv)


def _moment(n, b, c, k):
    f = lambda s: s ** n / (s * s + b * s + c) ** (k / 2.0)
    v, err = integrate.quad(f, -1, 1, epsabs=1e-15, epsrel=1e-13, limit=400)
    return v


GEOMS = {
    # name: (xb, pb, xbar)
    "near": ([0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.6, 0.2, 0.3]),           # c < 10, forward
    "mid": ([0.1, -0.2, 0.3], [0.6, 0.0, 0.8], [2.1, 1.5, -0.7]),         # c < 10
    "far": ([0.0, 0.0, 0.0], [0.48, 0.6, 0.64], [3.0, -1.5, 1.2]),        # c >= 10, backward
    "very_far": ([1.0, 1.0, 1.0], [0.0, 1.0, 0.0], [7.0, 2.0, -4.0]),     # c >= 10
    # Collinear points (d = 0 branch).  Only points behind the source fiber
    # (R0 . pb < 0) are finite in the reference: ahead of it the I_0 formula
    # (Int_analytical.f90:113) evaluates log(0) - log(0).
    "axis_near": ([0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, -2.5]),    # d = 0 branch, forward
    "axis_far": ([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [-4.0, 0.0, 0.0]),     # d = 0 branch, backward
}


@pytest.mark.parametrize("geom", list(GEOMS))
def test_analytic_moments_vs_adaptive_quadrature(geom):
    """fortran/Int_analytical.f90:72-207: I_n, J_n, S_n for n <= M+2 equal the
    integrals of s^n/|R|^{1,3,5} (forward, backward and collinear branches)."""
    xb, pb, xbar = (np.array(v, float) for v in GEOMS[geom])
    R0 = xbar - xb
    b = -2 * pb.dot(R0)
    c = R0.dot(R0)
    M = 5
    L, I, J, S = oracle.analytical_int(xb, pb, xbar, M)
    # The backward branch (c >= 10) starts Miller-style at n = 27 with zeros
    # (Int_analytical.f90:138-158); its truncation error is ~1e-10 relative.
    tol = 2e-11 if c < 10 else 1e-9
    for n in range(M + 3):
        for arr, k in ((I, 1), (J, 3), (S, 5)):
            ref = _moment(n, b, c, k)
            scale = max(abs(_moment(0, b, c, k)), 1e-300)
            assert abs(arr[n] - ref) <= tol * scale, (geom, n, k, arr[n], ref)


@pytest.mark.parametrize("geom", ["near", "mid", "far", "very_far"])
@pytest.mark.parametrize("l", [0, 1, 3, 5])
def test_analytic_G_vs_kernel_quadrature(geom, l):
    """G_compute_GQ_kg (Int_analytical.f90:663-767) equals the integral of the
    Stokeslet+doublet kernel (quadrature.f90:137-148) times P_l."""
    xb, pb, xbar = (np.array(v, float) for v in GEOMS[geom])
    G = oracle.G_analytic(xb, pb, xbar, EPS, l)
    Pl = np.polynomial.legendre.Legendre.basis(l)
    comps = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]

    def kern(s, a, bb):
        R = xbar - (xb + s * pb)
        r = np.linalg.norm(R)
        d = 1.0 if a == bb else 0.0
        return (d / r + R[a] * R[bb] / r ** 3 + 2 * EPS ** 2 * (d / r ** 3 - 3 * R[a] * R[bb] / r ** 5)) * Pl(s)

    for q, (a, bb) in enumerate(comps):
        ref, _ = integrate.quad(kern, -1, 1, args=(a, bb), epsabs=1e-14, epsrel=1e-13, limit=400)
        assert abs(G[q] - ref) < 1e-11, (geom, l, q, G[q], ref)


def test_numeric_G_converges_to_analytic():
    xb, pb, xbar = (np.array(v, float) for v in GEOMS["mid"])
    for l in (1, 2, 4):
        ga = oracle.G_analytic(xb, pb, xbar, EPS, l)
        pv, _ = oracle.quadrature(64)
        lv = oracle.legendre(5, pv)[l - 1]
        gn = oracle.G_numeric(xb, pb, xbar, EPS, 64, lv)
        np.testing.assert_allclose(gn, ga, atol=1e-10)


def _rand_layout(N, box, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, box, size=(N, 3))
    t = rng.normal(size=(N, 3))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    return pos.ravel(), t.ravel()


@pytest.mark.parametrize("analytic", [0, 1])
def test_matrix_structure(analytic):
    N, M = 7, 4
    pos, orient = _rand_layout(N, 5.0, 3)
    A = oracle.assemble_matrix(analytic, EPS, pos, orient, M, 6)
    n = 3 * M * N
    for i in range(N):
        blk = A[i * 3 * M:(i + 1) * 3 * M, i * 3 * M:(i + 1) * 3 * M]
        assert np.array_equal(blk, np.eye(3 * M))  # matrix_assembly.f90:308-311
    assert np.abs(A - np.eye(n)).max() > 1e-3


def test_mat_prod_equals_dense_operator():
    """mat_prod (matrix_assembly.f90:159-265) is A x without forming A."""
    N, M, noqi = 6, 5, 4
    pos, orient = _rand_layout(N, 4.0, 5)
    A = oracle.assemble_matrix(0, EPS, pos, orient, M, noqi)
    x = np.random.default_rng(1).normal(size=3 * M * N)
    y = oracle.mat_prod(EPS, pos, orient, M, noqi, x)
    np.testing.assert_allclose(y, A @ x, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("analytic", [1, 0])
def test_apply_rows_equals_dense_rows(analytic):
    """apply_rows gives the rows of A x for chosen fibers without forming A;
    a dropped block, a transposed block or a wrong source index changes
    them at O(1)."""
    N, M, noqi = 9, 4, 3
    pos, orient = _rand_layout(N, 3.0, 17)
    A = oracle.assemble_matrix(analytic, EPS, pos, orient, M, noqi)
    x = np.random.default_rng(3).normal(size=3 * M * N)
    targets = np.array([0, 4, 8, 2])
    y = oracle.apply_rows(analytic, EPS, pos, orient, M, noqi, x, targets)
    m = 3 * M
    want = np.concatenate([(A @ x)[t * m:(t + 1) * m] for t in targets])
    np.testing.assert_allclose(y, want, rtol=0, atol=1e-13 * np.abs(want).max())
    # blocks transposed in place would not do
    At = A.copy()
    for i in range(N):
        for j in range(N):
            At[i * m:(i + 1) * m, j * m:(j + 1) * m] = A[i * m:(i + 1) * m, j * m:(j + 1) * m].T
    wrong = np.concatenate([(At @ x)[t * m:(t + 1) * m] for t in targets])
    assert np.abs(wrong - y).max() > 1e-3


def _rot(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


def test_equivariance_numeric():
    """Rotating all positions, orientations and forces by R maps A to
    (I x R) A (I x R)^T and b, V, Omega to their rotations; translations
    leave them unchanged."""
    N, M, noqi = 6, 3, 4
    pos, orient = _rand_layout(N, 4.0, 11)
    force = np.tile([0.3, -0.2, -1.0], N)
    R = _rot(2)
    Rb = np.kron(np.eye(M * N), R)
    Rf = np.kron(np.eye(N), R)
    A = oracle.assemble_matrix(0, EPS, pos, orient, M, noqi)
    b = oracle.assemble_rhs(0, EPS, pos, orient, force, M, noqi)
    pos2 = (Rf @ pos) + np.tile([1.5, -2.0, 0.25], N)
    A2 = oracle.assemble_matrix(0, EPS, pos2, Rf @ orient, M, noqi)
    b2 = oracle.assemble_rhs(0, EPS, pos2, Rf @ orient, Rf @ force, M, noqi)
    np.testing.assert_allclose(A2, Rb @ A @ Rb.T, atol=1e-12)
    np.testing.assert_allclose(b2, Rb @ b, atol=1e-12)
    x = np.linalg.solve(A, b)
    v, w = oracle.compute_velocities(0, EPS, pos, orient, force, M, noqi, x)
    v2, w2 = oracle.compute_velocities(0, EPS, pos2, Rf @ orient, Rf @ force, M, noqi, Rb @ x)
    np.testing.assert_allclose(v2, Rf @ v, atol=1e-12)
    np.testing.assert_allclose(w2, Rf @ w, atol=1e-12)


def test_equivariance_analytic_about_gravity_axis():
    # The analytic RHS reads only the z force component (matrix_assembly.f90:857-859),
    # so the symmetry group is rotations about z.
    N, M, noqi = 5, 5, 8
    pos, orient = _rand_layout(N, 4.0, 12)
    force = np.tile([0.0, 0.0, -1.0], N)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    Rb = np.kron(np.eye(M * N), R)
    Rf = np.kron(np.eye(N), R)
    A = oracle.assemble_matrix(1, EPS, pos, orient, M, noqi)
    b = oracle.assemble_rhs(1, EPS, pos, orient, force, M, noqi)
    A2 = oracle.assemble_matrix(1, EPS, Rf @ pos, Rf @ orient, M, noqi)
    b2 = oracle.assemble_rhs(1, EPS, Rf @ pos, Rf @ orient, force, M, noqi)
    np.testing.assert_allclose(A2, Rb @ A @ Rb.T, atol=1e-12)
    np.testing.assert_allclose(b2, Rb @ b, atol=1e-12)


def test_lu_and_gmres_vs_numpy():
    N, M = 8, 5
    pos, orient = _rand_layout(N, 4.0, 7)
    A = oracle.assemble_matrix(1, EPS, pos, orient, M, 8)
    b = oracle.assemble_rhs(1, EPS, pos, orient, np.tile([0, 0, -1.0], N), M, 8)
    ref = np.linalg.solve(A, b)
    np.testing.assert_allclose(oracle.lu_solve(A, b), ref, rtol=1e-12, atol=1e-14)
    x, st, its, be = oracle.gmres(A, b, np.zeros_like(b), 10, 200, 1e-13)
    assert st == 0 and be <= 1e-13
    np.testing.assert_allclose(x, ref, rtol=1e-10, atol=1e-13)
    # forced non-convergence: info = -4 (coeffs_and_vels.f90:331-335 falls back to LU)
    x, st, its, be = oracle.gmres(A, b, np.zeros_like(b), 3, 2, 1e-14)
    assert st == -4 and its == 2


@pytest.mark.parametrize("analytic", [0, 1])
def test_single_fiber_velocity_closed_form(analytic):
    """No interactions: V = [1/2((d+e)F + (d-e)(t.F)t)]/d, Omega = 0
    (coeffs_and_vels.f90:596,626-627)."""
    t = np.array([0.36, 0.48, 0.8])
    F = np.array([0.2, -0.1, -1.0])
    v, w = oracle.compute_velocities(analytic, EPS, np.zeros(3), t, F, 5, 8, np.zeros(15))
    d = -np.log(EPS ** 2 * np.e)
    np.testing.assert_allclose(v, 0.5 * ((d + 2) * F + (d - 2) * t.dot(F) * t) / d, rtol=1e-14)
    assert np.all(w == 0)


@pytest.mark.parametrize("analytic", [0, 1])
def test_single_fiber_in_shear_closed_form(analytic):
    """Background shear (sh y, 0, 0) (coeffs_and_vels.f90:597-600): a lone
    fiber translates with the flow at its centre plus its sedimentation
    velocity, V = sh y e_x + [1/2((d+e)F + (d-e)(t.F)t)]/d, and turns as an
    infinitely slender rod in Jeffery's equation, dt/dt = (I - t t)(grad u) t
    = sh t_y (e_x - t_x t)."""
    t = np.array([0.36, 0.48, 0.8])
    X = np.array([0.3, -1.7, 2.2])
    F = np.array([0.2, -0.1, -1.0])
    sh = 0.7
    v, w = oracle.compute_velocities(analytic, EPS, X, t, F, 5, 8, np.zeros(15), sh=sh)
    d = -np.log(EPS ** 2 * np.e)
    v_sed = 0.5 * ((d + 2) * F + (d - 2) * t.dot(F) * t) / d
    np.testing.assert_allclose(v, v_sed + sh * X[1] * np.array([1.0, 0.0, 0.0]), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(w, sh * t[1] * (np.array([1.0, 0.0, 0.0]) - t[0] * t), rtol=1e-14, atol=1e-16)
    assert abs(w.dot(t)) < 1e-15  # the rotation keeps |t| = 1 to first order


def test_single_fiber_jeffery_orbit():
    """Time integration of the shear term: a force-free rod in the shear
    plane follows cot(phi) = cot(phi0) + sh t (Jeffery's orbit for infinite
    aspect ratio), and its centre moves with x = x0 + sh y0 t; second-order
    multistep after an Euler step, so the angle error is O(dt^2)."""
    sh, dt, K, phi0 = 1.0, 0.01, 200, 1.2
    X0 = np.array([0.5, 0.8, -0.3])
    t0 = np.array([np.cos(phi0), np.sin(phi0), 0.0])
    out = oracle.simulate(1, 1, EPS, dt, K, X0, t0, force=np.zeros(3), sh=sh)
    T = out["orient"][-1]
    phi = np.arctan2(T[1], T[0])
    phi_exact = np.arctan2(1.0, 1.0 / np.tan(phi0) + sh * K * dt)
    assert abs(phi - phi_exact) < 1e-5, (phi, phi_exact)  # 3.5e-6 measured (1.4e-5 at dt 0.02)
    assert abs(T[2]) < 1e-15
    P = out["pos"][-1]
    np.testing.assert_allclose(P, X0 + np.array([sh * X0[1] * K * dt, 0.0, 0.0]), rtol=0, atol=1e-12)


def test_time_scheme_identities(layout100, sims):
    """Euler first step and 2nd-order multistep (advect_fibers.f90:498-523)."""
    pos0, _ = layout100
    o = sims["analytic_direct"]
    np.testing.assert_allclose(o["pos"][0], pos0 + 0.1 * o["vel"][0], atol=1e-14)
    exp = 4 / 3 * o["pos"][0] - 1 / 3 * pos0 + 2 / 3 * 0.1 * (2 * o["vel"][1] - o["vel"][0])
    np.testing.assert_allclose(o["pos"][1], exp, atol=1e-14)
    nrm = np.linalg.norm(o["orient"].reshape(2, -1, 3), axis=2)
    np.testing.assert_allclose(nrm, 1.0, atol=1e-15)
