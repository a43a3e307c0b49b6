"""HIP path vs the reference: golden fixtures, the FP64 oracle, and
size-independent properties at the full N=1000 configuration.

Every call goes through librigidfibers.so (include/rigidfibers.h).
Tolerances:
  * fixtures (F16.8 text, rounded at 5e-9): |gpu - ref| <= 5e-9 + 2e-12;
  * matrix / rhs / velocities vs oracle: <= 1e-12 of the largest entry for the
    numeric quadrature.  With the analytic inner integrals: matrix 1e-10,
    rhs 1e-11, velocities 1e-12.  The analytic moment recurrences
    (Int_analytical.f90:123-206) amplify last-bit differences: on the N=21
    layouts used here the FP64 oracle is itself 2.3e-11 (M=5) and 3.8e-11
    (M=8) of the matrix scale from the same code in 80-bit extended
    precision, and the GPU is 3.5e-11 / 2.6e-11 from the oracle (rhs
    3.5e-12, velocities 1.6e-13);
  * positions and orientations after K steps: <= 1e-9 relative (2-norm).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import rigidfibers_kth_amd as rf
from conftest import GOLDEN, QUANTITIES, ROOT, VARIANTS, load_ref, read_layout
from rigidfibers_kth_amd.layouts import layout

pytestmark = pytest.mark.gpu
EPS = 0.01


def _params(analytic, direct, **kw):
    base = dict(use_analytical_integration=int(analytic), use_direct_solver=int(direct), gmres_restart=10,
                gmres_max_iters=100, gmres_tol=1e-5)
    base.update(kw)
    return rf.default_params(**base)


def _run_phases(p, pos, ori, steps):
    out = {k: [] for _, k in QUANTITIES}
    out["iters"] = []
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        for _ in range(steps):
            sim.assemble_system()
            out["b"].append(sim.rhs())
            sim.solve_system()
            out["x"].append(sim.coefficients())
            out["iters"].append(sim.stats().last_gmres_iters)
            sim.update_velocities()
            v, w = sim.velocities()
            out["vel"].append(v)
            out["rot"].append(w)
            sim.update_fibers()
            P, T = sim.state()
            out["pos"].append(P)
            out["orient"].append(T)
    return out


def _layout(N, phi, seed, length=2.0, eps=EPS):
    """Seeded generator: centres uniform in a cube sized for volume fraction
    phi = N pi a^2 l / L^3 with a = eps*l; orientations uniform on S^2."""
    a = eps * length
    L = (N * np.pi * a * a * length / phi) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, L, size=(N, 3))
    z = rng.uniform(-1.0, 1.0, N)
    ph = rng.uniform(0.0, 2 * np.pi, N)
    s = np.sqrt(1 - z * z)
    ori = np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1)
    return pos.ravel(), ori.ravel()


@pytest.mark.parametrize("variant", VARIANTS)
def test_reference_fixtures(layout100, variant):
    """tests/reference/100_*: the Fortran's own outputs, both timesteps."""
    pos, ori = layout100
    out = _run_phases(_params(variant.startswith("analytic"), variant.endswith("direct")), pos, ori, 2)
    if variant.endswith("gmres"):
        assert out["iters"] == [10, 9]
    for step in (0, 1):
        for name, key in QUANTITIES:
            ref = load_ref(variant, step, name)
            mine = out[key][step]
            d = np.abs(mine - ref).max()
            assert d <= 5e-9 + 2e-12, (variant, step, name, d)


@pytest.mark.parametrize("analytic", [1, 0])
@pytest.mark.parametrize("M,noqi", [(5, 8), (3, 4), (8, 6), (1, 2)])
def test_matrix_rhs_vs_oracle(analytic, M, noqi):
    pos, ori = _layout(21, 2e-2, 100 + M)
    p = _params(analytic, 0, num_terms=M, num_quad_intervals=noqi, gmres_tol=1e-12, gmres_max_iters=300)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.assemble_system()
        A = sim.matrix()
        b = sim.rhs()
    Ao = oracle.assemble_matrix(analytic, EPS, pos, ori, M, noqi)
    bo = oracle.assemble_rhs(analytic, EPS, pos, ori, np.tile([0, 0, -1.0], 21), M, noqi)
    scale = np.abs(Ao - np.eye(Ao.shape[0])).max()
    tol_a, tol_b = (1e-10, 1e-11) if analytic else (1e-12, 1e-12)
    print("parity A %.2e b %.2e (analytic %d, M %d, NoQI %d)" % (
        np.abs(A - Ao).max() / scale, np.abs(b - bo).max() / np.abs(bo).max(), analytic, M, noqi))
    assert np.abs(A - Ao).max() <= tol_a * scale
    assert np.abs(b - bo).max() <= tol_b * np.abs(bo).max()


@pytest.mark.parametrize("analytic", [1, 0])
def test_velocities_vs_oracle(analytic):
    N, M, noqi = 30, 5, 8
    pos, ori = _layout(N, 1e-2, 7)
    force = np.tile([0.2, -0.1, -1.0], N)
    p = _params(analytic, 0, gmres_tol=1e-13, gmres_max_iters=300, external_force=(0.2, -0.1, -1.0))
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.assemble_system()
        sim.solve_system()
        x = sim.coefficients()
        sim.update_velocities()
        v, w = sim.velocities()
    vo, wo = oracle.compute_velocities(analytic, EPS, pos, ori, force, M, noqi, x)
    tol = 1e-12
    print("parity V %.2e Omega %.2e (analytic %d)" % (np.abs(v - vo).max() / np.abs(vo).max(),
                                                     np.abs(w - wo).max() / np.abs(wo).max(), analytic))
    assert np.abs(v - vo).max() <= tol * np.abs(vo).max()
    assert np.abs(w - wo).max() <= tol * max(np.abs(wo).max(), 1e-300)


@pytest.mark.parametrize("precision", [rf.PREC_FP64, rf.PREC_MIXED])
@pytest.mark.parametrize("analytic", [1, 0])
def test_shear_velocities_vs_oracle(analytic, precision):
    """Background shear (shear_rate y, 0, 0) in the velocities
    (fortran/coeffs_and_vels.f90:597-600): the stored velocity operator (FP64)
    and the matrix-free operator (mixed) against the oracle."""
    N, M, noqi, sh = 30, 5, 8, 0.8
    pos, ori = _layout(N, 1e-2, 9)
    force = np.tile([0.2, -0.1, -1.0], N)
    p = _params(analytic, 0, gmres_tol=1e-13, gmres_max_iters=300, external_force=(0.2, -0.1, -1.0),
                shear_rate=sh, precision=precision)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.assemble_system()
        sim.solve_system()
        x = sim.coefficients()
        sim.update_velocities()
        v, w = sim.velocities()
        P, T = sim.state()
    vo, wo = oracle.compute_velocities(analytic, EPS, P, T, force, M, noqi, x, sh=sh)
    v0, w0 = oracle.compute_velocities(analytic, EPS, P, T, force, M, noqi, x, sh=0.0)
    assert np.abs(v0 - vo).max() > 0.1 * np.abs(vo).max()  # the shear term is not negligible here
    tol = 1e-12
    print("parity shear V %.2e Omega %.2e (analytic %d, precision %d)" % (
        np.abs(v - vo).max() / np.abs(vo).max(), np.abs(w - wo).max() / np.abs(wo).max(), analytic, precision))
    assert np.abs(v - vo).max() <= tol * np.abs(vo).max()
    assert np.abs(w - wo).max() <= tol * np.abs(wo).max()


def test_shear_k_steps_vs_oracle():
    """K steps in a background shear against the oracle, 1e-9 relative."""
    N, K, sh = 48, 6, 0.5
    pos, ori = _layout(N, 1e-2, 23)
    p = _params(1, 0, gmres_tol=1e-12, gmres_max_iters=300, shear_rate=sh)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        for _ in range(K):
            sim.step()
        P, T = sim.state()
    ref = oracle.simulate(1, 0, EPS, 0.1, K, pos, ori, M=5, noqi=8, restart=10, max_iters=300, tol=1e-12, sh=sh)
    still = oracle.simulate(1, 0, EPS, 0.1, K, pos, ori, M=5, noqi=8, restart=10, max_iters=300, tol=1e-12)
    assert np.linalg.norm(still["pos"][-1] - ref["pos"][-1]) > 1e-3 * np.linalg.norm(ref["pos"][-1])
    assert np.linalg.norm(P - ref["pos"][-1]) / np.linalg.norm(ref["pos"][-1]) <= 1e-9
    assert np.linalg.norm(T - ref["orient"][-1]) / np.linalg.norm(ref["orient"][-1]) <= 1e-9


def test_mfma_16x16x4_lane_layout():
    """v_mfma_{f64,f32}_16x16x4: lane l holds A[l%16][l/16] and B[l/16][l%16];
    accumulator r of lane l is D[l/16 + 4r][l%16] (f64) and D[4(l/16) + r][l%16]
    (f32) -- the layouts the LU trailing update uses."""
    lanes = np.arange(64)
    r = np.arange(4)[None, :]
    g = (lanes // 16)[:, None]
    rows_f64 = g + 4 * r
    rows_f32 = 4 * g + r
    cols = np.repeat((lanes % 16)[:, None], 4, axis=1)
    got = {m: rf.selftest_mfma_layout(m) for m in (0, 1, 2)}
    report = {
        "f64 rows": np.array_equal(got[0][0], rows_f64 + 1),
        "f32 rows": np.array_equal(got[0][1], rows_f32 + 1),
        "f64 cols": np.array_equal(got[1][0], cols + 1),
        "f32 cols": np.array_equal(got[1][1], cols + 1),
        "f64 k": np.array_equal(got[2][0], np.ones((64, 4))),
        "f32 k": np.array_equal(got[2][1], np.ones((64, 4))),
    }
    assert all(report.values()), (report, got[0][1][:20:4], got[1][1][:20:4])


@pytest.mark.parametrize("valu,nb,column_panel", [(1, 0, 0), (0, 0, 0), (0, 40, 0), (0, 13, 0), (0, 0, 1),
                                                 (1, 13, 1)])
def test_lu_solution_vs_oracle(valu, nb, column_panel):
    """Blocked LU (panel width nb, 0 = 64; partial last panels and trailing
    tiles) on the vector ALUs or the matrix cores, block-pivoted or
    column-by-column panels, vs the oracle's dgesv."""
    N = 30
    pos, ori = _layout(N, 2e-2, 13)
    p = _params(1, 1, lu_valu_update=valu, lu_block=nb, lu_column_panel=column_panel)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.assemble_system()
        sim.solve_system()
        x = sim.coefficients()
        st = sim.stats()
        A, b = sim.matrix(), sim.rhs()
    xo = oracle.lu_solve(A, b)
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
    assert st.last_backward_error <= 1e-14
    assert st.lu_panel_fallbacks == 0  # fiber systems never pivot


def test_lu_and_gmres_agree_with_oracle():
    N = 40
    pos, ori = _layout(N, 2e-2, 11)
    sols = {}
    for direct in (1, 0):
        with rf.Simulation(rf.Configuration(_params(1, direct, gmres_tol=1e-13, gmres_max_iters=400), pos,
                                            ori)) as sim:
            sim.assemble_system()
            sim.solve_system()
            sols[direct] = sim.coefficients()
            st = sim.stats()
            assert st.last_backward_error <= (1e-13 if not direct else 1e-14)
            assert st.last_solver == (rf.SOLVER_LU if direct else rf.SOLVER_GMRES)
            if direct:
                A, b = sim.matrix(), sim.rhs()
    xo = oracle.lu_solve(A, b)
    np.testing.assert_allclose(sols[1], xo, rtol=0, atol=1e-12 * np.abs(xo).max())
    np.testing.assert_allclose(sols[0], xo, rtol=0, atol=1e-11 * np.abs(xo).max())


def test_gmres_fallback_to_lu():
    """GMRES info=-4 falls back to the direct solver (coeffs_and_vels.f90:331-335)."""
    pos, ori = _layout(20, 2e-2, 5)
    with rf.Simulation(rf.Configuration(_params(1, 0, gmres_max_iters=2, gmres_tol=1e-15), pos, ori)) as sim:
        sim.assemble_system()
        sim.solve_system()
        st = sim.stats()
        assert st.last_solver == rf.SOLVER_LU_FALLBACK
        x = sim.coefficients()
        A, b = sim.matrix(), sim.rhs()
    np.testing.assert_allclose(x, oracle.lu_solve(A, b), atol=1e-12)


@pytest.mark.parametrize("analytic,direct", [(1, 0), (1, 1), (0, 0)])
def test_k_steps_vs_oracle(analytic, direct):
    """Positions and orientations after K steps within 1e-9 relative."""
    N, K = 48, 6
    pos, ori = _layout(N, 1e-2, 21)
    p = _params(analytic, direct, gmres_tol=1e-12, gmres_max_iters=300)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        for _ in range(K):
            sim.step()
        P, T = sim.state()
    ref = oracle.simulate(analytic, direct, EPS, 0.1, K, pos, ori, M=5, noqi=8, restart=10, max_iters=300,
                          tol=1e-12)
    assert np.linalg.norm(P - ref["pos"][-1]) / np.linalg.norm(ref["pos"][-1]) <= 1e-9
    assert np.linalg.norm(T - ref["orient"][-1]) / np.linalg.norm(ref["orient"][-1]) <= 1e-9


def test_warm_start_modes_agree():
    N = 40
    pos, ori = _layout(N, 1e-2, 3)
    res = {}
    for ws in (rf.WARM_PREV, rf.WARM_ZERO, rf.WARM_EXTRAP):
        p = _params(1, 0, gmres_tol=1e-12, gmres_max_iters=300, warm_start=ws)
        with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
            its = []
            for _ in range(4):
                sim.step()
                its.append(sim.stats().last_gmres_iters)
            res[ws] = (sim.state()[0], its)
    for ws in (rf.WARM_ZERO, rf.WARM_EXTRAP):
        d = np.linalg.norm(res[ws][0] - res[rf.WARM_PREV][0]) / np.linalg.norm(res[rf.WARM_PREV][0])
        assert d <= 1e-10
    # later steps start closer to the solution than x0 = 0
    assert sum(res[rf.WARM_PREV][1][1:]) < sum(res[rf.WARM_ZERO][1][1:])


def test_cli_reproduces_fixtures(tmp_path):
    """fibers run100.in XcT_init100.in: POS.out etc. after two steps are the
    1_*.ref files of 100_analytic_direct."""
    exe = os.path.join(ROOT, "rigidfibers_kth_amd", "fibers")
    inp = os.path.join(GOLDEN, "inputs")
    subprocess.run([exe, "--steps", "2", "--out", str(tmp_path), os.path.join(inp, "run100.in"),
                    os.path.join(inp, "XcT_init100.in")], check=True, capture_output=True, timeout=120)
    for name in ("POS", "ORIENT", "TRANSVel", "ROTVel", "XVec", "BVec"):
        mine = np.loadtxt(tmp_path / (name + ".out"))
        ref = load_ref("analytic_direct", 1, name)
        assert np.abs(mine - ref).max() <= 5e-9 + 2e-12, name
    snap = np.loadtxt(tmp_path / "XcTData1_p1")
    assert snap.size == 3 * 3 * 2 * 100  # initial + 2 saves, positions and orientations


def test_c1_as_configured_vs_oracle():
    """The CPU-runnable configuration as specified: N=100, the seeded
    generator at volume fraction 1e-3, M=5, 8x3 quadrature, GMRES tol 1e-10,
    dt 0.1, 10 second-order multistep steps: positions and orientations
    within 1e-9 relative of the oracle after every step."""
    N, K = 100, 10
    pos, ori = layout(N, 1e-3, 0x5EED0000)
    p = _params(1, 0, gmres_tol=1e-10, gmres_max_iters=300)
    traj = []
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        for _ in range(K):
            sim.step()
            traj.append(sim.state())
    ref = oracle.simulate(1, 0, EPS, 0.1, K, pos, ori, M=5, noqi=8, restart=10, max_iters=300, tol=1e-10)
    for k, (P, T) in enumerate(traj):
        ep = np.linalg.norm(P - ref["pos"][k]) / np.linalg.norm(ref["pos"][k])
        eo = np.linalg.norm(T - ref["orient"][k]) / np.linalg.norm(ref["orient"][k])
        print("step %d: positions %.2e orientations %.2e" % (k + 1, ep, eo))
        assert ep <= 1e-9 and eo <= 1e-9


def test_full_size_c2_properties():
    """N=1000, volume fraction 1e-2 (the plain FP64 GPU configuration):
    sampled matrix blocks, rhs rows and velocities against the oracle, identity
    diagonal blocks, and the GMRES backward error."""
    N, M, noqi = 1000, 5, 8
    pos, ori = layout(N, 1e-2, 0x5EED0001)
    p = _params(1, 0, gmres_tol=1e-10, gmres_max_iters=500)
    rng = np.random.default_rng(2)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.assemble_system()
        A = sim.matrix()
        b = sim.rhs()
        sim.solve_system()
        st = sim.stats()
        x = sim.coefficients()
        sim.update_velocities()
        v, w = sim.velocities()
    assert st.last_backward_error <= 1e-10
    r = b - A @ x
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)
    m = 3 * M
    P3, T3 = pos.reshape(-1, 3), ori.reshape(-1, 3)
    scale = np.abs(A - np.eye(A.shape[0])).max()
    for i in rng.choice(N, 8, replace=False):
        assert np.array_equal(A[i * m:(i + 1) * m, i * m:(i + 1) * m], np.eye(m))
    worst = 0.0
    for _ in range(24):
        i, j = rng.choice(N, 2, replace=False)
        blk = oracle.matrix_block(1, EPS, P3[i], T3[i], P3[j], T3[j], M, noqi)
        mine = A[i * m:(i + 1) * m, j * m:(j + 1) * m]
        worst = max(worst, np.abs(mine - blk).max() / scale)
        assert np.abs(mine - blk).max() <= 1e-10 * scale
    targets = rng.choice(N, 6, replace=False)
    force = np.tile([0, 0, -1.0], N)
    bo = oracle.assemble_rhs_rows(1, EPS, pos, ori, force, M, noqi, targets)
    bm = np.concatenate([b[t * m:(t + 1) * m] for t in targets])
    vo, wo = oracle.compute_velocities_rows(1, EPS, pos, ori, force, M, noqi, x, targets)
    vm = np.concatenate([v[3 * t:3 * t + 3] for t in targets])
    wm = np.concatenate([w[3 * t:3 * t + 3] for t in targets])
    print("parity N=1000 blocks %.2e rhs %.2e V %.2e Omega %.2e" % (
        worst, np.abs(bm - bo).max() / np.abs(bo).max(), np.abs(vm - vo).max() / np.abs(vo).max(),
        np.abs(wm - wo).max() / np.abs(wo).max()))
    assert np.abs(bm - bo).max() <= 1e-11 * np.abs(bo).max()
    assert np.abs(vm - vo).max() <= 1e-12 * np.abs(vo).max()
    assert np.abs(wm - wo).max() <= 1e-12 * np.abs(wo).max()


# Blocks of the FP32 matrix (far-field moments from FP32 backward recurrences,
# FP32 storage) against the FP64 oracle, relative to the largest sampled
# entry: 2.4e-6 measured at N=4000 (volume fraction 5e-2), 2.0e-5 at N=10000
# (volume fraction 1e-2, where sampled pairs are far apart and their blocks
# small).  The matrix-vector error that matters to the solver is 2.2e-7.
FP32_BLOCK_TOL = 5e-5


def _spot_checks(sim, pos, ori, x, v, w, b, M, noqi, tol, rng, pairs=16, rows=4, tol_vel=None):
    """Sampled matrix blocks, identity diagonal blocks, rhs rows and velocity
    rows against the oracle (relative to the largest entry of each check)."""
    N = pos.size // 3
    m = 3 * M
    P3, T3 = pos.reshape(-1, 3), ori.reshape(-1, 3)
    blocks = []
    for _ in range(pairs):
        i, j = rng.choice(N, 2, replace=False)
        blk = oracle.matrix_block(1, EPS, P3[i], T3[i], P3[j], T3[j], M, noqi)
        blocks.append((sim.matrix_block(i * m, j * m, m, m), blk))
    scale = max(np.abs(o).max() for _, o in blocks)
    print("spot checks: blocks %.2e of the largest sampled entry" % (
        max(np.abs(mine - blk).max() for mine, blk in blocks) / scale))
    for mine, blk in blocks:
        assert np.abs(mine - blk).max() <= tol * scale
    for i in rng.choice(N, 4, replace=False):
        assert np.array_equal(sim.matrix_block(i * m, i * m, m, m), np.eye(m))
    targets = rng.choice(N, rows, replace=False)
    force = np.tile([0, 0, -1.0], N)
    bo = oracle.assemble_rhs_rows(1, EPS, pos, ori, force, M, noqi, targets)
    bm = np.concatenate([b[t * m:(t + 1) * m] for t in targets])
    assert np.abs(bm - bo).max() <= tol * np.abs(bo).max()
    tv = tol if tol_vel is None else tol_vel
    vo, wo = oracle.compute_velocities_rows(1, EPS, pos, ori, force, M, noqi, x, targets)
    vm = np.concatenate([v[3 * t:3 * t + 3] for t in targets])
    wm = np.concatenate([w[3 * t:3 * t + 3] for t in targets])
    print("spot checks: rhs %.2e, V %.2e, Omega %.2e" % (np.abs(bm - bo).max() / np.abs(bo).max(),
                                                       np.abs(vm - vo).max() / np.abs(vo).max(),
                                                       np.abs(wm - wo).max() / np.abs(wo).max()))
    assert np.abs(vm - vo).max() <= tv * np.abs(vo).max()
    assert np.abs(wm - wo).max() <= tv * np.abs(wo).max()


def _fp64_residual_checks(sim, pos, ori, x, M, noqi, tol, rng, rows=3):
    """Backward error of x against the true FP64 system: the matrix-free GPU
    operator over all rows, and the oracle's rows b_i - (A x)_i for a few
    fibers (no stored matrix on either side).  Returns the GPU figure.

    The FP64 analytic operator is itself evaluated with rounding errors of up
    to ~1e-9 ||b|| near fiber contacts at volume fraction 5e-2 (the
    reference's formula in double against an extended-precision evaluation
    of the same rows), so:
      * FP64 mode (stored FP64 matrix, solved to tol against it): the
        operator's backward error is at most tol plus the measured
        difference between the operator and the stored matrix on x;
      * mixed mode (flexible GMRES on the operator): tol, or 16x the
        operator noise the solve measured where it reports a noise-limited
        result."""
    N = pos.size // 3
    m = 3 * M
    st = sim.stats()
    r = sim.operator(rf.OP_RESIDUAL, x)
    b64 = sim.operator(rf.OP_RHS)
    be = np.linalg.norm(r) / np.linalg.norm(b64)
    targets = rng.choice(N, rows, replace=False)
    force = np.tile([0, 0, -1.0], N)
    bo = oracle.assemble_rhs_rows(1, EPS, pos, ori, force, M, noqi, targets)
    yo = oracle.apply_rows(1, EPS, pos, ori, M, noqi, x, targets)
    ro = bo - yo
    rg = np.concatenate([r[t * m:(t + 1) * m] for t in targets])
    bg = np.concatenate([b64[t * m:(t + 1) * m] for t in targets])
    print("FP64 backward error %.2e; oracle rows: |r| %.2e of |b| (GPU rows %.2e); |r_gpu - r_oracle| %.2e of |b|; "
          "rhs rows %.2e" % (be, np.linalg.norm(ro) / np.linalg.norm(bo), np.linalg.norm(rg) / np.linalg.norm(bg),
                             np.abs(rg - ro).max() / np.abs(bo).max(), np.abs(bg - bo).max() / np.abs(bo).max()))
    if st.matrix_elem_bytes == 8:
        noise = np.linalg.norm(sim.operator(rf.OP_AX, x) - sim.matvec(x)) / np.linalg.norm(b64)
        print("operator vs stored FP64 matrix on x: %.2e of |b|" % noise)
        assert noise <= 2e-9
        assert be <= 1.01 * tol + noise
    elif st.last_noise_limited:
        assert be <= 1.01 * 16 * st.last_operator_noise
    else:
        assert be <= 1.01 * tol
    assert np.abs(bg - bo).max() <= 1e-11 * np.abs(bo).max()  # 1.7e-12 measured
    # the oracle's own rows of the FP64 residual: at the solver's tolerance
    # (an FP32-operator solution leaves ~1e-7 here)
    assert np.linalg.norm(ro) <= 20 * tol * np.linalg.norm(bo)
    return be


def _meets_tol(st, tol):
    """The solve's backward error meets tol, or (mixed mode) it was accepted
    within 16x the FP64 operator's measured rounding noise, which near
    interpenetrating fibers lies above tol."""
    if st.last_backward_error <= tol:
        return True
    return bool(st.last_noise_limited) and st.last_backward_error <= 16 * st.last_operator_noise


def _solve_reporting_refinements(sim):
    """rf_solve; prints the mixed mode's FP64 backward errors (initial guess,
    outer estimates, final) and inner iterations (also when it fails)."""
    try:
        sim.solve_system()
    finally:
        st = sim.stats()
        k = min(8, st.last_outer_iters + 1)
        print("FP64 backward error: guess %.2e, outer estimates %s, final %.2e (operator noise %.1e, noise "
              "limited %d); %d outer iterations, inner iterations %s; precond %d" % (
                  st.refine_history[0], ["%.2e" % v for v in list(st.refine_history)[1:k]], st.last_backward_error,
                  st.last_operator_noise, st.last_noise_limited, st.last_outer_iters,
                  list(st.refine_inner_iters)[:k - 1], st.last_precond), flush=True)


def test_full_size_n4000_properties():
    """N=4000, volume fraction 5e-2 (n = 60000; 28.8 GB FP64 matrix) on the
    FP64 operator with the neighbour preconditioner switching to FP32-LU
    GMRES-IR: sampled blocks, rhs and velocity rows against the oracle,
    identity diagonal blocks, and the backward error recomputed with the GPU
    matvec, the matrix-free operator and the oracle's rows."""
    N, M, noqi = 4000, 5, 8
    pos, ori = layout(N, 5e-2, 0x5EED0003)
    p = _params(1, 0, gmres_restart=30, gmres_tol=1e-10, gmres_max_iters=3000, precond=rf.PRECOND_AUTO,
                lu_factor_fp32=1, lu_update_fp16=1, precond_switch_iters=10)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.assemble_system()
        b = sim.rhs()
        sim.solve_system()
        st = sim.stats()
        x = sim.coefficients()
        r = b - sim.matvec(x)
        sim.update_velocities()
        v, w = sim.velocities()
        assert st.last_backward_error <= 1e-10
        assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b)
        _fp64_residual_checks(sim, pos, ori, x, M, noqi, 1e-10, np.random.default_rng(3))
        # measured: blocks 2.4e-11, rhs 5.6e-14, V 4.1e-14, Omega 3.4e-12
        _spot_checks(sim, pos, ori, x, v, w, b, M, noqi, 1e-10, np.random.default_rng(4), tol_vel=1e-11)


def test_full_size_n4000_mixed_learned():
    """N=4000, volume fraction 5e-2 on the mixed path with the learned
    initial guess: FP32 matrix (FP32 GEMV inside GMRES, FP64 Arnoldi), FP32 LU,
    refinement against the FP64 operator.  The network is trained briefly on
    the solved coefficients of the first configuration; two steps with
    RF_WARM_LEARNED then meet the FP64 tolerance against the true operator and
    the oracle's rows, with blocks at the FP32 storage tolerance and FP64
    velocities."""
    N, M, noqi = 4000, 5, 8
    pos, ori = layout(N, 5e-2, 0x5EED0003)
    p = _params(1, 0, gmres_restart=30, gmres_tol=1e-10, gmres_max_iters=3000, precond=rf.PRECOND_AUTO,
                precision=rf.PREC_MIXED, lu_update_fp16=1, precond_switch_iters=10, warm_start=rf.WARM_LEARNED)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.assemble_system()
        sim.solve_system()
        X, Y = sim.mlp_samples()
        sim.mlp_train(X, Y, epochs=4, batch=256, lr=2e-3, seed=1)
        sim.update_velocities()
        sim.update_fibers()
        for k in range(2):
            sim.assemble_system()
            b = sim.rhs()
            _solve_reporting_refinements(sim)
            st = sim.stats()
            x = sim.coefficients()
            P, T = sim.state()
            sim.update_velocities()
            v, w = sim.velocities()
            print("step %d: %d iterations, %d refinements, guess alpha %.3f beta %.3f" % (
                k + 2, st.last_gmres_iters, st.last_refinements, st.last_guess_alpha, st.last_guess_beta))
            assert st.last_guess_beta != 0.0
            assert _meets_tol(st, 1e-10)
            _fp64_residual_checks(sim, P, T, x, M, noqi, 1e-10, np.random.default_rng(40 + k))
            _spot_checks(sim, P, T, x, v, w, b, M, noqi, FP32_BLOCK_TOL, np.random.default_rng(50 + k), pairs=8,
                         tol_vel=5e-11)
            sim.update_fibers()


def test_full_size_n10000_mixed_steps():
    """N=10000, volume fraction 1e-2 (n = 150000): FP32 matrix (90 GB) on one
    GPU, neighbour preconditioner, refinement against the FP64 operator to
    1e-8, dt = 0.05, three steps: every solve meets 1e-8 against the true FP64
    system (operator and oracle rows), blocks at the FP32 storage tolerance,
    FP64 velocities, Euler then multistep advance."""
    N, M, noqi = 10000, 5, 8
    pos, ori = layout(N, 1e-2, 0x5EED0004)
    p = _params(1, 0, gmres_restart=30, gmres_tol=1e-8, gmres_max_iters=1000, precision=rf.PREC_MIXED,
                precond=rf.PRECOND_NEIGHBOUR, timestep=0.05)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        hist = [sim.state()]
        vels = []
        for k in range(3):
            sim.assemble_system()
            b = sim.rhs()
            _solve_reporting_refinements(sim)
            st = sim.stats()
            x = sim.coefficients()
            P, T = sim.state()
            sim.update_velocities()
            v, w = sim.velocities()
            vels.append((v, w))
            print("step %d: %d iterations, %d refinements" % (k + 1, st.last_gmres_iters, st.last_refinements))
            assert st.last_solver == rf.SOLVER_GMRES
            assert _meets_tol(st, 1e-8)
            _fp64_residual_checks(sim, P, T, x, M, noqi, 1e-8, np.random.default_rng(60 + k), rows=2)
            if k == 0:
                _spot_checks(sim, P, T, x, v, w, b, M, noqi, FP32_BLOCK_TOL, np.random.default_rng(5), pairs=8,
                             tol_vel=5e-11)
            sim.update_fibers()
            hist.append(sim.state())
    X0, X1, X2 = hist[0][0], hist[1][0], hist[2][0]
    dt = 0.05
    assert np.allclose(X1, X0 + dt * vels[0][0], rtol=0, atol=1e-12 * np.abs(X0).max())  # Euler
    want = 4.0 / 3.0 * X1 - 1.0 / 3.0 * X0 + 2.0 / 3.0 * dt * (2.0 * vels[1][0] - vels[0][0])
    assert np.allclose(X2, want, rtol=0, atol=1e-12 * np.abs(X0).max())  # second-order multistep
    for _, T in hist:
        assert np.allclose(np.linalg.norm(T.reshape(-1, 3), axis=1), 1.0, atol=1e-14)


def test_full_size_n2000_mixed_lu_properties():
    """N=2000, volume fraction 2e-2 (n = 30000): FP32 matrix with FP64 near
    field, FP32 LU on the matrix cores, refinement against the FP64
    matrix-free operator to 1e-10."""
    N, M, noqi = 2000, 5, 8
    pos, ori = layout(N, 2e-2, 0x5EED0002)
    p = _params(1, 1, precision=rf.PREC_MIXED, gmres_tol=1e-10)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.assemble_system()
        b = sim.rhs()
        sim.solve_system()
        st = sim.stats()
        x = sim.coefficients()
        sim.update_velocities()
        v, w = sim.velocities()
        print("refinements", st.last_refinements, "backward error", st.last_backward_error)
        assert st.last_solver == rf.SOLVER_LU and st.lu_factorizations == 1
        assert st.last_backward_error <= 1e-10
        _fp64_residual_checks(sim, pos, ori, x, M, noqi, 1e-10, np.random.default_rng(7))
        _spot_checks(sim, pos, ori, x, v, w, b, M, noqi, FP32_BLOCK_TOL, np.random.default_rng(6), pairs=12, tol_vel=5e-11)
