"""Solver extensions beyond the reference, checked on the GPU through the C ABI:

* fiber neighbourhoods (k nearest by segment-segment distance) vs numpy;
* the neighbour-coupled block-Jacobi preconditioner vs a numpy restatement
  (cluster = fiber + K nearest, rows of the cluster inverse), and GMRES with it
  reaching the same solution as the reference's unpreconditioned GMRES;
* mixed precision (FP32 matrix with FP64 near field and FP64 closed forms,
  FP32 LU, refinement against the FP64 matrix-free operator, FP64 Arnoldi):
  error against the FP64 path.
"""
import numpy as np
import pytest

import rigidfibers_kth_amd as rf
from bench import layout, segment_distances

pytestmark = pytest.mark.gpu


def _sim(pos, ori, **kw):
    base = dict(gmres_restart=30, gmres_max_iters=3000, gmres_tol=1e-10)
    base.update(kw)
    return rf.Simulation(rf.Configuration(rf.default_params(**base), pos, ori))


def _hard_state():
    """The configuration after two steps of the volume-fraction 5e-2 dynamics
    (N=300): sedimentation has brought fibers into near contact and the
    reference's unpreconditioned GMRES(30) needs ~1000 iterations on it."""
    pos, ori = layout(300, 5e-2, 0x5EED0003)
    with _sim(pos, ori) as sim:
        sim.step()
        sim.step()
        return sim.state()


def test_neighbours_vs_numpy():
    N, K = 180, 6
    pos, ori = layout(N, 2e-2, 41)
    with _sim(pos, ori, precond=1, precond_neighbours=K) as sim:
        nbr, dist = sim.neighbours(K)
    P, T = pos.reshape(-1, 3), ori.reshape(-1, 3)
    for i in range(N):
        d = segment_distances(P[i], T[i], P, T)
        d[i] = np.inf
        order = np.lexsort((np.arange(N), d))[:K]
        np.testing.assert_array_equal(nbr[i], order)
        np.testing.assert_allclose(dist[i], d[order], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("precision", [rf.PREC_FP64, rf.PREC_MIXED])
def test_precond_apply_vs_numpy(precision):
    N, K, M = 60, 4, 5
    pos, ori = layout(N, 2e-2, 43)
    with _sim(pos, ori, precond=1, precond_neighbours=K, precision=precision) as sim:
        sim.assemble_system()
        A = sim.matrix()
        sim.solve_system()
        nbr, _ = sim.neighbours(K)
        r = np.random.default_rng(0).normal(size=sim.n)
        z = sim.precond_apply(r)
    m = 3 * M
    zr = np.empty_like(r)
    for i in range(N):
        C = np.concatenate([[i], nbr[i]])
        idx = np.concatenate([np.arange(c * m, (c + 1) * m) for c in C])
        zr[i * m:(i + 1) * m] = np.linalg.solve(A[np.ix_(idx, idx)], r[idx])[:m]
    assert np.abs(z - zr).max() <= 1e-10 * np.abs(zr).max()


def test_precond_gmres_same_solution_fewer_iterations():
    pos, ori = _hard_state()
    out = {}
    for pc in (0, 1):
        with _sim(pos, ori, precond=pc, precond_neighbours=4) as sim:
            sim.assemble_system()
            sim.solve_system()
            st = sim.stats()
            out[pc] = (sim.coefficients(), st.last_gmres_iters, st.last_backward_error, st.last_solver)
    x0, it0, be0, s0 = out[0]
    x1, it1, be1, s1 = out[1]
    assert s1 == rf.SOLVER_GMRES and be1 <= 1e-10
    assert it1 * 5 <= it0, (it0, it1)
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)


def test_mixed_precision_matrix_and_solution():
    N = 200
    pos, ori = layout(N, 2e-2, 53)
    res = {}
    for prec in (rf.PREC_FP64, rf.PREC_MIXED):
        with _sim(pos, ori, precision=prec, precond=1) as sim:
            sim.assemble_system()
            A, b = sim.matrix(), sim.rhs()
            sim.solve_system()
            x = sim.coefficients()
            sim.update_velocities()
            v, w = sim.velocities()
            res[prec] = (A, b, x, v, w)
    A64, b64, x64, v64, w64 = res[rf.PREC_FP64]
    A32, b32, x32, v32, w32 = res[rf.PREC_MIXED]
    scale = np.abs(A64 - np.eye(A64.shape[0])).max()
    assert np.abs(A32 - A64).max() <= 1e-6 * scale
    # the rhs shares the matrix's far-field integrals (FP32 backward
    # recurrences, FP64 closed forms): ~5e-7 of its scale
    assert np.abs(b32 - b64).max() <= 2e-6 * np.abs(b64).max()
    # the solve is refined against the FP64 operator, the velocities come
    # from it: both are FP64 results at the solver tolerance (1e-10)
    assert np.linalg.norm(x32 - x64) <= 1e-8 * np.linalg.norm(x64)
    assert np.linalg.norm(v32 - v64) <= 1e-8 * np.linalg.norm(v64)
    assert np.linalg.norm(w32 - w64) <= 1e-8 * np.linalg.norm(w64)


def test_mixed_lu_with_refinement():
    """Direct solver in the mixed mode: FP32 LU of the FP32 matrix, each
    correction refined on the FP32 matrix, the iterate refined against the
    FP64 residual of the matrix-free operator: the result solves the FP64
    system, not the FP32 one."""
    N = 120
    pos, ori = layout(N, 2e-2, 59)
    with _sim(pos, ori, use_direct_solver=1) as sim:
        sim.assemble_system()
        A64 = sim.matrix()
        sim.solve_system()
        x64 = sim.coefficients()
    with _sim(pos, ori, use_direct_solver=1, precision=rf.PREC_MIXED, gmres_tol=1e-12) as sim:
        sim.assemble_system()
        A32, b = sim.matrix(), sim.rhs()
        sim.solve_system()
        x32 = sim.coefficients()
        st = sim.stats()
        r64 = sim.operator(rf.OP_RESIDUAL, x32)
        b64 = sim.operator(rf.OP_RHS)
    print("refinements", st.last_refinements, "backward error", st.last_backward_error)
    assert st.last_backward_error <= 1e-12 and st.last_refinements >= 2
    assert np.linalg.norm(r64) <= 1.01e-12 * np.linalg.norm(b64)
    assert np.linalg.norm(b64 - A64 @ x32) <= 2e-12 * np.linalg.norm(b64)
    # the FP32 system is a different one: its residual stays at FP32 rounding
    assert np.linalg.norm(b - A32 @ x32) > 1e-10 * np.linalg.norm(b)
    assert np.linalg.norm(x32 - x64) <= 1e-10 * np.linalg.norm(x64)


@pytest.mark.parametrize("precision", [rf.PREC_FP64, rf.PREC_MIXED])
def test_lu_preconditioned_gmres(precision):
    """GMRES right-preconditioned by the LU factors of the stored matrix
    (FP32 factors in the mixed mode, i.e. GMRES-based iterative refinement
    with FP64 residuals) reaches the reference's GMRES solution of the same
    stored operator in a few iterations, also on the near-contact state."""
    pos, ori = _hard_state()
    out = {}
    for pc in (rf.PRECOND_NONE, rf.PRECOND_LU):
        # unpreconditioned inner solves take ~1000 iterations each on this state
        with _sim(pos, ori, precond=pc, precision=precision, gmres_max_iters=20000) as sim:
            sim.assemble_system()
            sim.solve_system()
            st = sim.stats()
            out[pc] = (sim.coefficients(), st.last_gmres_iters, st.last_backward_error, st.last_solver)
    x0, it0, _, _ = out[rf.PRECOND_NONE]
    x1, it1, be1, s1 = out[rf.PRECOND_LU]
    assert s1 == rf.SOLVER_GMRES and be1 <= 1e-10
    assert it1 <= (3 if precision == rf.PREC_FP64 else 20), it1
    assert it1 * 10 <= it0, (it0, it1)
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)


def test_auto_preconditioner_switches_to_lu_on_hard_systems():
    """RF_PRECOND_AUTO: the neighbour preconditioner within the iteration
    budget, then GMRES on A (LU)^-1 from the iterate reached; the result is
    the reference GMRES solution of the same operator."""
    pos, ori = _hard_state()
    res = {}
    for pc, extra in ((rf.PRECOND_NONE, {}), (rf.PRECOND_AUTO, dict(precond_switch_iters=5)),
                      (rf.PRECOND_AUTO, dict(precond_switch_iters=2000))):
        with _sim(pos, ori, precond=pc, **extra) as sim:
            sim.assemble_system()
            sim.solve_system()
            st = sim.stats()
            res[(pc, extra.get("precond_switch_iters"))] = (sim.coefficients(), st.last_gmres_iters,
                                                            st.last_precond, st.last_backward_error,
                                                            st.lu_factorizations)
    x0 = res[(rf.PRECOND_NONE, None)][0]
    xs, its, kind, be, nlu = res[(rf.PRECOND_AUTO, 5)]
    assert kind == rf.PRECOND_LU and nlu == 1 and be <= 1e-10 and its <= 5 + 10
    assert np.linalg.norm(xs - x0) <= 1e-8 * np.linalg.norm(x0)
    xl, its, kind, be, nlu = res[(rf.PRECOND_AUTO, 2000)]
    assert kind == rf.PRECOND_NEIGHBOUR and nlu == 0 and be <= 1e-10
    assert np.linalg.norm(xl - x0) <= 1e-8 * np.linalg.norm(x0)


def test_auto_preconditioner_starts_on_the_lu_after_a_switch():
    """After a solve that ended on the LU, RF_PRECOND_AUTO starts the next
    ones on the LU (one factorisation per step, no neighbour cycle) and
    every 8th goes back to the neighbour preconditioner first."""
    pos, ori = _hard_state()
    nbr, kinds, nlu = [], [], []
    with _sim(pos, ori, precond=rf.PRECOND_AUTO, precond_switch_iters=5, lu_factor_fp32=1) as sim:
        for _ in range(10):
            sim.step()
            st = sim.stats()
            nbr.append(st.last_neighbour_iters)
            kinds.append(st.last_precond)
            nlu.append(st.lu_factorizations)
            assert st.last_backward_error <= 1e-10
    assert all(k == rf.PRECOND_LU for k in kinds), kinds
    assert nlu == list(range(1, 11))
    # steps 2..8 skip the neighbour-preconditioned iterations; step 9 retries them
    assert nbr[0] == 5 and nbr[1:8] == [0] * 7 and nbr[8] == 5 and nbr[9] == 0, nbr


def test_gmres_long_restart_matches_direct():
    """GMRES(40): the fused update-and-dots kernel of the Gram-Schmidt steps
    serves k <= 32 basis vectors and the two-kernel path the rest; both reach
    the direct solver's solution."""
    pos, ori = _hard_state()
    with _sim(pos, ori, use_direct_solver=1) as sim:
        sim.assemble_system()
        sim.solve_system()
        x64 = sim.coefficients()
    with _sim(pos, ori, gmres_restart=40, gmres_tol=1e-11, warm_start=rf.WARM_ZERO) as sim:
        sim.assemble_system()
        sim.solve_system()
        xg, st = sim.coefficients(), sim.stats()
    print("GMRES(40) iterations", st.last_gmres_iters)
    assert st.last_backward_error <= 1e-11 and st.last_gmres_iters > 40
    assert np.linalg.norm(xg - x64) <= 1e-8 * np.linalg.norm(x64)


@pytest.mark.parametrize("nb,fp16", [(0, 0), (100, 0), (24, 0), (0, 1), (100, 1), (13, 1)])
def test_fp32_factors_of_fp64_matrix(nb, fp16):
    """lu_factor_fp32: FP32 LU of the FP64 matrix (panel width nb, 0 = 128)
    as the direct solver (with FP64 refinement) and as the GMRES
    preconditioner, where lu_update_fp16 rounds the trailing-update operands
    to fp16 (the direct solver keeps FP32 updates: its refinement needs
    them on this ill-conditioned near-contact matrix)."""
    pos, ori = _hard_state()
    with _sim(pos, ori, use_direct_solver=1) as sim:
        sim.assemble_system()
        sim.solve_system()
        x64 = sim.coefficients()
    kw = dict(lu_factor_fp32=1, lu_block=nb, lu_update_fp16=fp16)
    with _sim(pos, ori, use_direct_solver=1, gmres_tol=1e-12, **kw) as sim:
        sim.assemble_system()
        sim.solve_system()
        xr, st = sim.coefficients(), sim.stats()
    assert st.last_backward_error <= 1e-12
    assert np.linalg.norm(xr - x64) <= 1e-9 * np.linalg.norm(x64)
    with _sim(pos, ori, precond=rf.PRECOND_LU, gmres_tol=1e-12, **kw) as sim:
        sim.assemble_system()
        sim.solve_system()
        xg, st = sim.coefficients(), sim.stats()
    print("nb", nb, "fp16", fp16, "LU-preconditioned GMRES iterations", st.last_gmres_iters)
    # FP32 factors of this near-contact matrix leave ||I - A (LU)^-1|| well
    # below 1 but not tiny: a few restart cycles instead of ~1000 iterations
    assert st.last_backward_error <= 1e-12 and st.last_gmres_iters <= (60 if not fp16 else 120)
    assert np.linalg.norm(xg - x64) <= 1e-9 * np.linalg.norm(x64)


def test_refine_tol_applies_once_the_lu_preconditioner_is_in_use():
    """refine_tol tightens the stopping test only for the LU-preconditioned
    iterations (the neighbour-preconditioned ones keep gmres_tol)."""
    pos, ori = _hard_state()
    with _sim(pos, ori, precond=rf.PRECOND_AUTO, precond_switch_iters=5, lu_factor_fp32=1,
              gmres_tol=1e-10, refine_tol=1e-13) as sim:
        sim.assemble_system()
        sim.solve_system()
        st = sim.stats()
    assert st.last_precond == rf.PRECOND_LU and st.last_backward_error <= 1e-13
    with _sim(pos, ori, precond=rf.PRECOND_NEIGHBOUR, gmres_tol=1e-10, refine_tol=1e-13) as sim:
        sim.assemble_system()
        sim.solve_system()
        st = sim.stats()
    assert st.last_precond == rf.PRECOND_NEIGHBOUR and 1e-13 < st.last_backward_error <= 1e-10
