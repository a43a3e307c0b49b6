"""The learned initial guess on the GPU, through the C ABI, against the numpy
oracle (oracle/mlp.py): features, targets, forward loss and gradients, one
Adam step, prediction; then the A/B test of the trained network against the
previous-step warm start the reference uses (fortran/advect_fibers.f90:464).
"""
import numpy as np
import pytest

import rigidfibers_kth_amd as rf
from bench import layout
from oracle import mlp

pytestmark = pytest.mark.gpu


def _sim(pos, ori, **kw):
    base = dict(gmres_restart=30, gmres_max_iters=3000, gmres_tol=1e-10, precond=1)
    base.update(kw)
    return rf.Simulation(rf.Configuration(rf.default_params(**base), pos, ori))


def _random_blob(seed):
    rng = np.random.default_rng(seed)
    w = mlp.init_weights(seed)
    L = mlp.layout()
    w[L["in_mean"]:L["in_mean"] + 113] = rng.normal(size=113) * 0.1
    w[L["in_std"]:L["in_std"] + 113] = rng.uniform(0.5, 2.0, 113)
    w[L["out_mean"]:L["out_mean"] + 15] = rng.normal(size=15)
    w[L["out_std"]:L["out_std"] + 15] = rng.uniform(0.5, 2.0, 15)
    for l in range(4):
        w[L["b%d" % l]:L["b%d" % l] + mlp.DIMS[l + 1]] = rng.normal(size=mlp.DIMS[l + 1]) * 0.1
    return w


def test_weight_count_matches_oracle_layout():
    assert rf.mlp_num_weights() == mlp.layout()["total"]


def test_features_targets_and_prediction_vs_oracle():
    N = 70
    pos, ori = layout(N, 2e-2, 71)
    with _sim(pos, ori) as sim:
        sim.assemble_system()
        sim.solve_system()
        X, Y = sim.mlp_samples()
        nbr, dist = sim.neighbours(16)
        x = sim.coefficients()
        P, T = sim.state()
        w = _random_blob(3).astype(np.float32)
        sim.mlp_set_weights(w)
        np.testing.assert_array_equal(sim.mlp_weights(), w)
        xp = sim.mlp_predict()
    Xo = mlp.features(P, T, nbr, dist)
    np.testing.assert_allclose(X, Xo, rtol=0, atol=2e-6 * np.abs(Xo).max())
    Yo = mlp.to_local(T, x)
    np.testing.assert_allclose(Y, Yo, rtol=0, atol=2e-6 * np.abs(Yo).max())
    xo = mlp.to_global(T, mlp.predict(w.astype(np.float64), Xo))
    assert np.abs(xp - xo).max() <= 2e-5 * np.abs(xo).max()


@pytest.mark.parametrize("B", [1, 37, 300])
def test_forward_backward_vs_oracle(B):
    rng = np.random.default_rng(B)
    pos, ori = layout(20, 2e-2, 5)
    X = rng.normal(size=(B, 113)).astype(np.float32)
    Y = rng.normal(size=(B, 15)).astype(np.float32)
    w = _random_blob(11).astype(np.float32)
    with _sim(pos, ori) as sim:
        sim.mlp_set_weights(w)
        loss, g = sim.mlp_forward_backward(X, Y)
    lo, go = mlp.backward(w.astype(np.float64), X.astype(np.float64), Y.astype(np.float64))
    assert loss == pytest.approx(lo, rel=2e-5)
    L = mlp.layout()
    for l in range(4):
        for key in ("W%d" % l, "b%d" % l):
            n = mlp.DIMS[l + 1] * (mlp.DIMS[l] if key[0] == "W" else 1)
            a, b = g[L[key]:L[key] + n], go[L[key]:L[key] + n]
            assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max() + 1e-9, key
    assert not g[:L["first_param"]].any() and not g[L["out_mean"]:].any()


def test_one_adam_step_vs_oracle():
    """One epoch with the whole set as one batch = one Adam step at lr."""
    rng = np.random.default_rng(8)
    B = 128
    X = rng.normal(size=(B, 113)).astype(np.float32)
    Y = rng.normal(size=(B, 15)).astype(np.float32)
    w = _random_blob(13).astype(np.float32)
    pos, ori = layout(20, 2e-2, 5)
    with _sim(pos, ori) as sim:
        sim.mlp_set_weights(w)
        _, g = sim.mlp_forward_backward(X, Y)
        hist = sim.mlp_train(X, Y, epochs=1, batch=B, lr=1e-3, init=False)
        w1 = sim.mlp_weights()
    L = mlp.layout()
    s = slice(L["first_param"], L["first_param"] + L["num_params"])
    step = -1e-3 * g[s] / (np.abs(g[s]) + 1e-8)
    np.testing.assert_allclose(w1[s] - w[s], step, rtol=0, atol=2e-6)
    assert hist[0] == pytest.approx(mlp.loss(w.astype(np.float64), X, Y), rel=1e-4)


def _collect(seeds, N, phi, steps, **kw):
    Xs, Ys = [], []
    for sd in seeds:
        pos, ori = layout(N, phi, sd)
        with _sim(pos, ori, **kw) as sim:
            for _ in range(steps):
                sim.assemble_system()
                sim.solve_system()
                X, Y = sim.mlp_samples()
                Xs.append(X)
                Ys.append(Y)
                sim.update_velocities()
                sim.update_fibers()
    return np.concatenate(Xs), np.concatenate(Ys)


def test_learned_guess_beats_previous_step_guess():
    """Train on plain-solver pairs from 8 trajectories, then on a held-out
    trajectory compare the initial GMRES residual of the reference's guess
    (previous solution) with the learned guess, state by state; iterations
    of both are reported."""
    N, phi = 300, 2e-2
    X, Y = _collect(range(900, 908), N, phi, 6)
    pos, ori = layout(N, phi, 999)
    prev_res, learn_res, it_prev, it_learn = [], [], [], []
    with _sim(pos, ori, warm_start=rf.WARM_LEARNED) as sim:
        hist = sim.mlp_train(X, Y, epochs=40, batch=256, lr=1e-3, seed=1)
        assert hist[-1] < 0.8 * hist[0]
        w = sim.mlp_weights()
        for step in range(6):
            sim.assemble_system()
            x_prev, b = sim.coefficients(), sim.rhs()
            r_prev = np.linalg.norm(b - sim.matvec(x_prev)) / np.linalg.norm(b)
            P, T = sim.state()
            sim.solve_system()
            st = sim.stats()
            if step > 0:
                prev_res.append(r_prev)
                learn_res.append(st.last_initial_residual)
                it_learn.append(st.last_gmres_iters)
                with _sim(P, T) as twin:  # same system, reference guess
                    twin.assemble_system()
                    twin.set_coefficients(x_prev)
                    twin.solve_system()
                    it_prev.append(twin.stats().last_gmres_iters)
            sim.update_velocities()
            sim.update_fibers()
    prev_res, learn_res = np.array(prev_res), np.array(learn_res)
    print("initial residual prev", prev_res, "learned", learn_res, "iterations prev", it_prev, "learned", it_learn)
    # the minimum-residual combination can only improve on x_prev ...
    assert np.all(learn_res <= prev_res * (1 + 1e-9))
    # ... and the trained network makes it improve on average
    assert learn_res.mean() <= 0.95 * prev_res.mean()
    assert np.all(np.isfinite(w))


def test_learned_guess_at_dense_volume_fraction():
    """Volume fraction 5e-2, the density of the N=4000 benchmark: the network
    trained on 20k pairs solved by the plain FP64 method (10 seeded layouts
    of 1000 fibers, two states each) predicts an unseen layout with a
    held-out loss well below 1 (1 = predicting the mean).  In the mixed path
    (FP32 matrix, flexible GMRES on the FP64 operator) its guess starts from
    a lower FP64 backward error than the previous step's coefficients, which
    at this density start worse than zero (fibers move ~0.1 per step against
    gaps of ~0.04).  Inner iterations are reported: the LU-preconditioned
    inner solves take a few iterations from either start."""
    N, phi = 1000, 5e-2
    X, Y = _collect(range(0x5EED0200, 0x5EED0200 + 10), N, phi, 2)
    assert X.shape[0] >= 20000
    pos, ori = layout(N, phi, 0x5EED02FF)
    mixed = dict(precision=rf.PREC_MIXED, precond=rf.PRECOND_AUTO, lu_update_fp16=1)
    with _sim(pos, ori) as sim:
        sim.assemble_system()
        sim.solve_system()
        Xh, Yh = sim.mlp_samples()
        hist = sim.mlp_train(X, Y, epochs=40, batch=256, lr=1e-3, seed=1)
        held_out, _ = sim.mlp_forward_backward(Xh, Yh)
        w = sim.mlp_weights()
    res = {}
    for ws in (rf.WARM_PREV, rf.WARM_LEARNED):
        with _sim(pos, ori, warm_start=ws, **mixed) as sim:
            if ws == rf.WARM_LEARNED:
                sim.mlp_set_weights(w)
            sim.step()
            be, it = [], []
            for _ in range(3):
                sim.assemble_system()
                sim.solve_system()
                st = sim.stats()
                be.append(st.refine_history[0])
                it.append(st.last_gmres_iters)
                assert st.last_backward_error <= 1e-10 or st.last_noise_limited
                sim.update_velocities()
                sim.update_fibers()
            res[ws] = (np.array(be), it)
    print("training loss %.3f -> %.3f, held-out %.3f; initial FP64 backward error previous-step %s learned %s; "
          "inner iterations previous-step %s learned %s" % (hist[0], hist[-1], held_out, res[rf.WARM_PREV][0],
                                                           res[rf.WARM_LEARNED][0], res[rf.WARM_PREV][1],
                                                           res[rf.WARM_LEARNED][1]))
    # 0.675 measured (a box of 1000 fibers has many near its faces); the
    # guess: 0.50-0.66 against 0.66-0.95, inner iterations equal
    assert held_out < 0.8
    assert res[rf.WARM_LEARNED][0].mean() <= 0.85 * res[rf.WARM_PREV][0].mean()
