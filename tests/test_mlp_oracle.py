"""The CPU restatement of the learned initial guess (oracle/mlp.py), pinned by
checks that do not restate its formulas: finite differences for the gradient,
orthonormality and round trips for the local frames, rotation and sign
invariances for the features, and the closed form of a first Adam step."""
import numpy as np
import pytest

from bench import layout, segment_distances
from oracle import mlp


def _knn(pos, ori, K=16):
    P, T = pos.reshape(-1, 3), ori.reshape(-1, 3)
    N = len(P)
    nbr = np.empty((N, K), dtype=np.int64)
    dist = np.empty((N, K))
    for i in range(N):
        d = segment_distances(P[i], T[i], P, T)
        d[i] = np.inf
        o = np.lexsort((np.arange(N), d))[:K]
        nbr[i], dist[i] = o, d[o]
    return nbr, dist


def test_layout_sizes():
    L = mlp.layout()
    assert L["num_params"] == 113 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 15 + 15
    assert L["total"] == L["num_params"] + 2 * 113 + 2 * 15


def test_local_frames_orthonormal_and_right_handed():
    rng = np.random.default_rng(3)
    T = rng.normal(size=(200, 3))
    T[:3] = [[0, 0, 1], [0, 0, -1], [1e-9, 0, 1]]  # parallel to gravity: fallback axis
    T /= np.linalg.norm(T, axis=1)[:, None]
    E = mlp.local_frames(T.ravel())
    for e in E:
        np.testing.assert_allclose(e @ e.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(e) == pytest.approx(1.0, abs=1e-12)
    np.testing.assert_allclose(E[:, 0], T, atol=1e-15)
    # e2 is horizontal whenever t is not parallel to z
    assert np.abs(E[3:, 1, 2]).max() < 1e-12


def test_coefficient_round_trip():
    rng = np.random.default_rng(4)
    T = rng.normal(size=(50, 3))
    T /= np.linalg.norm(T, axis=1)[:, None]
    x = rng.normal(size=50 * 15)
    np.testing.assert_allclose(mlp.to_global(T.ravel(), mlp.to_local(T.ravel(), x)), x, atol=1e-13)
    # the component along e1 is the projection on t
    Y = mlp.to_local(T.ravel(), x)
    np.testing.assert_allclose(Y[:, 0], np.einsum("ij,ij->i", x.reshape(50, 5, 3)[:, 0], T), atol=1e-13)


def test_features_invariant_under_rotation_about_gravity_and_sign_flips():
    pos, ori = layout(40, 2e-2, 11)
    nbr, dist = _knn(pos, ori)
    X0 = mlp.features(pos, ori, nbr, dist)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    P2 = (pos.reshape(-1, 3) @ R.T).ravel() + np.tile([3.0, -1.0, 2.0], 40)
    T2 = (ori.reshape(-1, 3) @ R.T).ravel()
    np.testing.assert_allclose(mlp.features(P2, T2, nbr, dist), X0, atol=1e-12)
    # a neighbour's orientation enters only up to sign
    T3 = ori.reshape(-1, 3).copy()
    T3[nbr[0, 0]] *= -1
    X3 = mlp.features(pos, T3.ravel(), nbr, dist)
    np.testing.assert_allclose(X3[0], X0[0], atol=1e-12)
    # proximity is 0.1/(d+0.1), centres are halved
    f = X0[5].reshape(-1)[:112].reshape(16, 7)
    np.testing.assert_allclose(f[:, 6], 0.1 / (dist[5] + 0.1))
    np.testing.assert_allclose(np.linalg.norm(f[:, :3], axis=1),
                               0.5 * np.linalg.norm(pos.reshape(-1, 3)[nbr[5]] - pos.reshape(-1, 3)[5], axis=1))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    w = mlp.init_weights(7)
    L = mlp.layout()
    w[L["in_mean"]:L["in_mean"] + 113] = rng.normal(size=113) * 0.1
    w[L["in_std"]:L["in_std"] + 113] = rng.uniform(0.5, 2.0, 113)
    w[L["out_mean"]:L["out_mean"] + 15] = rng.normal(size=15)
    w[L["out_std"]:L["out_std"] + 15] = rng.uniform(0.5, 2.0, 15)
    for l in range(4):  # nonzero biases so their gradient path is exercised
        w[L["b%d" % l]:L["b%d" % l] + mlp.DIMS[l + 1]] = rng.normal(size=mlp.DIMS[l + 1]) * 0.1
    X = rng.normal(size=(9, 113))
    Y = rng.normal(size=(9, 15))
    loss, g = mlp.backward(w, X, Y)
    assert loss == pytest.approx(mlp.loss(w, X, Y), rel=1e-14)
    picks = []
    for l in range(4):
        picks += list(L["W%d" % l] + rng.choice(mlp.DIMS[l + 1] * mlp.DIMS[l], 6, replace=False))
        picks += list(L["b%d" % l] + rng.choice(mlp.DIMS[l + 1], 3, replace=False))
    h = 1e-6
    for p in picks:
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (mlp.loss(wp, X, Y) - mlp.loss(wm, X, Y)) / (2 * h)
        assert g[p] == pytest.approx(fd, rel=1e-6, abs=1e-10), p
    # standardisation entries are data, not parameters
    assert not g[:L["first_param"]].any() and not g[L["out_mean"]:].any()


def test_first_adam_step_is_signed_learning_rate():
    w = mlp.init_weights(1)
    g = np.zeros_like(w)
    L = mlp.layout()
    s = slice(L["first_param"], L["first_param"] + L["num_params"])
    g[s] = np.random.default_rng(2).normal(size=L["num_params"])
    w1 = w.copy()
    mlp.adam_step(w1, g, np.zeros_like(w), np.zeros_like(w), 1, 1e-3)
    # bias-corrected first moment g, second moment g^2: step = lr g / (|g| + eps)
    np.testing.assert_allclose(w1[s] - w[s], -1e-3 * g[s] / (np.abs(g[s]) + 1e-8), rtol=1e-12)
    np.testing.assert_array_equal(w1[:L["first_param"]], w[:L["first_param"]])
