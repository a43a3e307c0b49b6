"""FP64 numpy restatement of the learned initial guess (features, MLP, training).

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in
``rigidfibers_kth_amd/csrc/rf_mlp.hip``; only ``tests/`` import it.

The reference has no learned component, so nothing here follows a reference
file.  What pins it: ``backward`` is checked against central finite
differences of ``loss`` (tests/test_mlp_oracle.py), the local frame against
orthonormality and the coefficient round trip, and the features against
rotation invariance about gravity.

Conventions (shared with the kernels):

* local frame of a fiber with unit orientation t: e1 = t, e2 = z x t / |z x t|
  (x x t / |x x t| when |z x t| < 1e-6), e3 = e1 x e2;
* features (113): for each of the 16 nearest fibers j (segment-segment
  distance, ascending, ties by index) the half relative centre
  (x_j - x_i)/2 and the sign-canonicalised orientation +-t_j (first nonzero
  of its local components made positive), both in the local frame of i, and
  the proximity 0.1/(d_ij + 0.1); then t_i . z;
* targets (15): the force coefficients a_i^k, k=1..5, in the local frame;
* network: 113 -> 128 -> 128 -> 128 -> 15, tanh on hidden layers, inputs and
  outputs standardised with stored means and deviations; weight blob layout
  in_mean, in_std, (W1, b1), ..., (W4, b4), out_mean, out_std with W[out][in].
"""
import numpy as np

NEIGHBOURS = 16
FEATURES_PER_NEIGHBOUR = 7
N_IN = NEIGHBOURS * FEATURES_PER_NEIGHBOUR + 1
HIDDEN = 128
N_OUT = 15
DIMS = (N_IN, HIDDEN, HIDDEN, HIDDEN, N_OUT)


def layout():
    """Offsets into the flat weight blob."""
    o = 0
    L = {"in_mean": o}
    o += N_IN
    L["in_std"] = o
    o += N_IN
    L["first_param"] = o
    for l in range(4):
        L["W%d" % l] = o
        o += DIMS[l + 1] * DIMS[l]
        L["b%d" % l] = o
        o += DIMS[l + 1]
    L["num_params"] = o - L["first_param"]
    L["out_mean"] = o
    o += N_OUT
    L["out_std"] = o
    o += N_OUT
    L["total"] = o
    return L


def unpack(w):
    L = layout()
    W = [w[L["W%d" % l]:L["W%d" % l] + DIMS[l + 1] * DIMS[l]].reshape(DIMS[l + 1], DIMS[l]) for l in range(4)]
    b = [w[L["b%d" % l]:L["b%d" % l] + DIMS[l + 1]] for l in range(4)]
    norm = {k: w[L[k]:L[k] + (N_IN if k.startswith("in") else N_OUT)]
            for k in ("in_mean", "in_std", "out_mean", "out_std")}
    return W, b, norm


def init_weights(seed=0):
    """Glorot-uniform weights, zero biases, identity standardisation."""
    L = layout()
    rng = np.random.default_rng(seed)
    w = np.zeros(L["total"])
    w[L["in_std"]:L["in_std"] + N_IN] = 1.0
    w[L["out_std"]:L["out_std"] + N_OUT] = 1.0
    for l in range(4):
        lim = np.sqrt(6.0 / (DIMS[l] + DIMS[l + 1]))
        w[L["W%d" % l]:L["W%d" % l] + DIMS[l + 1] * DIMS[l]] = rng.uniform(-lim, lim, DIMS[l + 1] * DIMS[l])
    return w


def local_frames(orient):
    """(N,3,3) array, rows e1, e2, e3."""
    T = np.asarray(orient, dtype=np.float64).reshape(-1, 3)
    e1 = T.copy()
    a = np.stack([-T[:, 1], T[:, 0], np.zeros(len(T))], axis=1)
    nrm = np.hypot(a[:, 0], a[:, 1])
    alt = nrm < 1e-6
    a[alt] = np.stack([np.zeros(alt.sum()), -T[alt, 2], T[alt, 1]], axis=1)
    nrm = np.linalg.norm(a, axis=1)
    e2 = a / nrm[:, None]
    e3 = np.cross(e1, e2)
    return np.stack([e1, e2, e3], axis=1)


def features(pos, orient, nbr, dist):
    """(N, 113) features from the 16-nearest lists (nbr, dist: (N, >=16))."""
    P = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    T = np.asarray(orient, dtype=np.float64).reshape(-1, 3)
    E = local_frames(T)
    N = len(P)
    X = np.empty((N, N_IN))
    for i in range(N):
        j = nbr[i, :NEIGHBOURS]
        rel = 0.5 * (P[j] - P[i]) @ E[i].T
        u = T[j] @ E[i].T
        sg = np.where((u[:, 0] > 0) | ((u[:, 0] == 0) & ((u[:, 1] > 0) | ((u[:, 1] == 0) & (u[:, 2] >= 0)))),
                      1.0, -1.0)
        f = np.concatenate([rel, sg[:, None] * u, (0.1 / (dist[i, :NEIGHBOURS] + 0.1))[:, None]], axis=1)
        X[i, :NEIGHBOURS * FEATURES_PER_NEIGHBOUR] = f.ravel()
        X[i, -1] = E[i, 0, 2]
    return X


def to_local(orient, x, M=5):
    """Global fiber-major coefficients (3MN) -> (N, 3M) in the local frames."""
    E = local_frames(orient)
    A = np.asarray(x, dtype=np.float64).reshape(len(E), M, 3)
    return np.einsum("nab,nkb->nka", E, A).reshape(len(E), 3 * M)


def to_global(orient, y, M=5):
    E = local_frames(orient)
    Y = np.asarray(y, dtype=np.float64).reshape(len(E), M, 3)
    return np.einsum("nab,nka->nkb", E, Y).reshape(-1)


def forward(w, X):
    """Returns (yhat standardised, cache)."""
    W, b, nm = unpack(w)
    Z = (X - nm["in_mean"]) / nm["in_std"]
    H = []
    h = Z
    for l in range(3):
        h = np.tanh(h @ W[l].T + b[l])
        H.append(h)
    yhat = h @ W[3].T + b[3]
    return yhat, (Z, H)


def predict(w, X):
    """Coefficients in the local frame (de-standardised)."""
    _, _, nm = unpack(w)
    yhat, _ = forward(w, X)
    return yhat * nm["out_std"] + nm["out_mean"]


def loss(w, X, Y):
    """Mean over all B*15 entries of (yhat - (Y - out_mean)/out_std)^2."""
    _, _, nm = unpack(w)
    yhat, _ = forward(w, X)
    r = yhat - (Y - nm["out_mean"]) / nm["out_std"]
    return float(np.mean(r * r))


def backward(w, X, Y):
    """(loss, gradient blob) with zeros in the standardisation entries."""
    L = layout()
    W, b, nm = unpack(w)
    yhat, (Z, H) = forward(w, X)
    r = yhat - (Y - nm["out_mean"]) / nm["out_std"]
    B = X.shape[0]
    delta = 2.0 * r / (B * N_OUT)
    g = np.zeros_like(w)
    ins = [Z] + H
    for l in range(3, -1, -1):
        g[L["W%d" % l]:L["W%d" % l] + DIMS[l + 1] * DIMS[l]] = (delta.T @ ins[l]).ravel()
        g[L["b%d" % l]:L["b%d" % l] + DIMS[l + 1]] = delta.sum(axis=0)
        if l > 0:
            delta = (delta @ W[l]) * (1.0 - H[l - 1] ** 2)
    return float(np.mean(r * r)), g


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update of the parameter range (in place)."""
    L = layout()
    s = slice(L["first_param"], L["first_param"] + L["num_params"])
    m[s] = b1 * m[s] + (1 - b1) * g[s]
    v[s] = b2 * v[s] + (1 - b2) * g[s] ** 2
    mh = m[s] / (1 - b1 ** step)
    vh = v[s] / (1 - b2 ** step)
    w[s] -= lr * mh / (np.sqrt(vh) + eps)


def set_standardisation(w, X, Y):
    L = layout()
    w[L["in_mean"]:L["in_mean"] + N_IN] = X.mean(axis=0)
    w[L["in_std"]:L["in_std"] + N_IN] = np.maximum(X.std(axis=0), 1e-6)
    w[L["out_mean"]:L["out_mean"] + N_OUT] = Y.mean(axis=0)
    w[L["out_std"]:L["out_std"] + N_OUT] = np.maximum(Y.std(axis=0), 1e-12)
