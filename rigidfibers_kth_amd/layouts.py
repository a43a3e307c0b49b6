"""Seeded synthetic fiber layouts (no counterpart in the reference, whose
layouts come from files: src/parameters.cc:186-227).

Pure numpy; importing this module does not load the HIP library.
"""
import numpy as np


def segment_distances(c, t, C, T, half=1.0):
    """Minimum distance between the segment c + s t (|s| <= half) and each
    segment C[k] + s T[k] (closest points with clamping to both segments)."""
    w0 = c[None, :] - C
    a = np.dot(t, t)
    b = T @ t
    cc = np.einsum("ij,ij->i", T, T)
    d = w0 @ t
    e = np.einsum("ij,ij->i", T, w0)
    den = a * cc - b * b
    s = np.where(den > 1e-12, (b * e - cc * d) / np.where(den > 1e-12, den, 1.0), 0.0)
    s = np.clip(s, -half, half)
    u = (b * s + e) / cc
    u = np.clip(u, -half, half)
    s = np.clip((b * u - d) / a, -half, half)
    diff = w0 + s[:, None] * t[None, :] - u[:, None] * T
    return np.sqrt(np.einsum("ij,ij->i", diff, diff))


def layout(N, phi, seed, length=2.0, eps=0.01, min_gap=None):
    """Seeded generator: centres uniform in [0,L)^3 with phi = N pi a^2 l / L^3,
    a = eps*l; orientations uniform on S^2 (z uniform in [-1,1], azimuth
    uniform).  Overlap policy: a fiber is redrawn while its segment-segment
    distance to an already placed fiber is below min_gap (default 2a, i.e.
    the fibers' cylinders would intersect)."""
    a = eps * length
    L = (N * np.pi * a * a * length / phi) ** (1.0 / 3.0)
    gap = 2.0 * a if min_gap is None else min_gap
    rng = np.random.default_rng(seed)
    pos = np.empty((N, 3))
    ori = np.empty((N, 3))
    for i in range(N):
        while True:
            p = rng.uniform(0.0, L, size=3)
            z = rng.uniform(-1.0, 1.0)
            ph = rng.uniform(0.0, 2 * np.pi)
            s = np.sqrt(1 - z * z)
            t = np.array([s * np.cos(ph), s * np.sin(ph), z])
            if i == 0 or gap <= 0:
                break
            # only fibers whose centres are closer than length + gap can come
            # within gap (the margin keeps every decision that of the full check)
            d = pos[:i] - p
            near = np.einsum("ij,ij->i", d, d) < (length + gap + 1e-9) ** 2
            if not near.any() or segment_distances(p, t, pos[:i][near], ori[:i][near], 0.5 * length).min() >= gap:
                break
        pos[i] = p
        ori[i] = t
    return pos.ravel(), ori.ravel()
