#!/usr/bin/env python3
"""Throughput of the rigid-fiber slender-body timestep on one MI355X.

Workload (the plain FP64 GPU configuration): N=1000 rigid fibers of length 2,
slenderness 1e-2, centres uniform in a cube sized for volume fraction 1e-2,
orientations uniform on the sphere (seeded), gravity (0,0,-1); M=5 Legendre
terms, 8x3 Gauss points, analytic inner integrals, FP64, GMRES(10) to a
backward error of 1e-10 warm-started from the previous step, dt=0.1.
A "step" is one full timestep (assemble, solve, velocities, advance).

Prints one JSON line.  `value` is steps/s of K timed steps on one GPU with all
state resident in HBM.  `roofline` describes the dominant kernel (the
system-matrix GEMV of the GMRES iterations), timed with HIP events on the
library's stream inside the timed region.  `cpu_baseline` times the FP64 CPU
oracle (a C restatement of the reference's Fortran) on the host cores on one
step of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "steps/sec at N=1000/4000 fibers: plain FP64 vs mixed+learned at ≤1e-6 rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gemv_f64_n15000_traffic.json")


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
            if i == 0 or gap <= 0 or segment_distances(p, t, pos[:i], ori[:i], 0.5 * length).min() >= gap:
                break
        pos[i] = p
        ori[i] = t
    return pos.ravel(), ori.ravel()


def analytic_assembly_flops(pos, ori, M=5, noqi=8):
    """Algorithmic FP64 flops of one analytic assembly: per ordered pair
    sum_p (816 + R_p) + LQ (12 M^2 + 6 M) + 45 M^2 with R_p = 20(M+1) when the
    point p has c = |x_i + s_p t_i - x_j|^2 < 10 (forward recurrences) and
    26*22 otherwise (backward recurrences)."""
    x0, _ = np.polynomial.legendre.leggauss(3)
    edges = np.linspace(-1.0, 1.0, noqi + 1)
    pv = np.concatenate([(a + b) / 2 + (b - a) / 2 * x0 for a, b in zip(edges[:-1], edges[1:])])
    P = pos.reshape(-1, 3)
    T = ori.reshape(-1, 3)
    N = P.shape[0]
    LQ = pv.size
    near = 0
    for i in range(N):
        X = P[i][None, :] + pv[:, None] * T[i][None, :]          # LQ x 3
        d = X[None, :, :] - P[:, None, :]                         # N x LQ x 3
        c = np.einsum("npk,npk->np", d, d)
        c[i, :] = np.inf
        near += int(np.count_nonzero(c < 10.0))
    total_pts = N * (N - 1) * LQ
    far = total_pts - near
    flops = near * (816 + 20 * (M + 1)) + far * (816 + 26 * 22)
    flops += N * (N - 1) * (LQ * (12 * M * M + 6 * M) + 45 * M * M)
    return float(flops), near / total_pts


def run_gpu(N, phi, seed, steps, warmup, **kw):
    import rigidfibers_kth_amd as rf

    pos, ori = layout(N, phi, seed)
    p = rf.default_params(gmres_restart=10, gmres_max_iters=2000, gmres_tol=1e-10, profile=1, **kw)
    sim = rf.Simulation(rf.Configuration(p, pos, ori))
    for _ in range(warmup):
        sim.step()
    sim.synchronize()
    s0 = sim.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    sim.synchronize()
    t1 = time.perf_counter()
    s1 = sim.stats()
    wall = t1 - t0
    d = lambda f: getattr(s1, f) - getattr(s0, f)  # noqa: E731
    out = {
        "steps_per_sec": steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "phases_ms_per_step": {
            "assemble": d("total_ms_assemble") / steps,
            "solve": d("total_ms_solve") / steps,
            "velocities": d("total_ms_velocities") / steps,
            "advance": d("total_ms_advance") / steps,
        },
        "gmres_iters_per_step": d("total_gmres_iters") / steps,
        "last_backward_error": s1.last_backward_error,
        "gemv_launches": d("gemv_launches"),
        "gemv_avg_ms": d("gemv_ms_total") / max(1, d("gemv_launches")),
        "gemv_total_ms": d("gemv_ms_total"),
        "assemble_avg_ms": d("assemble_ms_total") / max(1, d("assemble_launches")),
        "velocity_gemv_avg_ms": d("velocity_gemv_ms_total") / max(1, d("velocity_gemv_launches")),
        "lu_avg_ms": d("lu_ms_total") / max(1, d("lu_factorizations")),
        "n": sim.n,
        "lda": s1.lda,
        "elem_bytes": s1.matrix_elem_bytes,
    }
    sim.close()
    return out, pos, ori


def cpu_baseline(N, phi, seed, threads):
    import oracle

    os.environ["OMP_NUM_THREADS"] = str(threads)
    pos, ori = layout(N, phi, seed)
    t0 = time.perf_counter()
    oracle.simulate(True, False, 0.01, 0.1, 1, pos, ori, M=5, noqi=8, restart=10, max_iters=2000, tol=1e-10)
    dt = time.perf_counter() - t0
    return {
        "value": 1.0 / dt,
        "unit": "steps/s",
        "cores": int(oracle.num_threads()),
        "kind": "port",
        "sample": "1 timestep of the same N=1000 workload (analytic assembly, GMRES(10) to 1e-10 from x0=0, "
                  "velocities, advance) with the FP64 C oracle, OpenMP over target fibers; %.1f s" % dt,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-lu", action="store_true", help="skip the direct-solver timing")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the multi-GPU row sharding is a design only "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    N, phi, seed = 1000, 1e-2, 0x5EED0001
    g, pos, ori = run_gpu(N, phi, seed, args.steps, args.warmup)
    n, lda, eb = g["n"], g["lda"], g["elem_bytes"]
    gemv_bytes = float(n) * n * eb + 2.0 * n * 8  # matrix once + x and y
    achieved = gemv_bytes / (g["gemv_avg_ms"] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            t = json.load(open(TRAFFIC_FILE))
            if t.get("n") == n and t.get("elem_bytes") == eb:
                traffic = t.get("bytes_per_launch")
        except Exception:
            traffic = None
    flops, near_frac = analytic_assembly_flops(pos, ori)
    asm_tflops = flops / (g["assemble_avg_ms"] * 1e-3) / 1e12
    extra = {}
    if not args.no_lu:
        lu, _, _ = run_gpu(N, phi, seed, 2, 1, use_direct_solver=1)
        extra["lu_steps_per_sec"] = lu["steps_per_sec"]
        extra["lu_factor_ms"] = lu["lu_avg_ms"]
        extra["lu_factor_tflops"] = (2.0 / 3.0) * n ** 3 / (lu["lu_avg_ms"] * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": g["steps_per_sec"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded fiber layout (seed 0x5EED0001)",
        "config": {
            "workload": "plain FP64, N=1000 fibers, volume fraction 1e-2, M=5, 8x3 Gauss points, analytic, "
                        "GMRES(10) tol 1e-10, dt=0.1",
            "fibers": N, "unknowns": n, "volume_fraction": phi, "solver": "gmres", "tol": 1e-10,
        },
        "roofline": {
            "kernel": "gemv_parts_kernel<double> (GMRES matvec, n=%d)" % n,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "phases_ms_per_step": g["phases_ms_per_step"],
        "gmres_iters_per_step": g["gmres_iters_per_step"],
        "gemv_avg_ms": g["gemv_avg_ms"],
        "gemv_share_of_step": g["gemv_total_ms"] / (g["ms_per_step"] * args.steps),
        "assembly": {
            "avg_ms": g["assemble_avg_ms"],
            "tflops": asm_tflops,
            "frac_fp64_peak": asm_tflops / FP64_PEAK_TFLOPS,
            "near_point_fraction": near_frac,
        },
        "velocity_gemv_avg_ms": g["velocity_gemv_avg_ms"],
    }
    line.update(extra)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(N, phi, seed, min(16, os.cpu_count() or 1))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
