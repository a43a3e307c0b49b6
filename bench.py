#!/usr/bin/env python3
"""Throughput of the rigid-fiber slender-body timestep on one MI355X.

Headline workload (`value`): N=1000 rigid fibers of length 2, slenderness
1e-2, centres uniform in a cube sized for volume fraction 1e-2, orientations
uniform on the sphere (seeded, overlaps rejected), gravity (0,0,-1); M=5
Legendre terms, 8x3 Gauss points, analytic inner integrals, dt=0.1; solved the
reference's way in FP64: GMRES(10) to a backward error of 1e-10, warm-started
from the previous step, LU fallback after 300 iterations
(fortran/coeffs_and_vels.f90:253-339).
A "step" is one full timestep (assemble, solve, velocities, advance).

Beside it, on the same initial state, the mixed-precision + learned path:
the matrix assembled and stored in FP32 (FP64 near-field integrals), FP32 LU
factors with fp16-operand trailing updates on the matrix cores as the
preconditioner of hard systems, flexible GMRES on the matrix-free FP64
operator (FP64 residuals and Arnoldi) around the FP32 solves, and the learned
initial guess (network trained on the GPU on solved pairs).  Also timed: the
same without the network, an FP64-matrix variant with FP32 LU factors, the
FP64 direct solver, and the same comparison at N=4000, volume fraction 5e-2.

Accuracy: the per-step ("shadow") position error.  The mixed + learned run's
state before every timed step is restored into a handle of the plain FP64
method (and of a 1e-13-tolerance FP64 method) and stepped once; the relative
position difference after that step is the solver's own contribution,
free of the chaotic amplification that near fiber contacts turns any 1e-10
difference into O(1) after many steps.  `meets_accuracy` is its maximum over
the timed steps <= 1e-6.  Position differences after all steps are reported
beside it.

Prints one JSON line.  `roofline` describes the dominant kernel of the
headline workload (the system-matrix GEMV of the GMRES iterations), timed with
HIP events on the library's stream inside the timed region.  `cpu_baseline`
times the FP64 CPU oracle (a C restatement of the reference's Fortran) on the
host cores on warm-started steps of the headline workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from rigidfibers_kth_amd.layouts import layout, segment_distances  # noqa: E402,F401  (re-exported for tests)

METRIC = "steps/sec at N=1000/4000 fibers: plain FP64 vs mixed+learned at ≤1e-6 rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP16_MFMA_PEAK_TFLOPS = 2500.0  # dense
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gemv_f64_n15000_traffic.json")
ACCURACY = 1e-6

C2 = dict(N=1000, phi=1e-2, seed=0x5EED0001)
C4 = dict(N=4000, phi=5e-2, seed=0x5EED0003)
# the reference's method: FP64 GMRES(10), previous-step guess, LU fallback
# (fortran/coeffs_and_vels.f90:331-335) after 300 iterations: about the cost
# of one FP64 LU and solve at N=1000 (112 ms + 5 ms = 300 iterations of
# 0.38 ms); beyond it more GMRES iterations cannot pay off.  Measured at
# N=1000: cap 200 12.75, 300 12.86, 500 10.84, 1000 8.66 steps/s.  At N=4000
# every step needs more and falls back (the direct solver is timed beside it)
PLAIN = dict(gmres_restart=10, gmres_max_iters=300, gmres_tol=1e-10)
PLAIN_DIRECT = dict(use_direct_solver=1)
# mixed precision + learned: FP32 matrix, FP32 LU factors (fp16-operand
# trailing updates) preconditioning the inner FP32 solves of hard systems,
# flexible GMRES on the matrix-free FP64 operator, learned initial guess
MIXED = dict(precision=1, gmres_restart=30, gmres_max_iters=3000, gmres_tol=1e-10, precond=3, lu_update_fp16=1,
             warm_start=3)
MIXED_NO_MLP = dict(MIXED, warm_start=0)
# FP64 matrix with FP32 LU factors as GMRES-IR preconditioner, learned guess
FP64_FP32LU = dict(gmres_restart=30, gmres_max_iters=3000, gmres_tol=1e-10, precond=3, lu_factor_fp32=1,
                   lu_update_fp16=1, warm_start=3)
# FP64 reference for the error figures: every solve to a 1e-13 backward error
TIGHT = dict(gmres_restart=30, gmres_max_iters=3000, gmres_tol=1e-13, refine_tol=1e-13, precond=3,
             lu_factor_fp32=1, lu_update_fp16=1)


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


def _stats_delta(s0, s1, steps):
    d = lambda f: getattr(s1, f) - getattr(s0, f)  # noqa: E731
    return {
        "phases_ms_per_step": {
            "assemble": d("total_ms_assemble") / steps,
            "solve": d("total_ms_solve") / steps,
            "velocities": d("total_ms_velocities") / steps,
            "advance": d("total_ms_advance") / steps,
        },
        "gmres_iters_per_step": d("total_gmres_iters") / steps,
        "gemv_launches": d("gemv_launches"),
        "gemv_avg_ms": d("gemv_ms_total") / max(1, d("gemv_launches")),
        "gemv_total_ms": d("gemv_ms_total"),
        "assemble_avg_ms": d("assemble_ms_total") / max(1, d("assemble_launches")),
        "velocity_gemv_avg_ms": d("velocity_gemv_ms_total") / max(1, d("velocity_gemv_launches")),
        "lu_factorizations": d("lu_factorizations"),
        "lu_avg_ms": d("lu_ms_total") / max(1, d("lu_factorizations")),
        "lu_panels_redone": d("lu_panel_fallbacks"),
        "operator_launches_per_step": d("operator_launches") / steps,
        "operator_avg_ms": d("operator_ms_total") / max(1, d("operator_launches")),
    }


def timed_run(rf, pos, ori, steps, warmup=0, weights=None, **kw):
    """K timed steps after W untimed ones; returns (record, positions after
    all steps)."""
    sim = rf.Simulation(rf.Configuration(rf.default_params(profile=1, **kw), pos, ori))
    if weights is not None:
        sim.mlp_set_weights(weights)
    for _ in range(warmup):
        sim.step()
    sim.synchronize()
    s0 = sim.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    sim.synchronize()
    wall = time.perf_counter() - t0
    s1 = sim.stats()
    out = {"steps_per_sec": steps / wall, "ms_per_step": 1e3 * wall / steps, "steps": steps,
           "last_backward_error": s1.last_backward_error, "n": sim.n, "elem_bytes": s1.matrix_elem_bytes}
    out.update(_stats_delta(s0, s1, steps))
    P, _ = sim.state()
    sim.close()
    return out, P


def shadow(rf, pos, ori, steps, warmup, weights, kw, refs):
    """Per-step position error of the method `kw` against each reference
    method restarted from the same state: the method's run is replayed
    (untimed) and its state before every step recorded; each reference
    handle then restores every recorded state and steps once.  Returns
    {ref: [relative position difference after each step]} and the same
    between the references (the references' own disagreement)."""
    sim = rf.Simulation(rf.Configuration(rf.default_params(**kw), pos, ori))
    if weights is not None:
        sim.mlp_set_weights(weights)
    for _ in range(warmup):
        sim.step()
    states, after = [], []
    for _ in range(steps):
        states.append(sim.restart_state())
        sim.step()
        after.append(sim.state()[0])
    sim.close()
    stepped = {}
    for name, rkw in refs.items():
        h = rf.Simulation(rf.Configuration(rf.default_params(**rkw), pos, ori))
        out = []
        for S in states:
            h.restore_state(S)
            h.step()
            out.append(h.state()[0])
        h.close()
        stepped[name] = out
    names = list(refs)
    res = {"method_vs_" + n: [rel(a, b) for a, b in zip(after, stepped[n])] for n in names}
    for i, a in enumerate(names):
        for b in names[i + 1:]:
            res[a + "_vs_" + b] = [rel(x, y) for x, y in zip(stepped[a], stepped[b])]
    return res


# the plain FP64 solution of the training states (FP64 matrix; FP32 LU factors
# only precondition GMRES, which still solves the FP64 system to 1e-10)
TRAIN_SOLVE = dict(gmres_restart=30, gmres_max_iters=3000, gmres_tol=1e-10, precond=3, lu_factor_fp32=1,
                   lu_update_fp16=1)


def train_network(rf, N, phi, seeds, states, held_pos, held_ori, epochs=40, held_kw=TRAIN_SOLVE, solve_kw=TRAIN_SOLVE):
    """Solved (configuration, coefficients) pairs from `states` steps of each
    seeded layout (N fibers, volume fraction phi), GPU training of the
    network, and its loss on the pairs of a held-out configuration (the
    benchmark's own initial state, never trained on).  Returns (weights,
    info)."""
    Xs, Ys = [], []
    t0 = time.perf_counter()
    for sd in seeds:
        pos, ori = layout(N, phi, sd)
        sim = rf.Simulation(rf.Configuration(rf.default_params(**solve_kw), pos, ori))
        for _ in range(states):
            sim.assemble_system()
            sim.solve_system()
            X, Y = sim.mlp_samples()
            Xs.append(X)
            Ys.append(Y)
            sim.update_velocities()
            sim.update_fibers()
        sim.close()
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    t1 = time.perf_counter()
    sim = rf.Simulation(rf.Configuration(rf.default_params(**held_kw), held_pos, held_ori))
    sim.assemble_system()
    sim.solve_system()
    Xh, Yh = sim.mlp_samples()
    t2 = time.perf_counter()
    hist = sim.mlp_train(X, Y, epochs=epochs, batch=256, lr=1e-3, seed=1)
    sim.synchronize()
    t3 = time.perf_counter()
    held_out, _ = sim.mlp_forward_backward(Xh, Yh)
    w = sim.mlp_weights()
    sim.close()
    return w, {"pairs": int(X.shape[0]), "layouts": len(seeds), "states_per_layout": states,
               "layout_fibers": N, "volume_fraction": phi, "held_out_pairs": int(Xh.shape[0]),
               "epochs": epochs, "sample_s": t1 - t0, "train_s": t3 - t2,
               "loss_first_epoch": float(hist[0]), "loss_last_epoch": float(hist[-1]),
               "held_out_loss": float(held_out),
               "loss_unit": "mean squared error of standardised coefficients (1 = predicting the mean)"}


def guess_quality(rf, pos, ori, steps, weights, kw):
    """Per step of `kw` (mixed path) with the learned guess and with the
    previous-step guess: the FP64 backward error of the initial guess and the
    inner iterations (untimed)."""
    out = {}
    for name, ws in (("learned", 3), ("previous_step", 0)):
        sim = rf.Simulation(rf.Configuration(rf.default_params(**dict(kw, warm_start=ws)), pos, ori))
        if ws == 3:
            sim.mlp_set_weights(weights)
        be, it = [], []
        for _ in range(steps):
            sim.step()
            st = sim.stats()
            be.append(float(st.refine_history[0]))
            it.append(int(st.last_gmres_iters))
        sim.close()
        out[name] = {"initial_backward_error": be, "inner_iterations": it}
    out["mean_ratio_learned_to_previous_step"] = float(np.mean(out["learned"]["initial_backward_error"]) /
                                                       np.mean(out["previous_step"]["initial_backward_error"]))
    return out


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def accuracy_record(sh, final_pos):
    """Shadow maxima and the positions after all steps."""
    mx = {k: max(v) for k, v in sh.items()}
    out = {
        "shadow_max_rel_err": mx["method_vs_plain"],
        "shadow_max_rel_err_vs_tight": mx["method_vs_tight"],
        "shadow_plain_vs_tight_max_rel_err": mx["plain_vs_tight"],
        "shadow_per_step": {k: v for k, v in sh.items()},
        "meets_accuracy": bool(mx["method_vs_plain"] <= ACCURACY),
        "accuracy_bound": ACCURACY,
    }
    Pm, Pp = final_pos
    out["position_rel_err_after_all_steps"] = rel(Pm, Pp)
    return out


def compare(rf, cfg, steps, warmup, train, label, plain_ref):
    """Plain FP64 and mixed + learned on the same initial state: steps/s,
    per-step shadow errors, guess quality and positions after all steps."""
    P0, T0 = layout(cfg["N"], cfg["phi"], cfg["seed"])
    tag = label.split(",")[0]
    w, tinfo = train_network(rf, train["N"], cfg["phi"], train["seeds"], 2, P0, T0)
    log("%s: network trained on %d pairs, held-out loss %.3f" % (tag, tinfo["pairs"], tinfo["held_out_loss"]))
    out = {"workload": label, "training": tinfo}
    plain, Pp = timed_run(rf, P0, T0, steps, warmup, **PLAIN)
    log("%s: plain FP64 %.3f steps/s" % (tag, plain["steps_per_sec"]))
    direct, Pd = timed_run(rf, P0, T0, steps, warmup, **PLAIN_DIRECT)
    log("%s: plain FP64 direct solver %.3f steps/s" % (tag, direct["steps_per_sec"]))
    mixed, Pm = timed_run(rf, P0, T0, steps, warmup, weights=w, **MIXED)
    log("%s: mixed + learned %.3f steps/s" % (tag, mixed["steps_per_sec"]))
    nomlp, _ = timed_run(rf, P0, T0, steps, warmup, **MIXED_NO_MLP)
    log("%s: mixed without the network %.3f steps/s" % (tag, nomlp["steps_per_sec"]))
    f64, _ = timed_run(rf, P0, T0, steps, warmup, weights=w, **FP64_FP32LU)
    log("%s: FP64 matrix + FP32 LU + learned %.3f steps/s" % (tag, f64["steps_per_sec"]))
    ref = PLAIN_DIRECT if plain_ref == "direct" else PLAIN
    sh = shadow(rf, P0, T0, steps, warmup, w, MIXED, {"plain": ref, "tight": TIGHT})
    log("%s: shadow error max %.2e" % (tag, max(sh["method_vs_plain"])))
    out["guess_quality"] = guess_quality(rf, P0, T0, steps, w, MIXED)
    best_plain = max(plain["steps_per_sec"], direct["steps_per_sec"])
    out.update({
        "plain_fp64": plain,
        "plain_fp64_direct": direct,
        "mixed_learned": mixed,
        "mixed_without_network": {k: nomlp[k] for k in ("steps_per_sec", "gmres_iters_per_step",
                                                        "operator_launches_per_step", "lu_factorizations")},
        "fp64_matrix_fp32_lu_learned": f64,
        "speedup": mixed["steps_per_sec"] / plain["steps_per_sec"],
        "speedup_vs_best_plain": mixed["steps_per_sec"] / best_plain,
        "shadow_reference": "plain FP64 %s" % ("direct solver" if plain_ref == "direct" else "GMRES(10), tol 1e-10"),
    })
    out.update(accuracy_record(sh, (Pm, Pd if plain_ref == "direct" else Pp)))
    return out


# the long runs (`--long c3|c4|c5 K`): the reference configurations over
# their full step counts
C3 = dict(N=2000, phi=2e-2, seed=0x5EED0002)
C5 = dict(N=10000, phi=1e-2, seed=0x5EED0004)
# N=10000: FP32 matrix (90 GB) with the neighbour block-Jacobi preconditioner
# (an FP32 LU copy would not fit beside it), flexible GMRES on the FP64
# operator to 1e-8, learned guess, dt=0.05
C5_MIXED = dict(MIXED, gmres_tol=1e-8, precond=1, timestep=0.05)
LONG = {
    "c3": dict(cfg=C3, steps=50,
               label="N=2000, volume fraction 2e-2, M=5, dt=0.1, tol 1e-10, from the seeded layout (seed 0x5EED0002)",
               methods=(("mixed", MIXED_NO_MLP), ("plain_fp64", PLAIN), ("plain_fp64_direct", PLAIN_DIRECT),
                        ("tight_fp64", TIGHT)),
               train=None, shadow_ref="tight_fp64"),
    "c4": dict(cfg=C4, steps=100,
               label="N=4000, volume fraction 5e-2, M=5, dt=0.1, tol 1e-10, from the seeded layout (seed 0x5EED0003)",
               methods=(("mixed_learned", MIXED), ("plain_fp64_direct", PLAIN_DIRECT), ("tight_fp64", TIGHT)),
               train=dict(N=2000, seeds=[C4["seed"] + 0x100 + k for k in range(5)], held=TRAIN_SOLVE),
               shadow_ref="tight_fp64"),
    "c5": dict(cfg=C5, steps=100,
               label="N=10000, volume fraction 1e-2, M=5, dt=0.05, tol 1e-8, from the seeded layout (seed 0x5EED0004)",
               methods=(("mixed_learned", C5_MIXED),),
               # solved pairs of two other N=10000 layouts (the FP64 matrix and
               # its FP32 factors would not fit beside each other at this size;
               # the mixed path solves them to 1e-10 against the FP64 operator)
               train=dict(N=10000, seeds=[C5["seed"] + 0x100 + k for k in range(2)],
                          held=dict(C5_MIXED, warm_start=0, gmres_tol=1e-10),
                          solve=dict(C5_MIXED, warm_start=0, gmres_tol=1e-10)),
               shadow_ref=None),
}


def fiber_spread(a, b, bound=ACCURACY):
    """Relative position difference of two runs, overall and per fiber (each
    fiber's displacement difference over the rms fiber position)."""
    d = np.linalg.norm((a - b).reshape(-1, 3), axis=1)
    scale = np.linalg.norm(b) / np.sqrt(d.size)
    r = d / scale
    return {"rel": rel(a, b), "fiber_median": float(np.median(r)), "fiber_p90": float(np.percentile(r, 90)),
            "fiber_max": float(r.max()), "fibers_within_bound": float(np.mean(r <= bound))}


def long_run(rf, name, steps):
    """`--long c3|c4|c5 K`: K steps of each method of the configuration from
    its seeded layout: steps/s, LU factorisations, the largest FP64 backward
    error of any step, positions compared at checkpoints (overall and per
    fiber), and the per-step shadow error of the first method against the
    reference method over all K steps (progress on stderr every 10 steps)."""
    spec = LONG[name]
    cfg = spec["cfg"]
    P0, T0 = layout(cfg["N"], cfg["phi"], cfg["seed"])
    w, tinfo = None, None
    if spec["train"]:
        tr = spec["train"]
        w, tinfo = train_network(rf, tr["N"], cfg["phi"], tr["seeds"], 2, P0, T0, held_kw=tr["held"],
                                 solve_kw=tr.get("solve", TRAIN_SOLVE))
        log("%s: network trained on %d pairs, held-out loss %.3f" % (name, tinfo["pairs"], tinfo["held_out_loss"]))
    marks = sorted({k for k in (1, 2, 3, 5, 10, 20, 30, 50, 70, 100, steps) if k <= steps})
    runs = {}
    for mname, kw in spec["methods"]:
        sim = rf.Simulation(rf.Configuration(rf.default_params(profile=1, **kw), P0, T0))
        if kw.get("warm_start") == 3:
            sim.mlp_set_weights(w)
        snaps, worst_be, noise_limited, t0 = {}, 0.0, 0, time.perf_counter()
        for k in range(1, steps + 1):
            sim.step()
            st = sim.stats()  # host copy of counters: no wait on the stream
            worst_be = max(worst_be, float(st.last_backward_error))
            noise_limited += int(st.last_noise_limited)
            if k in marks:
                snaps[k] = sim.state()[0]
            if k % 10 == 0:
                log("%s %s: %d steps, %.1f s" % (name, mname, k, time.perf_counter() - t0))
        sim.synchronize()
        wall = time.perf_counter() - t0
        st = sim.stats()
        sim.close()
        runs[mname] = {"snaps": snaps, "steps_per_sec": steps / wall, "lu_factorizations": int(st.lu_factorizations),
                       "lu_panels_redone": int(st.lu_panel_fallbacks),
                       "gmres_iters": int(st.total_gmres_iters), "max_backward_error": worst_be,
                       "noise_limited_steps": noise_limited}
    names = [m for m, _ in spec["methods"]]
    diffs = {}
    for i, a in enumerate(names):
        for b in names[i + 1:]:
            diffs["%s_vs_%s" % (a, b)] = {str(k): fiber_spread(runs[a]["snaps"][k], runs[b]["snaps"][k])
                                          for k in marks}
    out = {
        "workload": spec["label"] + ", %d steps" % steps,
        "steps": steps,
        "training": tinfo,
        "methods": {m: {k: v for k, v in r.items() if k != "snaps"} for m, r in runs.items()},
        "position_diff_at_step": diffs,
        "position_diff_unit": "rel: |X_a - X_b| / |X_b| over all fibers; fiber_*: one fiber's displacement "
                              "difference over the rms fiber position; fibers_within_bound: share <= 1e-6",
    }
    plain = [m for m in names if m.startswith("plain")]
    if plain:
        best = max(runs[m]["steps_per_sec"] for m in plain)
        out["speedup_vs_best_plain"] = runs[names[0]]["steps_per_sec"] / best
    if spec["shadow_ref"]:
        ref_kw = dict(spec["methods"])[spec["shadow_ref"]]
        sh = shadow(rf, P0, T0, steps, 0, w, spec["methods"][0][1], {"ref": ref_kw})
        per = sh["method_vs_ref"]
        out["shadow"] = {"reference": spec["shadow_ref"], "max_rel_err": max(per), "per_step": per,
                         "meets_accuracy": bool(max(per) <= ACCURACY)}
        log("%s: shadow error over %d steps, max %.2e" % (name, steps, max(per)))
    return out


def cpu_baseline(N, phi, seed, threads, warmup=2, steps=5):
    """The FP64 C oracle on warm-started steps of the headline workload:
    time(warmup + steps) - time(warmup) from the same initial state."""
    import oracle

    os.environ["OMP_NUM_THREADS"] = str(threads)
    pos, ori = layout(N, phi, seed)
    kw = dict(M=5, noqi=8, restart=10, max_iters=300, tol=1e-10)
    t0 = time.perf_counter()
    oracle.simulate(True, False, 0.01, 0.1, warmup, pos, ori, **kw)
    t1 = time.perf_counter()
    oracle.simulate(True, False, 0.01, 0.1, warmup + steps, pos, ori, **kw)
    t2 = time.perf_counter()
    dt = (t2 - t1) - (t1 - t0)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": steps / dt,
        "unit": "steps/s",
        "cores": int(oracle.num_threads()),
        "kind": "port",
        "cpu_model": model,
        "sample": "steps %d..%d of the N=1000 headline workload (analytic assembly, GMRES(10) to 1e-10 warm-started "
                  "from the previous step with LU fallback after 300 iterations, velocities, advance) with the FP64 C "
                  "oracle, OpenMP over target fibers: time of %d steps minus time of the first %d, %.1f s for the %d "
                  "steps" % (warmup + 1, warmup + steps, warmup + steps, warmup, dt, steps),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-lu", action="store_true", help="skip the direct-solver timing")
    ap.add_argument("--no-n4000", action="store_true", help="skip the N=4000 comparison")
    ap.add_argument("--n4000-steps", type=int, default=3)
    ap.add_argument("--long", nargs=2, metavar=("CONFIG", "K"),
                    help="instead of the benchmark: K steps of configuration c3 (N=2000), c4 (N=4000) or c5 "
                         "(N=10000) for each of its methods, with position differences at checkpoints and the "
                         "per-step shadow error over all K steps (one JSON line)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; the multi-GPU row sharding is a design only "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import rigidfibers_kth_amd as rf

    if args.long:
        if args.long[0] not in LONG or int(args.long[1]) < 1:
            ap.error("--long takes c3, c4 or c5 and a positive step count")
        print(json.dumps(long_run(rf, args.long[0], int(args.long[1]))), flush=True)
        return
    pos, ori = layout(C2["N"], C2["phi"], C2["seed"])
    g, Pp = timed_run(rf, pos, ori, args.steps, args.warmup, **PLAIN)
    log("N=1000 plain FP64: %.2f steps/s" % g["steps_per_sec"])
    n, eb = g["n"], g["elem_bytes"]
    gemv_bytes = float(n) * n * eb + 2.0 * n * 8  # matrix once + x and y
    achieved = gemv_bytes / (g["gemv_avg_ms"] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        try:
            t = json.load(open(TRAFFIC_FILE))
            if t.get("n") == n and t.get("elem_bytes") == eb:
                traffic = t.get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    flops, near_frac = analytic_assembly_flops(pos, ori)
    asm_tflops = flops / (g["assemble_avg_ms"] * 1e-3) / 1e12

    # mixed + learned at N=1000 on the same initial state and step count; the
    # network is trained on solved pairs of ten other seeded layouts
    w, tinfo = train_network(rf, C2["N"], C2["phi"], [C2["seed"] + 100 + k for k in range(10)], 2, pos, ori)
    log("N=1000: network trained on %d pairs, held-out loss %.3f" % (tinfo["pairs"], tinfo["held_out_loss"]))
    mixed, Pm = timed_run(rf, pos, ori, args.steps, args.warmup, weights=w, **MIXED)
    log("N=1000 mixed + learned: %.2f steps/s" % mixed["steps_per_sec"])
    nomlp, _ = timed_run(rf, pos, ori, args.steps, args.warmup, **MIXED_NO_MLP)
    f64, _ = timed_run(rf, pos, ori, args.steps, args.warmup, weights=w, **FP64_FP32LU)
    sh = shadow(rf, pos, ori, args.steps, args.warmup, w, MIXED, {"plain": PLAIN, "tight": TIGHT})
    log("N=1000 shadow error max %.2e" % max(sh["method_vs_plain"]))
    n1000 = dict(mixed, speedup=mixed["steps_per_sec"] / g["steps_per_sec"], training=tinfo)
    n1000.update(accuracy_record(sh, (Pm, Pp)))
    n1000["guess_quality"] = guess_quality(rf, pos, ori, 10, w, MIXED)
    extra = {
        "n1000_mixed_learned": n1000,
        "n1000_mixed_without_network": {k: nomlp[k] for k in ("steps_per_sec", "gmres_iters_per_step",
                                                              "operator_launches_per_step")},
        "n1000_fp64_matrix_fp32_lu_learned": {k: f64[k] for k in ("steps_per_sec", "gmres_iters_per_step",
                                                                  "lu_factorizations")},
    }
    if not args.no_lu:
        lu, _ = timed_run(rf, pos, ori, 2, 1, use_direct_solver=1)
        log("N=1000 FP64 LU: %.1f ms per factorisation" % lu["lu_avg_ms"])
        extra["n1000_lu_fp64"] = {
            "steps_per_sec": lu["steps_per_sec"],
            "lu_factor_ms": lu["lu_avg_ms"],
            "lu_tflops": (2.0 / 3.0) * n ** 3 / (lu["lu_avg_ms"] * 1e-3) / 1e12,
            "frac_fp64_peak": (2.0 / 3.0) * n ** 3 / (lu["lu_avg_ms"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        }
    if not args.no_n4000:
        c4 = compare(rf, C4, args.n4000_steps, 0, dict(N=2000, seeds=[C4["seed"] + 0x100 + k for k in range(5)]),
                     "N=4000, volume fraction 5e-2, M=5, dt=0.1, tol 1e-10, from the seeded initial layout; network "
                     "trained on 2 states of 5 other seeded layouts of 2000 fibers at the same volume fraction",
                     "direct")
        mi = c4["mixed_learned"]
        if mi["lu_factorizations"]:
            n4 = mi["n"]
            tf = (2.0 / 3.0) * n4 ** 3 / (mi["lu_avg_ms"] * 1e-3) / 1e12
            c4["fp32_lu"] = {"factor_ms": mi["lu_avg_ms"], "tflops": tf, "update": "fp16 operands, FP32 accumulation",
                             "frac_fp16_mfma_peak": tf / FP16_MFMA_PEAK_TFLOPS}
        extra["n4000"] = c4
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
        "data": "synthetic: seeded fiber layouts (N=1000 seed 0x5EED0001, N=4000 seed 0x5EED0003), overlaps rejected",
        "config": {
            "workload": "plain FP64, N=1000 fibers, volume fraction 1e-2, M=5, 8x3 Gauss points, analytic, "
                        "GMRES(10) tol 1e-10 with LU fallback after 300 iterations, previous-step warm start, dt=0.1",
            "fibers": C2["N"], "unknowns": n, "volume_fraction": C2["phi"], "solver": "gmres(10)", "tol": 1e-10,
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
        log("CPU oracle steps")
        threads = min(16, len(os.sched_getaffinity(0)))
        line["cpu_baseline"] = cpu_baseline(C2["N"], C2["phi"], C2["seed"], threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
