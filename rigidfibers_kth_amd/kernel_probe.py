"""Runs one kernel family of the timestep in isolation, for profilers.

    python3 -m rigidfibers_kth_amd.kernel_probe gemv --fibers 1000 --reps 20
    python3 -m rigidfibers_kth_amd.kernel_probe copy --bytes 1073741824 --reps 5
    python3 -m rigidfibers_kth_amd.kernel_probe lu --fibers 1000 [--fp32 [--fp16]] [--column-panel] [--block NB]
    python3 -m rigidfibers_kth_amd.kernel_probe assemble --fibers 1000 --reps 5
    python3 -m rigidfibers_kth_amd.kernel_probe steps --fibers 4000 --warmup 20 --reps 10

The layouts are the seeded ones of the benchmark (volume fraction 1e-2 for
N <= 2000, 5e-2 above).  Each run prints what it launched, so a profile of
the command can be read against it: `gemv` launches the system-matrix GEMV
(rf_matvec) `reps` times after one assembly, `copy` streams a buffer through
the plain copy kernel (HBM byte-counter calibration), `lu` factorises the
assembled matrix once (as the direct solver; with --fp16 as the GMRES
preconditioner, whose FP32 factors take the fp16-operand trailing update),
`assemble` assembles `reps` times, `steps` runs `warmup` then `reps` full
timesteps of the mixed-precision path (FP32 matrix, automatic neighbour / fp16
LU preconditioner, flexible GMRES on the FP64 operator to 1e-10, previous-step
guess) and prints the phases of the `reps` steps.
"""
import argparse
import sys
import time

import numpy as np

from . import (PREC_FP64, PREC_MIXED, PRECOND_AUTO, PRECOND_LU, Configuration, Simulation, default_params,
               selftest_copy)
from .layouts import layout


def _sim(fibers, **kw):
    phi, seed = (1e-2, 0x5EED0001) if fibers <= 2000 else (5e-2, 0x5EED0003)
    pos, ori = layout(fibers, phi, seed)
    return Simulation(Configuration(default_params(**kw), pos, ori))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("kernel", choices=["gemv", "copy", "lu", "assemble", "steps"])
    ap.add_argument("--warmup", type=int, default=20, help="steps: untimed steps first")
    ap.add_argument("--switch-iters", type=int, default=0,
                    help="steps: neighbour-preconditioned iterations before the fp16 LU (0 = library default)")
    ap.add_argument("--fibers", type=int, default=1000)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--fp32", action="store_true", help="FP32 matrix (gemv, assemble) or FP32 factors (lu)")
    ap.add_argument("--fp16", action="store_true", help="lu: FP32 factors with fp16-operand trailing updates, "
                    "factorised as the preconditioner of one GMRES solve")
    ap.add_argument("--column-panel", action="store_true", help="lu: pivot search over whole columns in every panel")
    ap.add_argument("--block", type=int, default=0, help="lu: panel width (0 = default)")
    a = ap.parse_args(argv)
    if a.kernel == "copy":
        ms = selftest_copy(a.bytes, a.reps)
        print("copy: %d bytes x %d launches, %.3f ms each, %.1f GB/s (read + write)"
              % (a.bytes, a.reps, ms, 2 * a.bytes / (ms * 1e-3) / 1e9), flush=True)
        return 0
    if a.kernel == "steps":
        kw = dict(precision=PREC_MIXED, profile=1, gmres_restart=30, gmres_max_iters=3000, gmres_tol=1e-10,
                  precond=PRECOND_AUTO, lu_update_fp16=1, precond_switch_iters=a.switch_iters)
        with _sim(a.fibers, **kw) as sim:
            for _ in range(a.warmup):
                sim.step()
            sim.synchronize()
            s0 = sim.stats()
            t0 = time.perf_counter()
            for _ in range(a.reps):
                sim.step()
            sim.synchronize()
            wall = time.perf_counter() - t0
            s1 = sim.stats()
            d = {f: getattr(s1, f) - getattr(s0, f) for f in ("total_ms_assemble", "total_ms_solve",
                                                               "total_ms_velocities", "total_ms_advance",
                                                               "total_gmres_iters", "lu_factorizations",
                                                               "operator_launches", "operator_ms_total")}
            print("steps: N=%d, %d steps after %d, %.1f ms per step (wall); per step: assemble %.1f ms, solve %.1f "
                  "ms, velocities %.1f ms, advance %.2f ms; %.1f inner iterations, %.2f LU factorisations, %.1f "
                  "operator evaluations of %.2f ms" % (
                      a.fibers, a.reps, a.warmup, 1e3 * wall / a.reps, d["total_ms_assemble"] / a.reps,
                      d["total_ms_solve"] / a.reps, d["total_ms_velocities"] / a.reps, d["total_ms_advance"] / a.reps,
                      d["total_gmres_iters"] / a.reps, d["lu_factorizations"] / a.reps,
                      d["operator_launches"] / a.reps, d["operator_ms_total"] / max(1, d["operator_launches"])),
                  flush=True)
        return 0
    prec = PREC_MIXED if (a.fp32 and a.kernel != "lu") else PREC_FP64
    kw = dict(precision=prec, profile=1)
    if a.kernel == "lu":
        kw.update(lu_factor_fp32=int(a.fp32 or a.fp16), lu_column_panel=int(a.column_panel), lu_block=a.block)
        if a.fp16:
            kw.update(precond=PRECOND_LU, lu_update_fp16=1, gmres_restart=30, gmres_max_iters=300, gmres_tol=1e-10)
        else:
            kw.update(use_direct_solver=1)
    with _sim(a.fibers, **kw) as sim:
        sim.assemble_system()
        sim.synchronize()
        if a.kernel == "assemble":
            for _ in range(a.reps - 1):
                sim.assemble_system()
            sim.synchronize()
            st = sim.stats()
            print("assemble: N=%d, %d launches, %.3f ms each" % (a.fibers, st.assemble_launches,
                                                                 st.assemble_ms_total / st.assemble_launches))
        elif a.kernel == "gemv":
            x = np.random.default_rng(0).normal(size=sim.n)
            t0 = time.perf_counter()
            for _ in range(a.reps):
                sim.matvec(x)
            sim.synchronize()
            st = sim.stats()
            print("gemv: n=%d, element bytes %d, %d launches, %.3f ms each (events), %.3f ms wall each"
                  % (sim.n, st.matrix_elem_bytes, st.gemv_launches, st.gemv_ms_total / max(1, st.gemv_launches),
                     1e3 * (time.perf_counter() - t0) / a.reps), flush=True)
        else:
            sim.solve_system()
            sim.synchronize()
            st = sim.stats()
            kind = "fp32, fp16-operand updates" if a.fp16 else ("fp32" if a.fp32 else "fp64")
            print("lu: n=%d, factors %s, %.1f ms, panels redone %d, GMRES iterations %d, backward error %.2e"
                  % (sim.n, kind, st.lu_ms_total, st.lu_panel_fallbacks, st.last_gmres_iters, st.last_backward_error),
                  flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
