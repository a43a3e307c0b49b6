"""The benchmark's host-side pieces, on the CPU: every configuration it runs
builds a valid parameter block, the per-fiber position spread and the
relative difference are what they claim, and a multi-GPU request is refused
with a message (one GPU is measured)."""
import subprocess
import sys

import numpy as np
import pytest

import bench
import rigidfibers_kth_amd as rf


def test_every_benchmark_configuration_builds_parameters():
    kws = [bench.PLAIN, bench.PLAIN_DIRECT, bench.MIXED, bench.MIXED_NO_MLP, bench.FP64_FP32LU, bench.TIGHT,
           bench.TRAIN_SOLVE, bench.C5_MIXED]
    for spec in bench.LONG.values():
        kws += [kw for _, kw in spec["methods"]]
        if spec["train"]:
            kws += [spec["train"]["held"], spec["train"].get("solve", bench.TRAIN_SOLVE)]
    for kw in kws:
        p = rf.default_params(**kw)
        for k, v in kw.items():
            assert getattr(p, k) == pytest.approx(v), k
    # the shadow reference of each long run is one of its methods
    for spec in bench.LONG.values():
        if spec["shadow_ref"]:
            assert spec["shadow_ref"] in dict(spec["methods"])


def test_fiber_spread():
    rng = np.random.default_rng(1)
    b = rng.uniform(0.0, 10.0, 3 * 50)
    a = b.copy()
    a[3 * 7:3 * 7 + 3] += np.array([3e-6, 4e-6, 0.0])  # one fiber moved by 5e-6
    s = bench.fiber_spread(a, b)
    scale = np.linalg.norm(b) / np.sqrt(50)
    assert s["fiber_max"] == pytest.approx(5e-6 / scale)
    assert s["fiber_median"] == 0.0
    assert s["fibers_within_bound"] == pytest.approx(49 / 50 if 5e-6 / scale > 1e-6 else 1.0)
    assert s["rel"] == pytest.approx(5e-6 / np.linalg.norm(b))
    assert bench.rel(b, b) == 0.0


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
