"""Restart from saved levels (no reader in the reference, whose Fortran
driver writes XcTData / VelRotData / Coeffvec: fortran/advect_fibers.f90:
395-424,546-591).

* In memory: K steps straight through equal K1 steps, rf_get_restart_state,
  a new handle, rf_restore_state, K - K1 steps -- bit for bit, for GMRES with
  the previous-step guess and for the direct solver (the multistep update
  reads both position levels and the previous velocities, so a restore that
  dropped any of them would differ at O(dt)).
* From the snapshot files: `fibers --resume DIR` continues a run written with
  save interval 1; its final POS.out / ORIENT.out equal the straight run's to
  the files' 16 decimals propagated over the remaining steps.
"""
import os
import subprocess

import numpy as np
import pytest

import rigidfibers_kth_amd as rf
from conftest import GOLDEN, ROOT
from rigidfibers_kth_amd.layouts import layout

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("direct", [0, 1])
def test_restore_continues_bit_for_bit(direct):
    N, K, K1 = 60, 6, 3
    pos, ori = layout(N, 1e-2, 77)
    p = rf.default_params(use_direct_solver=direct, gmres_restart=10, gmres_max_iters=300, gmres_tol=1e-11)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        for _ in range(K):
            sim.step()
        want = sim.restart_state()
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        for _ in range(K1):
            sim.step()
        saved = sim.restart_state()
    assert saved["steps_taken"] == K1
    # a fresh handle made from some other layout: everything comes from the restore
    other_pos, other_ori = layout(N, 1e-2, 78)
    with rf.Simulation(rf.Configuration(p, other_pos, other_ori)) as sim:
        sim.restore_state(saved)
        assert sim.stats().steps_taken == K1
        for _ in range(K - K1):
            sim.step()
        got = sim.restart_state()
    for k in ("pos", "orient", "pos_prev", "orient_prev", "vel", "rot", "coeffs"):
        assert np.array_equal(got[k], want[k]), k
    assert got["steps_taken"] == K


def test_restore_of_a_dropped_level_differs():
    """The check above can fail: restoring with the previous level replaced by
    the current one (what a reader that keeps one level would do) moves the
    positions after one step by O(dt |V|)."""
    N = 60
    pos, ori = layout(N, 1e-2, 77)
    p = rf.default_params(use_direct_solver=1)
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        for _ in range(3):
            sim.step()
        saved = sim.restart_state()
        sim.step()
        want = sim.restart_state()
    bad = dict(saved, pos_prev=saved["pos"], orient_prev=saved["orient"])
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        sim.restore_state(bad)
        sim.step()
        got = sim.restart_state()
    assert np.abs(got["pos"] - want["pos"]).max() > 1e-4


def _params_file(tmp_path, steps):
    src = open(os.path.join(GOLDEN, "inputs", "run100.in")).read().splitlines()
    src[5] = "%d     !No. of timesteps" % steps
    src[6] = "1     !Save interval"
    src[7] = "2     !Number of timesteps saved in one file"
    f = tmp_path / ("run_%d.in" % steps)
    f.write_text("\n".join(src) + "\n")
    return str(f)


def test_cli_resume_from_snapshot_files(tmp_path):
    exe = os.path.join(ROOT, "rigidfibers_kth_amd", "fibers")
    layout_file = os.path.join(GOLDEN, "inputs", "XcT_init100.in")
    a, b = tmp_path / "straight", tmp_path / "resumed"
    a.mkdir()
    b.mkdir()
    subprocess.run([exe, "--out", str(a), _params_file(tmp_path, 5), layout_file], check=True, capture_output=True,
                   timeout=300)
    subprocess.run([exe, "--out", str(b), _params_file(tmp_path, 3), layout_file], check=True, capture_output=True,
                   timeout=300)
    # the three-step run wrote parts 1 and 2 (two saves per file)
    assert (b / "XcTData1_p2").exists()
    out = subprocess.run([exe, "--out", str(b), "--resume", str(b), _params_file(tmp_path, 5), layout_file],
                         check=True, capture_output=True, timeout=300, text=True)
    assert "Resumed after timestep 3" in out.stdout
    for name in ("POS", "ORIENT", "TRANSVel", "ROTVel"):
        x = np.loadtxt(a / (name + ".out"))
        y = np.loadtxt(b / (name + ".out"))
        assert np.abs(x - y).max() <= 1e-10 * max(1.0, np.abs(x).max()), name
    # the continuation went into new parts
    assert (b / "XcTData1_p3").exists()
