"""C ABI checks that need no GPU: symbols, parsers, error paths, CLI.

Parsers follow src/parameters.cc:124-227 and fortran/advect_fibers.f90:182-222.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import rigidfibers_kth_amd as rf
from conftest import GOLDEN, ROOT, read_layout

INPUTS = os.path.join(GOLDEN, "inputs")


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "rigidfibers.h")).read()
    return sorted(set(re.findall(r"\b(rf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    L = rf.lib()
    syms = _header_symbols()
    assert len(syms) >= 24
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(rf.exported_symbols())


def test_version_string():
    assert b"gfx950" in rf.lib().rf_version()


def test_parse_run100_direct():
    p = rf.Parameters.parse_parameter_file(os.path.join(INPUTS, "run100.in"))
    assert (p.run_label, p.layout_label, p.num_terms) == (1, 100, 5)
    assert p.slenderness == 0.01 and p.timestep == 0.1
    assert (p.num_timesteps, p.save_interval, p.saves_per_file, p.num_quad_intervals) == (10, 1, 10, 8)
    assert (p.use_analytical_integration, p.use_direct_solver) == (1, 1)
    assert p.num_quad_points_per_interval == 3
    assert list(p.external_force) == [0.0, 0.0, -1.0]
    assert p.precision == rf.PREC_FP64 and p.warm_start == rf.WARM_PREV


def test_parse_run100_iter():
    p = rf.Parameters.parse_parameter_file(os.path.join(INPUTS, "run100_iter.in"))
    assert (p.run_label, p.use_direct_solver) == (2, 0)
    assert (p.gmres_restart, p.gmres_max_iters, p.gmres_tol) == (10, 100, 1e-5)


def test_parse_extension_lines(tmp_path):
    """"key value" lines after the mandatory ones select the extensions;
    unknown keys and values are errors."""
    src = open(os.path.join(INPUTS, "run100_iter.in")).read()
    f = tmp_path / "run.in"
    f.write_text(src + "precision mixed\nwarm_start learned\nprofile 1\n")
    p = rf.Parameters.parse_parameter_file(str(f))
    assert p.precision == rf.PREC_MIXED and p.warm_start == rf.WARM_LEARNED and p.profile == 1
    f.write_text(src + "precond auto\nprecond_switch_iters 45\nlu_factor_fp32 1\nrefine_tol 1d-13\n"
                       "precond_neighbours 6\nmixed_inner_tol 1e-11\nlu_block 96\nlu_update_fp16 1\n"
                       "lu_column_panel 1\nshear_rate 0.25\n")
    p = rf.Parameters.parse_parameter_file(str(f))
    assert (p.precond, p.precond_switch_iters, p.lu_factor_fp32, p.precond_neighbours, p.lu_block) == \
        (rf.PRECOND_AUTO, 45, 1, 6, 96)
    assert (p.lu_update_fp16, p.lu_column_panel) == (1, 1)
    assert p.refine_tol == 1e-13 and p.mixed_inner_tol == 1e-11 and p.shear_rate == 0.25
    for bad in ("precond sometimes\n", "warm_start later\n", "precision fp16\n", "lu_reuse_iters 4\n",
                "refine_floor 1e-8\n"):
        f.write_text(src + bad)
        with pytest.raises(rf.RigidFibersError):
            rf.Parameters.parse_parameter_file(str(f))
    f.write_text(src + "bogus 3\n")
    with pytest.raises(rf.RigidFibersError) as e:
        rf.Parameters.parse_parameter_file(str(f))
    assert e.value.code == rf.RF_EIO


def test_parse_truncated_parameter_file(tmp_path):
    f = tmp_path / "short.in"
    f.write_text("1\n100\n5\n")
    with pytest.raises(rf.RigidFibersError) as e:
        rf.Parameters.parse_parameter_file(str(f))
    assert e.value.code == rf.RF_EIO and "expected" in str(e.value)


def test_parse_missing_file():
    with pytest.raises(rf.RigidFibersError) as e:
        rf.Parameters.parse_parameter_file("/nonexistent/run.in")
    assert e.value.code == rf.RF_EIO and "Could not open" in str(e.value)


@pytest.mark.parametrize("name,N", [("XcT_init1.in", 2), ("XcT_init100.in", 100), ("XcT_init1000.in", 1000)])
def test_parse_layout(name, N):
    path = os.path.join(INPUTS, name)
    pos, ori = rf.Parameters.parse_initial_layout_file(path)
    rp, ro = read_layout(path)
    assert pos.size == 3 * N
    np.testing.assert_array_equal(pos, rp)
    np.testing.assert_array_equal(ori, ro)


def test_parse_layout_truncated(tmp_path):
    f = tmp_path / "bad.in"
    f.write_text("3\n0 0 0\n0 0 1\n")
    with pytest.raises(rf.RigidFibersError) as e:
        rf.Parameters.parse_initial_layout_file(str(f))
    assert e.value.code == rf.RF_EIO


def test_parse_configuration_files():
    cfg = rf.Parameters.parse_configuration_files(os.path.join(INPUTS, "run100.in"),
                                                  os.path.join(INPUTS, "XcT_init100.in"))
    assert cfg.num_fibers == 100 and cfg.parameters.num_terms == 5


def _has_gpu():
    return os.path.exists("/dev/kfd")


def test_create_rejects_bad_parameters():
    p = rf.default_params(num_terms=9)
    cfg = rf.Configuration(p, np.zeros(3), np.array([0, 0, 1.0]))
    with pytest.raises(rf.RigidFibersError) as e:
        rf.Simulation(cfg)
    assert e.value.code == rf.RF_EINVAL


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_create_without_gpu_fails_loudly():
    cfg = rf.Configuration(rf.default_params(), np.zeros(3), np.array([0, 0, 1.0]))
    with pytest.raises(rf.RigidFibersError) as e:
        rf.Simulation(cfg)
    assert e.value.code == rf.RF_EHIP


def test_cli_help_and_version():
    exe = os.path.join(ROOT, "rigidfibers_kth_amd", "fibers")
    out = subprocess.run([exe, "--help"], capture_output=True, text=True, check=True).stdout
    assert "fibers [options] <parameters> <layout>" in out
    out = subprocess.run([exe, "--version"], capture_output=True, text=True, check=True).stdout
    assert "gfx950" in out
    r = subprocess.run([exe, "only-one-arg"], capture_output=True, text=True)
    assert r.returncode == 1


def test_read_snapshot_records(tmp_path):
    """rf_read_snapshot: records of '(2F24.16)' lines with blank separators
    (fortran/advect_fibers.f90:395-424,546-591), the last by default, an
    incomplete trailing record ignored, a record beyond the end an error."""
    N = 2
    f = tmp_path / "XcTData1_p1"
    vals = np.arange(3 * 6 * N, dtype=float) * 0.125 - 1.0
    with open(f, "w") as fh:
        for r in range(3):
            for i in range(N):
                for half in range(2):
                    for c in range(3):
                        fh.write("%24.16f\n" % vals[r * 6 * N + 6 * i + 3 * half + c])
                    fh.write(" \n")
        fh.write("%24.16f\n" % 7.0)  # incomplete record
    last, n = rf.read_snapshot(str(f), 6 * N)
    assert n == 3 and np.array_equal(last, vals[12 * 2:12 * 3])
    first, _ = rf.read_snapshot(str(f), 6 * N, 0)
    assert np.array_equal(first, vals[:12])
    with pytest.raises(rf.RigidFibersError):
        rf.read_snapshot(str(f), 6 * N, 3)
    with pytest.raises(rf.RigidFibersError):
        rf.read_snapshot(str(tmp_path / "missing"), 6 * N)
