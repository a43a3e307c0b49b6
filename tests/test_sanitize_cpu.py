"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code, run
on the CPU: the FP64 oracle (all entry points, cross-checked against each
other) and the library's host-side file readers (parameter, layout and
snapshot files, well-formed, truncated and malformed).  Built from
tests/sanitize/Makefile with gcc/g++; a sanitizer report aborts the program.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="gcc/g++ not found")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", SAN, "-j2"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(SAN, "_build")


def _run(cmd, **env):
    e = dict(os.environ)
    e.update(env)
    # leak checking on; the first report ends the run with a non-zero status
    e["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1:verify_asan_link_order=0"
    e["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run(cmd, capture_output=True, text=True, env=e, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out
    return out


def test_host_file_readers_under_sanitizers(built, tmp_path):
    out = _run([os.path.join(built, "io_check"), str(tmp_path)])
    assert "io_check: 0 failed checks" in out


def test_oracle_under_sanitizers(built):
    out = _run([os.path.join(built, "oracle_check")], OMP_NUM_THREADS="4")
    assert "oracle_check: 0 failed checks" in out
