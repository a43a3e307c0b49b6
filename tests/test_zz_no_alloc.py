"""rf_operator_noalloc: the matrix-free operator on caller device buffers
with its scratch in a file-scope device array (no allocation), against
rf_operator through a handle (the existing entry point, same kernels) for
every mode, analytic and numeric inner integrals, on two calls in a row with
the second smaller; sizes beyond the scratch are rejected on the host."""
import ctypes

import numpy as np
import pytest

import rigidfibers_kth_amd as rf
from rigidfibers_kth_amd.layouts import layout

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _noalloc(p, pos, ori, mode, x):
    dev = torch.device("cuda:0")
    N = pos.size // 3
    n = 3 * p.num_terms * N
    force = np.tile(np.asarray(list(p.external_force)), N)
    tp, to, tf = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (pos, ori, force))
    tx = torch.from_numpy(np.ascontiguousarray(x)).to(dev) if x is not None else None
    out = torch.full((6 * N if mode == rf.OP_VELOCITY else n,), float("nan"), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    rc = rf.lib().rf_operator_noalloc(ctypes.byref(p), N, mode, tp.data_ptr(), to.data_ptr(), tf.data_ptr(),
                                      tx.data_ptr() if tx is not None else None, out.data_ptr(),
                                      ctypes.c_void_p(stream.cuda_stream))
    assert rc == rf.RF_OK, rf.lib().rf_last_error().decode()
    torch.cuda.synchronize(dev)
    return out.cpu().numpy()


def _handle(p, pos, ori, mode, x):
    """rf_operator through a handle, and the handle's state (rf_create
    renormalises the orientations)."""
    with rf.Simulation(rf.Configuration(p, pos, ori)) as sim:
        y = sim.operator(mode, x)
        P, T = sim.state()
    return (np.concatenate(y) if mode == rf.OP_VELOCITY else y), P, T


@pytest.mark.parametrize("analytic", [1, 0])
def test_noalloc_operator_matches_rf_operator(analytic):
    p = rf.default_params(use_analytical_integration=analytic, external_force=(0.1, -0.2, -1.0), shear_rate=0.3)
    for N, seed in ((60, 71), (25, 72)):  # the second call smaller than the first
        pos, ori = layout(N, 2e-2, seed)
        rng = np.random.default_rng(seed)
        x = rng.normal(size=3 * p.num_terms * N)
        for mode in (rf.OP_AX, rf.OP_RESIDUAL, rf.OP_RHS, rf.OP_VELOCITY):
            xm = None if mode == rf.OP_RHS else x
            print("N=%d analytic=%d mode=%d" % (N, analytic, mode), flush=True)
            want, P, T = _handle(p, pos, ori, mode, xm)
            got = _noalloc(p, P, T, mode, xm)
            assert np.all(np.isfinite(got))
            err = np.abs(got - want).max() / np.abs(want).max()
            print("  max rel diff %.2e (bit-identical: %s)" % (err, np.array_equal(got, want)))
            # same kernels, tables and launch plan on the same inputs
            assert np.array_equal(got, want)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    p = rf.default_params()
    N = 200000
    rc = rf.lib().rf_operator_noalloc(ctypes.byref(p), N, rf.OP_AX, 1, 1, 1, 1, 1, None)
    assert rc == rf.RF_EINVAL
    assert b"maximum" in rf.lib().rf_last_error()
