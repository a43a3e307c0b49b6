"""FP64 CPU oracle for the rigid-fiber timestep (ctypes wrapper of liboracle.so).

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.
The product package ``rigidfibers_kth_amd`` never imports it and has no CPU
fallback.

The C sources (``rf_oracle.c``) restate the Fortran reference
(fortran/quadrature.f90, fortran/Int_analytical.f90,
fortran/matrix_assembly.f90, fortran/coeffs_and_vels.f90, fortran/dGMRES.f,
fortran/advect_fibers.f90); each C function cites the routine it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile liboracle.so with gcc (Makefile beside this file)."""
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_quad_pts_wts.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_legendre.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_legendre.restype = ctypes.c_int
        L.orc_analytical_int.argtypes = [_dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_G_analytic.argtypes = [_dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_G_numeric.argtypes = [_dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_assemble_matrix.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_assemble_rhs.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_mat_prod.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                   ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_compute_velocities.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                             _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                             ctypes.c_double]
        L.orc_assemble_rhs_rows.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                            _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip, _dp]
        L.orc_apply_rows.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                     ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _ip, _dp]
        L.orc_compute_velocities_rows.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                  _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp,
                                                  ctypes.c_int, _ip, _dp, _dp, ctypes.c_double]
        L.orc_lu_solve.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_lu_solve.restype = ctypes.c_int
        L.orc_gmres.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, _ip, _dp]
        L.orc_gmres.restype = ctypes.c_int
        L.orc_simulate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, ctypes.c_double]
        L.orc_simulate.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads():
    return lib().orc_num_threads()


def quadrature(noqi):
    lq = 3 * noqi
    pv = np.empty(lq)
    wv = np.empty(lq)
    lib().orc_quad_pts_wts(noqi, _p(pv), _p(wv))
    return pv, wv


def legendre(M, pv):
    """P_1..P_M at the points; shape (M, LQ)."""
    pv = _f64(pv)
    out = np.empty((M, pv.size))
    rc = lib().orc_legendre(M, pv.size, _p(pv), _p(out))
    if rc:
        raise ValueError("M must be 1..8")
    return out


def analytical_int(xb, pb, xbar, lmax):
    """Returns (L[7, lmax+1], I[30], J[30], S[30])."""
    L = np.empty(7 * (lmax + 1))
    I, J, S = np.empty(30), np.empty(30), np.empty(30)
    lib().orc_analytical_int(_p(_f64(xb)), _p(_f64(pb)), _p(_f64(xbar)), lmax, _p(L), _p(I), _p(J), _p(S))
    return L.reshape(7, lmax + 1), I, J, S


def G_analytic(xb, pb, xbar, eps, l):
    out = np.empty(6)
    lib().orc_G_analytic(_p(_f64(xb)), _p(_f64(pb)), _p(_f64(xbar)), eps, l, _p(out))
    return out


def G_numeric(xb, pb, xbar, eps, noqi, lvec_m):
    pv, wv = quadrature(noqi)
    out = np.empty(6)
    lv = _f64(lvec_m)
    lib().orc_G_numeric(_p(_f64(xb)), _p(_f64(pb)), _p(_f64(xbar)), eps, pv.size, _p(pv), _p(wv), _p(lv), _p(out))
    return out


def _setup(M, noqi):
    pv, wv = quadrature(noqi)
    lv = legendre(M, pv)
    return pv, wv, lv


def assemble_matrix(analytic, eps, pos, orient, M, noqi):
    pos, orient = _f64(pos).ravel(), _f64(orient).ravel()
    N = pos.size // 3
    n = 3 * M * N
    pv, wv, lv = _setup(M, noqi)
    A = np.empty((n, n), order="F")
    lib().orc_assemble_matrix(int(analytic), eps, N, M, _p(pos), _p(orient), pv.size, _p(pv), _p(wv), _p(lv),
                              A.ctypes.data_as(_dp))
    return A


def assemble_rhs(analytic, eps, pos, orient, force, M, noqi):
    pos, orient, force = _f64(pos).ravel(), _f64(orient).ravel(), _f64(force).ravel()
    N = pos.size // 3
    pv, wv, lv = _setup(M, noqi)
    b = np.empty(3 * M * N)
    lib().orc_assemble_rhs(int(analytic), eps, N, M, _p(pos), _p(orient), _p(force), pv.size, _p(pv), _p(wv),
                           _p(lv), _p(b))
    return b


def mat_prod(eps, pos, orient, M, noqi, x):
    pos, orient, x = _f64(pos).ravel(), _f64(orient).ravel(), _f64(x)
    N = pos.size // 3
    pv, wv, lv = _setup(M, noqi)
    y = np.empty_like(x)
    lib().orc_mat_prod(eps, N, M, _p(pos), _p(orient), pv.size, _p(pv), _p(wv), _p(lv), _p(x), _p(y))
    return y


def compute_velocities(analytic, eps, pos, orient, force, M, noqi, coeff, sh=0.0):
    pos, orient, force, coeff = (_f64(a).ravel() for a in (pos, orient, force, coeff))
    N = pos.size // 3
    pv, wv, lv = _setup(M, noqi)
    vel = np.empty(3 * N)
    rot = np.empty(3 * N)
    lib().orc_compute_velocities(int(analytic), eps, N, M, _p(pos), _p(orient), _p(force), pv.size, _p(pv),
                                 _p(wv), _p(lv), _p(coeff), _p(vel), _p(rot), float(sh))
    return vel, rot


def assemble_rhs_rows(analytic, eps, pos, orient, force, M, noqi, targets):
    pos, orient, force = _f64(pos).ravel(), _f64(orient).ravel(), _f64(force).ravel()
    N = pos.size // 3
    pv, wv, lv = _setup(M, noqi)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    b = np.empty(3 * M * t.size)
    lib().orc_assemble_rhs_rows(int(analytic), eps, N, M, _p(pos), _p(orient), _p(force), pv.size, _p(pv),
                                _p(wv), _p(lv), t.size, t.ctypes.data_as(_ip), _p(b))
    return b


def apply_rows(analytic, eps, pos, orient, M, noqi, x, targets):
    """Rows (A x)_i of the target fibers (3M values each), without forming A."""
    pos, orient, x = _f64(pos).ravel(), _f64(orient).ravel(), _f64(x).ravel()
    N = pos.size // 3
    pv, wv, lv = _setup(M, noqi)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    y = np.empty(3 * M * t.size)
    lib().orc_apply_rows(int(analytic), eps, N, M, _p(pos), _p(orient), pv.size, _p(pv), _p(wv), _p(lv), _p(x),
                         t.size, t.ctypes.data_as(_ip), _p(y))
    return y


def compute_velocities_rows(analytic, eps, pos, orient, force, M, noqi, coeff, targets, sh=0.0):
    pos, orient, force, coeff = (_f64(a).ravel() for a in (pos, orient, force, coeff))
    N = pos.size // 3
    pv, wv, lv = _setup(M, noqi)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    vel = np.empty(3 * t.size)
    rot = np.empty(3 * t.size)
    lib().orc_compute_velocities_rows(int(analytic), eps, N, M, _p(pos), _p(orient), _p(force), pv.size, _p(pv),
                                      _p(wv), _p(lv), _p(coeff), t.size, t.ctypes.data_as(_ip), _p(vel), _p(rot),
                                      float(sh))
    return vel, rot


def matrix_block(analytic, eps, xi, ti, xj, tj, M, noqi):
    """Block (i, j) of A: the off-diagonal block of the two-fiber system."""
    pos = np.concatenate([_f64(xi), _f64(xj)])
    ori = np.concatenate([_f64(ti), _f64(tj)])
    A = assemble_matrix(analytic, eps, pos, ori, M, noqi)
    m = 3 * M
    return A[:m, m:]


def lu_solve(A, b):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    x = _f64(b).copy()
    rc = lib().orc_lu_solve(x.size, A.ctypes.data_as(_dp), _p(x))
    if rc:
        raise np.linalg.LinAlgError("singular matrix")
    return x


def gmres(A, b, x0, restart, max_iters, tol):
    A = np.asfortranarray(A, dtype=np.float64)
    b = _f64(b)
    x = _f64(x0).copy()
    it = ctypes.c_int(0)
    be = ctypes.c_double(0.0)
    st = lib().orc_gmres(b.size, A.ctypes.data_as(_dp), _p(b), _p(x), restart, max_iters, tol,
                         ctypes.byref(it), ctypes.byref(be))
    return x, st, it.value, be.value


def simulate(analytic, direct, eps, dt, steps, pos0, orient0, M=5, noqi=8, restart=10, max_iters=100,
             tol=1e-5, force=None, sh=0.0):
    """Runs the reference time loop; returns dict of per-step arrays."""
    pos0, orient0 = _f64(pos0).ravel(), _f64(orient0).ravel()
    N = pos0.size // 3
    n = 3 * M * N
    if force is None:
        force = np.tile([0.0, 0.0, -1.0], N)
    force = _f64(force).ravel()
    out = {k: np.empty((steps, n)) for k in ("b", "x")}
    out.update({k: np.empty((steps, 3 * N)) for k in ("vel", "rot", "pos", "orient")})
    iters = np.zeros(steps, dtype=np.int32)
    rc = lib().orc_simulate(int(analytic), int(direct), eps, dt, steps, N, M, noqi, restart, max_iters, tol,
                            _p(force), _p(pos0), _p(orient0), _p(out["b"]), _p(out["x"]), _p(out["vel"]),
                            _p(out["rot"]), _p(out["pos"]), _p(out["orient"]),
                            iters.ctypes.data_as(_ip), float(sh))
    if rc:
        raise RuntimeError("oracle simulate failed rc=%d" % rc)
    out["iters"] = iters
    return out
