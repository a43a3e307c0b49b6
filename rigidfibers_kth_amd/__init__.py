"""MI355X-native rigid-fiber slender-body timestep (Python host mirror).

The compute path is ``librigidfibers.so`` (HIP for gfx950, C ABI declared in
``include/rigidfibers.h``).  This module binds it with ctypes and mirrors the
reference's host interface:

* ``Parameters.parse_configuration_files(parameters, layout)`` --
  src/parameters.cc:24-43 (layout first, then the v1 parameter file);
* ``Simulation(configuration)`` / ``Simulation.step()`` --
  src/simulation.h:36-87, src/simulation.cu:279-301.

There is no CPU fallback: every call goes to the HIP library, and the first
call raises if the library has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librigidfibers.so")

RF_OK = 0
RF_EINVAL, RF_EIO, RF_ENOMEM, RF_EHIP, RF_ENOCONV = -1, -2, -3, -4, -5
PREC_FP64, PREC_MIXED = 0, 1
WARM_PREV, WARM_ZERO, WARM_EXTRAP, WARM_LEARNED = 0, 1, 2, 3
FMT_F16_8, FMT_F24_16 = 0, 1
SOLVER_GMRES, SOLVER_LU, SOLVER_LU_FALLBACK = 0, 1, 2
PRECOND_NONE, PRECOND_NEIGHBOUR, PRECOND_LU, PRECOND_AUTO = 0, 1, 2, 3
OP_AX, OP_RESIDUAL, OP_RHS, OP_VELOCITY = 0, 1, 2, 3


class rf_params(ctypes.Structure):
    _fields_ = [
        ("run_label", ctypes.c_int32),
        ("layout_label", ctypes.c_int32),
        ("num_terms", ctypes.c_int32),
        ("slenderness", ctypes.c_double),
        ("timestep", ctypes.c_double),
        ("num_timesteps", ctypes.c_int32),
        ("save_interval", ctypes.c_int32),
        ("saves_per_file", ctypes.c_int32),
        ("num_quad_intervals", ctypes.c_int32),
        ("num_quad_points_per_interval", ctypes.c_int32),
        ("use_analytical_integration", ctypes.c_int32),
        ("use_direct_solver", ctypes.c_int32),
        ("gmres_restart", ctypes.c_int32),
        ("gmres_max_iters", ctypes.c_int32),
        ("gmres_tol", ctypes.c_double),
        ("external_force", ctypes.c_double * 3),
        ("precision", ctypes.c_int32),
        ("warm_start", ctypes.c_int32),
        ("profile", ctypes.c_int32),
        ("lu_block", ctypes.c_int32),
        ("mixed_inner_tol", ctypes.c_double),
        ("lu_valu_update", ctypes.c_int32),
        ("precond", ctypes.c_int32),
        ("precond_neighbours", ctypes.c_int32),
        ("precond_switch_iters", ctypes.c_int32),
        ("lu_factor_fp32", ctypes.c_int32),
        ("lu_column_panel", ctypes.c_int32),
        ("lu_update_fp16", ctypes.c_int32),
        ("refine_tol", ctypes.c_double),
        ("shear_rate", ctypes.c_double),
    ]


class rf_stats(ctypes.Structure):
    _fields_ = [
        ("steps_taken", ctypes.c_int64),
        ("last_gmres_iters", ctypes.c_int32),
        ("last_solver", ctypes.c_int32),
        ("last_backward_error", ctypes.c_double),
        ("total_gmres_iters", ctypes.c_int64),
        ("ms_assemble", ctypes.c_double),
        ("ms_solve", ctypes.c_double),
        ("ms_velocities", ctypes.c_double),
        ("ms_advance", ctypes.c_double),
        ("total_ms_assemble", ctypes.c_double),
        ("total_ms_solve", ctypes.c_double),
        ("total_ms_velocities", ctypes.c_double),
        ("total_ms_advance", ctypes.c_double),
        ("gemv_launches", ctypes.c_int64),
        ("gemv_ms_total", ctypes.c_double),
        ("assemble_launches", ctypes.c_int64),
        ("assemble_ms_total", ctypes.c_double),
        ("velocity_gemv_launches", ctypes.c_int64),
        ("velocity_gemv_ms_total", ctypes.c_double),
        ("lu_factorizations", ctypes.c_int64),
        ("lu_ms_total", ctypes.c_double),
        ("matrix_bytes", ctypes.c_double),
        ("matrix_elem_bytes", ctypes.c_int32),
        ("lda", ctypes.c_int32),
        ("last_initial_residual", ctypes.c_double),
        ("last_guess_alpha", ctypes.c_double),
        ("last_guess_beta", ctypes.c_double),
        ("last_precond", ctypes.c_int32),
        ("last_neighbour_iters", ctypes.c_int32),
        ("lu_panel_fallbacks", ctypes.c_int64),
        ("operator_launches", ctypes.c_int64),
        ("operator_ms_total", ctypes.c_double),
        ("last_refinements", ctypes.c_int32),
        ("refine_inner_iters", ctypes.c_int32 * 8),
        ("last_outer_iters", ctypes.c_int32),
        ("refine_history", ctypes.c_double * 8),
        ("last_operator_noise", ctypes.c_double),
        ("last_noise_limited", ctypes.c_int32),
        ("reserved_i", ctypes.c_int32),
        ("reserved", ctypes.c_double * 2),
    ]

    def as_dict(self):
        out = {}
        for f in self._fields_:
            if f[0].startswith("reserved"):
                continue
            v = getattr(self, f[0])
            out[f[0]] = list(v) if isinstance(v, ctypes.Array) else v
        return out


_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_lib = None

# learned initial guess: features per fiber, coefficients per fiber
MLP_INPUTS, MLP_OUTPUTS, MLP_NEIGHBOURS = 113, 15, 16


class RigidFibersError(RuntimeError):
    def __init__(self, code, where):
        msg = lib().rf_last_error().decode(errors="replace")
        super().__init__("%s failed (%d): %s" % (where, code, msg))
        self.code = code


# exported symbols and their argument types (include/rigidfibers.h)
_SIGNATURES = {
    "rf_params_default": (None, [ctypes.POINTER(rf_params)]),
    "rf_parse_parameters": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(rf_params)]),
    "rf_parse_layout": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(_dp), ctypes.POINTER(_dp)]),
    "rf_free": (None, [ctypes.c_void_p]),
    "rf_create": (ctypes.c_int, [ctypes.POINTER(rf_params), ctypes.c_int32, _dp, _dp,
                                 ctypes.POINTER(ctypes.c_void_p)]),
    "rf_destroy": (None, [ctypes.c_void_p]),
    "rf_step": (ctypes.c_int, [ctypes.c_void_p]),
    "rf_assemble": (ctypes.c_int, [ctypes.c_void_p]),
    "rf_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "rf_update_velocities": (ctypes.c_int, [ctypes.c_void_p]),
    "rf_advance": (ctypes.c_int, [ctypes.c_void_p]),
    "rf_get_state": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "rf_get_velocities": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "rf_get_coefficients": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "rf_get_rhs": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "rf_get_matrix": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int64]),
    "rf_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(rf_stats)]),
    "rf_get_matrix_block": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                           ctypes.c_int64, _dp, ctypes.c_int64]),
    "rf_set_coefficients": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "rf_matvec": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "rf_operator": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp, _dp]),
    "rf_operator_noalloc": (ctypes.c_int, [ctypes.POINTER(rf_params), ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p]),
    "rf_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "rf_get_neighbours": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _dp]),
    "rf_precond_apply": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "rf_selftest_copy": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _dp]),
    "rf_mlp_num_weights": (ctypes.c_int64, []),
    "rf_mlp_set_weights": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "rf_mlp_get_weights": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "rf_mlp_samples": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp]),
    "rf_mlp_predict": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "rf_mlp_forward_backward": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp, ctypes.c_int32, _fp, _dp]),
    "rf_mlp_train": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_float, ctypes.c_uint64, ctypes.c_int32, _fp]),
    "rf_write_step_outputs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32]),
    "rf_get_restart_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)] + [_dp] * 7),
    "rf_restore_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64] + [_dp] * 7),
    "rf_read_snapshot": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int32, _dp,
                                        ctypes.POINTER(ctypes.c_int32)]),
    "rf_selftest_mfma_layout": (ctypes.c_int, [ctypes.c_int32, _dp, ctypes.POINTER(ctypes.c_float)]),
    "rf_selftest_lu": (ctypes.c_int, [ctypes.c_int32] * 6 + [_dp, ctypes.POINTER(ctypes.c_int32),
                                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "rf_num_fibers": (ctypes.c_int32, [ctypes.c_void_p]),
    "rf_num_unknowns": (ctypes.c_int64, [ctypes.c_void_p]),
    "rf_last_error": (ctypes.c_char_p, []),
    "rf_version": (ctypes.c_char_p, []),
}


def lib():
    """Loads librigidfibers.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("librigidfibers.so is not built (run __graft_entry__.build() or "
                              "make -C rigidfibers_kth_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGNATURES)


def _check(rc, where):
    if rc != RF_OK:
        raise RigidFibersError(rc, where)


def _p(a):
    return a.ctypes.data_as(_dp)


def selftest_mfma_layout(mode):
    """Accumulators (64 lanes x 4) of the probe wavefront; see include/rigidfibers.h."""
    d = np.empty(256)
    f = np.empty(256, dtype=np.float32)
    _check(lib().rf_selftest_mfma_layout(mode, _p(d), f.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
           "rf_selftest_mfma_layout")
    return d.reshape(64, 4), f.reshape(64, 4).astype(np.float64)


def selftest_lu(A, elem_bytes=8, nb=0, column_panel=0, valu=0, update_fp16=0):
    """The library's LU of an arbitrary square matrix (rf_selftest_lu):
    returns (LU, ipiv, info, fallback) with LU the packed factors as LAPACK
    getrf's, ipiv 0-based."""
    a = np.array(A, dtype=np.float64, order="F", copy=True)
    n = a.shape[0]
    if a.shape != (n, n):
        raise ValueError("square matrix expected")
    ipiv = np.zeros(n, dtype=np.int32)
    info, fb = ctypes.c_int32(0), ctypes.c_int32(0)
    _check(lib().rf_selftest_lu(n, elem_bytes, nb, column_panel, valu, update_fp16, _p(a),
                                ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(info),
                                ctypes.byref(fb)), "rf_selftest_lu")
    return a, ipiv, int(info.value), int(fb.value)


def _f(a):
    return a.ctypes.data_as(_fp)


def mlp_num_weights():
    """Floats in a weight blob of the learned-guess network."""
    return int(lib().rf_mlp_num_weights())


def selftest_copy(nbytes, reps=10):
    """Average time (ms) of a plain device copy kernel streaming nbytes."""
    ms = ctypes.c_double(0.0)
    _check(lib().rf_selftest_copy(int(nbytes), int(reps), ctypes.byref(ms)), "rf_selftest_copy")
    return ms.value


def read_snapshot(path, values_per_record, record=-1):
    """(record values, number of complete records) of an XcTData /
    VelRotData / Coeffvec file (fortran/advect_fibers.f90:395-424,546-591);
    record -1 is the last."""
    nrec = ctypes.c_int32(0)
    _check(lib().rf_read_snapshot(os.fsencode(path), int(values_per_record), -1, None, ctypes.byref(nrec)),
           "rf_read_snapshot")
    out = np.empty(int(values_per_record))
    _check(lib().rf_read_snapshot(os.fsencode(path), int(values_per_record), int(record), _p(out),
                                  ctypes.byref(nrec)), "rf_read_snapshot")
    return out, int(nrec.value)


def default_params(**kw):
    p = rf_params()
    lib().rf_params_default(ctypes.byref(p))
    names = {f[0] for f in rf_params._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError("unknown run parameter %r" % k)
        if k == "external_force":
            for i in range(3):
                p.external_force[i] = v[i]
        else:
            setattr(p, k, v)
    return p


class Configuration:
    """FiberParams + initial layout (src/parameters.h:27-31)."""

    def __init__(self, parameters, positions, orientations):
        self.parameters = parameters
        self.initial_positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1)
        self.initial_orientations = np.ascontiguousarray(orientations, dtype=np.float64).reshape(-1)

    @property
    def num_fibers(self):
        return self.initial_positions.size // 3


class Parameters:
    """src/parameters.{h,cc}: v1 parameter and layout files."""

    @staticmethod
    def parse_parameter_file(path):
        p = rf_params()
        _check(lib().rf_parse_parameters(os.fsencode(path), ctypes.byref(p)), "rf_parse_parameters")
        return p

    @staticmethod
    def parse_initial_layout_file(path):
        n = ctypes.c_int32(0)
        pos = _dp()
        ori = _dp()
        _check(lib().rf_parse_layout(os.fsencode(path), ctypes.byref(n), ctypes.byref(pos), ctypes.byref(ori)),
               "rf_parse_layout")
        N = n.value
        P = np.ctypeslib.as_array(pos, shape=(3 * N,)).copy()
        T = np.ctypeslib.as_array(ori, shape=(3 * N,)).copy()
        lib().rf_free(ctypes.cast(pos, ctypes.c_void_p))
        lib().rf_free(ctypes.cast(ori, ctypes.c_void_p))
        return P, T

    @staticmethod
    def parse_configuration_files(parameters_filename, layout_filename):
        pos, ori = Parameters.parse_initial_layout_file(layout_filename)
        params = Parameters.parse_parameter_file(parameters_filename)
        params.num_quad_points_per_interval = 3
        return Configuration(params, pos, ori)


class Simulation:
    """One rf_sim handle (src/simulation.h:36-87)."""

    def __init__(self, configuration):
        self.configuration = configuration
        self._h = ctypes.c_void_p()
        _check(lib().rf_create(ctypes.byref(configuration.parameters), configuration.num_fibers,
                               _p(configuration.initial_positions), _p(configuration.initial_orientations),
                               ctypes.byref(self._h)), "rf_create")
        self.N = lib().rf_num_fibers(self._h)
        self.n = lib().rf_num_unknowns(self._h)

    def close(self):
        if self._h:
            lib().rf_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- phases (Simulation::step, src/simulation.cu:279-301) ---
    def step(self, current_timestep=None):
        _check(lib().rf_step(self._h), "rf_step")

    def assemble_system(self):
        _check(lib().rf_assemble(self._h), "rf_assemble")

    def solve_system(self):
        _check(lib().rf_solve(self._h), "rf_solve")

    def update_velocities(self):
        _check(lib().rf_update_velocities(self._h), "rf_update_velocities")

    def update_fibers(self):
        _check(lib().rf_advance(self._h), "rf_advance")

    # --- state ---
    def state(self):
        pos = np.empty(3 * self.N)
        ori = np.empty(3 * self.N)
        _check(lib().rf_get_state(self._h, _p(pos), _p(ori)), "rf_get_state")
        return pos, ori

    def velocities(self):
        v = np.empty(3 * self.N)
        w = np.empty(3 * self.N)
        _check(lib().rf_get_velocities(self._h, _p(v), _p(w)), "rf_get_velocities")
        return v, w

    def coefficients(self):
        x = np.empty(self.n)
        _check(lib().rf_get_coefficients(self._h, _p(x)), "rf_get_coefficients")
        return x

    def set_coefficients(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().rf_set_coefficients(self._h, _p(x)), "rf_set_coefficients")

    def rhs(self):
        b = np.empty(self.n)
        _check(lib().rf_get_rhs(self._h, _p(b)), "rf_get_rhs")
        return b

    def matrix(self):
        A = np.empty((self.n, self.n), order="F")
        _check(lib().rf_get_matrix(self._h, A.ctypes.data_as(_dp), self.n), "rf_get_matrix")
        return A

    def matrix_block(self, row0, col0, rows, cols):
        """A[row0:row0+rows, col0:col0+cols] promoted to FP64."""
        out = np.empty((cols, rows))
        _check(lib().rf_get_matrix_block(self._h, row0, col0, rows, cols, out.ctypes.data_as(_dp), rows),
               "rf_get_matrix_block")
        return out.T.copy()

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        _check(lib().rf_matvec(self._h, _p(x), _p(y)), "rf_matvec")
        return y

    def operator(self, mode, x=None):
        """The matrix-free FP64 operator (rf_operator): OP_AX -> A x,
        OP_RESIDUAL -> b - A x, OP_RHS -> b, OP_VELOCITY -> (V, Omega) of the
        coefficients x."""
        if x is None:
            if mode != OP_RHS:
                raise ValueError("x is needed for this mode")
            xp = None
        else:
            x = np.ascontiguousarray(x, dtype=np.float64)
            xp = _p(x)
        y = np.empty(6 * self.N if mode == OP_VELOCITY else self.n)
        _check(lib().rf_operator(self._h, mode, xp, _p(y)), "rf_operator")
        if mode == OP_VELOCITY:
            return y[:3 * self.N].copy(), y[3 * self.N:].copy()
        return y

    def mat_prod(self, x):
        """A x without the stored matrix (mat_prod, fortran/matrix_assembly.f90:159-265)."""
        return self.operator(OP_AX, x)

    def neighbours(self, k):
        nbr = np.empty(self.N * k, dtype=np.int32)
        dist = np.empty(self.N * k)
        _check(lib().rf_get_neighbours(self._h, k, nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _p(dist)),
               "rf_get_neighbours")
        return nbr.reshape(self.N, k), dist.reshape(self.N, k)

    def precond_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.n)
        _check(lib().rf_precond_apply(self._h, _p(r), _p(z)), "rf_precond_apply")
        return z

    # -- learned initial guess (no counterpart in the reference) -------------

    def mlp_samples(self):
        """(features [N,113], local-frame coefficients [N,15]) of the current
        configuration and its solved coefficients (after solve_system)."""
        X = np.empty((self.N, MLP_INPUTS), dtype=np.float32)
        Y = np.empty((self.N, MLP_OUTPUTS), dtype=np.float32)
        _check(lib().rf_mlp_samples(self._h, _f(X), _f(Y)), "rf_mlp_samples")
        return X, Y

    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != mlp_num_weights():
            raise ValueError("weight blob must hold %d floats" % mlp_num_weights())
        _check(lib().rf_mlp_set_weights(self._h, _f(w)), "rf_mlp_set_weights")

    def mlp_weights(self):
        w = np.empty(mlp_num_weights(), dtype=np.float32)
        _check(lib().rf_mlp_get_weights(self._h, _f(w)), "rf_mlp_get_weights")
        return w

    def mlp_predict(self):
        """Network prediction of the coefficient vector (n doubles, global frame)."""
        x = np.empty(self.n)
        _check(lib().rf_mlp_predict(self._h, _p(x)), "rf_mlp_predict")
        return x

    def mlp_forward_backward(self, X, Y):
        """(loss, gradient blob) of the loaded network on one batch."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.empty(mlp_num_weights(), dtype=np.float32)
        loss = ctypes.c_double(0.0)
        _check(lib().rf_mlp_forward_backward(self._h, _f(X), _f(Y), X.shape[0], _f(g), ctypes.byref(loss)),
               "rf_mlp_forward_backward")
        return loss.value, g

    def mlp_train(self, X, Y, epochs=30, batch=256, lr=1e-3, seed=0, init=True):
        """Trains the network on the GPU (Adam, cosine schedule); returns the
        mean training loss of every epoch."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        hist = np.empty(epochs, dtype=np.float32)
        _check(lib().rf_mlp_train(self._h, _f(X), _f(Y), X.shape[0], epochs, batch, lr, seed, int(init), _f(hist)),
               "rf_mlp_train")
        return hist

    # -- restart (the reference saves these levels but has no reader) ---------

    def restart_state(self):
        """Everything the next step reads: steps taken, current and previous
        positions / orientations, last velocities, last coefficients."""
        st = ctypes.c_int64(0)
        a = {k: np.empty(3 * self.N) for k in ("pos", "orient", "pos_prev", "orient_prev", "vel", "rot")}
        a["coeffs"] = np.empty(self.n)
        _check(lib().rf_get_restart_state(self._h, ctypes.byref(st), _p(a["pos"]), _p(a["orient"]),
                                          _p(a["pos_prev"]), _p(a["orient_prev"]), _p(a["vel"]), _p(a["rot"]),
                                          _p(a["coeffs"])), "rf_get_restart_state")
        a["steps_taken"] = int(st.value)
        return a

    def restore_state(self, state):
        """Loads a restart_state() dict (or one read from snapshot files)."""
        args = [np.ascontiguousarray(state[k], dtype=np.float64) if state.get(k) is not None else None
                for k in ("pos", "orient", "pos_prev", "orient_prev", "vel", "rot", "coeffs")]
        ptrs = [_p(a) if a is not None else None for a in args]
        _check(lib().rf_restore_state(self._h, int(state["steps_taken"]), *ptrs), "rf_restore_state")

    def synchronize(self):
        _check(lib().rf_synchronize(self._h), "rf_synchronize")

    def stats(self):
        s = rf_stats()
        _check(lib().rf_get_stats(self._h, ctypes.byref(s)), "rf_get_stats")
        return s

    def write_step_outputs(self, directory, fmt=FMT_F16_8):
        _check(lib().rf_write_step_outputs(self._h, os.fsencode(directory), fmt), "rf_write_step_outputs")
