/*
 * rigidfibers.h -- C ABI of the MI355X-native rigid-fiber slender-body
 * timestep (librigidfibers.so).
 *
 * Drop-in for the simulation surface of ericwolter/rigidfibers_kth:
 *   - the line-ordered run-parameter file  (src/parameters.cc:124-180,
 *     fortran/advect_fibers.f90:182-222),
 *   - the XcT_init layout file             (src/parameters.cc:186-227,
 *     fortran/advect_fibers.f90:192-196,316-322),
 *   - the per-step text outputs            (fortran/advect_fibers.f90:529-539,
 *     fortran/coeffs_and_vels.f90:55-72,633-642),
 *   - the Simulation object and its step   (src/simulation.h:36-87,
 *     src/simulation.cu:279-301; the Fortran loop body
 *     fortran/advect_fibers.f90:437-593).
 *
 * Plain C types only: no HIP, torch or C++ types cross this boundary.
 * All arrays are FP64 and caller-owned unless stated otherwise.
 *   positions / orientations / velocities : 3N doubles, fiber-major xyz
 *   coefficient vector x and rhs b        : n = 3*M*N doubles, index
 *                                           3*M*i + 3*(k-1) + c
 *   matrix A                               : n x n column-major
 *
 * Return codes: RF_OK (0) or a negative RF_E* code; rf_last_error() gives a
 * message for the calling thread.  One rf_sim handle per thread; each handle
 * owns its device memory and one HIP stream on the current device.
 */
#ifndef RIGIDFIBERS_H_
#define RIGIDFIBERS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_OK 0
#define RF_EINVAL (-1)  /* bad argument / parameter */
#define RF_EIO (-2)     /* file could not be opened or parsed */
#define RF_ENOMEM (-3)  /* host or device allocation failed */
#define RF_EHIP (-4)    /* HIP runtime error (incl. no GPU / kernel image missing) */
#define RF_ENOCONV (-5) /* solver failed, including the LU fallback */

/* precision modes */
#define RF_PREC_FP64 0  /* FP64 assembly, FP64 matrix, FP64 solve (reference) */
#define RF_PREC_MIXED 1 /* FP32 matrix (FP64 near field) and FP32 LU; solves refined against the FP64
                           residual of the matrix-free operator; FP64 Arnoldi and velocities */

/* initial guess for GMRES */
#define RF_WARM_PREV 0    /* previous step's coefficients (fortran/advect_fibers.f90:464) */
#define RF_WARM_ZERO 1    /* x0 = 0 every step */
#define RF_WARM_EXTRAP 2  /* 2 x_n - x_{n-1} (linear extrapolation in time) */
#define RF_WARM_LEARNED 3 /* min-residual combination of x_prev and the MLP prediction */

/* GMRES preconditioners (the reference runs GMRES without one) */
#define RF_PRECOND_NONE 0
#define RF_PRECOND_NEIGHBOUR 1 /* block-Jacobi on each fiber + its K nearest fibers (restricted additive Schwarz) */
#define RF_PRECOND_LU 2        /* LU factors of the stored matrix (FP32 in the mixed mode: GMRES-IR) */
#define RF_PRECOND_AUTO 3      /* neighbour preconditioner; the LU once its convergence rate cannot meet the budget */

/* output number formats */
#define RF_FMT_F16_8 0  /* '(*(F16.8))' as the Fortran's *.out / tests/reference *.ref */
#define RF_FMT_F24_16 1 /* '(2F24.16)' as XcTData / VelRotData / Coeffvec */

/* solver recorded in rf_stats.last_solver */
#define RF_SOLVER_GMRES 0
#define RF_SOLVER_LU 1
#define RF_SOLVER_LU_FALLBACK 2

/* Run parameters.  Lines 0..13 of the v1 parameter file, in order, then the
 * extensions this library adds (all default to the reference behaviour). */
typedef struct rf_params {
    int32_t run_label;                    /* line 0 */
    int32_t layout_label;                 /* line 1 */
    int32_t num_terms;                    /* line 2: M, 1..8 */
    double slenderness;                   /* line 3: eps */
    double timestep;                      /* line 4: dt */
    int32_t num_timesteps;                /* line 5 */
    int32_t save_interval;                /* line 6 */
    int32_t saves_per_file;               /* line 7 */
    int32_t num_quad_intervals;           /* line 8: NoQI */
    int32_t num_quad_points_per_interval; /* fixed 3 (src/parameters.cc:36) */
    int32_t use_analytical_integration;   /* line 9 */
    int32_t use_direct_solver;            /* line 10: 1 LU, 0 GMRES */
    int32_t gmres_restart;                /* line 11 */
    int32_t gmres_max_iters;              /* line 12 */
    double gmres_tol;                     /* line 13 */
    double external_force[3];             /* per fiber; (0,0,-1) (advect_fibers.f90:307-312) */
    /* extensions */
    int32_t precision;  /* RF_PREC_* */
    int32_t warm_start; /* RF_WARM_* */
    int32_t profile;    /* 1: time every GEMV / assembly launch with HIP events */
    int32_t lu_block;   /* LU panel width (0 = default) */
    double mixed_inner_tol; /* mixed mode: relative tolerance of each inner solve on the FP32 matrix between
                               two FP64 residuals of the matrix-free operator (0 = default 1e-6) */
    int32_t lu_valu_update; /* 1: LU trailing update on the vector ALUs instead of the matrix cores */
    int32_t precond;        /* RF_PRECOND_*: right preconditioner of GMRES */
    int32_t precond_neighbours; /* K nearest fibers per preconditioner cluster (1..6, 0 = 4) */
    int32_t precond_switch_iters; /* RF_PRECOND_AUTO: neighbour-preconditioned iterations before the LU (0 = 2 restarts);
                                     after a solve that ended on the LU the next start on it, every 8th
                                     tries the neighbour preconditioner first again */
    int32_t lu_factor_fp32;       /* 1: FP32 LU factors also for the FP64 matrix (preconditioner / refinement) */
    int32_t lu_column_panel;      /* 1: LU pivot search over whole columns in every panel (0: inside the
                                     panel's diagonal block first; a panel where partial pivoting would
                                     choose another row is redone over whole columns) */
    int32_t lu_update_fp16;       /* 1: FP32 factors (mixed mode or lu_factor_fp32) that precondition GMRES
                                     take trailing updates on fp16-rounded operands (FP32 accumulation);
                                     the direct solver's factors keep FP32 updates */
    double refine_tol;            /* GMRES tolerance once the LU preconditioner is in use (0 = gmres_tol) */
    double shear_rate;            /* background shear flow (shear_rate * y, 0, 0) added to the velocities: the
                                     reference's `sh`, hard-coded 0 there (fortran/coeffs_and_vels.f90:574,
                                     597-600); like the reference it enters the velocities only */
} rf_params;

typedef struct rf_stats {
    int64_t steps_taken;
    int32_t last_gmres_iters;      /* iterations of the last solve (0 for LU) */
    int32_t last_solver;           /* RF_SOLVER_* */
    double last_backward_error;    /* ||b - A x|| / ||b|| of the last solve */
    int64_t total_gmres_iters;
    /* phase times of the last step and totals, milliseconds (HIP events) */
    double ms_assemble, ms_solve, ms_velocities, ms_advance;
    double total_ms_assemble, total_ms_solve, total_ms_velocities, total_ms_advance;
    /* per-kernel timing, filled when rf_params.profile != 0 */
    int64_t gemv_launches;          /* system-matrix GEMV launches (GMRES + residuals) */
    double gemv_ms_total;
    int64_t assemble_launches;      /* assembly kernel launches */
    double assemble_ms_total;
    int64_t velocity_gemv_launches;
    double velocity_gemv_ms_total;
    int64_t lu_factorizations;
    double lu_ms_total;
    double matrix_bytes;            /* bytes of one system-matrix copy (n*lda*elem) */
    int32_t matrix_elem_bytes;      /* 8 FP64, 4 FP32 */
    int32_t lda;
    double last_initial_residual;   /* ||b - A x0|| / ||b|| of the last GMRES solve */
    double last_guess_alpha;        /* learned guess: x0 = alpha x_prev + beta MLP */
    double last_guess_beta;
    int32_t last_precond;           /* RF_PRECOND_* in use at the end of the last GMRES solve */
    int32_t last_neighbour_iters;   /* iterations of the last GMRES solve under the neighbour preconditioner */
    int64_t lu_panel_fallbacks;     /* LU panels redone with the pivot search over whole columns */
    int64_t operator_launches;      /* matrix-free operator evaluations (profile: timed) */
    double operator_ms_total;
    int32_t last_refinements;       /* mixed mode: FP64 residuals (matrix-free operator) of the last solve */
    int32_t refine_inner_iters[8];  /* mixed mode: inner iterations (FP32 matrix) of the first 8 outer iterations */
    int32_t last_outer_iters;       /* mixed mode: outer flexible-GMRES iterations on the FP64 operator */
    double refine_history[8];       /* mixed mode: [0] FP64 backward error ||b - A x0|| / ||b|| of the initial
                                       guess, [i] the outer Arnoldi estimate after outer iteration i */
    double last_operator_noise;     /* mixed mode: rounding noise of the FP64 operator measured in the last
                                       solve, ||A x - A x1 - A x2|| / ||b|| for x = x1 + x2 (0: not measured;
                                       measured only when the outer Arnoldi estimate met gmres_tol and the
                                       FP64 residual did not) */
    int32_t last_noise_limited;     /* mixed mode: 1 when the last solve was accepted at the operator's noise
                                       floor (backward error above gmres_tol, at most 16x the noise) */
    int32_t reserved_i;
    double reserved[2];
} rf_stats;

typedef struct rf_sim rf_sim;

/* Defaults: M=5, eps=0.01, dt=0.1, NoQI=8, analytic, GMRES(10,100,1e-5),
 * force (0,0,-1), FP64, previous-step warm start. */
void rf_params_default(rf_params *p);

/* v1 parameter file (src/parameters.cc:124-180; fortran/advect_fibers.f90:182-222).
 * Lines after the last mandatory one may hold "key value" extensions:
 * precision fp64|mixed, warm_start prev|zero|extrap|learned, profile 0|1,
 * lu_block NB, precond none|neighbour|lu|auto, precond_neighbours K, precond_switch_iters I,
 * lu_factor_fp32 0|1, lu_column_panel 0|1, lu_update_fp16 0|1, refine_tol T, mixed_inner_tol T,
 * shear_rate S. */
int rf_parse_parameters(const char *path, rf_params *out);

/* XcT_init layout (src/parameters.cc:186-227): N, then x y z and tx ty tz per
 * fiber, any whitespace.  *pos and *orient are library-allocated (3N doubles
 * each); release with rf_free. */
int rf_parse_layout(const char *path, int32_t *n, double **pos, double **orient);
void rf_free(void *p);

/* Simulation(Configuration) (src/simulation.cu:40-58).  Copies the inputs,
 * normalises the orientations (fortran/advect_fibers.f90:332-340) and
 * allocates all device memory. */
int rf_create(const rf_params *params, int32_t num_fibers, const double *pos, const double *orient,
              rf_sim **out);
void rf_destroy(rf_sim *sim);

/* One timestep: assemble, solve, velocities, advance (Euler on the first
 * call, second-order multistep afterwards).  Simulation::step
 * (src/simulation.cu:279-301); fortran/advect_fibers.f90:437-523. */
int rf_step(rf_sim *sim);

/* The four phases of rf_step, for tests and timing. */
int rf_assemble(rf_sim *sim);            /* assemble_matrix[_an] + assemble_rhs[_an] */
int rf_solve(rf_sim *sim);               /* dgesv / dGMRES (coeffs_and_vels.f90:10-532) */
int rf_update_velocities(rf_sim *sim);   /* compute_velocities[_an] (coeffs_and_vels.f90:543-746) */
int rf_advance(rf_sim *sim);             /* update_fibers[_firststep] (advect_fibers.f90:498-523) */

/* Copies out of the current state (positions of the newest time level). */
int rf_get_state(const rf_sim *sim, double *pos, double *orient);
int rf_get_velocities(const rf_sim *sim, double *vel, double *rot);
int rf_get_coefficients(const rf_sim *sim, double *x);
int rf_get_rhs(const rf_sim *sim, double *b);
/* The assembled system matrix, promoted to FP64, into a caller buffer with
 * leading dimension lda >= n (test hook; the reference's AMat.out). */
int rf_get_matrix(const rf_sim *sim, double *A, int64_t lda);
/* Rows [row0, row0+rows) x columns [col0, col0+cols) of the same matrix into
 * out (column-major, leading dimension ldo >= rows): spot checks of large
 * systems without copying the whole matrix.  No counterpart in the reference. */
int rf_get_matrix_block(const rf_sim *sim, int64_t row0, int64_t col0, int64_t rows, int64_t cols, double *out,
                        int64_t ldo);
int rf_get_stats(const rf_sim *sim, rf_stats *stats);
/* Sets the coefficient vector (used as the next GMRES initial guess). */
int rf_set_coefficients(rf_sim *sim, const double *x);

/* The k nearest fibers of every fiber by segment-segment distance at the
 * current configuration (ascending, ties by index; nbr and dist hold N*k
 * entries), 1 <= k <= min(16, N-1).  No counterpart in the reference. */
int rf_get_neighbours(rf_sim *sim, int32_t k, int32_t *nbr, double *dist);
/* z = M^-1 r with the neighbour preconditioner built by the last rf_solve
 * (n doubles each).  No counterpart in the reference. */
int rf_precond_apply(rf_sim *sim, const double *r, double *z);

/* Learned initial guess.  No counterpart in the reference (the reference
 * warm-starts GMRES from the previous step only, fortran/advect_fibers.f90:464).
 * A per-fiber network (FP32, 113 -> 128 -> 128 -> 128 -> 15, tanh) reads the
 * relative centres, orientations and proximities of the fiber's 16 nearest
 * fibers in the fiber's own frame (plus its tilt against gravity) and
 * predicts its 3M = 15 coefficients (needs M = 5 and N >= 17).  With
 * warm_start = RF_WARM_LEARNED and weights loaded, each GMRES solve starts
 * from alpha x_prev + beta x_mlp, (alpha, beta) minimising the residual.
 * Weight blob: rf_mlp_num_weights() floats, layout in_mean[113] in_std[113]
 * W1[128][113] b1[128] W2[128][128] b2 W3[128][128] b3 W4[15][128] b4[15]
 * out_mean[15] out_std[15]. */
int64_t rf_mlp_num_weights(void);
int rf_mlp_set_weights(rf_sim *sim, const float *w);
int rf_mlp_get_weights(rf_sim *sim, float *w);
/* Training pairs of the current configuration after rf_solve: X[N][113]
 * features, Y[N][15] solved coefficients in each fiber's frame. */
int rf_mlp_samples(rf_sim *sim, float *X, float *Y);
/* Network prediction of the coefficient vector (n doubles, global frame). */
int rf_mlp_predict(rf_sim *sim, double *x);
/* Mean squared error (standardised targets) and its gradient (blob layout,
 * zero in the standardisation entries) of the loaded network on B pairs. */
int rf_mlp_forward_backward(rf_sim *sim, const float *X, const float *Y, int32_t B, float *grad, double *loss);
/* Trains on `count` pairs on the GPU: minibatch Adam, cosine learning rate
 * from lr, one shuffle per epoch from seed.  init = 1 starts from
 * Glorot-uniform weights and takes the standardisation from the data;
 * init = 0 continues from the loaded weights.  loss_per_epoch (epochs
 * floats, may be NULL) receives the mean training loss of every epoch. */
int rf_mlp_train(rf_sim *sim, const float *X, const float *Y, int64_t count, int32_t epochs, int32_t batch,
                 float lr, uint64_t seed, int32_t init, float *loss_per_epoch);

/* Waits for all work queued on the handle's stream. */
int rf_synchronize(rf_sim *sim);

/* y = A x with the assembled device matrix (FP64 arithmetic), n doubles. */
int rf_matvec(rf_sim *sim, const double *x, double *y);

/* The matrix-free FP64 operator at the current configuration, without the
 * stored matrix: mat_prod (fortran/matrix_assembly.f90:159-265, numeric inner
 * integrals), here for both the numeric and the analytic inner integrals of
 * rf_params.use_analytical_integration, plus the force term of the
 * right-hand side (fortran/matrix_assembly.f90:632-905) and of the
 * velocities (fortran/coeffs_and_vels.f90:543-746).
 *   RF_OP_AX        y = A x                  (x, y: n doubles)
 *   RF_OP_RESIDUAL  y = b - A x              (x, y: n doubles)
 *   RF_OP_RHS       y = b                    (x unused, may be NULL; y: n doubles)
 *   RF_OP_VELOCITY  the velocities for the coefficients x: y[0:3N] = V, y[3N:6N] = Omega
 *                   (fiber-major xyz, as rf_get_velocities) */
#define RF_OP_AX 0
#define RF_OP_RESIDUAL 1
#define RF_OP_RHS 2
#define RF_OP_VELOCITY 3
int rf_operator(rf_sim *sim, int32_t mode, const double *x, double *y);

/* rf_operator on caller-provided device buffers, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  pos, orient, force (3N doubles each, the per-fiber
 * external force), x and y (n doubles; 6N for RF_OP_VELOCITY: V then Omega)
 * are device pointers; the work is queued on `stream` (a hipStream_t, NULL =
 * the default stream) and the call returns without waiting.  Scratch is one
 * 32 MiB device array per GPU, held until the process ends; a system that
 * does not fit (roughly N > 40000 with M = 5) is rejected with RF_EINVAL.
 * All calls on one GPU share that array: calls that overlap on one GPU (other
 * host threads, other streams) race. */
int rf_operator_noalloc(const rf_params *params, int32_t num_fibers, int32_t mode, const double *pos,
                        const double *orient, const double *force, const double *x, double *y, void *stream);

/* Restart.  Everything the next step reads, as the two-level scheme of
 * fortran/advect_fibers.f90:498-523 needs it: the current positions and
 * orientations (X^n, t^n), those of the previous level (X^{n-1}, t^{n-1}),
 * the velocities of the last step (V^{n-1}, Omega^{n-1}; the multistep
 * update's old level) and the last solved coefficients (the next initial
 * guess, fortran/advect_fibers.f90:464).  The reference saves these in
 * XcTData / VelRotData / Coeffvec (fortran/advect_fibers.f90:395-424,
 * 546-591) but has no reader.  rf_get_restart_state copies them out (3N or n
 * doubles each; at steps_taken = 0 the previous level equals the current one
 * and the velocities are zero); rf_restore_state loads them into a handle of
 * the same N and parameters, so the next rf_step continues the run: bit for
 * bit with the previous-step or zero initial guess.  pos_prev, orient_prev,
 * vel, rot may be NULL when steps_taken = 0; coeffs NULL means zero. */
int rf_get_restart_state(const rf_sim *sim, int64_t *steps_taken, double *pos, double *orient, double *pos_prev,
                         double *orient_prev, double *vel, double *rot, double *coeffs);
int rf_restore_state(rf_sim *sim, int64_t steps_taken, const double *pos, const double *orient,
                     const double *pos_prev, const double *orient_prev, const double *vel, const double *rot,
                     const double *coeffs);

/* Reader of the snapshot files (fortran/advect_fibers.f90:395-424,546-591:
 * '(2F24.16)' one value per line, a blank line after each 3-vector or
 * coefficient block): record `record` (0-based; -1 = the last) of
 * values_per_record numbers into out (may be NULL), *num_records = the number
 * of complete records in the file.  XcTData and VelRotData records hold 6N
 * values (per fiber x y z then t_x t_y t_z, resp. V then Omega), Coeffvec
 * records n values. */
int rf_read_snapshot(const char *path, int64_t values_per_record, int32_t record, double *out,
                     int32_t *num_records);

/* POS.out ORIENT.out TRANSVel.out ROTVel.out XVec.out BVec.out of the last
 * step into dir, RF_FMT_F16_8 or RF_FMT_F24_16. */
int rf_write_step_outputs(const rf_sim *sim, const char *dir, int32_t format);

/* Diagnostic with no counterpart in the reference: lane layout of the
 * 16x16x4 matrix-core instructions used by the LU trailing update.  One
 * wavefront runs v_mfma_f64_16x16x4f64 and v_mfma_f32_16x16x4f32 with
 * operands that encode row (mode 0), column (mode 1) or the K slot (mode 2);
 * out64 / out32 receive the 4 accumulator values of each of the 64 lanes
 * (lane-major, 256 values each). */
int rf_selftest_mfma_layout(int32_t mode, double *out64, float *out32);

/* Diagnostic with no counterpart in the reference: streams `bytes` bytes
 * through a plain device copy kernel `reps` times and returns the average
 * kernel time in *ms (HBM calibration for the byte counters). */
int rf_selftest_copy(int64_t bytes, int32_t reps, double *ms);

/* Diagnostic with no counterpart in the reference: the library's LU
 * factorisation (the dgesv factorisation of fortran/coeffs_and_vels.f90:62,158)
 * of an arbitrary n x n column-major matrix a, overwritten with L\U (unit L
 * below the diagonal).  elem_bytes 8: FP64; 4: a rounded to FP32 and
 * factorised in FP32.  nb: panel width (0 = 64; at most 64 in FP64, 128 in FP32).
 * column_panel 1: pivot search over whole columns; 0: inside each panel's
 * diagonal block first, and a panel where partial pivoting would have
 * chosen another row is redone column by column (*fallback = number of such
 * panels); the pivots are partial pivoting's either way.  valu 1: trailing update on
 * the vector ALUs; update_fp16 1 (FP32 only): on the matrix cores with
 * fp16-rounded operands (rf_params.lu_update_fp16).  ipiv: n 0-based rows as LAPACK's (row k swapped with
 * ipiv[k]); *info = k+1 for the first U(k,k) = 0. */
int rf_selftest_lu(int32_t n, int32_t elem_bytes, int32_t nb, int32_t column_panel, int32_t valu,
                   int32_t update_fp16, double *a, int32_t *ipiv, int32_t *info, int32_t *fallback);

/* Number of fibers / unknowns of a handle. */
int32_t rf_num_fibers(const rf_sim *sim);
int64_t rf_num_unknowns(const rf_sim *sim);

/* Message of the last failing call on this thread. */
const char *rf_last_error(void);
/* Library version string. */
const char *rf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* RIGIDFIBERS_H_ */
