/*
 * rf_oracle.h -- FP64 CPU restatement of the rigid-fiber slender-body
 * timestep of ericwolter/rigidfibers_kth (the Fortran code under fortran/).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (rigidfibers_kth_amd) never links or calls it.
 *
 * Conventions (all FP64, 0-based, caller-allocated):
 *   N fibers, M Legendre force terms, LQ = 3*NoQI quadrature points,
 *   n = 3*M*N unknowns.
 *   pos/orient/force/vel/rot : 3N, fiber-major xyz (Fortran XcVecs(3M)).
 *   coefficient vector       : n, index 3*M*i + 3*(k-1) + c
 *                              (fortran/matrix_assembly.f90:312-313).
 *   matrix                   : n x n column-major, A[r + c*n]
 *                              (fortran/matrix_assembly.f90:352-395).
 */
#ifndef RF_ORACLE_H_
#define RF_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* fortran/quadrature.f90:6-96 */
void orc_quad_pts_wts(int noqi, double *pv, double *wv);
/* fortran/matrix_assembly.f90:8-57 ; lvec is LQ x M column-major (P_1..P_M). */
int orc_legendre(int M, int LQ, const double *pv, double *lvec);

/* fortran/Int_analytical.f90:5-293.  Writes L11..L24 for l = 0..lmax into
 * L[7*(lmax+1)] ordered L11[0..lmax], L12[..], L13, L14, L22, L23, L24.
 * Also returns the raw moments I, J, S (30 each, moment index 0..29). */
void orc_analytical_int(const double xb[3], const double pb[3], const double xbar[3],
                        int lmax, double *L, double *I, double *J, double *S);

/* fortran/Int_analytical.f90:663-767: 6 comps (G11,G22,G33,G12,G13,G23) of
 * the analytic inner integral against P_l, l>=0 (l=0 is G_f_GQ_kg's tensor). */
void orc_G_analytic(const double xb[3], const double pb[3], const double xbar[3],
                    double eps, int l, double out6[6]);
/* fortran/quadrature.f90:100-164 (numeric; lvec_m = P_l at the points). */
void orc_G_numeric(const double xb[3], const double pb[3], const double xbar[3],
                   double eps, int LQ, const double *pv, const double *wv,
                   const double *lvec_m, double out6[6]);

/* Dense matrix A (identity + off-diagonal blocks).
 * analytic=1: fortran/matrix_assembly.f90:407-629; 0: :267-404. */
void orc_assemble_matrix(int analytic, double eps, int N, int M,
                         const double *pos, const double *orient,
                         int LQ, const double *pv, const double *wv,
                         const double *lvec, double *A);
/* RHS b.  analytic=1: fortran/matrix_assembly.f90:796-905; 0: :632-794. */
void orc_assemble_rhs(int analytic, double eps, int N, int M,
                      const double *pos, const double *orient, const double *force,
                      int LQ, const double *pv, const double *wv,
                      const double *lvec, double *b);
/* Rows of b for the fibers targets[0..ntargets) only (b: 3M*ntargets). */
void orc_assemble_rhs_rows(int analytic, double eps, int N, int M,
                           const double *pos, const double *orient, const double *force,
                           int LQ, const double *pv, const double *wv,
                           const double *lvec, int ntargets, const int *targets, double *b);
/* Rows of y = A x for the fibers targets[0..ntargets) (y: 3M*ntargets),
 * block by block without forming A (analytic or numeric inner integrals). */
void orc_apply_rows(int analytic, double eps, int N, int M,
                    const double *pos, const double *orient,
                    int LQ, const double *pv, const double *wv, const double *lvec,
                    const double *x, int ntargets, const int *targets, double *y);
/* Matrix-free y = A x (numeric kernel): fortran/matrix_assembly.f90:159-265. */
void orc_mat_prod(double eps, int N, int M, const double *pos, const double *orient,
                  int LQ, const double *pv, const double *wv, const double *lvec,
                  const double *x, double *y);

/* Velocities. analytic=1: fortran/coeffs_and_vels.f90:647-746; 0: :543-644.
 * sh: rate of the background shear flow (sh y, 0, 0), the reference's `sh`
 * (hard-coded 0 there, fortran/coeffs_and_vels.f90:574,597-600). */
void orc_compute_velocities(int analytic, double eps, int N, int M,
                            const double *pos, const double *orient, const double *force,
                            int LQ, const double *pv, const double *wv, const double *lvec,
                            const double *coeff, double *vel, double *rot, double sh);

/* V, Omega of the fibers targets[0..ntargets) only (vel, rot: 3*ntargets). */
void orc_compute_velocities_rows(int analytic, double eps, int N, int M,
                                 const double *pos, const double *orient, const double *force,
                                 int LQ, const double *pv, const double *wv, const double *lvec,
                                 const double *coeff, int ntargets, const int *targets,
                                 double *vel, double *rot, double sh);

/* LU with partial pivoting; overwrites A with factors and b with x.
 * Stands in for LAPACK dgesv (fortran/coeffs_and_vels.f90:62,158).
 * Returns 0, or k>0 if U(k,k) is exactly zero. */
int orc_lu_solve(int n, double *A, double *b);

/* Restarted GMRES with the CERFACS dGMRES semantics used by the reference
 * (fortran/coeffs_and_vels.f90:253-339, fortran/dGMRES.f:491-1495):
 * MGS, Givens rotations, stop when the Arnoldi backward error
 * |g_{j+1}|/||b|| <= tol and the true one ||b-Ax||/||b|| <= tol, true
 * residual at each restart, x holds the initial guess on entry.
 * Returns 0 (converged) or -4 (max_iters reached without convergence);
 * *iters = iterations, *be = final true backward error. */
int orc_gmres(int n, const double *A, const double *b, double *x,
              int restart, int max_iters, double tol, int *iters, double *be);

/* The time-stepping driver: fortran/advect_fibers.f90:437-593 with the
 * K-step loop (no_ts steps), Euler first step, 2nd-order multistep after,
 * warm start init_guess = previous coefficients (:464).
 * direct=1 -> LU, 0 -> GMRES(restart,max_iters,tol), with LU fallback on -4.
 * Per-step outputs (may be NULL): b_out/x_out [steps*n], vel_out/rot_out/
 * pos_out/orient_out [steps*3N]; iters_out[steps].
 * pos/orient are inputs (orient normalised on entry like :332-340). */
int orc_simulate(int analytic, int direct, double eps, double dt, int steps,
                 int N, int M, int noqi, int restart, int max_iters, double tol,
                 const double *force, const double *pos0, const double *orient0,
                 double *b_out, double *x_out, double *vel_out, double *rot_out,
                 double *pos_out, double *orient_out, int *iters_out, double sh);

/* Threads OpenMP will use (for reporting the CPU baseline core count). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
