// rf_sim.h -- definition of the opaque rf_sim handle (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "rf_internal.h"
#include "rf_mlp.h"

struct rf_sim {
    rf_params p{};
    int N = 0, M = 0, LQ = 0, n = 0;
    int mixed = 0;
    int restart = 10;
    int num_cus = 256;
    int64_t lda = 0, ldw = 0;
    int wcols = 0;
    int asm_splits = 1;
    rf::Tables tab;
    rf::GemvPlan planA, planW;
    rf::OperatorPlan planOp;
    hipStream_t stream = nullptr;
    hipEvent_t ev_p0 = nullptr, ev_p1 = nullptr;
    std::vector<hipEvent_t> events;
    std::vector<void*> allocs;

    // time levels (fortran/advect_fibers.f90:294-298): lvl_three is current
    int lvl_one = 0, lvl_two = 1, lvl_three = 2;
    int vel_old = 0, vel_new = 1;
    int64_t steps_taken = 0;
    int64_t solves_done = 0;
    bool assembled = false, solved = false, have_velocities = false;
    // the operator's split partials hold the final FP64 residual of the last
    // mixed solve at the current x, with source vectors equal to the
    // velocity's: update_velocities needs only the epilogue
    bool op_velocity_ready = false;
    double last_bea = 0.0;

    double* d_pos[3] = {nullptr, nullptr, nullptr};
    double* d_orient[3] = {nullptr, nullptr, nullptr};
    double* d_vel[2] = {nullptr, nullptr};
    double* d_rot[2] = {nullptr, nullptr};
    double* d_force = nullptr;
    double* d_wp = nullptr;
    double* d_pv = nullptr;
    void* d_A = nullptr;  // n x n, lda, double or float
    void* d_W = nullptr;  // 6N x 3(M+1)N, ldw
    double* d_b = nullptr;
    double* d_bpart = nullptr;
    double* d_x = nullptr;
    double* d_xprev = nullptr;
    double* d_guess = nullptr;
    double* d_fext = nullptr;
    double* d_y6 = nullptr;
    // matrix-free operator: scratch, and the mixed mode's FP64 residual and correction
    double* d_op = nullptr;
    double* d_res64 = nullptr;
    double* d_dx = nullptr;
    double cur_tol = 1e-10;  // stopping tolerance of the solve in progress
    // GMRES workspace
    double* d_V = nullptr;
    double* d_w = nullptr;
    double* d_r = nullptr;
    double* d_xc = nullptr;
    double* d_gparts = nullptr;
    double* d_dparts = nullptr;
    double* d_dots = nullptr;
    unsigned int* d_counter = nullptr;
    // mixed mode: basis and preconditioned directions of the outer flexible
    // GMRES on the FP64 operator
    static constexpr int kOuterRestart = 10;   // outer basis vectors (at most the inner restart)
    static constexpr int kMaxOuterIters = 40;  // outer iterations per solve
    static constexpr double kNoiseFactor = 16.0; // accepted backward error / measured operator noise
    int operator_noise(double bn, double* eta);
    int outer_m = 0;
    double* d_fV = nullptr;  // [outer_m + 1][lda]
    double* d_fZ = nullptr;  // [outer_m][lda]
    double* d_fw = nullptr;  // [lda]
    double* h_dots = nullptr;  // pinned; its last double holds the copy's sequence word
    unsigned host_seq = 0;     // sequence number of the last copy into h_dots
    unsigned* host_seq_word() const { return reinterpret_cast<unsigned*>(h_dots + 3 * (restart + 2) + 7); }
    int await_host_copy(bool poll = false);
    double* h_y = nullptr;     // pinned
    // neighbourhoods and preconditioner (allocated on first use)
    int knn = 0;           // neighbours kept per fiber
    int ras_k = 0;         // neighbours per preconditioner cluster
    int* d_nbr = nullptr;  // [N][knn]
    double* d_nbr_dist = nullptr;
    double* d_B = nullptr;  // [N][3M(K+1)][3M]
    double* d_z = nullptr;
    double* d_t = nullptr;
    int* d_pinfo = nullptr;
    bool prec_ready = false;
    bool ras_valid = false;  // d_B holds the neighbour preconditioner of the current matrix
    int prec_kind = 0;       // RF_PRECOND_NEIGHBOUR or RF_PRECOND_LU once prec_ready
    int lu_es = 8;           // element bytes of the LU factors
    bool nbr_valid = false;  // neighbour lists match the current configuration
    // learned initial guess (allocated on first use)
    float* d_mlp_w = nullptr;    // weight blob
    float* d_mlp_X = nullptr;    // [N][113] features
    float* d_mlp_Y = nullptr;    // [N][15] local-frame targets
    float* d_mlp_work = nullptr; // activations for N rows
    rf::MlpBuffers mlp_buf;
    bool mlp_weights = false;    // a trained network is loaded
    bool r0_ready = false;       // d_r holds b - A x for the next GMRES start
    // LU workspace (allocated on first use)
    void* d_LU = nullptr;
    int* d_ipiv = nullptr;
    int* d_info = nullptr;
    void* d_lu_ws = nullptr;   // pivot-search partials and counter
    void* d_lu_aux = nullptr;  // FP32 factors: permutation and diagonal-block inverses
    double* d_lu_tmp = nullptr;
    bool lu_factored = false;  // d_LU holds the factors of the current matrix
    bool lu_have = false;      // d_LU holds factors (of the current or an earlier matrix)
    int auto_lu_run = 0;       // RF_PRECOND_AUTO: consecutive solves that ended on the LU
    static constexpr int kAutoRetry = 8;
    bool lu_fp16 = false;      // ... with fp16-rounded trailing updates (preconditioner use only)

    rf_stats stats{};

    // per-kernel timing without synchronisation: events are recorded around
    // launches and resolved at the next point where the stream is idle.
    enum TimedKind { kGemvA = 0, kAssemble = 1, kGemvW = 2, kLU = 3, kOperator = 4 };
    struct Timed {
        hipEvent_t a, b;
        int kind;
    };
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_free;
    std::vector<Timed> ev_pending;
    int timed_pair(hipEvent_t* a, hipEvent_t* b);
    void timed_push(hipEvent_t a, hipEvent_t b, int kind);
    void resolve_timed();

    ~rf_sim();
    int alloc_all();
    int assemble();
    int solve();
    int update_velocities();
    int advance();
    int step();
    // y = A x (mode 0) or b - A x (mode 1); with x_nrm2, A (x / sqrt(*x_nrm2))
    // and the scaled vector stored to x_out
    int gemv_A(const double* x, double* y, const double* b, int mode, const double* x_nrm2 = nullptr,
               double* x_out = nullptr);
    int fetch_dots(int count);
    int norm2(const double* v, double* out);
    int solve_gmres(int* iters, double* be, int* status, int maxit);
    int solve_system(int maxit, double tol, int* iters, double* be);
    int solve_refined(int* iters, double* be);
    int apply_operator(int mode, const double* x, double* y);
    int solve_lu(double* be);
    int factorize(bool* done, bool precond_only);
    int lu_apply(double* x);
    int setup_neighbours();
    int mlp_alloc();
    int mlp_infer(double* x_out);
    int learned_guess();
    int setup_precond(int kind);
    int precond_apply(const double* r, double* z);
    int update_iterate(int k);
    int phase_begin();
    double phase_end();
    static double event_ms(hipEvent_t a, hipEvent_t b);
};
