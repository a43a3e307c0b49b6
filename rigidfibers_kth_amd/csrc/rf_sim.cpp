// rf_sim.cpp -- the rf_sim handle: device memory, the timestep phases and the
// solvers, behind the C ABI of include/rigidfibers.h.
//
// One handle = one HIP stream on the current device.  All state lives in HBM
// for the whole run; the host sees O(restart) scalars per GMRES iteration
// (the Hessenberg column) and nothing else inside a step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rf_internal.h"
#include "rf_lu.h"
#include "rf_precond.h"
#include "rf_sim.h"

namespace rf {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }

int assemble_targets_per_block(int LQ, int M, bool analytic);

// fortran/quadrature.f90:6-96, fortran/matrix_assembly.f90:8-57,294-305,370
void build_tables(int M, int noqi, double eps, Tables* t)
{
    t->M = M;
    t->LQ = 3 * noqi;
    t->eps = eps;
    const double c = std::log(eps * eps * std::exp(1.0));
    t->d = -c;
    t->e = 2.0;
    t->D1 = 0.75 / (t->d - 2.0);
    double lambda = 2.0;
    for (int k = 1; k <= M; ++k) {
        if (k > 1) lambda += 2.0 / k;
        t->Ek[k] = (t->d - t->e - lambda) / 2.0 / (t->d - lambda);
        t->gammak[k] = 0.5 * (2.0 * k + 1.0) / (t->d + t->e - lambda);
    }
    const double p0[3] = {-std::sqrt(15.0) / 5.0, 0.0, std::sqrt(15.0) / 5.0};
    const double w0[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    double a = -1.0;
    const double iv = 2.0 / noqi;
    for (int i = 0; i < noqi; ++i) {
        for (int q = 0; q < 3; ++q) {
            t->pv[3 * i + q] = (2.0 * a + iv + p0[q] * iv) / 2.0;
            t->wv[3 * i + q] = w0[q] / noqi;
        }
        a += iv;
    }
    for (int p = 0; p < t->LQ; ++p) {
        // Legendre three-term recurrence for P_0..P_M at the point
        const double x = t->pv[p];
        double Pkm1 = 1.0, Pk = x;
        t->wp[0 * t->LQ + p] = t->wv[p];
        if (M >= 1) t->wp[1 * t->LQ + p] = t->wv[p] * x;
        for (int k = 1; k < M; ++k) {
            const double Pk1 = ((2.0 * k + 1.0) * x * Pk - k * Pkm1) / (k + 1.0);
            Pkm1 = Pk;
            Pk = Pk1;
            t->wp[(k + 1) * t->LQ + p] = t->wv[p] * Pk;
        }
    }
}

}  // namespace rf

using namespace rf;

#define HIP_TRY(call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            rf::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                   \
            return RF_EHIP;                                                                       \
        }                                                                                         \
    } while (0)

namespace {

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

template <typename T>
int dev_alloc(T** p, size_t elems)
{
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(elems, 1) * sizeof(T));
    if (e != hipSuccess) {
        rf::set_error(std::string("hipMalloc(") + std::to_string(elems * sizeof(T)) + " bytes): " +
                      hipGetErrorString(e));
        return RF_ENOMEM;
    }
    *p = static_cast<T*>(q);
    return RF_OK;
}

struct PhaseTimer {
    hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace

rf_sim::~rf_sim()
{
    for (void* q : allocs) (void)hipFree(q);
    if (h_dots) (void)hipHostFree(h_dots);
    if (h_y) (void)hipHostFree(h_y);
    for (auto& e : events) (void)hipEventDestroy(e);
    for (auto& pr : ev_free) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (auto& t : ev_pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    if (stream) (void)hipStreamDestroy(stream);
}

int rf_sim::alloc_all()
{
#define ALLOC(ptr, count)                                 \
    do {                                                  \
        int rc_ = dev_alloc(&(ptr), (count));             \
        if (rc_) return rc_;                              \
        allocs.push_back((void*)(ptr));                   \
    } while (0)
    const size_t N3 = 3 * (size_t)N;
    for (int i = 0; i < 3; ++i) {
        ALLOC(d_pos[i], N3);
        ALLOC(d_orient[i], N3);
    }
    for (int i = 0; i < 2; ++i) {
        ALLOC(d_vel[i], N3);
        ALLOC(d_rot[i], N3);
    }
    ALLOC(d_force, N3);
    ALLOC(d_wp, (size_t)(kMaxTerms + 1) * kMaxQuad);
    ALLOC(d_pv, (size_t)kMaxQuad);
    const size_t es = mixed ? sizeof(float) : sizeof(double);
    {
        void* q = nullptr;
        if (hipMalloc(&q, (size_t)lda * n * es) != hipSuccess) {
            rf::set_error("hipMalloc of the system matrix (" + std::to_string((size_t)lda * n * es) +
                          " bytes) failed");
            return RF_ENOMEM;
        }
        allocs.push_back(q);
        d_A = q;
        // the mixed mode takes its velocities from the matrix-free operator
        if (!mixed) {
            if (hipMalloc(&q, (size_t)ldw * wcols * es) != hipSuccess) {
                rf::set_error("hipMalloc of the velocity operator failed");
                return RF_ENOMEM;
            }
            allocs.push_back(q);
            d_W = q;
        }
    }
    ALLOC(d_b, (size_t)lda);
    ALLOC(d_bpart, (size_t)asm_splits * n);
    ALLOC(d_x, (size_t)lda);
    ALLOC(d_xprev, (size_t)lda);
    ALLOC(d_guess, (size_t)lda);
    ALLOC(d_fext, (size_t)wcols);
    ALLOC(d_y6, (size_t)ldw);
    const int m = restart;
    ALLOC(d_V, (size_t)lda * (m + 1));
    ALLOC(d_w, (size_t)lda);
    ALLOC(d_r, (size_t)lda);
    ALLOC(d_xc, (size_t)lda);
    ALLOC(d_gparts, std::max(gemv_parts_elems(planA), gemv_parts_elems(planW)));
    ALLOC(d_dparts, (size_t)kMaxDotBlocks * (m + 2) + 64);
    ALLOC(d_dots, (size_t)(3 * (m + 2) + 8));
    ALLOC(d_counter, 1);
    ALLOC(d_op, operator_scratch_doubles(planOp));
    ALLOC(d_res64, (size_t)lda);
    ALLOC(d_dx, (size_t)lda);
    if (mixed) {
        outer_m = std::min(kOuterRestart, m);
        ALLOC(d_fV, (size_t)lda * (outer_m + 1));
        ALLOC(d_fZ, (size_t)lda * outer_m);
        ALLOC(d_fw, (size_t)lda);
    }
#undef ALLOC
    HIP_TRY(hipMemsetAsync(d_counter, 0, sizeof(unsigned int), stream));
    // zero whole buffers once so the padding rows of A and W read as 0
    HIP_TRY(hipMemsetAsync(d_A, 0, (size_t)lda * n * es, stream));
    if (d_W) HIP_TRY(hipMemsetAsync(d_W, 0, (size_t)ldw * wcols * es, stream));
    HIP_TRY(hipMemsetAsync(d_x, 0, sizeof(double) * lda, stream));
    HIP_TRY(hipMemsetAsync(d_xprev, 0, sizeof(double) * lda, stream));
    HIP_TRY(hipMemsetAsync(d_b, 0, sizeof(double) * lda, stream));
    HIP_TRY(hipMemsetAsync(d_V, 0, sizeof(double) * lda * (m + 1), stream));
    HIP_TRY(hipMemsetAsync(d_w, 0, sizeof(double) * lda, stream));
    HIP_TRY(hipMemsetAsync(d_r, 0, sizeof(double) * lda, stream));
    HIP_TRY(hipMemsetAsync(d_xc, 0, sizeof(double) * lda, stream));
    return RF_OK;
}

double rf_sim::event_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventSynchronize(b) != hipSuccess) return 0.0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

int rf_sim::timed_pair(hipEvent_t* a, hipEvent_t* b)
{
    if (ev_pending.size() >= 256) {  // bound the pool: drain
        HIP_TRY(hipStreamSynchronize(stream));
        resolve_timed();
    }
    if (!ev_free.empty()) {
        *a = ev_free.back().first;
        *b = ev_free.back().second;
        ev_free.pop_back();
        return RF_OK;
    }
    HIP_TRY(hipEventCreate(a));
    HIP_TRY(hipEventCreate(b));
    return RF_OK;
}

void rf_sim::timed_push(hipEvent_t a, hipEvent_t b, int kind) { ev_pending.push_back({a, b, kind}); }

// Call only when the stream is idle (after a synchronisation).
void rf_sim::resolve_timed()
{
    std::vector<Timed> running;  // pairs whose end event has not completed yet
    for (const Timed& t : ev_pending) {
        if (hipEventQuery(t.b) != hipSuccess) {
            running.push_back(t);
            continue;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess) ms = 0.f;
        switch (t.kind) {
        case kGemvA:
            stats.gemv_ms_total += ms;
            stats.gemv_launches += 1;
            break;
        case kAssemble:
            stats.assemble_ms_total += ms;
            stats.assemble_launches += 1;
            break;
        case kGemvW:
            stats.velocity_gemv_ms_total += ms;
            stats.velocity_gemv_launches += 1;
            break;
        case kLU:
            stats.lu_ms_total += ms;
            break;
        case kOperator:
            stats.operator_ms_total += ms;
            break;
        }
        ev_free.push_back({t.a, t.b});
    }
    ev_pending.swap(running);
}

int rf_sim::gemv_A(const double* x, double* y, const double* b, int mode, const double* x_nrm2, double* x_out)
{
    hipEvent_t a = nullptr, e = nullptr;
    if (p.profile) {
        int rc = timed_pair(&a, &e);
        if (rc) return rc;
    }
    HIP_TRY(launch_gemv(planA, d_A, x, d_gparts, y, b, mode, stream, a, e, x_nrm2, x_out));
    if (p.profile) timed_push(a, e, kGemvA);
    return RF_OK;
}

// The copy into h_dots whose sequence number is host_seq has landed: its
// sequence word is written after its data at system scope.  Normally it has
// by the time the stream reports completion; otherwise wait for the stream
// and look again, and report a stale copy as an error.  poll: the stream
// holds more work after the copy, so watch the sequence word itself (up to
// a second) before waiting for the whole stream.
int rf_sim::await_host_copy(bool poll)
{
    const unsigned* w = host_seq_word();
    for (int spin = 0; spin < 1024; ++spin)
        if (__atomic_load_n(w, __ATOMIC_ACQUIRE) == host_seq) return RF_OK;
    if (poll) {
        const auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            for (int spin = 0; spin < 4096; ++spin)
                if (__atomic_load_n(w, __ATOMIC_ACQUIRE) == host_seq) return RF_OK;
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) break;
        }
    }
    HIP_TRY(hipStreamSynchronize(stream));
    if (__atomic_load_n(w, __ATOMIC_ACQUIRE) == host_seq) return RF_OK;
    rf::set_error("the dot block copied to host memory is stale (sequence word " +
                  std::to_string(__atomic_load_n(w, __ATOMIC_ACQUIRE)) + ", expected " + std::to_string(host_seq) +
                  ")");
    return RF_EHIP;
}

int rf_sim::fetch_dots(int count)
{
    HIP_TRY(launch_to_host(d_dots, h_dots, sizeof(double) * count, stream, host_seq_word(), ++host_seq));
    HIP_TRY(hipStreamSynchronize(stream));
    resolve_timed();
    return await_host_copy();
}

// ||v||^2 into h_dots[0]
int rf_sim::norm2(const double* v, double* out)
{
    HIP_TRY(launch_multidot(nullptr, 0, 0, v, n, d_dparts, d_counter, d_dots, stream));
    int rc = fetch_dots(1);
    if (rc) return rc;
    *out = h_dots[0];
    return RF_OK;
}

// Upper-triangular solve H(0:k,0:k) y = g(0:k) (dtrsv 'U','N','N'), host.
static void hess_solve(const std::vector<double>& H, int ldh, const std::vector<double>& g, int k,
                       std::vector<double>& y)
{
    for (int i = 0; i < k; ++i) y[i] = g[i];
    for (int i = k - 1; i >= 0; --i) {
        y[i] = y[i] / H[i + (size_t)i * ldh];
        for (int r = 0; r < i; ++r) y[r] -= H[r + (size_t)i * ldh] * y[i];
    }
}

// drotg / drot of fortran/blas_rot.f:78-148
static void givens(double da, double db, double* c, double* s)
{
    const double roe = (std::fabs(da) > std::fabs(db)) ? da : db;
    const double scale = std::fabs(da) + std::fabs(db);
    if (scale == 0.0) {
        *c = 1.0;
        *s = 0.0;
        return;
    }
    double r = scale * std::sqrt((da / scale) * (da / scale) + (db / scale) * (db / scale));
    r = std::copysign(1.0, roe) * r;
    *c = da / r;
    *s = db / r;
}
static void rot1(double* x, double* y, double c, double s)
{
    const double t = c * (*x) + s * (*y);
    *y = c * (*y) - s * (*x);
    *x = t;
}

// xc = x + V y (no preconditioner) or x + M^-1 (V y) (right preconditioner)
int rf_sim::update_iterate(int k)
{
    if (!prec_ready) {
        HIP_TRY(launch_vy_plus_x(d_V, lda, k, h_y, d_x, n, d_xc, stream));
        return RF_OK;
    }
    HIP_TRY(launch_vy_plus_x(d_V, lda, k, h_y, nullptr, n, d_t, stream));
    int rc = precond_apply(d_t, d_z);
    if (rc) return rc;
    HIP_TRY(launch_axpby(n, 1.0, d_x, 1.0, d_z, d_xc, stream));
    return RF_OK;
}

// Restarted GMRES with the CERFACS dGMRES semantics the reference uses
// (fortran/coeffs_and_vels.f90:253-339, fortran/dGMRES.f:736-1495):
// no preconditioner, stop when the Arnoldi backward error |g_{j+1}|/||b||
// and then the true one ||b - A x||/||b|| are <= tol, true residual at each
// restart, d_x holds the initial guess.  Orthogonalisation is classical
// Gram-Schmidt applied twice (one fused multi-dot per pass instead of j
// sequential dots); in exact arithmetic it builds the same Krylov basis and
// the same iterates as the reference's modified Gram-Schmidt.  With
// rf_params.precond the same iteration runs on A M^-1 (right preconditioning,
// the stopping test still uses the true residual b - A x).
// iterations done after step k of restart cycle `cycle`
static inline int total_before(int cycle, int m, int k) { return cycle * m + k; }

int rf_sim::solve_gmres(int* iters_out, double* be_out, int* status_out, int maxit)
{
    const int m = restart;
    double tol = cur_tol;
    if (!mixed && prec_ready && prec_kind == RF_PRECOND_LU && p.refine_tol > 0.0) tol = p.refine_tol;
    const int ldh = m + 1;
    std::vector<double> H((size_t)(m + 1) * (m + 2), 0.0), g(m + 1, 0.0), cs(m + 1), sn(m + 1), y(m + 1);
    int rc;
    double bn2;
    if ((rc = norm2(d_b, &bn2))) return rc;
    const double bn = std::sqrt(bn2);
    *iters_out = 0;
    *status_out = 0;
    if (bn == 0.0) {
        HIP_TRY(hipMemsetAsync(d_x, 0, sizeof(double) * n, stream));
        *be_out = 0.0;
        return RF_OK;
    }
    // r0 = b - A x (the learned guess leaves it in d_r)
    if (r0_ready)
        r0_ready = false;
    else if ((rc = gemv_A(d_x, d_r, d_b, 1)))
        return rc;
    double beta2;
    if ((rc = norm2(d_r, &beta2))) return rc;
    double beta = std::sqrt(beta2);
    stats.last_initial_residual = beta / bn;
    if (beta == 0.0) {
        *be_out = 0.0;
        return RF_OK;
    }
    // norm2 left ||r0||^2 in d_dots[0]
    HIP_TRY(launch_scale_copy(d_r, d_dots, 0, n, d_V, stream));
    int iterOut = 0;
    const int64_t ldv = lda;
    double* D1 = d_dots;
    double* D2 = d_dots + (m + 2);
    double* D3 = d_dots + 2 * (m + 2);
    // w = A v_j, or A M^-1 v_j with the right preconditioner
    auto matvec = [&](int j) -> int {
        const double* vj = d_V + (size_t)j * lda;
        if (!prec_ready) return gemv_A(vj, d_w, nullptr, 0);
        const int r = precond_apply(vj, d_z);
        return r ? r : gemv_A(d_z, d_w, nullptr, 0);
    };
    for (;;) {
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        double bea = 0.0;
        bool have_xc = false;
        bool w_ready = false;  // d_w already holds the product for v_jH
        for (int jH = 0; jH < m; ++jH) {
            const int k = jH + 1;
            if (!w_ready && (rc = matvec(jH))) return rc;
            w_ready = false;
            HIP_TRY(launch_multidot(d_V, ldv, k, d_w, n, d_dparts, d_counter, D1, stream));
            HIP_TRY(launch_orth_dots(d_V, ldv, k, D1, d_w, n, 1, d_dparts, d_counter, D2, stream));
            // The second pass ends by copying the Hessenberg column to the
            // host.  The next basis vector is then on the device, so its
            // product is queued before the host waits (it is discarded if
            // this iteration ends the solve); without a preconditioner the
            // product forms v_k = w / |w| itself and stores it.
            HostCopy hc;
            hc.src = reinterpret_cast<const unsigned*>(d_dots);
            hc.dst = reinterpret_cast<unsigned*>(h_dots);
            hc.words = (int)(sizeof(double) / sizeof(unsigned)) * 3 * (m + 2);
            hc.seq_word = host_seq_word();
            hc.seq = ++host_seq;
            HIP_TRY(launch_orth_dots(d_V, ldv, k, D2, d_w, n, 0, d_dparts, d_counter, D3, stream, hc));
            const bool next = k < m && total_before(iterOut, m, k) < maxit;
            if (next && !prec_ready) {
                if ((rc = gemv_A(d_w, d_w, nullptr, 0, D3, d_V + (size_t)k * ldv))) return rc;
                w_ready = true;
            } else if (k < m) {
                HIP_TRY(launch_scale_copy(d_w, D3, 0, n, d_V + (size_t)k * ldv, stream));
                if (next) {
                    if ((rc = matvec(k))) return rc;
                    w_ready = true;
                }
            }
            if ((rc = await_host_copy(true))) return rc;
            for (int j = 0; j < k; ++j) H[j + (size_t)jH * ldh] = h_dots[j] + h_dots[(m + 2) + j];
            const double dnormw = std::sqrt(h_dots[2 * (m + 2)]);
            H[k + (size_t)jH * ldh] = dnormw;
            for (int j = 0; j < jH; ++j) rot1(&H[j + (size_t)jH * ldh], &H[j + 1 + (size_t)jH * ldh], cs[j], sn[j]);
            givens(H[jH + (size_t)jH * ldh], H[k + (size_t)jH * ldh], &cs[jH], &sn[jH]);
            rot1(&g[jH], &g[k], cs[jH], sn[jH]);
            rot1(&H[jH + (size_t)jH * ldh], &H[k + (size_t)jH * ldh], cs[jH], sn[jH]);
            H[k + (size_t)jH * ldh] = 0.0;
            bea = std::fabs(g[k]) / bn;
            const int total = iterOut * m + k;
            last_bea = bea;
            if (bea <= tol || total >= maxit) {
                hess_solve(H, ldh, g, k, y);
                std::copy(y.begin(), y.begin() + k, h_y);
                if ((rc = update_iterate(k))) return rc;
                have_xc = true;
                if ((rc = gemv_A(d_xc, d_r, d_b, 1))) return rc;
                double rr;
                if ((rc = norm2(d_r, &rr))) return rc;
                const double be = std::sqrt(rr) / bn;
                if (be <= tol || total >= maxit) {
                    HIP_TRY(launch_copy_vec(d_xc, d_x, n, stream));
                    *iters_out = total;
                    *be_out = be;
                    *status_out = (be <= tol) ? 0 : -4;
                    return RF_OK;
                }
            }
        }
        ++iterOut;
        if (bea > tol || !have_xc) {
            hess_solve(H, ldh, g, m, y);
            std::copy(y.begin(), y.begin() + m, h_y);
            if ((rc = update_iterate(m))) return rc;
        }
        HIP_TRY(launch_copy_vec(d_xc, d_x, n, stream));
        if ((rc = gemv_A(d_x, d_r, d_b, 1))) return rc;
        if ((rc = norm2(d_r, &beta2))) return rc;
        beta = std::sqrt(beta2);
        if (beta == 0.0) {
            *iters_out = iterOut * m;
            *be_out = 0.0;
            return RF_OK;
        }
        HIP_TRY(launch_scale_copy(d_r, d_dots, 0, n, d_V, stream));
    }
}

// LU factors of the stored matrix in d_LU / d_ipiv (once per assembled
// system: *done is set and checked).  precond_only: the factors only
// precondition GMRES, and FP32 factors may then take the fp16 trailing
// update (lu_update_fp16); the direct solver's refinement needs
// FP32-accurate factors and refactors factors made that way.
int rf_sim::factorize(bool* done, bool precond_only)
{
    const int es = lu_es;
    const bool fp16 = precond_only && es == 4 && p.lu_update_fp16;
    if (*done && (precond_only || !lu_fp16)) return RF_OK;
    if (!d_LU) {
        const int ws_fp16 = (es == 4 && p.lu_update_fp16) ? 1 : 0;
        void* q = nullptr;
        if (hipMalloc(&q, (size_t)lda * n * es) != hipSuccess) {
            rf::set_error("cannot allocate the LU workspace");
            return RF_ENOMEM;
        }
        allocs.push_back(q);
        d_LU = q;
        const size_t piv_bytes = (sizeof(int) * (n + 2) + 255) / 256 * 256;
        if (hipMalloc(&q, piv_bytes + lu_workspace_bytes(n, ws_fp16)) != hipSuccess) {
            rf::set_error("cannot allocate pivots");
            return RF_ENOMEM;
        }
        allocs.push_back(q);
        d_ipiv = static_cast<int*>(q);
        d_info = d_ipiv + n;
        d_lu_ws = static_cast<char*>(q) + piv_bytes;
        HIP_TRY(hipMemsetAsync(d_lu_ws, 0, lu_workspace_bytes(n, ws_fp16), stream));
    }
    // panel width 64 (one wavefront factorises the diagonal block); FP32
    // factors accept up to 128
    const int nb_max = (es == 4) ? 128 : 64;
    const int nb = (p.lu_block > 0) ? std::min(nb_max, p.lu_block) : 64;
    hipEvent_t ea = nullptr, eb = nullptr;
    int rc0;
    if (p.profile) {
        if ((rc0 = timed_pair(&ea, &eb))) return rc0;
        HIP_TRY(hipEventRecord(ea, stream));
    }
    // block-pivoted panels, each checked against partial pivoting and redone
    // column by column when the check fails
    int info[2] = {0, 0};
    if (es == (mixed ? 4 : 8))
        HIP_TRY(hipMemcpyAsync(d_LU, d_A, (size_t)lda * n * es, hipMemcpyDeviceToDevice, stream));
    else  // FP32 factors of the FP64 matrix
        HIP_TRY(launch_f64_to_f32(static_cast<const double*>(d_A), static_cast<float*>(d_LU), (size_t)lda * n, stream));
    HIP_TRY(hipMemsetAsync(d_info, 0, 2 * sizeof(int), stream));
    lu_have = false;
    int redone = 0;
    HIP_TRY(lu_factor(d_LU, es, lda, n, nb, d_ipiv, d_info, p.lu_valu_update ? 0 : 1, p.lu_column_panel ? 0 : 1,
                      fp16 ? 1 : 0, d_lu_ws, &redone, reinterpret_cast<int*>(h_dots), stream));
    stats.lu_panel_fallbacks += redone;
    HIP_TRY(hipMemcpyAsync(h_dots, d_info, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
    if (p.profile) {
        HIP_TRY(hipEventRecord(eb, stream));
        timed_push(ea, eb, kLU);
    }
    stats.lu_factorizations += 1;
    HIP_TRY(hipStreamSynchronize(stream));
    std::memcpy(info, h_dots, 2 * sizeof(int));
    resolve_timed();
    if (info[0] != 0) {
        rf::set_error("LU: matrix is singular (U(" + std::to_string(info[0]) + "," + std::to_string(info[0]) + ") = 0)");
        return RF_ENOCONV;
    }
    {  // permutation and diagonal-block inverses for the block sweeps
        if (!d_lu_aux) {
            void* q = nullptr;
            if (hipMalloc(&q, lu_solve_aux_bytes(n, es) + sizeof(double) * (size_t)lda) != hipSuccess) {
                rf::set_error("cannot allocate the LU solve workspace");
                return RF_ENOMEM;
            }
            allocs.push_back(q);
            d_lu_aux = q;
            d_lu_tmp = reinterpret_cast<double*>(static_cast<char*>(q) + lu_solve_aux_bytes(n, es));
        }
        HIP_TRY(lu_prepare_solve(d_LU, es, lda, n, d_ipiv, d_lu_aux, stream));
    }
    *done = true;
    lu_have = true;
    lu_fp16 = fp16;
    return RF_OK;
}

// x <- (LU)^-1 x with the current factors
int rf_sim::lu_apply(double* x)
{
    HIP_TRY(lu_solve_blocks(d_LU, lu_es, lda, n, d_lu_aux, x, d_lu_tmp, stream));
    return RF_OK;
}

// Direct solve (dgesv, fortran/coeffs_and_vels.f90:62,158): LU with partial
// pivoting of a copy of A, then the two triangular solves.  In the mixed mode
// the FP32 matrix is factorised in FP32 (matrix-core trailing updates) and the
// solution is refined against the FP64-arithmetic residual b - A x.
int rf_sim::solve_lu(double* be_out)
{
    const int es = lu_es;
    {
        int rcf = factorize(&lu_factored, false);
        if (rcf) return rcf;
    }
    int rc;
    double bn2;
    if ((rc = norm2(d_b, &bn2))) return rc;
    const double bn = std::sqrt(bn2);
    if (es == 8) {
        HIP_TRY(launch_copy_vec(d_b, d_x, n, stream));
        if ((rc = lu_apply(d_x))) return rc;
        if ((rc = gemv_A(d_x, d_r, d_b, 1))) return rc;
        double rr;
        if ((rc = norm2(d_r, &rr))) return rc;
        *be_out = (bn > 0.0) ? std::sqrt(rr) / bn : 0.0;
        return RF_OK;
    }
    // iterative refinement: x += LU^-1 (b - A x)
    const double tol = cur_tol;
    HIP_TRY(hipMemsetAsync(d_x, 0, sizeof(double) * n, stream));
    HIP_TRY(launch_copy_vec(d_b, d_r, n, stream));
    double be = 1.0;
    for (int it = 0; it < 30; ++it) {
        if ((rc = lu_apply(d_r))) return rc;
        HIP_TRY(launch_axpby(n, 1.0, d_x, 1.0, d_r, d_x, stream));
        if ((rc = gemv_A(d_x, d_r, d_b, 1))) return rc;
        double rr;
        if ((rc = norm2(d_r, &rr))) return rc;
        be = (bn > 0.0) ? std::sqrt(rr) / bn : 0.0;
        stats.last_gmres_iters = it + 1;
        if (be <= tol) break;
    }
    *be_out = be;
    return RF_OK;
}

int rf_sim::setup_neighbours()
{
    if (knn < 1) return RF_OK;
    if (!d_nbr) {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(int) * (size_t)N * knn) != hipSuccess) {
            rf::set_error("cannot allocate the neighbour lists");
            return RF_ENOMEM;
        }
        allocs.push_back(q);
        d_nbr = static_cast<int*>(q);
        if (hipMalloc(&q, sizeof(double) * (size_t)N * knn) != hipSuccess) {
            rf::set_error("cannot allocate the neighbour distances");
            return RF_ENOMEM;
        }
        allocs.push_back(q);
        d_nbr_dist = static_cast<double*>(q);
    }
    if (nbr_valid) return RF_OK;
    HIP_TRY(launch_knn(N, knn, d_pos[lvl_three], d_orient[lvl_three], d_nbr, d_nbr_dist, stream));
    nbr_valid = true;
    return RF_OK;
}

// ---------------------------------------------------------------------------
// Learned initial guess: features of the 16 nearest fibers -> MLP -> the
// fiber's coefficients, combined with the previous solution by a
// two-vector minimum-residual projection.
// ---------------------------------------------------------------------------

int rf_sim::mlp_alloc()
{
    if (M != kMlpTerms || knn < kMlpNeighbours) {
        rf::set_error("learned guess: needs M = 5 and at least 17 fibers");
        return RF_EINVAL;
    }
    if (d_mlp_w) return RF_OK;
    const MlpLayout L = mlp_layout();
    const int64_t floats = L.total + (int64_t)N * (kMlpIn + kMlpOut) + mlp_buffer_floats(N);
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(float) * floats) != hipSuccess) {
        rf::set_error("learned guess: cannot allocate network buffers");
        return RF_ENOMEM;
    }
    allocs.push_back(q);
    d_mlp_w = static_cast<float*>(q);
    d_mlp_X = d_mlp_w + L.total;
    d_mlp_Y = d_mlp_X + (int64_t)N * kMlpIn;
    d_mlp_work = d_mlp_Y + (int64_t)N * kMlpOut;
    mlp_buffer_bind(d_mlp_work, N, &mlp_buf);
    return RF_OK;
}

// x_out (device, n doubles) = network prediction at the current configuration.
int rf_sim::mlp_infer(double* x_out)
{
    int rc = mlp_alloc();
    if (rc) return rc;
    if (!mlp_weights) {
        rf::set_error("learned guess: no network weights loaded");
        return RF_EINVAL;
    }
    if ((rc = setup_neighbours())) return rc;
    HIP_TRY(mlp_features(N, knn, d_pos[lvl_three], d_orient[lvl_three], d_nbr, d_nbr_dist, d_mlp_X, stream));
    HIP_TRY(mlp_forward(d_mlp_w, N, d_mlp_X, mlp_buf, stream));
    HIP_TRY(mlp_to_global(N, M, d_orient[lvl_three], d_mlp_w, mlp_buf.Yhat, x_out, stream));
    return RF_OK;
}

// d_x <- alpha d_xprev + beta g with g the prediction and (alpha, beta)
// minimising ||b - A(alpha x_prev + beta g)||; costs two GEMVs.  The result
// is never worse than the previous-step guess alone (alpha = 1, beta = 0 is
// a candidate) up to rounding.
int rf_sim::learned_guess()
{
    int rc = mlp_infer(d_guess);
    if (rc) return rc;
    const int64_t ldv = lda;
    double* u = d_V;            // A x_prev
    double* v = d_V + ldv;      // A g
    double* bb = d_V + 2 * ldv; // b
    if (restart < 2) {
        rf::set_error("learned guess: needs gmres_restart >= 2");
        return RF_EINVAL;
    }
    if ((rc = gemv_A(d_xprev, u, nullptr, 0))) return rc;
    if ((rc = gemv_A(d_guess, v, nullptr, 0))) return rc;
    HIP_TRY(launch_copy_vec(d_b, bb, n, stream));
    // D[0..2] = (u.b, v.b, b.b), D[3..4] = (u.v, v.v), D[5] = u.u
    HIP_TRY(launch_multidot(d_V, ldv, 2, bb, n, d_dparts, d_counter, d_dots, stream));
    HIP_TRY(launch_multidot(d_V, ldv, 1, v, n, d_dparts, d_counter, d_dots + 3, stream));
    HIP_TRY(launch_multidot(nullptr, 0, 0, u, n, d_dparts, d_counter, d_dots + 5, stream));
    if ((rc = fetch_dots(6))) return rc;
    const double ub = h_dots[0], vb = h_dots[1], uv = h_dots[3], vv = h_dots[4], uu = h_dots[5];
    double alpha = 0.0, beta = 0.0;
    const double det = uu * vv - uv * uv;
    if (uu > 0.0 && vv > 0.0 && det > 1e-12 * uu * vv) {
        alpha = (ub * vv - vb * uv) / det;
        beta = (vb * uu - ub * uv) / det;
    } else if (vv > 0.0 && (uu == 0.0 || vb * vb / vv > ub * ub / uu)) {
        beta = vb / vv;
    } else if (uu > 0.0) {
        alpha = ub / uu;
    }
    stats.last_guess_alpha = alpha;
    stats.last_guess_beta = beta;
    HIP_TRY(launch_axpby(n, alpha, d_xprev, beta, d_guess, d_x, stream));
    // r0 = b - alpha A x_prev - beta A g, by linearity, for the GMRES start
    HIP_TRY(launch_axpby(n, 1.0, bb, -alpha, u, d_r, stream));
    HIP_TRY(launch_axpby(n, 1.0, d_r, -beta, v, d_r, stream));
    r0_ready = true;
    return RF_OK;
}

// Right preconditioners of GMRES, built from the assembled matrix of this
// step: the neighbour-coupled block-Jacobi (restricted additive Schwarz)
// inverse, or the LU factors of the stored matrix (FP32 factors in the
// mixed mode: GMRES on A (LU)^-1 with FP64 residuals is then the
// iterative refinement of the FP32 factorisation).
int rf_sim::setup_precond(int kind)
{
    if (!d_z) {
        void* q = nullptr;
        for (double** v : {&d_z, &d_t}) {
            if (hipMalloc(&q, sizeof(double) * (size_t)lda) != hipSuccess) {
                rf::set_error("cannot allocate preconditioner vectors");
                return RF_ENOMEM;
            }
            allocs.push_back(q);
            *v = static_cast<double*>(q);
        }
    }
    int rc;
    if (kind == RF_PRECOND_LU) {
        if ((rc = factorize(&lu_factored, true))) return rc;
        prec_kind = RF_PRECOND_LU;
        prec_ready = true;
        return RF_OK;
    }
    const int bs = 3 * M;
    const int dim = bs * (ras_k + 1);
    if (!d_B) {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(double) * (size_t)N * dim * bs) != hipSuccess) {
            rf::set_error("cannot allocate the preconditioner");
            return RF_ENOMEM;
        }
        allocs.push_back(q);
        d_B = static_cast<double*>(q);
        if (hipMalloc(&q, sizeof(int)) != hipSuccess) {
            rf::set_error("cannot allocate preconditioner flag");
            return RF_ENOMEM;
        }
        allocs.push_back(q);
        d_pinfo = static_cast<int*>(q);
    }
    if ((rc = setup_neighbours())) return rc;
    if (ras_valid) {
        prec_kind = RF_PRECOND_NEIGHBOUR;
        prec_ready = true;
        return RF_OK;
    }
    HIP_TRY(hipMemsetAsync(d_pinfo, 0, sizeof(int), stream));
    HIP_TRY(launch_ras_setup(N, ras_k, knn, bs, d_nbr, d_A, mixed ? 4 : 8, lda, d_B, d_pinfo, stream));
    int info = 0;
    HIP_TRY(hipMemcpyAsync(h_dots, d_pinfo, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    resolve_timed();
    std::memcpy(&info, h_dots, sizeof(int));
    if (info != 0) {
        rf::set_error("preconditioner: singular cluster matrix at fiber " + std::to_string(info - 1));
        return RF_ENOCONV;
    }
    prec_kind = RF_PRECOND_NEIGHBOUR;
    prec_ready = true;
    ras_valid = true;
    return RF_OK;
}

int rf_sim::precond_apply(const double* r, double* z)
{
    if (prec_kind == RF_PRECOND_LU) {
        if (r != z) HIP_TRY(launch_copy_vec(r, z, n, stream));
        return lu_apply(z);
    }
    HIP_TRY(launch_ras_apply(N, ras_k, knn, 3 * M, d_nbr, d_B, r, z, stream));
    return RF_OK;
}

int rf_sim::phase_begin()
{
    HIP_TRY(hipEventRecord(ev_p0, stream));
    return RF_OK;
}
double rf_sim::phase_end()
{
    if (hipEventRecord(ev_p1, stream) != hipSuccess) return 0.0;
    const double ms = event_ms(ev_p0, ev_p1);  // waits for the phase
    resolve_timed();
    return ms;
}

int rf_sim::assemble()
{
    op_velocity_ready = false;
    int rc = phase_begin();
    if (rc) return rc;
    AssembleArgs a;
    a.N = N;
    a.n = n;
    a.lda = lda;
    a.ldw = ldw;
    a.analytic = p.use_analytical_integration;
    a.rhs_z_only = 1;
    a.splits = asm_splits;
    a.pos = d_pos[lvl_three];
    a.orient = d_orient[lvl_three];
    a.force = d_force;
    a.wp = d_wp;
    a.pv = d_pv;
    a.A = d_A;
    a.W = d_W;  // nullptr in the mixed mode
    a.bpart = d_bpart;
    a.mixed = mixed;
    hipEvent_t ea = nullptr, eb = nullptr;
    if (p.profile) {
        if ((rc = timed_pair(&ea, &eb))) return rc;
        HIP_TRY(hipEventRecord(ea, stream));
    }
    HIP_TRY(launch_assemble(tab, a, stream));
    if (p.profile) {
        HIP_TRY(hipEventRecord(eb, stream));
        timed_push(ea, eb, kAssemble);
    }
    HIP_TRY(launch_reduce_splits(d_bpart, asm_splits, n, n, d_b, stream));
    stats.ms_assemble = phase_end();
    stats.total_ms_assemble += stats.ms_assemble;
    assembled = true;
    lu_factored = false;
    ras_valid = false;
    return RF_OK;
}

// The matrix-free FP64 operator at the current configuration (kOp* modes).
int rf_sim::apply_operator(int mode, const double* x, double* y)
{
    hipEvent_t ea = nullptr, eb = nullptr;
    int rc;
    op_velocity_ready = false;
    if (p.profile) {
        if ((rc = timed_pair(&ea, &eb))) return rc;
        HIP_TRY(hipEventRecord(ea, stream));
    }
    HIP_TRY(launch_operator(planOp, tab, mode, d_pos[lvl_three], d_orient[lvl_three], d_force, x, d_wp, d_pv, d_op, y,
                            stream));
    if (p.profile) {
        HIP_TRY(hipEventRecord(eb, stream));
        timed_push(ea, eb, kOperator);
    }
    stats.operator_launches += 1;
    return RF_OK;
}

// Solve of the stored system A x = b (d_b, d_x holding the initial guess) to
// the relative tolerance `tol` with the configured method: the direct
// solver, or GMRES (optionally preconditioned) with the reference's LU
// fallback.  *be is the backward error against the stored matrix.
int rf_sim::solve_system(int maxit, double tol, int* iters_out, double* be_out)
{
    int rc;
    int iters = 0, status = 0;
    double be = 0.0;
    cur_tol = tol;
    if (p.use_direct_solver) {
        r0_ready = false;
        if ((rc = solve_lu(&be))) return rc;
        stats.last_solver = RF_SOLVER_LU;
        *iters_out = 0;
        *be_out = be;
        return RF_OK;
    }
    prec_ready = false;
    prec_kind = RF_PRECOND_NONE;
    if (p.precond == RF_PRECOND_AUTO) {
        // Neighbour preconditioner within a budget (two restart cycles by
        // default), then GMRES on A (LU)^-1 from the iterate reached.  After
        // the first cycle the observed residual reduction per iteration
        // predicts the iterations still needed; if they exceed the budget the
        // switch happens at once.  A solve that ended on the LU makes the next
        // ones start on it (the dynamics change the matrix little per step),
        // and every kAutoRetry-th of them tries the neighbour preconditioner
        // again; factors of this very matrix are always used at once.
        const int budget = std::min(maxit, p.precond_switch_iters > 0 ? p.precond_switch_iters : 2 * restart);
        const bool lu_first = lu_factored || (auto_lu_run > 0 && auto_lu_run % kAutoRetry != 0);
        if (lu_first) {
            if ((rc = setup_precond(RF_PRECOND_LU))) return rc;
            if ((rc = solve_gmres(&iters, &be, &status, maxit))) return rc;
        } else {
            if ((rc = setup_precond(RF_PRECOND_NEIGHBOUR))) return rc;
            if ((rc = solve_gmres(&iters, &be, &status, std::min(budget, restart)))) return rc;
        }
        const double r0 = stats.last_initial_residual;
        if (!lu_first && status != 0 && iters < budget && r0 > 0.0 && be > 0.0 && be < r0) {
            const double rate = std::pow(be / r0, 1.0 / iters);
            const double needed = std::log(tol / be) / std::log(rate);
            if (iters + needed <= budget) {
                int it2 = 0;
                if ((rc = solve_gmres(&it2, &be, &status, budget - iters))) return rc;
                iters += it2;
            }
        }
        stats.last_neighbour_iters = lu_first ? 0 : iters;
        if (!lu_first && status != 0 && iters < maxit) {
            int it2 = 0;
            if ((rc = setup_precond(RF_PRECOND_LU))) return rc;
            if ((rc = solve_gmres(&it2, &be, &status, maxit - iters))) return rc;
            iters += it2;
        }
        stats.last_initial_residual = r0;
    } else {
        if (p.precond != RF_PRECOND_NONE && (rc = setup_precond(p.precond))) return rc;
        if ((rc = solve_gmres(&iters, &be, &status, maxit))) return rc;
        stats.last_neighbour_iters = (p.precond == RF_PRECOND_NEIGHBOUR) ? iters : 0;
    }
    stats.last_precond = prec_kind;
    stats.last_solver = RF_SOLVER_GMRES;
    if (status == -4) {  // fortran/coeffs_and_vels.f90:331-335
        if ((rc = solve_lu(&be))) return rc;
        stats.last_solver = RF_SOLVER_LU_FALLBACK;
    }
    *iters_out = iters;
    *be_out = be;
    return RF_OK;
}

// Rounding noise of the matrix-free FP64 operator at the current
// configuration, from residual evaluations: with r(v) = b - A v and
// x1 + x2 = 2 x, 2 r(x) - r(x1) - r(x2) = A x1 + A x2 - 2 A x is 0 in exact
// arithmetic; x1 = 2 u x and x2 = 2 (1 - u) x with u pseudo-random per
// entry.  d_res64 holds r(x) on entry and keeps it.  Where two fibers come
// within ~1e-3 of each other's axis the analytic moments reach ~1e12 and the
// FP64 evaluation of those rows carries rounding errors far above
// 1e-16 ||b||.  Between restart cycles the flexible GMRES buffers are free:
// d_fV slots 0 and 1, d_fZ slot 0, d_dx and d_fw serve as scratch.
int rf_sim::operator_noise(double bn, double* eta)
{
    int rc;
    double* v0 = d_fV;
    double* v1 = d_fV + lda;
    double* z0 = d_fZ;
    HIP_TRY(launch_random_split(n, d_x, v0, v1, stream));  // v0 + v1 = x
    HIP_TRY(launch_axpby(n, 2.0, v0, 0.0, v0, z0, stream));  // x1
    HIP_TRY(launch_axpby(n, 2.0, v1, 0.0, v1, v0, stream));  // x2
    if ((rc = apply_operator(kOpResidual, z0, d_fw))) return rc;
    if ((rc = apply_operator(kOpResidual, v0, d_dx))) return rc;
    HIP_TRY(launch_axpby(n, 2.0, d_res64, -1.0, d_fw, v1, stream));
    HIP_TRY(launch_axpby(n, 1.0, v1, -1.0, d_dx, z0, stream));
    double e2, rr;
    if ((rc = norm2(z0, &e2))) return rc;
    *eta = std::sqrt(e2) / bn;
    return norm2(d_res64, &rr);  // ||r||^2 back in d_dots[0] for the next cycle
}

// Mixed precision: flexible GMRES on the FP64 system, with the FP32 matrix's
// solves as the preconditioner (GMRES-based iterative refinement):
//     r = b - A x with the matrix-free FP64 operator (no stored FP64 matrix),
//     v_0 = r / ||r||;  for j = 0, 1, ...:
//         z_j = (solve A32 z = v_j to mixed_inner_tol with the configured method
//                on the FP32 matrix: GMRES with FP64 Arnoldi, its preconditioner
//                and FP32 LU factors, or the direct solver),
//         w = A z_j (FP64 operator), orthogonalised against v_0..v_j (classical
//         Gram-Schmidt applied twice), Givens rotations on the host;
//     x += Z y once |g_{j+1}| / ||b|| <= gmres_tol, then the true FP64 residual.
// Plain refinement (x += z_0, repeated) contracts by the spectral radius of
// I - A A32^-1, which reaches ~1 for a few modes of the dense N=4000 systems
// (FP32 storage and FP32 far-field recurrences leave A32 2e-7 from A, and
// those systems are ill-conditioned enough); the Krylov outer iteration
// removes such modes in a few steps at the same cost per step (one operator
// evaluation and one inner solve).  When a restart cycle's Arnoldi estimate
// meets the tolerance and the FP64 residual does not, the operator's own
// rounding noise is measured once (operator_noise); a backward error within
// kNoiseFactor of it is accepted and reported (last_noise_limited): no
// FP64 evaluation of the system can resolve a smaller residual there.
int rf_sim::solve_refined(int* iters_out, double* be_out)
{
    const int mo = outer_m;
    const double tol = p.gmres_tol;
    const double inner_tol = (p.mixed_inner_tol > 0.0) ? p.mixed_inner_tol : 1e-6;
    const int maxit = p.gmres_max_iters;
    int rc;
    double bn2;
    if ((rc = norm2(d_b, &bn2))) return rc;  // the scale only
    const double bn = std::sqrt(bn2);
    *iters_out = 0;
    stats.last_refinements = 0;
    stats.last_outer_iters = 0;
    stats.last_operator_noise = 0.0;
    stats.last_noise_limited = 0;
    for (int i = 0; i < 8; ++i) {
        stats.refine_history[i] = 0.0;
        stats.refine_inner_iters[i] = 0;
    }
    if (bn == 0.0) {
        HIP_TRY(launch_fill(d_x, 0.0, (size_t)n, stream));
        *be_out = 0.0;
        return RF_OK;
    }
    r0_ready = false;  // a residual of the learned guess is on the FP32 matrix
    const int ldh = mo + 1;
    std::vector<double> H((size_t)(mo + 1) * (mo + 1)), g(mo + 1), cs(mo + 1), sn(mo + 1), y(mo + 1);
    const int m = restart;  // slot layout of d_dots (inner GMRES)
    double* D1 = d_dots;
    double* D2 = d_dots + (m + 2);
    double* D3 = d_dots + 2 * (m + 2);
    int total = 0, outer = 0;
    double be = 0.0;
    // the FP64 residual b - A x; ||r||^2 stays in d_dots[0] for the scaling
    // the residual's source vectors equal the velocity's unless the analytic
    // rhs drops an x/y force component (fortran/matrix_assembly.f90:857-859)
    const bool fusable = !p.use_analytical_integration ||
                         (p.external_force[0] == 0.0 && p.external_force[1] == 0.0);
    auto residual = [&](double* beta) -> int {
        int r = apply_operator(kOpResidual, d_x, d_res64);
        if (r) return r;
        op_velocity_ready = fusable;  // until the next evaluation
        double rr;
        if ((r = norm2(d_res64, &rr))) return r;
        *beta = std::sqrt(rr);
        return RF_OK;
    };
    double beta;
    if ((rc = residual(&beta))) return rc;
    be = beta / bn;
    stats.refine_history[0] = be;
    stats.last_initial_residual = be;
    stats.last_refinements = 1;
    while (be > tol) {
        if (outer >= kMaxOuterIters || total >= maxit) {
            char msg[200];
            snprintf(msg, sizeof(msg),
                     "mixed precision: flexible GMRES on the FP64 operator reached backward error %.3e after %d "
                     "outer and %d inner iterations (tolerance %.1e)", be, outer, total, tol);
            rf::set_error(msg);
            stats.last_outer_iters = outer;
            stats.last_backward_error = be;
            *iters_out = total;
            *be_out = be;
            return RF_ENOCONV;
        }
        HIP_TRY(launch_scale_copy(d_res64, d_dots, 0, n, d_fV, stream));
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        double bea = 1.0;
        for (int j = 0; j < mo; ++j) {
            // z_j = A32^-1 v_j on the buffers of the plain solve
            HIP_TRY(launch_copy_vec(d_fV + (size_t)j * lda, d_res64, n, stream));
            std::swap(d_b, d_res64);
            std::swap(d_x, d_dx);
            int it = 0;
            double be_in = 0.0;
            rc = launch_fill(d_x, 0.0, (size_t)n, stream) == hipSuccess ? RF_OK : RF_EHIP;
            if (rc == RF_OK) rc = solve_system(std::max(1, maxit - total), inner_tol, &it, &be_in);
            std::swap(d_b, d_res64);
            std::swap(d_x, d_dx);
            if (rc) return rc;
            if (outer < 8) stats.refine_inner_iters[outer] = it;
            total += it;
            double* zj = d_fZ + (size_t)j * lda;
            HIP_TRY(launch_copy_vec(d_dx, zj, n, stream));
            if ((rc = apply_operator(kOpAx, zj, d_fw))) return rc;
            k = j + 1;
            HIP_TRY(launch_multidot(d_fV, lda, k, d_fw, n, d_dparts, d_counter, D1, stream));
            HIP_TRY(launch_orth_dots(d_fV, lda, k, D1, d_fw, n, 1, d_dparts, d_counter, D2, stream));
            HIP_TRY(launch_orth_dots(d_fV, lda, k, D2, d_fw, n, 0, d_dparts, d_counter, D3, stream));
            if (k < mo) HIP_TRY(launch_scale_copy(d_fw, D3, 0, n, d_fV + (size_t)k * lda, stream));
            if ((rc = fetch_dots(3 * (m + 2)))) return rc;
            for (int i = 0; i < k; ++i) H[i + (size_t)j * ldh] = h_dots[i] + h_dots[(m + 2) + i];
            H[k + (size_t)j * ldh] = std::sqrt(h_dots[2 * (m + 2)]);
            for (int i = 0; i < j; ++i) rot1(&H[i + (size_t)j * ldh], &H[i + 1 + (size_t)j * ldh], cs[i], sn[i]);
            givens(H[j + (size_t)j * ldh], H[k + (size_t)j * ldh], &cs[j], &sn[j]);
            rot1(&g[j], &g[k], cs[j], sn[j]);
            rot1(&H[j + (size_t)j * ldh], &H[k + (size_t)j * ldh], cs[j], sn[j]);
            H[k + (size_t)j * ldh] = 0.0;
            ++outer;
            bea = std::fabs(g[k]) / bn;
            if (outer < 8) stats.refine_history[outer] = bea;
            if (bea <= tol || outer >= kMaxOuterIters || total >= maxit) break;
        }
        // x += Z y, then the true FP64 residual
        hess_solve(H, ldh, g, k, y);
        std::copy(y.begin(), y.begin() + k, h_y);
        HIP_TRY(launch_vy_plus_x(d_fZ, lda, k, h_y, d_x, n, d_xc, stream));
        HIP_TRY(launch_copy_vec(d_xc, d_x, n, stream));
        if ((rc = residual(&beta))) return rc;
        be = beta / bn;
        stats.last_refinements += 1;
        if (be > tol && bea <= tol) {
            // the Krylov model converged, the FP64 residual did not: noise?
            if (stats.last_operator_noise == 0.0) {
                double eta = 0.0;
                if ((rc = operator_noise(bn, &eta))) return rc;
                stats.last_operator_noise = eta;
            }
            if (be <= kNoiseFactor * stats.last_operator_noise) {
                stats.last_noise_limited = 1;
                break;
            }
        }
    }
    stats.last_outer_iters = outer;
    *iters_out = total;
    *be_out = be;
    return RF_OK;
}

int rf_sim::solve()
{
    op_velocity_ready = false;
    if (!assembled) {
        rf::set_error("rf_solve: system not assembled");
        return RF_EINVAL;
    }
    int rc = phase_begin();
    if (rc) return rc;
    // initial guess (fortran/advect_fibers.f90:464: init_guess = coeffvec)
    if (p.warm_start == RF_WARM_ZERO) {
        HIP_TRY(launch_fill(d_x, 0.0, (size_t)n, stream));
    } else if (p.warm_start == RF_WARM_EXTRAP && solves_done >= 2) {
        HIP_TRY(launch_axpby(n, 2.0, d_x, -1.0, d_xprev, d_guess, stream));
        HIP_TRY(launch_copy_vec(d_x, d_xprev, n, stream));
        HIP_TRY(launch_copy_vec(d_guess, d_x, n, stream));
    } else {
        HIP_TRY(launch_copy_vec(d_x, d_xprev, n, stream));
        if (p.warm_start == RF_WARM_LEARNED && mlp_weights && !p.use_direct_solver && (rc = learned_guess()))
            return rc;
    }
    int iters = 0;
    double be = 0.0;
    if (mixed)
        rc = solve_refined(&iters, &be);
    else
        rc = solve_system(p.gmres_max_iters, p.gmres_tol, &iters, &be);
    if (rc) return rc;
    if (!p.use_direct_solver) auto_lu_run = (prec_kind == RF_PRECOND_LU) ? auto_lu_run + 1 : 0;
    stats.last_gmres_iters = iters;
    stats.total_gmres_iters += iters;
    stats.last_backward_error = be;
    ++solves_done;
    stats.ms_solve = phase_end();
    stats.total_ms_solve += stats.ms_solve;
    solved = true;
    return RF_OK;
}

int rf_sim::update_velocities()
{
    int rc = RF_OK;
    if (!solved) {
        rf::set_error("rf_update_velocities: no solution for this configuration");
        return RF_EINVAL;
    }
    rc = phase_begin();
    if (rc) return rc;
    // old=new; new=3-old (fortran/advect_fibers.f90:465-466)
    vel_old = vel_new;
    vel_new = 1 - vel_old;
    if (mixed) {
        // FP64 velocity sums from the matrix-free operator (no stored W):
        // from the final residual's partials when they hold them
        if (op_velocity_ready) {
            HIP_TRY(launch_operator_epilogue(planOp, tab, kOpVelocity, d_orient[lvl_three], d_x, d_op, d_y6, stream));
            op_velocity_ready = false;
        } else if ((rc = apply_operator(kOpVelocity, d_x, d_y6))) {
            return rc;
        }
    } else {
        HIP_TRY(launch_build_fext(N, M, d_force, d_x, d_fext, stream));
        hipEvent_t ea = nullptr, eb = nullptr;
        if (p.profile && (rc = timed_pair(&ea, &eb))) return rc;
        HIP_TRY(launch_gemv(planW, d_W, d_fext, d_gparts, d_y6, nullptr, 0, stream, ea, eb));
        if (p.profile) timed_push(ea, eb, kGemvW);
    }
    HIP_TRY(launch_velocity_epilogue(N, tab.d, tab.e, p.shear_rate, d_pos[lvl_three], d_orient[lvl_three], d_force, d_y6,
                                     d_vel[vel_new], d_rot[vel_new], stream));
    stats.ms_velocities = phase_end();
    stats.total_ms_velocities += stats.ms_velocities;
    have_velocities = true;
    return RF_OK;
}

int rf_sim::advance()
{
    op_velocity_ready = false;
    if (!have_velocities) {
        rf::set_error("rf_advance: velocities not computed");
        return RF_EINVAL;
    }
    int rc = phase_begin();
    if (rc) return rc;
    // one=two; two=three; three=6-(one+two)  (fortran/advect_fibers.f90:479-481)
    lvl_one = lvl_two;
    lvl_two = lvl_three;
    lvl_three = 3 - (lvl_one + lvl_two);
    const int first = (steps_taken == 0) ? 1 : 0;
    HIP_TRY(launch_advance(N, first, p.timestep, d_pos[lvl_one], d_pos[lvl_two], d_orient[lvl_one],
                           d_orient[lvl_two], d_vel[vel_old], d_vel[vel_new], d_rot[vel_old], d_rot[vel_new],
                           d_pos[lvl_three], d_orient[lvl_three], stream));
    stats.ms_advance = phase_end();
    stats.total_ms_advance += stats.ms_advance;
    ++steps_taken;
    stats.steps_taken = steps_taken;
    assembled = solved = have_velocities = false;
    prec_ready = false;
    nbr_valid = false;
    return RF_OK;
}

int rf_sim::step()
{
    int rc;
    if ((rc = assemble())) return rc;
    if ((rc = solve())) return rc;
    if ((rc = update_velocities())) return rc;
    return advance();
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" const char* rf_last_error(void) { return g_error.c_str(); }
extern "C" const char* rf_version(void) { return "rigidfibers-mi355x 0.1 (gfx950)"; }

extern "C" int rf_create(const rf_params* params, int32_t num_fibers, const double* pos, const double* orient,
                         rf_sim** out)
{
    if (!params || !pos || !orient || !out) {
        set_error("rf_create: null argument");
        return RF_EINVAL;
    }
    *out = nullptr;
    const rf_params& P = *params;
    if (num_fibers < 1 || P.num_terms < 1 || P.num_terms > kMaxTerms || P.num_quad_intervals < 1 ||
        3 * P.num_quad_intervals > kMaxQuad || !(P.slenderness > 0.0) || P.gmres_restart < 1 ||
        P.gmres_restart > 60 || P.gmres_max_iters < 1 || (P.precision != RF_PREC_FP64 && P.precision != RF_PREC_MIXED) ||
        P.precond < 0 || P.precond > 3 || P.precond_neighbours < 0 || P.precond_neighbours > 6 ||
        P.precond_switch_iters < 0 ||
        ((P.precond == RF_PRECOND_NEIGHBOUR || P.precond == RF_PRECOND_AUTO) && num_fibers < 2)) {
        set_error("rf_create: invalid parameters (need N>=1, 1<=M<=8, 1<=NoQI<=32, eps>0, 1<=restart<=60, "
                  "precond 0..3, 0<=precond_neighbours<=6, N>=2 for the neighbour preconditioner)");
        return RF_EINVAL;
    }
    int dev = 0, ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) {
        set_error(std::string("rf_create: no HIP device: ") + hipGetErrorString(e));
        return RF_EHIP;
    }
    rf_sim* s = new rf_sim();
    s->p = P;
    s->N = num_fibers;
    s->M = P.num_terms;
    s->n = 3 * s->M * s->N;
    s->mixed = (P.precision == RF_PREC_MIXED) ? 1 : 0;
    s->restart = P.gmres_restart;
    s->lu_es = (s->mixed || P.lu_factor_fp32) ? 4 : 8;
    s->ras_k = (P.precond == RF_PRECOND_NEIGHBOUR || P.precond == RF_PRECOND_AUTO) ? std::min(num_fibers - 1, P.precond_neighbours > 0 ? P.precond_neighbours : 4) : 0;
    // neighbour lists keep the 16 nearest fibers (the network's input); the
    // preconditioner uses the first ras_k of them
    s->knn = std::min(num_fibers - 1, kMlpNeighbours);
    build_tables(s->M, P.num_quad_intervals, P.slenderness, &s->tab);
    s->LQ = s->tab.LQ;
    s->lda = round_up(s->n, 256);
    s->ldw = round_up(6 * (int64_t)s->N, 256);
    s->wcols = 3 * (s->M + 1) * s->N;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&s->num_cus, hipDeviceAttributeMultiprocessorCount,
                                                                  dev) != hipSuccess) {
        delete s;
        set_error("rf_create: cannot query the device");
        return RF_EHIP;
    }
    const int T = assemble_targets_per_block(s->LQ, s->M, P.use_analytical_integration != 0);
    const int tiles = (s->N + T - 1) / T;
    s->asm_splits = std::max(1, std::min(std::min(64, s->N), (4096 + tiles - 1) / tiles));
    const int es = s->mixed ? 4 : 8;
    s->planA = plan_gemv(s->n, s->n, s->lda, es, s->num_cus);
    s->planW = plan_gemv(6 * s->N, s->wcols, s->ldw, es, s->num_cus);
    s->planOp = plan_operator(s->N, s->M, s->LQ, P.use_analytical_integration ? 1 : 0, s->num_cus);
    // coherent (fine-grained) pinned memory: the GPU's writes into it bypass
    // its caches, so the per-iteration event needs no system-scope fence
    if (hipHostMalloc((void**)&s->h_dots, sizeof(double) * (3 * (s->restart + 2) + 8), hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostMalloc((void**)&s->h_y, sizeof(double) * (s->restart + 2)) != hipSuccess) {
        delete s;
        set_error("rf_create: hipHostMalloc failed");
        return RF_ENOMEM;
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        set_error("rf_create: hipStreamCreate failed");
        return RF_EHIP;
    }
    for (hipEvent_t* ev : {&s->ev_p0, &s->ev_p1}) {
        if (hipEventCreateWithFlags(ev, 0u) != hipSuccess) {
            delete s;
            set_error("rf_create: hipEventCreate failed");
            return RF_EHIP;
        }
        s->events.push_back(*ev);
    }
    int rc = s->alloc_all();
    if (rc) {
        delete s;
        return rc;
    }
    // state: current level = index 2 ("three"), orientations normalised
    // (fortran/advect_fibers.f90:332-340)
    std::vector<double> t(3 * (size_t)s->N), F(3 * (size_t)s->N);
    for (int i = 0; i < s->N; ++i) {
        const double* o = orient + 3 * i;
        const double m = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        if (!(m > 0.0)) {
            delete s;
            set_error("rf_create: zero orientation vector for fiber " + std::to_string(i));
            return RF_EINVAL;
        }
        for (int c = 0; c < 3; ++c) {
            t[3 * i + c] = o[c] / m;
            F[3 * i + c] = P.external_force[c];
        }
    }
    const size_t bytes3 = sizeof(double) * 3 * s->N;
    e = hipMemcpyAsync(s->d_pos[s->lvl_three], pos, bytes3, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_orient[s->lvl_three], t.data(), bytes3, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_force, F.data(), bytes3, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_wp, s->tab.wp, sizeof(s->tab.wp), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_pv, s->tab.pv, sizeof(s->tab.pv), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        delete s;
        set_error(std::string("rf_create: upload failed: ") + hipGetErrorString(e));
        return RF_EHIP;
    }
    s->stats.matrix_elem_bytes = es;
    s->stats.lda = (int32_t)s->lda;
    s->stats.matrix_bytes = (double)s->lda * s->n * es;
    *out = s;
    return RF_OK;
}

extern "C" void rf_destroy(rf_sim* s)
{
    if (!s) return;
    (void)hipStreamSynchronize(s->stream);
    delete s;
}

#define CHECK_SIM(s)                          \
    do {                                      \
        if (!(s)) {                           \
            set_error("null rf_sim handle"); \
            return RF_EINVAL;                 \
        }                                     \
    } while (0)

extern "C" int rf_step(rf_sim* s)
{
    CHECK_SIM(s);
    return s->step();
}
extern "C" int rf_assemble(rf_sim* s)
{
    CHECK_SIM(s);
    return s->assemble();
}
extern "C" int rf_solve(rf_sim* s)
{
    CHECK_SIM(s);
    return s->solve();
}
extern "C" int rf_update_velocities(rf_sim* s)
{
    CHECK_SIM(s);
    return s->update_velocities();
}
extern "C" int rf_advance(rf_sim* s)
{
    CHECK_SIM(s);
    return s->advance();
}

static int copy_out(const rf_sim* s, double* dst, const double* src, size_t count)
{
    if (!dst) {
        set_error("null output buffer");
        return RF_EINVAL;
    }
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return RF_OK;
}

extern "C" int rf_get_state(const rf_sim* s, double* pos, double* orient)
{
    CHECK_SIM(s);
    int rc = copy_out(s, pos, s->d_pos[s->lvl_three], 3 * (size_t)s->N);
    if (rc) return rc;
    return copy_out(s, orient, s->d_orient[s->lvl_three], 3 * (size_t)s->N);
}
extern "C" int rf_get_velocities(const rf_sim* s, double* vel, double* rot)
{
    CHECK_SIM(s);
    int rc = copy_out(s, vel, s->d_vel[s->vel_new], 3 * (size_t)s->N);
    if (rc) return rc;
    return copy_out(s, rot, s->d_rot[s->vel_new], 3 * (size_t)s->N);
}
extern "C" int rf_get_coefficients(const rf_sim* s, double* x)
{
    CHECK_SIM(s);
    return copy_out(s, x, s->d_x, (size_t)s->n);
}
extern "C" int rf_get_rhs(const rf_sim* s, double* b)
{
    CHECK_SIM(s);
    return copy_out(s, b, s->d_b, (size_t)s->n);
}
extern "C" int rf_set_coefficients(rf_sim* s, const double* x)
{
    CHECK_SIM(s);
    if (!x) {
        set_error("rf_set_coefficients: null");
        return RF_EINVAL;
    }
    s->op_velocity_ready = false;
    HIP_TRY(hipMemcpyAsync(s->d_x, x, sizeof(double) * s->n, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return RF_OK;
}

// Restart state (the levels the next step reads; see the declaration).
extern "C" int rf_get_restart_state(const rf_sim* s, int64_t* steps_taken, double* pos, double* orient,
                                    double* pos_prev, double* orient_prev, double* vel, double* rot, double* coeffs)
{
    CHECK_SIM(s);
    if (!steps_taken || !pos || !orient || !pos_prev || !orient_prev || !vel || !rot || !coeffs) {
        set_error("rf_get_restart_state: null buffer");
        return RF_EINVAL;
    }
    const size_t N3 = 3 * (size_t)s->N;
    *steps_taken = s->steps_taken;
    int rc;
    if ((rc = copy_out(s, pos, s->d_pos[s->lvl_three], N3))) return rc;
    if ((rc = copy_out(s, orient, s->d_orient[s->lvl_three], N3))) return rc;
    if ((rc = copy_out(s, coeffs, s->d_x, (size_t)s->n))) return rc;
    if (s->steps_taken == 0) {
        std::memcpy(pos_prev, pos, sizeof(double) * N3);
        std::memcpy(orient_prev, orient, sizeof(double) * N3);
        std::memset(vel, 0, sizeof(double) * N3);
        std::memset(rot, 0, sizeof(double) * N3);
        return RF_OK;
    }
    if ((rc = copy_out(s, pos_prev, s->d_pos[s->lvl_two], N3))) return rc;
    if ((rc = copy_out(s, orient_prev, s->d_orient[s->lvl_two], N3))) return rc;
    if ((rc = copy_out(s, vel, s->d_vel[s->vel_new], N3))) return rc;
    return copy_out(s, rot, s->d_rot[s->vel_new], N3);
}

extern "C" int rf_restore_state(rf_sim* s, int64_t steps_taken, const double* pos, const double* orient,
                                const double* pos_prev, const double* orient_prev, const double* vel,
                                const double* rot, const double* coeffs)
{
    CHECK_SIM(s);
    if (steps_taken < 0 || !pos || !orient ||
        (steps_taken > 0 && (!pos_prev || !orient_prev || !vel || !rot))) {
        set_error("rf_restore_state: steps_taken < 0, or a null level the next step needs");
        return RF_EINVAL;
    }
    const size_t b3 = sizeof(double) * 3 * (size_t)s->N;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_pos[s->lvl_three], pos, b3, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_orient[s->lvl_three], orient, b3, hipMemcpyHostToDevice, s->stream));
    if (steps_taken > 0) {
        HIP_TRY(hipMemcpyAsync(s->d_pos[s->lvl_two], pos_prev, b3, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->d_orient[s->lvl_two], orient_prev, b3, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->d_vel[s->vel_new], vel, b3, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->d_rot[s->vel_new], rot, b3, hipMemcpyHostToDevice, s->stream));
    }
    if (coeffs) {
        HIP_TRY(hipMemcpyAsync(s->d_x, coeffs, sizeof(double) * s->n, hipMemcpyHostToDevice, s->stream));
    } else {
        HIP_TRY(launch_fill(s->d_x, 0.0, (size_t)s->n, s->stream));
    }
    HIP_TRY(launch_copy_vec(s->d_x, s->d_xprev, s->n, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->steps_taken = steps_taken;
    s->stats.steps_taken = steps_taken;
    s->solves_done = steps_taken;
    s->assembled = s->solved = s->have_velocities = false;
    s->op_velocity_ready = false;
    s->prec_ready = false;
    s->nbr_valid = false;
    s->ras_valid = false;
    s->lu_factored = false;
    s->auto_lu_run = 0;
    return RF_OK;
}

extern "C" int rf_get_matrix_block(const rf_sim* s, int64_t row0, int64_t col0, int64_t rows, int64_t cols,
                                   double* out, int64_t ldo)
{
    CHECK_SIM(s);
    if (!out || rows < 1 || cols < 1 || row0 < 0 || col0 < 0 || row0 + rows > s->n || col0 + cols > s->n ||
        ldo < rows) {
        set_error("rf_get_matrix_block: null buffer, block outside the n x n matrix, or ldo < rows");
        return RF_EINVAL;
    }
    if (!s->mixed) {
        const double* src = static_cast<const double*>(s->d_A) + col0 * s->lda + row0;
        HIP_TRY(hipMemcpy2DAsync(out, sizeof(double) * ldo, src, sizeof(double) * s->lda, sizeof(double) * rows, cols,
                                 hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        return RF_OK;
    }
    std::vector<float> tmp((size_t)rows * cols);
    const float* src = static_cast<const float*>(s->d_A) + col0 * s->lda + row0;
    HIP_TRY(hipMemcpy2DAsync(tmp.data(), sizeof(float) * rows, src, sizeof(float) * s->lda, sizeof(float) * rows, cols,
                             hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int64_t c = 0; c < cols; ++c)
        for (int64_t r = 0; r < rows; ++r) out[c * ldo + r] = tmp[(size_t)c * rows + r];
    return RF_OK;
}

extern "C" int rf_get_matrix(const rf_sim* s, double* A, int64_t lda)
{
    CHECK_SIM(s);
    if (!A || lda < s->n) {
        set_error("rf_get_matrix: null buffer or lda < n");
        return RF_EINVAL;
    }
    return rf_get_matrix_block(s, 0, 0, s->n, s->n, A, lda);
}

extern "C" int rf_get_stats(const rf_sim* s, rf_stats* st)
{
    CHECK_SIM(s);
    if (!st) {
        set_error("rf_get_stats: null");
        return RF_EINVAL;
    }
    *st = s->stats;
    return RF_OK;
}

extern "C" int rf_matvec(rf_sim* s, const double* x, double* y)
{
    CHECK_SIM(s);
    if (!x || !y) {
        set_error("rf_matvec: null");
        return RF_EINVAL;
    }
    HIP_TRY(hipMemcpyAsync(s->d_xc, x, sizeof(double) * s->n, hipMemcpyHostToDevice, s->stream));
    int rc = s->gemv_A(s->d_xc, s->d_r, nullptr, 0);
    if (rc) return rc;
    return copy_out(s, y, s->d_r, (size_t)s->n);
}

extern "C" int rf_operator(rf_sim* s, int32_t mode, const double* x, double* y)
{
    CHECK_SIM(s);
    if (!y || (mode != RF_OP_RHS && !x) || mode < RF_OP_AX || mode > RF_OP_VELOCITY) {
        set_error("rf_operator: null buffer or unknown mode");
        return RF_EINVAL;
    }
    int rc;
    if (mode != RF_OP_RHS) HIP_TRY(hipMemcpyAsync(s->d_xc, x, sizeof(double) * s->n, hipMemcpyHostToDevice, s->stream));
    const int kmode = (mode == RF_OP_AX) ? kOpAx : (mode == RF_OP_RESIDUAL) ? kOpResidual : (mode == RF_OP_RHS) ? kOpRhs : kOpVelocity;
    double* out = (mode == RF_OP_VELOCITY) ? s->d_y6 : s->d_t ? s->d_t : s->d_w;
    if ((rc = s->apply_operator(kmode, s->d_xc, out))) return rc;
    if (mode != RF_OP_VELOCITY) return copy_out(s, y, out, (size_t)s->n);
    // V and Omega of these coefficients (the handle's own velocities stay unchanged)
    double* v = s->d_r;
    double* w = s->d_r + 3 * (size_t)s->N;
    HIP_TRY(launch_velocity_epilogue(s->N, s->tab.d, s->tab.e, s->p.shear_rate, s->d_pos[s->lvl_three], s->d_orient[s->lvl_three],
                                     s->d_force, s->d_y6, v, w, s->stream));
    return copy_out(s, y, v, 6 * (size_t)s->N);
}

extern "C" int rf_get_neighbours(rf_sim* s, int32_t k, int32_t* nbr, double* dist)
{
    CHECK_SIM(s);
    if (!nbr || !dist || k < 1 || k > s->knn) {
        set_error("rf_get_neighbours: need 1 <= k <= " + std::to_string(s->knn));
        return RF_EINVAL;
    }
    int rc = s->setup_neighbours();
    if (rc) return rc;
    std::vector<int> tn((size_t)s->N * s->knn);
    std::vector<double> td((size_t)s->N * s->knn);
    HIP_TRY(hipMemcpyAsync(tn.data(), s->d_nbr, sizeof(int) * tn.size(), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(td.data(), s->d_nbr_dist, sizeof(double) * td.size(), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int i = 0; i < s->N; ++i)
        for (int r = 0; r < k; ++r) {
            nbr[(size_t)i * k + r] = tn[(size_t)i * s->knn + r];
            dist[(size_t)i * k + r] = td[(size_t)i * s->knn + r];
        }
    return RF_OK;
}

extern "C" int rf_precond_apply(rf_sim* s, const double* r, double* z)
{
    CHECK_SIM(s);
    if (!r || !z || !s->prec_ready) {
        set_error("rf_precond_apply: no preconditioner (needs precond and a previous rf_solve)");
        return RF_EINVAL;
    }
    HIP_TRY(hipMemcpyAsync(s->d_t, r, sizeof(double) * s->n, hipMemcpyHostToDevice, s->stream));
    int rc = s->precond_apply(s->d_t, s->d_z);
    if (rc) return rc;
    return copy_out(s, z, s->d_z, (size_t)s->n);
}

// ---- learned initial guess ------------------------------------------------

extern "C" int64_t rf_mlp_num_weights(void) { return mlp_layout().total; }

extern "C" int rf_mlp_set_weights(rf_sim* s, const float* w)
{
    CHECK_SIM(s);
    if (!w) {
        set_error("rf_mlp_set_weights: null");
        return RF_EINVAL;
    }
    int rc = s->mlp_alloc();
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(s->d_mlp_w, w, sizeof(float) * mlp_layout().total, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->mlp_weights = true;
    return RF_OK;
}

extern "C" int rf_mlp_get_weights(rf_sim* s, float* w)
{
    CHECK_SIM(s);
    if (!w || !s->mlp_weights) {
        set_error("rf_mlp_get_weights: no weights loaded");
        return RF_EINVAL;
    }
    HIP_TRY(hipMemcpyAsync(w, s->d_mlp_w, sizeof(float) * mlp_layout().total, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return RF_OK;
}

extern "C" int rf_mlp_samples(rf_sim* s, float* X, float* Y)
{
    CHECK_SIM(s);
    if (!X || !Y || !s->solved) {
        set_error("rf_mlp_samples: needs output buffers and a solved system (rf_solve) at this configuration");
        return RF_EINVAL;
    }
    int rc = s->mlp_alloc();
    if (rc) return rc;
    if ((rc = s->setup_neighbours())) return rc;
    HIP_TRY(mlp_features(s->N, s->knn, s->d_pos[s->lvl_three], s->d_orient[s->lvl_three], s->d_nbr, s->d_nbr_dist,
                         s->d_mlp_X, s->stream));
    HIP_TRY(mlp_targets(s->N, s->M, s->d_orient[s->lvl_three], s->d_x, s->d_mlp_Y, s->stream));
    HIP_TRY(hipMemcpyAsync(X, s->d_mlp_X, sizeof(float) * (size_t)s->N * kMlpIn, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(Y, s->d_mlp_Y, sizeof(float) * (size_t)s->N * kMlpOut, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return RF_OK;
}

extern "C" int rf_mlp_predict(rf_sim* s, double* x)
{
    CHECK_SIM(s);
    if (!x) {
        set_error("rf_mlp_predict: null");
        return RF_EINVAL;
    }
    int rc = s->mlp_infer(s->d_guess);
    if (rc) return rc;
    return copy_out(s, x, s->d_guess, (size_t)s->n);
}

extern "C" int rf_mlp_forward_backward(rf_sim* s, const float* X, const float* Y, int32_t B, float* grad,
                                       double* loss)
{
    CHECK_SIM(s);
    if (!X || !Y || !grad || !loss || B < 1 || !s->mlp_weights) {
        set_error("rf_mlp_forward_backward: bad arguments or no weights loaded");
        return RF_EINVAL;
    }
    const MlpLayout L = mlp_layout();
    const int64_t used = L.total + (int64_t)B * (kMlpIn + kMlpOut) + mlp_buffer_floats(B);
    const int64_t floats = used + 4;  // + one 8-byte aligned double after `used`
    float* q = nullptr;
    if (hipMalloc(&q, sizeof(float) * floats) != hipSuccess) {
        set_error("rf_mlp_forward_backward: cannot allocate");
        return RF_ENOMEM;
    }
    float* g = q;
    float* dX = g + L.total;
    float* dY = dX + (int64_t)B * kMlpIn;
    float* work = dY + (int64_t)B * kMlpOut;
    double* acc = reinterpret_cast<double*>(q + (used + 1) / 2 * 2);
    MlpBuffers buf;
    mlp_buffer_bind(work, B, &buf);
    hipError_t e = hipMemcpyAsync(dX, X, sizeof(float) * (size_t)B * kMlpIn, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dY, Y, sizeof(float) * (size_t)B * kMlpOut, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g, 0, sizeof(float) * L.total, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(acc, 0, sizeof(double), s->stream);
    if (e == hipSuccess) e = mlp_forward(s->d_mlp_w, B, dX, buf, s->stream);
    if (e == hipSuccess) e = mlp_loss(s->d_mlp_w, B, dY, buf, s->stream);
    if (e == hipSuccess) e = mlp_accumulate_loss(B, buf, acc, s->stream);
    if (e == hipSuccess) e = mlp_backward(s->d_mlp_w, B, buf, g, s->stream);
    double a = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(grad, g, sizeof(float) * L.total, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&a, acc, sizeof(double), hipMemcpyDeviceToHost, s->stream);
    hipError_t e2 = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = e2;
    (void)hipFree(q);
    if (e != hipSuccess) {
        set_error(std::string("rf_mlp_forward_backward: ") + hipGetErrorString(e));
        return RF_EHIP;
    }
    *loss = a / ((double)B * kMlpOut);
    return RF_OK;
}

extern "C" int rf_mlp_train(rf_sim* s, const float* X, const float* Y, int64_t count, int32_t epochs,
                            int32_t batch, float lr, uint64_t seed, int32_t init, float* loss_per_epoch)
{
    CHECK_SIM(s);
    if (!X || !Y || count < 1 || epochs < 1 || batch < 1 || !(lr > 0.f) || (!init && !s->mlp_weights)) {
        set_error("rf_mlp_train: bad arguments (or init = 0 without loaded weights)");
        return RF_EINVAL;
    }
    int rc = s->mlp_alloc();
    if (rc) return rc;
    const MlpLayout L = mlp_layout();
    if (init) {
        std::vector<float> w(L.total, 0.f);
        mlp_init_weights(seed, w.data());
        mlp_set_standardisation(X, Y, count, w.data());
        HIP_TRY(hipMemcpyAsync(s->d_mlp_w, w.data(), sizeof(float) * L.total, hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    float* q = nullptr;
    if (hipMalloc(&q, sizeof(float) * count * (kMlpIn + kMlpOut)) != hipSuccess) {
        set_error("rf_mlp_train: cannot allocate the training set on the device");
        return RF_ENOMEM;
    }
    hipError_t e = hipMemcpyAsync(q, X, sizeof(float) * count * kMlpIn, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(q + count * kMlpIn, Y, sizeof(float) * count * kMlpOut, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = mlp_train(s->d_mlp_w, q, q + count * kMlpIn, count, epochs, batch, lr, seed, loss_per_epoch, s->stream);
    hipError_t e2 = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = e2;
    (void)hipFree(q);
    if (e != hipSuccess) {
        set_error(std::string("rf_mlp_train: ") + hipGetErrorString(e));
        return RF_EHIP;
    }
    s->mlp_weights = true;
    return RF_OK;
}

extern "C" int rf_synchronize(rf_sim* s)
{
    CHECK_SIM(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->resolve_timed();
    return RF_OK;
}

extern "C" int rf_selftest_copy(int64_t bytes, int32_t reps, double* ms)
{
    if (bytes < 1024 || reps < 1 || !ms) {
        set_error("rf_selftest_copy: bad arguments");
        return RF_EINVAL;
    }
    bytes = bytes / 32 * 32;
    void *src = nullptr, *dst = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipMalloc(&src, bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, bytes);
    if (e == hipSuccess) e = hipMemset(src, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    double total = 0.0;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        e = hipEventRecord(a, nullptr);
        if (e == hipSuccess) e = launch_copy(src, dst, (size_t)bytes, nullptr);
        if (e == hipSuccess) e = hipEventRecord(b, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
        total += t;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e != hipSuccess) {
        set_error(std::string("rf_selftest_copy: ") + hipGetErrorString(e));
        return RF_EHIP;
    }
    *ms = total / reps;
    return RF_OK;
}

extern "C" int rf_selftest_mfma_layout(int32_t mode, double* out64, float* out32)
{
    if (!out64 || !out32 || mode < 0 || mode > 2) {
        set_error("rf_selftest_mfma_layout: bad arguments");
        return RF_EINVAL;
    }
    double* d64 = nullptr;
    float* d32 = nullptr;
    HIP_TRY(hipMalloc(&d64, 256 * sizeof(double)));
    hipError_t e = hipMalloc(&d32, 256 * sizeof(float));
    if (e == hipSuccess) e = mfma_probe(mode, d64, d32, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out64, d64, 256 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out32, d32, 256 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d64);
    if (d32) (void)hipFree(d32);
    if (e != hipSuccess) {
        set_error(std::string("rf_selftest_mfma_layout: ") + hipGetErrorString(e));
        return RF_EHIP;
    }
    return RF_OK;
}

extern "C" int rf_selftest_lu(int32_t n, int32_t elem_bytes, int32_t nb, int32_t column_panel, int32_t valu,
                              int32_t update_fp16, double* a, int32_t* ipiv, int32_t* info, int32_t* fallback)
{
    const int nb_max = (elem_bytes == 4) ? 128 : 64;
    if (n < 1 || n > 65536 || (elem_bytes != 4 && elem_bytes != 8) || nb < 0 || nb > nb_max || !a || !ipiv || !info ||
        !fallback) {
        set_error("rf_selftest_lu: bad arguments");
        return RF_EINVAL;
    }
    if (nb == 0) nb = 64;
    const int64_t lda = (n + 3) / 4 * 4;
    const size_t es = (size_t)elem_bytes;
    std::vector<char> host((size_t)lda * n * es, 0);
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = 0; r < n; ++r) {
            const double v = a[c * n + r];
            if (elem_bytes == 8)
                reinterpret_cast<double*>(host.data())[c * lda + r] = v;
            else
                reinterpret_cast<float*>(host.data())[c * lda + r] = (float)v;
        }
    void *dA = nullptr, *dP = nullptr, *dW = nullptr;
    hipError_t e = hipMalloc(&dA, host.size());
    if (e == hipSuccess) e = hipMalloc(&dP, sizeof(int) * ((size_t)n + 2));
    if (e == hipSuccess) e = hipMalloc(&dW, lu_workspace_bytes(n, update_fp16));
    if (e == hipSuccess) e = hipMemcpy(dA, host.data(), host.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dP, 0, sizeof(int) * ((size_t)n + 2));
    if (e == hipSuccess) e = hipMemset(dW, 0, lu_workspace_bytes(n, update_fp16));
    int* d_ipiv = static_cast<int*>(dP);
    int redone = 0;
    if (e == hipSuccess)
        e = lu_factor(dA, elem_bytes, lda, n, nb, d_ipiv, d_ipiv + n, valu ? 0 : 1, column_panel ? 0 : 1,
                      update_fp16, dW, &redone, nullptr, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    std::vector<int> piv((size_t)n + 2);
    if (e == hipSuccess) e = hipMemcpy(piv.data(), dP, sizeof(int) * ((size_t)n + 2), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(host.data(), dA, host.size(), hipMemcpyDeviceToHost);
    if (dA) (void)hipFree(dA);
    if (dP) (void)hipFree(dP);
    if (dW) (void)hipFree(dW);
    if (e != hipSuccess) {
        set_error(std::string("rf_selftest_lu: ") + hipGetErrorString(e));
        return RF_EHIP;
    }
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = 0; r < n; ++r)
            a[c * n + r] = (elem_bytes == 8) ? reinterpret_cast<double*>(host.data())[c * lda + r]
                                             : (double)reinterpret_cast<float*>(host.data())[c * lda + r];
    for (int i = 0; i < n; ++i) ipiv[i] = piv[i];
    *info = piv[n];
    *fallback = redone;
    return RF_OK;
}

extern "C" int32_t rf_num_fibers(const rf_sim* s) { return s ? s->N : 0; }
extern "C" int64_t rf_num_unknowns(const rf_sim* s) { return s ? s->n : 0; }

static int write_vec(const std::string& path, const std::vector<double>& v, int fmt)
{
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        set_error("cannot open " + path);
        return RF_EIO;
    }
    const char* spec = (fmt == RF_FMT_F24_16) ? "%24.16f\n" : "%16.8f\n";
    for (double x : v) std::fprintf(f, spec, x);
    std::fclose(f);
    return RF_OK;
}

// fortran/advect_fibers.f90:529-539, fortran/coeffs_and_vels.f90:55-72,633-642
extern "C" int rf_write_step_outputs(const rf_sim* s, const char* dir, int32_t format)
{
    CHECK_SIM(s);
    if (!dir) {
        set_error("rf_write_step_outputs: null dir");
        return RF_EINVAL;
    }
    const size_t N3 = 3 * (size_t)s->N;
    std::vector<double> pos(N3), ori(N3), vel(N3), rot(N3), x(s->n), b(s->n);
    int rc;
    if ((rc = rf_get_state(s, pos.data(), ori.data()))) return rc;
    if ((rc = rf_get_velocities(s, vel.data(), rot.data()))) return rc;
    if ((rc = rf_get_coefficients(s, x.data()))) return rc;
    if ((rc = rf_get_rhs(s, b.data()))) return rc;
    const std::string d(dir);
    if ((rc = write_vec(d + "/POS.out", pos, format))) return rc;
    if ((rc = write_vec(d + "/ORIENT.out", ori, format))) return rc;
    if ((rc = write_vec(d + "/TRANSVel.out", vel, format))) return rc;
    if ((rc = write_vec(d + "/ROTVel.out", rot, format))) return rc;
    if ((rc = write_vec(d + "/XVec.out", x, format))) return rc;
    return write_vec(d + "/BVec.out", b, format);
}
