// fibers -- command line driver: `fibers <parameters> <layout>`
// (src/main.cu:26-72, src/fiberopt.h:43-62) on top of librigidfibers.
//
// Runs the number of timesteps of the parameter file and writes, like the
// Fortran driver (fortran/advect_fibers.f90:395-424,529-591):
//   POS.out ORIENT.out TRANSVel.out ROTVel.out XVec.out BVec.out  (F16.8,
//     overwritten each step),
//   XcTData<label>_p<k>, VelRotData<label>_p<k>, Coeffvec<label>_p<k>
//     (F24.16, the initial state and every save_interval steps, a new file
//     every save_interval*saves_per_file steps).
// With --resume DIR it continues a run from the last two XcTData records, the
// last VelRotData and Coeffvec records found in DIR (written with
// save_interval 1, so that both levels of the multistep scheme are there),
// writing its snapshots to the parts after the last one.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rigidfibers.h"

static void usage()
{
    std::printf(
        "fibers - rigid fiber slender-body simulation (MI355X)\n\n"
        "Usage:\n"
        "  fibers [options] <parameters> <layout>\n"
        "  fibers --help | --version\n\n"
        "Options:\n"
        "  --out DIR        directory for output files (default .)\n"
        "  --no-output      do not write any files\n"
        "  --steps K        override the number of timesteps\n"
        "  --resume DIR     continue the run whose snapshots (save interval 1) are in DIR\n");
}

static int die(const char* what, int rc)
{
    std::fprintf(stderr, "fibers: %s failed (%d): %s\n", what, rc, rf_last_error());
    return 1;
}

struct Snap {
    FILE* x = nullptr;
    FILE* v = nullptr;
    FILE* c = nullptr;
    void close()
    {
        if (x) std::fclose(x);
        if (v) std::fclose(v);
        if (c) std::fclose(c);
        x = v = c = nullptr;
    }
    bool open(const std::string& dir, int label, int part)
    {
        close();
        const std::string sfx = std::to_string(label) + "_p" + std::to_string(part);
        x = std::fopen((dir + "/XcTData" + sfx).c_str(), "w");
        v = std::fopen((dir + "/VelRotData" + sfx).c_str(), "w");
        c = std::fopen((dir + "/Coeffvec" + sfx).c_str(), "w");
        return x && v && c;
    }
};

static bool file_exists(const std::string& path)
{
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    std::fclose(f);
    return true;
}

// Records of one snapshot series across its parts <prefix>_p1, _p2, ...
struct Series {
    std::vector<std::pair<std::string, int>> recs;  // (file, record index)
    int last_part = 0;
};

static int scan_series(const std::string& prefix, int64_t per, Series* out)
{
    for (int part = 1;; ++part) {
        const std::string f = prefix + "_p" + std::to_string(part);
        if (!file_exists(f)) break;
        int32_t nrec = 0;
        int rc = rf_read_snapshot(f.c_str(), per, -1, nullptr, &nrec);
        if (rc) return rc;
        for (int r = 0; r < nrec; ++r) out->recs.push_back({f, r});
        out->last_part = part;
    }
    return RF_OK;
}

static int read_rec(const std::pair<std::string, int>& at, int64_t per, std::vector<double>* v)
{
    v->assign((size_t)per, 0.0);
    int32_t nrec = 0;
    return rf_read_snapshot(at.first.c_str(), per, at.second, v->data(), &nrec);
}

// From the snapshots of run `label` in dir: restores the handle and returns
// the steps taken and the last part written.
static int resume_from(rf_sim* sim, const rf_params& p, const std::string& dir, int64_t* steps, int* last_part)
{
    if (p.save_interval != 1) {
        std::fprintf(stderr, "fibers: --resume needs save_interval 1 (the multistep scheme reads two levels)\n");
        return RF_EINVAL;
    }
    const int N = rf_num_fibers(sim);
    const int64_t n = rf_num_unknowns(sim);
    const std::string lab = std::to_string(p.run_label);
    Series xs, vs, cs;
    int rc;
    if ((rc = scan_series(dir + "/XcTData" + lab, 6 * (int64_t)N, &xs))) return rc;
    if ((rc = scan_series(dir + "/VelRotData" + lab, 6 * (int64_t)N, &vs))) return rc;
    if ((rc = scan_series(dir + "/Coeffvec" + lab, n, &cs))) return rc;
    if (xs.recs.size() < 2 || vs.recs.size() != xs.recs.size() || cs.recs.size() + 1 != xs.recs.size()) {
        std::fprintf(stderr, "fibers: %s holds %zu XcTData, %zu VelRotData and %zu Coeffvec records; resuming needs "
                     "at least two states with their velocities and coefficients\n", dir.c_str(), xs.recs.size(),
                     vs.recs.size(), cs.recs.size());
        return RF_EIO;
    }
    std::vector<double> cur, prev, vel, coef;
    if ((rc = read_rec(xs.recs.back(), 6 * (int64_t)N, &cur))) return rc;
    if ((rc = read_rec(xs.recs[xs.recs.size() - 2], 6 * (int64_t)N, &prev))) return rc;
    if ((rc = read_rec(vs.recs.back(), 6 * (int64_t)N, &vel))) return rc;
    if ((rc = read_rec(cs.recs.back(), n, &coef))) return rc;
    std::vector<double> X(3 * (size_t)N), T(X.size()), Xp(X.size()), Tp(X.size()), V(X.size()), W(X.size());
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) {
            X[3 * i + c] = cur[6 * i + c];
            T[3 * i + c] = cur[6 * i + 3 + c];
            Xp[3 * i + c] = prev[6 * i + c];
            Tp[3 * i + c] = prev[6 * i + 3 + c];
            V[3 * i + c] = vel[6 * i + c];
            W[3 * i + c] = vel[6 * i + 3 + c];
        }
    *steps = (int64_t)xs.recs.size() - 1;
    *last_part = std::max(xs.last_part, std::max(vs.last_part, cs.last_part));
    return rf_restore_state(sim, *steps, X.data(), T.data(), Xp.data(), Tp.data(), V.data(), W.data(), coef.data());
}

static void put3(FILE* f, const double* a)
{
    for (int k = 0; k < 3; ++k) std::fprintf(f, "%24.16f\n", a[k]);
    std::fprintf(f, " \n");
}

int main(int argc, char** argv)
{
    std::string out_dir = ".";
    std::string resume_dir;
    bool write = true;
    int steps_override = -1;
    std::vector<std::string> pos_args;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else if (a == "--version") {
            std::printf("%s\n", rf_version());
            return 0;
        } else if (a == "--out" && i + 1 < argc) {
            out_dir = argv[++i];
        } else if (a == "--no-output") {
            write = false;
        } else if (a == "--resume" && i + 1 < argc) {
            resume_dir = argv[++i];
        } else if (a == "--steps" && i + 1 < argc) {
            steps_override = std::atoi(argv[++i]);
        } else {
            pos_args.push_back(a);
        }
    }
    if (pos_args.size() != 2) {
        usage();
        return 1;
    }
    rf_params p;
    int rc = rf_parse_parameters(pos_args[0].c_str(), &p);
    if (rc) return die("parsing the parameters", rc);
    int32_t N = 0;
    double *pos = nullptr, *ori = nullptr;
    rc = rf_parse_layout(pos_args[1].c_str(), &N, &pos, &ori);
    if (rc) return die("parsing the layout", rc);
    if (steps_override >= 0) p.num_timesteps = steps_override;
    std::printf("**************************************************\n");
    std::printf("Number of fibers                   : %d\n", N);
    std::printf("Number of timesteps                : %d\n", p.num_timesteps);
    std::printf("Size of timesteps                  : %g\n", p.timestep);
    std::printf("Slenderness                        : %g\n", p.slenderness);
    std::printf("Number of terms in force expansion : %d\n", p.num_terms);
    std::printf("Number of quadrature intervals     : %d\n", p.num_quad_intervals);
    std::printf("**************************************************\n");
    rf_sim* sim = nullptr;
    rc = rf_create(&p, N, pos, ori, &sim);
    if (rc) return die("rf_create", rc);
    const size_t N3 = 3 * (size_t)N;
    const size_t n = (size_t)rf_num_unknowns(sim);
    std::vector<double> X(N3), T(N3), V(N3, 0.0), R(N3, 0.0), C(n);
    Snap snap;
    int part = 1;
    int64_t start = 0;
    if (!resume_dir.empty()) {
        int last_part = 0;
        rc = resume_from(sim, p, resume_dir, &start, &last_part);
        if (rc) return die("resuming from the snapshots", rc);
        part = last_part + 1;
        std::printf("Resumed after timestep %lld from %s\n", (long long)start, resume_dir.c_str());
        if (write && !snap.open(out_dir, p.run_label, part)) {
            std::fprintf(stderr, "fibers: cannot open snapshot files in %s\n", out_dir.c_str());
            return 1;
        }
    } else if (write) {
        if (!snap.open(out_dir, p.run_label, part)) {
            std::fprintf(stderr, "fibers: cannot open snapshot files in %s\n", out_dir.c_str());
            return 1;
        }
        rf_get_state(sim, X.data(), T.data());
        for (int i = 0; i < N; ++i) {
            put3(snap.x, &X[3 * i]);
            put3(snap.x, &T[3 * i]);
            put3(snap.v, &V[3 * i]);
            put3(snap.v, &R[3 * i]);
        }
    }
    for (int tt = (int)start + 1; tt <= p.num_timesteps; ++tt) {
        rc = rf_step(sim);
        if (rc) return die("rf_step", rc);
        rf_stats st;
        rf_get_stats(sim, &st);
        std::printf("     [GPU]      : Timestep %d of %d  assemble %.3f ms  solve %.3f ms (%s, %d it, be %.2e)  "
                    "velocities %.3f ms  update %.3f ms\n",
                    tt, p.num_timesteps, st.ms_assemble, st.ms_solve,
                    st.last_solver == RF_SOLVER_GMRES ? "gmres" : (st.last_solver == RF_SOLVER_LU ? "lu" : "lu-fallback"),
                    st.last_gmres_iters, st.last_backward_error, st.ms_velocities, st.ms_advance);
        if (!write) continue;
        rc = rf_write_step_outputs(sim, out_dir.c_str(), RF_FMT_F16_8);
        if (rc) return die("writing step outputs", rc);
        if (p.save_interval > 0 && tt % p.save_interval == 0) {
            rf_get_state(sim, X.data(), T.data());
            rf_get_velocities(sim, V.data(), R.data());
            rf_get_coefficients(sim, C.data());
            for (int i = 0; i < N; ++i) {
                put3(snap.x, &X[3 * i]);
                put3(snap.x, &T[3 * i]);
                put3(snap.v, &V[3 * i]);
                put3(snap.v, &R[3 * i]);
            }
            for (size_t k = 0; k < n; ++k) std::fprintf(snap.c, "%24.16f\n", C[k]);
            std::fprintf(snap.c, " \n");
        }
        if (p.save_interval > 0 && p.saves_per_file > 0 && tt % (p.save_interval * p.saves_per_file) == 0 &&
            tt < p.num_timesteps) {
            ++part;
            if (!snap.open(out_dir, p.run_label, part)) {
                std::fprintf(stderr, "fibers: cannot open snapshot files\n");
                return 1;
            }
        }
    }
    snap.close();
    rf_stats st;
    rf_get_stats(sim, &st);
    std::printf("DONE! steps %lld  total ms: assemble %.3f solve %.3f velocities %.3f update %.3f  gmres its %lld\n",
                (long long)st.steps_taken, st.total_ms_assemble, st.total_ms_solve, st.total_ms_velocities,
                st.total_ms_advance, (long long)st.total_gmres_iters);
    rf_destroy(sim);
    rf_free(pos);
    rf_free(ori);
    return 0;
}
