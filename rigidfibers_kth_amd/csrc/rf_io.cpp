// rf_io.cpp -- the reference's file formats: the line-ordered v1 run
// parameter file and the XcT_init layout file.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rf_internal.h"

extern "C" void rf_params_default(rf_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->run_label = 1;
    p->layout_label = 1;
    p->num_terms = 5;
    p->slenderness = 0.01;
    p->timestep = 0.1;
    p->num_timesteps = 10;
    p->save_interval = 1;
    p->saves_per_file = 10;
    p->num_quad_intervals = 8;
    p->num_quad_points_per_interval = 3;
    p->use_analytical_integration = 1;
    p->use_direct_solver = 0;
    p->gmres_restart = 10;   // run100_iter.in
    p->gmres_max_iters = 100;
    p->gmres_tol = 1e-5;     // init_dgmres default cntl(1) (dGMRES.f)
    p->external_force[0] = 0.0;
    p->external_force[1] = 0.0;
    p->external_force[2] = -1.0;  // gravity (advect_fibers.f90:307-312)
    p->precision = RF_PREC_FP64;
    p->warm_start = RF_WARM_PREV;
    p->profile = 0;
    p->lu_block = 0;
    p->mixed_inner_tol = 0.0;
    p->shear_rate = 0.0;  // fortran/coeffs_and_vels.f90:574
}

namespace {

std::string strip_comment(const std::string& line)
{
    const size_t bang = line.find('!');  // Fortran-style trailing comment
    std::string s = (bang == std::string::npos) ? line : line.substr(0, bang);
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

bool parse_int(const std::string& tok, int32_t* out)
{
    char* end = nullptr;
    const long v = std::strtol(tok.c_str(), &end, 10);
    if (end == tok.c_str()) return false;
    *out = (int32_t)v;
    return true;
}

bool parse_double(const std::string& tok, double* out)
{
    std::string t = tok;
    for (char& ch : t)
        if (ch == 'd' || ch == 'D') ch = 'e';  // Fortran double exponent
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str()) return false;
    *out = v;
    return true;
}

}  // namespace

// src/parameters.cc:124-180 and fortran/advect_fibers.f90:182-222: values are
// read by line position; lines 11..13 exist only for the iterative solver.
extern "C" int rf_parse_parameters(const char* path, rf_params* out)
{
    if (!path || !out) {
        rf::set_error("rf_parse_parameters: null argument");
        return RF_EINVAL;
    }
    std::ifstream f(path);
    if (!f) {
        rf::set_error(std::string("Could not open parameters file: ") + path);
        return RF_EIO;
    }
    rf_params p;
    rf_params_default(&p);
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(f, line)) lines.push_back(line);
    auto first_token = [](const std::string& s) {
        std::istringstream is(s);
        std::string t;
        is >> t;
        return t;
    };
    int idx = 0;      // index of the next mandatory value
    size_t li = 0;    // line cursor
    bool ok = true;
    int mandatory = 11;
    for (; li < lines.size() && idx < mandatory; ++li) {
        const std::string s = strip_comment(lines[li]);
        if (s.empty()) continue;
        const std::string t = first_token(s);
        switch (idx) {
        case 0: ok = parse_int(t, &p.run_label); break;
        case 1: ok = parse_int(t, &p.layout_label); break;
        case 2: ok = parse_int(t, &p.num_terms); break;
        case 3: ok = parse_double(t, &p.slenderness); break;
        case 4: ok = parse_double(t, &p.timestep); break;
        case 5: ok = parse_int(t, &p.num_timesteps); break;
        case 6: ok = parse_int(t, &p.save_interval); break;
        case 7: ok = parse_int(t, &p.saves_per_file); break;
        case 8: ok = parse_int(t, &p.num_quad_intervals); break;
        case 9: ok = parse_int(t, &p.use_analytical_integration); break;
        case 10:
            ok = parse_int(t, &p.use_direct_solver);
            if (ok && p.use_direct_solver == 0) mandatory = 14;
            break;
        case 11: ok = parse_int(t, &p.gmres_restart); break;
        case 12: ok = parse_int(t, &p.gmres_max_iters); break;
        case 13: ok = parse_double(t, &p.gmres_tol); break;
        }
        if (!ok) {
            rf::set_error(std::string("rf_parse_parameters: bad value on line ") + std::to_string(li + 1) + " of " +
                          path);
            return RF_EIO;
        }
        ++idx;
    }
    if (idx < mandatory) {
        rf::set_error(std::string("rf_parse_parameters: expected ") + std::to_string(mandatory) +
                      " parameter lines in " + path);
        return RF_EIO;
    }
    // optional "key value" extension lines
    for (; li < lines.size(); ++li) {
        const std::string s = strip_comment(lines[li]);
        if (s.empty()) continue;
        std::istringstream is(s);
        std::string key, val;
        is >> key >> val;
        if (key == "precision") {
            if (val == "fp64") p.precision = RF_PREC_FP64;
            else if (val == "mixed") p.precision = RF_PREC_MIXED;
            else ok = false;
        } else if (key == "warm_start") {
            if (val == "prev") p.warm_start = RF_WARM_PREV;
            else if (val == "zero") p.warm_start = RF_WARM_ZERO;
            else if (val == "extrap") p.warm_start = RF_WARM_EXTRAP;
            else if (val == "learned") p.warm_start = RF_WARM_LEARNED;
            else ok = false;
        } else if (key == "profile") {
            ok = parse_int(val, &p.profile);
        } else if (key == "lu_block") {
            ok = parse_int(val, &p.lu_block);
        } else if (key == "precond") {
            if (val == "none") p.precond = RF_PRECOND_NONE;
            else if (val == "neighbour") p.precond = RF_PRECOND_NEIGHBOUR;
            else if (val == "lu") p.precond = RF_PRECOND_LU;
            else if (val == "auto") p.precond = RF_PRECOND_AUTO;
            else ok = false;
        } else if (key == "precond_neighbours") {
            ok = parse_int(val, &p.precond_neighbours);
        } else if (key == "precond_switch_iters") {
            ok = parse_int(val, &p.precond_switch_iters);
        } else if (key == "lu_factor_fp32") {
            ok = parse_int(val, &p.lu_factor_fp32);
        } else if (key == "lu_column_panel") {
            ok = parse_int(val, &p.lu_column_panel);
        } else if (key == "lu_update_fp16") {
            ok = parse_int(val, &p.lu_update_fp16);
        } else if (key == "refine_tol") {
            ok = parse_double(val, &p.refine_tol);
        } else if (key == "mixed_inner_tol") {
            ok = parse_double(val, &p.mixed_inner_tol);
        } else if (key == "shear_rate") {
            ok = parse_double(val, &p.shear_rate);
        } else {
            ok = false;
        }
        if (!ok) {
            rf::set_error("rf_parse_parameters: unknown extension line '" + s + "'");
            return RF_EIO;
        }
    }
    p.num_quad_points_per_interval = 3;  // src/parameters.cc:36
    *out = p;
    return RF_OK;
}

// Free-format reader (fortran/advect_fibers.f90:195,318-321): N then six
// numbers per fiber; the line-based CUDA reader (src/parameters.cc:186-227)
// accepts a subset of these files.
extern "C" int rf_parse_layout(const char* path, int32_t* n, double** pos, double** orient)
{
    if (!path || !n || !pos || !orient) {
        rf::set_error("rf_parse_layout: null argument");
        return RF_EINVAL;
    }
    std::ifstream f(path);
    if (!f) {
        rf::set_error(std::string("Could not open layout file: ") + path);
        return RF_EIO;
    }
    std::string tok;
    if (!(f >> tok)) {
        rf::set_error(std::string("rf_parse_layout: empty file ") + path);
        return RF_EIO;
    }
    int32_t N = 0;
    if (!parse_int(tok, &N) || N < 1) {
        rf::set_error(std::string("rf_parse_layout: bad fiber count in ") + path);
        return RF_EIO;
    }
    double* P = static_cast<double*>(std::malloc(sizeof(double) * 3 * (size_t)N));
    double* T = static_cast<double*>(std::malloc(sizeof(double) * 3 * (size_t)N));
    if (!P || !T) {
        std::free(P);
        std::free(T);
        rf::set_error("rf_parse_layout: out of memory");
        return RF_ENOMEM;
    }
    for (int32_t i = 0; i < N; ++i) {
        double v[6];
        for (int c = 0; c < 6; ++c) {
            if (!(f >> tok) || !parse_double(tok, &v[c])) {
                std::free(P);
                std::free(T);
                rf::set_error(std::string("rf_parse_layout: file ended or bad number at fiber ") +
                              std::to_string(i) + " in " + path);
                return RF_EIO;
            }
        }
        for (int c = 0; c < 3; ++c) {
            P[3 * i + c] = v[c];
            T[3 * i + c] = v[3 + c];
        }
    }
    *n = N;
    *pos = P;
    *orient = T;
    return RF_OK;
}

extern "C" void rf_free(void* p) { std::free(p); }

// Snapshot files of fortran/advect_fibers.f90:395-424,546-591: '(2F24.16)'
// one value per line, blank lines between 3-vectors and after each
// coefficient block.  Records are consecutive groups of values_per_record
// numbers; blank lines carry no record boundary of their own.
extern "C" int rf_read_snapshot(const char* path, int64_t values_per_record, int32_t record, double* out,
                                int32_t* num_records)
{
    if (!path || !num_records || values_per_record < 1 || record < -1) {
        rf::set_error("rf_read_snapshot: null argument, values_per_record < 1 or record < -1");
        return RF_EINVAL;
    }
    std::ifstream f(path);
    if (!f) {
        rf::set_error(std::string("rf_read_snapshot: cannot open ") + path);
        return RF_EIO;
    }
    std::vector<double> cur;
    std::vector<double> want;
    cur.reserve((size_t)values_per_record);
    int32_t complete = 0;
    std::string tok;
    while (f >> tok) {
        double v = 0.0;
        if (!parse_double(tok, &v)) {
            rf::set_error(std::string("rf_read_snapshot: bad number '") + tok + "' in " + path);
            return RF_EIO;
        }
        cur.push_back(v);
        if ((int64_t)cur.size() == values_per_record) {
            if (record == -1 || record == complete) want.swap(cur);
            cur.clear();
            ++complete;
        }
    }
    *num_records = complete;
    if (!out) return RF_OK;
    if (complete == 0 || (record >= 0 && record >= complete)) {
        rf::set_error(std::string("rf_read_snapshot: ") + path + " holds " + std::to_string(complete) +
                      " complete records of " + std::to_string(values_per_record) + " values");
        return RF_EIO;
    }
    std::copy(want.begin(), want.end(), out);
    return RF_OK;
}
