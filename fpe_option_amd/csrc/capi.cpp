// C ABI of the engine (include/fpe_engine.h): the reference's 17 symbols with
// identical layouts and ownership rules, plus batched extensions.
#include "../../include/fpe_engine.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "domain.hpp"
#include "engine.hpp"

namespace {

thread_local std::string g_last_error;

fpe::OptionSpec from_batch(const FpeBatchOption& b) {
    fpe::OptionSpec o{};
    o.kappa = b.kappa, o.theta = b.theta, o.eta = b.sigma, o.rho = b.rho, o.r = b.r, o.T = b.T;
    o.S0 = b.S0, o.V0 = b.V0;
    o.s_min = b.s_min >= 0.0 ? b.s_min : 0.0;
    o.s_max = b.s_max > 0.0 ? b.s_max : 3.0 * b.S0;
    o.v_max = b.v_max > 0.0 ? b.v_max : 1.0;
    o.barrier = b.barrier, o.strike = b.strike, o.option_type = b.option_type;
    o.n_s = b.n_s, o.n_v = b.n_v, o.disc_mode = b.disc_mode, o.num_insert = b.num_insert;
    o.complex_mode = b.complex_mode, o.precision = b.precision, o.exact_constants = b.exact_constants;
    o.ic_sigma_s = b.ic_sigma_s, o.ic_sigma_v = b.ic_sigma_v;
    o.rtol = b.rtol > 0 ? b.rtol : 1e-4;
    o.atol = b.atol > 0 ? b.atol : 1e-6;
    o.knot_width_s = b.knot_width_s, o.knot_width_v = b.knot_width_v;
    o.knot_center_v = b.knot_center_v;
    o.lu_variant = b.lu_variant, o.max_steps = b.max_steps, o.refine_max = b.refine_max;
    o.ic_variant = b.ic_variant;
    return o;
}

const char* status_text(int st) {
    switch (st) {
        case 0: return "ok";
        case 1: return "invalid parameters";
        case 2: return "Radau: max steps exceeded";
        case 3: return "Radau: step size underflow";
        case 4: return "Radau: non-finite error norm";
        case 5: return "grid outside the kernels' limits";
        case 6: return "singular diagonal block in the stage-matrix factorisation";
        case 7: return "mixed-precision refinement did not reach its tolerance";
        default: return "unknown status";
    }
}

// False (with a message) when any solve of a batch did not succeed.
bool solves_ok(const std::vector<fpe::SolveResult>& res, std::string& err) {
    for (size_t k = 0; k < res.size(); ++k)
        if (res[k].status != 0) {
            err = std::string("solve ") + std::to_string(k) + " failed: " + status_text(res[k].status) + " (status " +
                  std::to_string(res[k].status) + ")";
            return false;
        }
    return true;
}

double* heap_copy(const std::vector<double>& v) {
    if (v.empty()) return nullptr;
    double* p = static_cast<double*>(std::malloc(v.size() * sizeof(double)));
    if (p) std::memcpy(p, v.data(), v.size() * sizeof(double));
    return p;
}

fpe::OptionSpec make_option(double kappa, double theta, double sigma, double rho, double r, double T, double S0,
                            double V0, int n_s, int n_v, double barrier, int type, int num_insert, double s_min,
                            double s_max) {
    fpe::OptionSpec o{};
    o.kappa = kappa, o.theta = theta, o.eta = sigma, o.rho = rho, o.r = r, o.T = T, o.S0 = S0, o.V0 = V0;
    // c_abi.mojo:127-128
    o.s_min = s_min >= 0.0 ? s_min : 0.0;
    o.s_max = s_max > 0.0 ? s_max : S0 * 3.0;
    o.v_max = 1.0;
    o.barrier = barrier, o.strike = S0, o.option_type = type;
    o.n_s = n_s, o.n_v = n_v;
    o.disc_mode = 0, o.num_insert = num_insert;
    o.complex_mode = 0, o.precision = 0, o.exact_constants = 0;
    o.ic_sigma_s = 0, o.ic_sigma_v = 0, o.rtol = 1e-4, o.atol = 1e-6;
    o.knot_width_s = 0.1, o.knot_width_v = 0.1;
    return o;
}

// Prices of one option at several strikes: one solve (two for in-options via
// parity), one pricing request per strike (src/server/pricer.mojo:82-155).
// Fails (false, message in err) when a grid does not fit the kernels or a
// solve does not succeed, as the reference raises from its solver
// (src/numerics/ode/radau.mojo:257-281,458-464).
bool price_strikes(const std::vector<fpe::OptionSpec>& opts, const std::vector<double>& K,
                   std::vector<double>& prices, std::string& err) {
    std::vector<fpe::SolveSpec> solves;
    std::vector<fpe::Domain> doms;
    std::vector<fpe::Request> reqs;
    const int nK = int(K.size());
    for (size_t a = 0; a < opts.size(); ++a) {
        const fpe::OptionSpec& o = opts[a];
        const int t = o.option_type;
        auto add = [&](int type, double sign) {
            fpe::SolveSpec s = fpe::solve_for(o, type);
            solves.push_back(s);
            doms.push_back(fpe::build_domain(s));
            if (!fpe::grid_fits(doms.back(), &err)) return false;
            for (int k = 0; k < nK; ++k)
                reqs.push_back({int(solves.size()) - 1, type, type >= 8 ? 0.0 : o.barrier, K[k], sign,
                                int(a) * nK + k});
            return true;
        };
        bool ok;
        if (fpe::is_in_option(t)) {
            const bool call = t % 2 == 0, down = t <= 1;
            ok = add(call ? 8 : 9, 1.0) && add((down ? 2 : 6) + (call ? 0 : 1), -1.0);
        } else {
            ok = add(t, 1.0);
        }
        if (!ok) return false;
    }
    std::vector<fpe::SolveResult> res;
    fpe::BatchTimes bt;
    fpe::RunOptions ro;
    if (!fpe::run_batch(solves, doms, reqs, int(opts.size()) * nK, ro, prices, res, bt, err)) return false;
    return solves_ok(res, err);
}

}  // namespace

// Pipeline state (src/server/compute_pipeline.mojo:17-48): domain built
// eagerly, initial condition / solve / PDF lazily on the GPU.
struct FpeCompute {
    fpe::OptionSpec opt;
    fpe::SolveSpec solve;
    fpe::Domain dom;
    bool have_q0 = false, have_solve = false, have_basis = false;
    std::vector<double> bas[2];
    std::vector<int> bas_c0[2];
    std::vector<double> q0, pdf, qT;
    std::vector<std::vector<double>> hist;
};

namespace {

bool ensure_solve(FpeCompute* h) {
    if (h->have_solve) return true;
    std::vector<fpe::SolveResult> res;
    fpe::BatchTimes bt;
    std::vector<double> dummy;
    int cap = 256;
    for (int attempt = 0; attempt < 8; ++attempt) {
        fpe::RunOptions ro;
        ro.want_q0 = true;
        ro.want_pdf = true;
        ro.hist_cap = cap;
        ro.want_qT = true;
        if (!fpe::run_batch({h->solve}, {h->dom}, {}, 0, ro, dummy, res, bt, g_last_error)) return false;
        if (!solves_ok(res, g_last_error)) return false;
        if (res[0].n_acc + 1 <= cap) break;
        cap = res[0].n_acc + 2;
    }
    h->q0 = res[0].q0;
    h->qT = res[0].qT;
    h->hist = std::move(res[0].hist);
    h->pdf = std::move(res[0].pdf);
    h->have_q0 = h->have_solve = true;
    return true;
}

}  // namespace

extern "C" {

FpeCompute* fpe_compute_create(double kappa, double theta, double sigma, double rho, double r, double T, double S0,
                               double V0, int32_t n_s, int32_t n_v, double barrier, int32_t option_type,
                               int32_t num_insert, double s_min, double s_max) {
    fpe::OptionSpec o =
        make_option(kappa, theta, sigma, rho, r, T, S0, V0, n_s, n_v, barrier, option_type, num_insert, s_min, s_max);
    std::string why;
    if (!fpe::option_is_valid(o, &why)) {
        g_last_error = "invalid FPE parameters: " + why;
        return nullptr;
    }
    FpeCompute* h = new FpeCompute();
    h->opt = o;
    h->solve = fpe::solve_for(o, option_type);
    h->dom = fpe::build_domain(h->solve);
    if (!fpe::grid_fits(h->dom, &g_last_error)) {
        delete h;
        return nullptr;
    }
    return h;
}

void fpe_compute_destroy(FpeCompute* ptr) { delete ptr; }

void fpe_compute_knots(FpeCompute* ptr, struct FpeVec2Result* result) {
    *result = FpeVec2Result{nullptr, 0, nullptr, 0};
    if (!ptr) return;
    result->s_data = heap_copy(ptr->dom.s_knots);
    result->s_len = int32_t(ptr->dom.s_knots.size());
    result->v_data = heap_copy(ptr->dom.v_knots);
    result->v_len = int32_t(ptr->dom.v_knots.size());
}

void fpe_compute_grid_points(FpeCompute* ptr, struct FpeGridPtsResult* result) {
    *result = FpeGridPtsResult{nullptr, 0, nullptr, 0, nullptr, nullptr};
    if (!ptr) return;
    result->s_data = heap_copy(ptr->dom.s_phys);
    result->s_len = int32_t(ptr->dom.s_phys.size());
    result->v_data = heap_copy(ptr->dom.v_x);
    result->v_len = int32_t(ptr->dom.v_x.size());
    result->sw_data = heap_copy(ptr->dom.s_w);
    result->vw_data = heap_copy(ptr->dom.v_w);
}

void fpe_compute_initial_condition(FpeCompute* ptr, struct FpeVecResult* result) {
    *result = FpeVecResult{nullptr, 0};
    if (!ptr) return;
    if (!ptr->have_q0) {
        std::vector<fpe::SolveResult> res;
        fpe::BatchTimes bt;
        std::vector<double> dummy;
        fpe::RunOptions ro;
        ro.want_q0 = true;
        ro.stop_after_ic = true;
        if (!fpe::run_batch({ptr->solve}, {ptr->dom}, {}, 0, ro, dummy, res, bt, g_last_error)) return;
        ptr->q0 = res[0].q0;
        ptr->have_q0 = true;
    }
    result->data = heap_copy(ptr->q0);
    result->len = int32_t(ptr->q0.size());
}

void fpe_compute_solve(FpeCompute* ptr, struct FpeMatResult* result) {
    *result = FpeMatResult{nullptr, 0, 0};
    if (!ptr || !ensure_solve(ptr) || ptr->hist.empty()) return;
    const size_t rows = ptr->hist.size(), cols = ptr->hist[0].size();
    std::vector<double> flat;
    flat.reserve(rows * cols);
    for (auto& h : ptr->hist) flat.insert(flat.end(), h.begin(), h.end());
    result->data = heap_copy(flat);
    result->n_rows = int32_t(rows);
    result->n_cols = int32_t(cols);
}

void fpe_compute_pdf(FpeCompute* ptr, struct FpeMatResult* result) {
    *result = FpeMatResult{nullptr, 0, 0};
    if (!ptr || !ensure_solve(ptr)) return;
    result->data = heap_copy(ptr->pdf);
    result->n_rows = int32_t(ptr->dom.s_x.size());
    result->n_cols = int32_t(ptr->dom.v_x.size());
}

// compute_pipeline.mojo:116-151: vanilla and out types price from the cached
// solve (its terminal state), in-types run the two-solve parity pricer.
void fpe_compute_price(FpeCompute* ptr, const double* K, int32_t n_K, struct FpeVecResult* result) {
    *result = FpeVecResult{nullptr, 0};
    if (!ptr || n_K < 0) return;
    std::vector<double> k(K, K + n_K), prices;
    const int t = ptr->opt.option_type;
    if (fpe::is_in_option(t)) {
        if (!price_strikes({ptr->opt}, k, prices, g_last_error)) return;
    } else {
        if (!ensure_solve(ptr)) return;
        std::vector<fpe::Request> reqs;
        for (int j = 0; j < n_K; ++j) reqs.push_back({0, t, t >= 8 ? 0.0 : ptr->opt.barrier, k[j], 1.0, j});
        std::vector<std::vector<double>> given(1, ptr->qT);
        std::vector<fpe::SolveResult> res;
        fpe::BatchTimes bt;
        fpe::RunOptions ro;
        ro.given_qT = &given;
        if (!fpe::run_batch({ptr->solve}, {ptr->dom}, reqs, n_K, ro, prices, res, bt, g_last_error)) return;
    }
    result->data = heap_copy(prices);
    result->len = n_K;
}

// compute_pipeline.mojo:258-276: bumps are applied to the truncated (revised)
// domain, as in the reference.
void fpe_compute_greeks(FpeCompute* ptr, const double* K, int32_t n_K, double rel_s, double rel_v,
                        struct FpeGreeksResult* result) {
    *result = FpeGreeksResult{nullptr, nullptr, nullptr, 0};
    if (!ptr || n_K < 0) return;
    std::vector<double> k(K, K + n_K);
    fpe::OptionSpec base = ptr->opt;
    fpe::OptionSpec rev = base;
    rev.s_min = ptr->solve.s_lo;
    rev.s_max = ptr->solve.s_hi;
    const double hs = rel_s * rev.S0, hv = rel_v * rev.V0;
    std::vector<fpe::OptionSpec> set(5, rev);
    set[0] = base;
    set[1].S0 += hs;
    set[2].S0 -= hs;
    set[3].V0 += hv;
    set[4].V0 -= hv;
    std::vector<double> p;
    if (!price_strikes(set, k, p, g_last_error)) return;
    std::vector<double> d(n_K), g(n_K), v(n_K);
    const double i2h = 1.0 / (2.0 * hs), ih2 = 1.0 / (hs * hs), i2v = 1.0 / (2.0 * hv);
    for (int j = 0; j < n_K; ++j) {
        const double p0 = p[j], us = p[n_K + j], ds = p[2 * n_K + j], uv = p[3 * n_K + j], dv = p[4 * n_K + j];
        d[j] = (us - ds) * i2h;
        g[j] = (us - 2.0 * p0 + ds) * ih2;
        v[j] = (uv - dv) * i2v;
    }
    result->delta = heap_copy(d);
    result->gamma = heap_copy(g);
    result->vega = heap_copy(v);
    result->len = n_K;
}

// c_abi.mojo:356-440 + src/server/pricing_engine.mojo:16-52: base price and the
// four bumped prices run as one GPU batch.
void fpe_price_oneshot(double kappa, double theta, double sigma, double rho, double r, double T, double S0, double V0,
                       const double* K, int32_t n_K, double barrier, int32_t option_type, int32_t n_s, int32_t n_v,
                       int32_t num_insert, double s_min, double s_max, struct FpeOneshotResult* result) {
    *result = FpeOneshotResult{nullptr, nullptr, nullptr, nullptr, 0};
    if (n_K < 0) return;
    fpe::OptionSpec o =
        make_option(kappa, theta, sigma, rho, r, T, S0, V0, n_s, n_v, barrier, option_type, num_insert, s_min, s_max);
    std::vector<double> k(K, K + n_K);
    bool valid = n_K > 0;
    for (int j = 0; j < n_K && valid; ++j) {
        o.strike = k[j];
        valid = fpe::option_is_valid(o, nullptr);
    }
    if (!valid) {  // pricing_engine.mojo:17-24: one failed result of zeros
        std::vector<double> z(1, 0.0);
        result->price = heap_copy(z), result->delta = heap_copy(z);
        result->gamma = heap_copy(z), result->vega = heap_copy(z);
        result->len = 1;
        return;
    }
    const double hs = 0.01 * S0, hv = 0.1 * V0;
    std::vector<fpe::OptionSpec> set(5, o);
    set[1].S0 += hs;
    set[2].S0 -= hs;
    set[3].V0 += hv;
    set[4].V0 -= hv;
    std::vector<double> p;
    if (!price_strikes(set, k, p, g_last_error)) return;
    std::vector<double> pr(n_K), d(n_K), g(n_K), v(n_K);
    const double i2h = 1.0 / (2.0 * hs), ih2 = 1.0 / (hs * hs), i2v = 1.0 / (2.0 * hv);
    for (int j = 0; j < n_K; ++j) {
        const double p0 = p[j], us = p[n_K + j], ds = p[2 * n_K + j], uv = p[3 * n_K + j], dv = p[4 * n_K + j];
        pr[j] = p0;
        d[j] = (us - ds) * i2h;
        g[j] = (us - 2.0 * p0 + ds) * ih2;
        v[j] = (uv - dv) * i2v;
    }
    result->price = heap_copy(pr), result->delta = heap_copy(d);
    result->gamma = heap_copy(g), result->vega = heap_copy(v);
    result->len = n_K;
}

void fpe_compute_free_vec(struct FpeVecResult* r) {
    if (r && r->data) std::free(r->data), r->data = nullptr;
}
void fpe_compute_free_vec2(struct FpeVec2Result* r) {
    if (!r) return;
    std::free(r->s_data), std::free(r->v_data);
    r->s_data = r->v_data = nullptr;
}
void fpe_compute_free_grid_pts(struct FpeGridPtsResult* r) {
    if (!r) return;
    std::free(r->s_data), std::free(r->v_data), std::free(r->sw_data), std::free(r->vw_data);
    r->s_data = r->v_data = r->sw_data = r->vw_data = nullptr;
}
void fpe_compute_free_mat(struct FpeMatResult* r) {
    if (r && r->data) std::free(r->data), r->data = nullptr;
}
void fpe_compute_free_greeks(struct FpeGreeksResult* r) {
    if (!r) return;
    std::free(r->delta), std::free(r->gamma), std::free(r->vega);
    r->delta = r->gamma = r->vega = nullptr;
}
void fpe_compute_free_oneshot(struct FpeOneshotResult* r) {
    if (!r) return;
    std::free(r->price), std::free(r->delta), std::free(r->gamma), std::free(r->vega);
    r->price = r->delta = r->gamma = r->vega = nullptr;
}

void fpe_batch_option_default(FpeBatchOption* o) {
    std::memset(o, 0, sizeof(*o));
    o->kappa = 1.2, o->theta = 0.05, o->sigma = 0.35, o->rho = -0.4, o->r = 0.05, o->T = 0.5;
    o->S0 = 100.0, o->V0 = 0.1, o->s_min = 0.0, o->s_max = 0.0, o->v_max = 1.0;
    o->barrier = 0.0, o->strike = 100.0, o->option_type = 8;
    o->n_s = 38, o->n_v = 38, o->disc_mode = FPE_DISC_REFERENCE, o->num_insert = 251;
    o->complex_mode = FPE_CPLX_GMRES, o->precision = FPE_PREC_FP64, o->exact_constants = 0;
    o->rtol = 1e-4, o->atol = 1e-6, o->knot_width_s = 0.1, o->knot_width_v = 0.1;
}

namespace {

void fill_stats(const std::vector<fpe::SolveResult>& res, const fpe::BatchTimes& bt, FpeBatchStats* stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->n_solves = int64_t(res.size());
    for (auto& r : res) {
        stats->n_accepted += r.n_acc, stats->n_rejected += r.n_rej, stats->n_lu += r.n_lu;
        stats->n_newton += r.n_newt, stats->n_err_solves += r.n_err, stats->n_inner_iters += r.n_inner;
        stats->n_admm_iters += r.n_admm;
        stats->n_refine += r.n_refine;
    }
    for (auto& r : res) {
        const double nb = r.nsb, m = r.nvb, N = nb * m;
        const double lu = 24.67 * nb * m * m * m, sol = 14.0 * nb * m * m, spmv = 98.0 * N;
        const double fac_bytes = 7.0 * nb * m * m * 8.0;
        // direct: real + complex LU per refresh; GMRES: real LU and the solves of the inner iterations
        const bool direct = r.direct;
        stats->alg_flops += r.n_lu * (direct ? 5.0 : 1.0) * lu +
                            r.n_newt * ((direct ? 5.0 : 3.0) * sol + 6.0 * spmv) +
                            r.n_inner * (2.0 * sol + 4.0 * spmv) + r.n_err * (sol + 2.0 * spmv);
        stats->alg_bytes += r.n_lu * (direct ? 3.0 : 1.0) * 2.0 * fac_bytes +
                            r.n_newt * (direct ? 3.0 : 1.0) * fac_bytes + r.n_inner * fac_bytes + r.n_err * fac_bytes;
    }
    for (auto& r : res) {
        stats->wg_s_assemble += r.tk[0], stats->wg_s_lu += r.tk[1], stats->wg_s_solve += r.tk[2];
        stats->wg_s_spmv += r.tk[3], stats->wg_s_other += r.tk[4], stats->wg_s_total += r.tk[5];
        stats->wg_s_lu_inverse += r.tk[6], stats->wg_s_lu_stage += r.tk[7];
        stats->wg_s_lu_lmul += r.tk[8], stats->wg_s_lu_schur += r.tk[9];
    }
    stats->ms_setup = bt.setup, stats->ms_assemble = bt.assemble, stats->ms_ic = bt.ic;
    stats->ms_radau = bt.radau, stats->ms_payoff = bt.payoff, stats->ms_total = bt.total;
    stats->n_radau_slots = bt.n_slots;
}

}  // namespace

int32_t fpe_batch_price_detail(const FpeBatchOption* opts, int32_t n, double* prices, int32_t* status,
                               FpeSolveInfo* info, double* qT, int64_t ldq, FpeBatchStats* stats) {
    g_last_error.clear();
    if (n < 0 || (n > 0 && (!opts || !prices || !status)) || (qT && ldq <= 0)) {
        g_last_error = "invalid arguments";
        return 1;
    }
    std::vector<fpe::OptionSpec> v(n);
    for (int i = 0; i < n; ++i) v[i] = from_batch(opts[i]);
    std::vector<double> p;
    std::vector<int> s;
    std::vector<fpe::SolveResult> res;
    std::vector<std::vector<int>> used;
    fpe::BatchTimes bt;
    if (!fpe::price_options(v, p, s, &res, &bt, g_last_error, &used, qT != nullptr)) return 2;
    for (int i = 0; i < n; ++i) prices[i] = p[i], status[i] = s[i];
    if (qT) {  // the terminal state of a single-solve option
        for (int i = 0; i < n; ++i) {
            if (used[i].size() != 1) continue;
            const std::vector<double>& q = res[used[i][0]].qT;
            if ((int64_t)q.size() <= ldq) std::copy(q.begin(), q.end(), qT + (int64_t)i * ldq);
        }
    }
    if (info) {
        for (int i = 0; i < n; ++i) {
            FpeSolveInfo& f = info[i];
            std::memset(&f, 0, sizeof(f));
            f.status = s[i];
            f.n_solves = int32_t(used[i].size());
            for (int k : used[i]) {
                const fpe::SolveResult& r = res[k];
                f.n_accepted += r.n_acc, f.n_rejected += r.n_rej, f.n_lu += r.n_lu, f.n_newton += r.n_newt;
                f.n_err_solves += r.n_err, f.n_inner_iters += r.n_inner, f.n_admm_iters += r.n_admm;
                f.n_refine += r.n_refine;
                if (f.n_sb == 0) f.n_sb = r.nsb, f.n_vb = r.nvb;
            }
        }
    }
    if (stats) fill_stats(res, bt, stats);
    return 0;
}

int32_t fpe_batch_price(const FpeBatchOption* opts, int32_t n, double* prices, int32_t* status,
                        FpeBatchStats* stats) {
    return fpe_batch_price_detail(opts, n, prices, status, nullptr, nullptr, 0, stats);
}

// fpe_batch_price without device allocation: every device buffer of the call is
// a slice of the library's fixed scratch array (see include/fpe_engine.h).
int32_t fpe_batch_price_noalloc(const FpeBatchOption* opts, int32_t n, double* prices, int32_t* status,
                                FpeBatchStats* stats) {
    g_last_error.clear();
    if (n < 0 || (n > 0 && (!opts || !prices || !status))) {
        g_last_error = "invalid arguments";
        return 1;
    }
    std::vector<fpe::OptionSpec> v(n);
    for (int i = 0; i < n; ++i) v[i] = from_batch(opts[i]);
    std::vector<double> p;
    std::vector<int> s;
    std::vector<fpe::SolveResult> res;
    fpe::BatchTimes bt;
    if (!fpe::price_options(v, p, s, &res, &bt, g_last_error, nullptr, false, true))
        return bt.scratch_exceeded ? 3 : 2;
    for (int i = 0; i < n; ++i) prices[i] = p[i], status[i] = s[i];
    if (stats) fill_stats(res, bt, stats);
    return 0;
}

void fpe_compute_basis_1d(FpeCompute* ptr, int32_t axis, int32_t deriv, struct FpeCsrResult* result) {
    *result = FpeCsrResult{nullptr, nullptr, nullptr, 0, 0, 0};
    if (!ptr || axis < 0 || axis > 1) return;
    if (!ptr->have_basis) {
        std::vector<fpe::SolveResult> res;
        fpe::BatchTimes bt;
        std::vector<double> dummy;
        fpe::RunOptions ro;
        ro.want_basis = true;
        ro.stop_after_ic = true;
        if (!fpe::run_batch({ptr->solve}, {ptr->dom}, {}, 0, ro, dummy, res, bt, g_last_error)) return;
        for (int a = 0; a < 2; ++a) ptr->bas[a] = res[0].bas[a], ptr->bas_c0[a] = res[0].bas_c0[a];
        ptr->have_basis = true;
    }
    const std::vector<double>& T = ptr->bas[axis];
    const std::vector<int>& c0 = ptr->bas_c0[axis];
    const int nq = int(c0.size()), nc = axis == 0 ? ptr->dom.n_sb : ptr->dom.n_vb;
    std::vector<double> val;
    std::vector<int32_t> idx, ip(1, 0);
    for (int p = 0; p < nq; ++p) {
        for (int t = 0; t < 4; ++t) {
            const double x = T[8LL * p + (deriv ? 4 : 0) + t];
            const int c = c0[p] + t;
            if (c < nc && x != 0.0) val.push_back(x), idx.push_back(c);
        }
        ip.push_back(int32_t(idx.size()));
    }
    result->data = heap_copy(val);
    result->indices = static_cast<int32_t*>(std::malloc(std::max<size_t>(1, idx.size()) * sizeof(int32_t)));
    result->indptr = static_cast<int32_t*>(std::malloc(ip.size() * sizeof(int32_t)));
    if (!idx.empty()) std::memcpy(result->indices, idx.data(), idx.size() * sizeof(int32_t));
    std::memcpy(result->indptr, ip.data(), ip.size() * sizeof(int32_t));
    result->n_rows = nq, result->n_cols = nc, result->nnz = int32_t(idx.size());
}

void fpe_compute_free_csr(struct FpeCsrResult* r) {
    if (!r) return;
    std::free(r->data), std::free(r->indices), std::free(r->indptr);
    r->data = nullptr, r->indices = nullptr, r->indptr = nullptr;
}

int32_t fpe_batch_terminal_states(const FpeBatchOption* opts, int32_t n, double* qT, int64_t ldq, int32_t* n_dof,
                                  int32_t* status, FpeBatchStats* stats) {
    g_last_error.clear();
    if (n < 0 || (n > 0 && (!opts || !qT || !n_dof || !status)) || ldq <= 0) {
        g_last_error = "invalid arguments";
        return 1;
    }
    std::vector<fpe::SolveSpec> solves;
    std::vector<fpe::Domain> doms;
    std::vector<int> which(n, -1);
    for (int i = 0; i < n; ++i) {
        n_dof[i] = 0;
        status[i] = 1;
        fpe::OptionSpec o = from_batch(opts[i]);
        std::string why;
        // one solve per option: vanilla on the full domain, out-types on the truncated one
        if (fpe::is_in_option(o.option_type) || !fpe::option_is_valid(o, &why)) continue;
        fpe::SolveSpec sp = fpe::solve_for(o, o.option_type);
        fpe::Domain d = fpe::build_domain(sp);
        if (!fpe::grid_fits(d, nullptr)) {
            status[i] = 5;
            continue;
        }
        if ((int64_t)d.n_sb * d.n_vb > ldq) continue;
        which[i] = int(solves.size());
        solves.push_back(sp);
        doms.push_back(std::move(d));
    }
    std::vector<fpe::SolveResult> res;
    fpe::BatchTimes bt;
    std::vector<double> dummy;
    fpe::RunOptions ro;
    ro.want_qT = true;
    if (!solves.empty() && !fpe::run_batch(solves, doms, {}, 0, ro, dummy, res, bt, g_last_error)) return 2;
    for (int i = 0; i < n; ++i) {
        if (which[i] < 0) continue;
        const fpe::SolveResult& r = res[which[i]];
        status[i] = r.status;
        n_dof[i] = int32_t(r.qT.size());
        std::copy(r.qT.begin(), r.qT.end(), qT + (int64_t)i * ldq);
    }
    if (stats) fill_stats(res, bt, stats);
    return 0;
}

int32_t fpe_batch_operators(const FpeBatchOption* opt, int32_t cap_dof, double* M49, double* K49, int32_t* n_sb,
                            int32_t* n_vb) {
    g_last_error.clear();
    if (!opt || !M49 || !K49 || !n_sb || !n_vb) {
        g_last_error = "invalid arguments";
        return 1;
    }
    fpe::OptionSpec o = from_batch(*opt);
    std::string why;
    if (fpe::is_in_option(o.option_type) || !fpe::option_is_valid(o, &why)) {
        g_last_error = "a valid vanilla or out option is required: " + why;
        return 1;
    }
    fpe::SolveSpec sp = fpe::solve_for(o, o.option_type);
    fpe::Domain d = fpe::build_domain(sp);
    *n_sb = d.n_sb, *n_vb = d.n_vb;
    if (!fpe::grid_fits(d, &g_last_error)) return 5;
    if ((int64_t)d.n_sb * d.n_vb > cap_dof) {
        g_last_error = "capacity too small";
        return 1;
    }
    std::vector<fpe::SolveResult> res;
    fpe::BatchTimes bt;
    std::vector<double> dummy;
    fpe::RunOptions ro;
    ro.want_stencils = true;
    ro.stop_after_ic = true;
    if (!fpe::run_batch({sp}, {d}, {}, 0, ro, dummy, res, bt, g_last_error)) return 2;
    std::copy(res[0].M49.begin(), res[0].M49.end(), M49);
    std::copy(res[0].K49.begin(), res[0].K49.end(), K49);
    return 0;
}

int32_t fpe_batch_apply_operators(const FpeBatchOption* opt, int32_t n_dof, const double* x, double* Kx,
                                  double* Mx) {
    g_last_error.clear();
    if (!opt || !x || !Kx || !Mx || n_dof <= 0) {
        g_last_error = "invalid arguments";
        return 1;
    }
    fpe::OptionSpec o = from_batch(*opt);
    std::string why;
    if (fpe::is_in_option(o.option_type) || !fpe::option_is_valid(o, &why)) {
        g_last_error = "a valid vanilla or out option is required: " + why;
        return 1;
    }
    fpe::SolveSpec sp = fpe::solve_for(o, o.option_type);
    fpe::Domain d = fpe::build_domain(sp);
    if (!fpe::grid_fits(d, &g_last_error)) return 5;
    if (d.n_sb * d.n_vb != n_dof) {
        g_last_error = "n_dof does not match the grid (" + std::to_string(d.n_sb * d.n_vb) + " dofs)";
        return 1;
    }
    std::vector<double> xv(x, x + n_dof);
    std::vector<fpe::SolveResult> res;
    fpe::BatchTimes bt;
    std::vector<double> dummy;
    fpe::RunOptions ro;
    ro.apply_x = &xv;
    ro.stop_after_ic = true;
    if (!fpe::run_batch({sp}, {d}, {}, 0, ro, dummy, res, bt, g_last_error)) return 2;
    std::copy(res[0].Kx.begin(), res[0].Kx.end(), Kx);
    std::copy(res[0].Mx.begin(), res[0].Mx.end(), Mx);
    return 0;
}

int32_t fpe_european_pricing_operator(const FpeBatchOption* opt, int32_t cap_qs, int32_t cap_dof, double* s_phys,
                                      double* marg, int32_t* n_qs, int32_t* n_dof) {
    g_last_error.clear();
    if (!opt || !s_phys || !marg || !n_qs || !n_dof) {
        g_last_error = "invalid arguments";
        return 1;
    }
    fpe::OptionSpec o = from_batch(*opt);
    std::string why;
    if ((o.option_type != 8 && o.option_type != 9) || !fpe::option_is_valid(o, &why)) {
        g_last_error = "a European option (type 8 or 9) with valid parameters is required: " + why;
        return 1;
    }
    fpe::SolveSpec sp = fpe::solve_for(o, o.option_type);
    fpe::Domain d = fpe::build_domain(sp);
    const int nqs = int(d.s_phys.size()), nqv = int(d.v_w.size()), ns = d.n_sb, nv = d.n_vb;
    *n_qs = nqs, *n_dof = ns * nv;
    if (nqs > cap_qs || ns * nv > cap_dof) {
        g_last_error = "capacity too small";
        return 1;
    }
    std::vector<fpe::SolveResult> res;
    fpe::BatchTimes bt;
    std::vector<double> dummy;
    fpe::RunOptions ro;
    ro.want_basis = true;
    ro.stop_after_ic = true;
    if (!fpe::run_batch({sp}, {d}, {}, 0, ro, dummy, res, bt, g_last_error)) return 2;
    const std::vector<double>& Ts = res[0].bas[0];
    const std::vector<double>& Tv = res[0].bas[1];
    const std::vector<int>& cs = res[0].bas_c0[0];
    const std::vector<int>& cv = res[0].bas_c0[1];
    std::vector<double> om(nv, 0.0);  // omega_l = sum_j Bv[j][l] w_v,j
    for (int j = 0; j < nqv; ++j)
        for (int t = 0; t < 4; ++t) {
            const int l = cv[j] + t;
            if (l >= 0 && l < nv) om[l] += Tv[8LL * j + t] * d.v_w[j];
        }
    std::fill(marg, marg + (int64_t)nqs * ns * nv, 0.0);
    for (int i = 0; i < nqs; ++i) {
        s_phys[i] = d.s_phys[i];
        for (int t = 0; t < 4; ++t) {
            const int k = cs[i] + t;
            if (k < 0 || k >= ns) continue;
            const double a = d.s_w[i] * Ts[8LL * i + t];
            for (int l = 0; l < nv; ++l) marg[(int64_t)i * ns * nv + (int64_t)k * nv + l] += a * om[l];
        }
    }
    return 0;
}

const char* fpe_last_error(void) { return g_last_error.c_str(); }

int32_t fpe_device_count(void) { return fpe::device_count(); }

}  // extern "C"

namespace fpe {
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace fpe
