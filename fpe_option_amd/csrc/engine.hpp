// Batched GPU engine: packs many independent Fokker-Planck solves into flat
// device arrays and runs the assemble -> eigen -> ADMM -> Radau -> payoff
// kernels over all of them.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "domain.hpp"

namespace fpe {

struct SolveResult {
    int status = 0, n_acc = 0, n_rej = 0, n_lu = 0, n_newt = 0, n_err = 0, n_inner = 0, n_admm = 0, n_refine = 0;
    int nsb = 0, nvb = 0;
    bool direct = false;
    double tk[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // seconds: assemble, lu, solve, mv, other, total, then factorisation parts: pivot inverse, staging, L multiply, Schur (solver workgroup)
    std::vector<double> q0, qT;              // filled when requested
    std::vector<std::vector<double>> hist;   // accepted states (stepwise API)
    std::vector<double> pdf;                 // n_qs x n_qv (stepwise API)
    std::vector<double> M49, K49;            // unscaled stencils, N x 49 (operator export)
    std::vector<double> Kx, Mx;              // sum-factorised K x, M x (RunOptions::apply_x)
    std::vector<double> bas[2];              // per-point basis tables (8 doubles) for s, v
    std::vector<int> bas_c0[2];              // first recombined column per point
};

struct BatchTimes {
    double setup = 0, assemble = 0, ic = 0, radau = 0, payoff = 0, total = 0;
    int n_slots = 0;  // Radau workgroups launched
    bool scratch_exceeded = false;  // noalloc: the batch does not fit the fixed scratch
};

// Bytes of the fixed device scratch array that serves every device buffer of an
// allocation-free call (RunOptions::noalloc, fpe_batch_price_noalloc).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;

struct Request {
    int solve;         // index into the solve list
    int option_type;   // payoff type of that solve
    double barrier, strike, sign;
    int out;           // output slot
};

struct RunOptions {
    bool want_q0 = false, want_qT = false, want_pdf = false, want_basis = false, want_stencils = false;
    int hist_cap = 0;  // > 0: keep up to this many accepted states per solve
    bool stop_after_ic = false;
    // when set (one state of N values per solve): no solve runs; prices and PDFs
    // are computed from these terminal states
    const std::vector<std::vector<double>>* given_qT = nullptr;
    // no device allocation: all buffers are slices of the fixed scratch array (pricing only)
    bool noalloc = false;
    // when set (N values for solve 0): K x and M x of solve 0 by the sum-factorised
    // products, unscaled (SolveResult::Kx, Mx)
    const std::vector<double>* apply_x = nullptr;
};

// Largest block-row count the kernels take (the v block width is bounded by MAX_BLOCK = 64).
constexpr int MAX_SB = 300;

// Whether a discretised domain fits the kernels' limits (block width, block
// rows, initial-condition LDS budget); why receives the reason when not.
bool grid_fits(const Domain& d, std::string* why);

// Runs solves + pricing requests; out_prices has n_out entries (accumulated).
// Returns false with a message in err on failure.
bool run_batch(const std::vector<SolveSpec>& solves, const std::vector<Domain>& domains,
               const std::vector<Request>& reqs, int n_out, const RunOptions& opt, std::vector<double>& out_prices,
               std::vector<SolveResult>& results, BatchTimes& times, std::string& err);

// Price options (with in-out parity); status per option.
// Options whose grid does not fit get status FPE_TOO_LARGE (5) and price 0;
// used (when given) receives the indices of each option's solves in results.
bool price_options(const std::vector<OptionSpec>& opts, std::vector<double>& prices, std::vector<int>& status,
                   std::vector<SolveResult>* results, BatchTimes* times, std::string& err,
                   std::vector<std::vector<int>>* used = nullptr, bool want_qT = false, bool noalloc = false);

int device_count();

// Sets the message fpe_last_error() returns in this thread.
void set_last_error(const std::string& msg);

}  // namespace fpe
