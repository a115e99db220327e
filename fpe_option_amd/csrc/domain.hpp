// Host-side problem setup for one Fokker-Planck solve: parameter validation,
// barrier truncation, boundary conditions, knot vectors and quadrature grids.
// These are O(n) serial steps (sorting, de-duplication) that feed the GPU
// kernels; everything O(n^2) and above runs on the device.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fpe {

// Model + option for one priced contract (the reference's FpeParams,
// src/server/option_types.mojo:4-11, with one strike).
struct OptionSpec {
    double kappa, theta, eta, rho, r, T, S0, V0;
    double s_min, s_max, v_max;
    double barrier, strike;
    int option_type;
    int n_s, n_v;
    int disc_mode, num_insert;
    int complex_mode, precision, exact_constants;
    double ic_sigma_s, ic_sigma_v, rtol, atol;
    double knot_width_s, knot_width_v;
    double knot_center_v;  // span mode: v breakpoints cluster here (<= 0: at V0)
    int lu_variant;        // factorisation code path: 0 tuned, 1 generic, 2 compiled k-steps
    int max_steps;         // Radau step limit (<= 0: 100000, the reference's)
    int refine_max;        // mixed precision: refinement rounds per stage solve (0: 10, < 0: none)
    int ic_variant;        // initial-condition ADMM launch form: 0 tuned, 1..3 forced (FpeBatchOption)
};

// One Fokker-Planck solve on a (possibly barrier-truncated) domain.
struct SolveSpec {
    double kappa, theta, eta, rho, r, T, S0, V0;
    double s_lo, s_hi, v_hi;  // physical domain after truncation
    int n_s, n_v, disc_mode, num_insert;
    bool s_left_neumann, s_right_neumann;  // v is Neumann at both ends
    double ic_sigma_s, ic_sigma_v, rtol, atol;
    int complex_mode, precision, exact_constants;
    double knot_width_s, knot_width_v;
    double knot_center_v;
    int lu_variant, max_steps, refine_max, ic_variant;
};

// Knots and quadrature of one solve.
struct Domain {
    std::vector<double> s_knots, v_knots;
    std::vector<double> s_x, s_w, s_phys;  // s: normalised points, weights (x Jacobian), physical
    std::vector<double> v_x, v_w;          // v: physical points and weights
    int n_sb = 0, n_vb = 0;                // recombined basis sizes (dofs per axis)
};

// Reference validity rules: src/engines/fpe/heston_params.mojo:22-36 and
// src/server/option_types.mojo:13-32.
bool option_is_valid(const OptionSpec& o, std::string* why);

// in-options (types 0,1,4,5) are priced by in-out parity
// (src/server/pricer.mojo:87-120).
bool is_in_option(int option_type);

// Solve for a non-in option type: truncation (src/server/option_types.mojo:34-54)
// and s boundary conditions (:56-70).
SolveSpec solve_for(const OptionSpec& o, int solve_type);

// Build knots and quadrature (reference: src/engines/fpe/domain.mojo:20-237 and
// src/numerics/bspline/knots.mojo:101-270; span mode: sinh-clustered breakpoints
// with 4-point Gauss-Legendre per span).
Domain build_domain(const SolveSpec& s);

// Undiscounted payoff (src/server/payoffs.mojo:19-50).
double payoff_value(int option_type, double barrier, double K, double S);

}  // namespace fpe
