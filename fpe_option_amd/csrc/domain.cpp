// Host-side setup of one solve's discretisation. See domain.hpp.
#include "domain.hpp"

#include <algorithm>
#include <cmath>

namespace fpe {
namespace {

// Evenly spaced points start + i*step (src/numerics/utils/helpers.mojo:39-53).
void append_linspace(std::vector<double>& out, double a, double b, int n) {
    if (n <= 0) return;
    if (n == 1) {
        out.push_back(a);
        return;
    }
    const double step = (b - a) / double(n - 1);
    for (int i = 0; i < n; ++i) out.push_back(a + double(i) * step);
}

// Affine map onto [0,1] by (x - min) * (1/range) (helpers.mojo:56-87).
std::vector<double> unit_scale(const std::vector<double>& x) {
    std::vector<double> y(x.size(), 0.0);
    if (x.empty()) return y;
    auto mm = std::minmax_element(x.begin(), x.end());
    const double lo = *mm.first, range = *mm.second - *mm.first;
    if (range == 0.0) return y;
    const double inv = 1.0 / range;
    std::transform(x.begin(), x.end(), y.begin(), [&](double v) { return (v - lo) * inv; });
    return y;
}

// 4-point rules on [-1,1]. exact = false: the reference's rule
// (src/numerics/bspline/knots.mojo:33-39), whose nodes use sqrt(5/6) where the
// Gauss-Legendre nodes need sqrt(6/5); kept for reference parity. exact = true:
// the Gauss-Legendre rule, used by the span discretisation.
struct GL4 {
    double x[4], w[4];
    explicit GL4(bool exact) {
        const double r = exact ? std::sqrt(6.0 / 5.0) : std::sqrt(5.0 / 6.0);
        const double a = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * r);
        const double b = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * r);
        const double wa = (18.0 - std::sqrt(30.0)) / 36.0, wb = (18.0 + std::sqrt(30.0)) / 36.0;
        x[0] = -a, x[1] = -b, x[2] = b, x[3] = a;
        w[0] = wa, w[1] = wb, w[2] = wb, w[3] = wa;
    }
};

// Parabolic internal knots on [lo, hi] clustered at c (knots.mojo:101-163): a
// downward parabola maps the lower branch, the upward one the part after the
// point nearest to c.
std::vector<double> parabolic_knots(int n, double lo, double hi, double c) {
    const double divide = std::sqrt((hi - c) / (c - lo)) + 1.0;
    int count = (std::fmod(double(n), divide) == 0.0) ? n : int(double(n) / divide) * int(divide);
    if (count < 1) count = n;
    const double f = 2.0;
    const double a = std::fabs(2.0 * (lo - c) / (f * f));
    const double top = std::sqrt(2.0 * (hi - c) / a) + f;
    std::vector<double> xs;
    append_linspace(xs, 0.0, top, count - 1);
    xs.push_back(f);
    std::sort(xs.begin(), xs.end());
    std::vector<double> ys(xs.size());
    size_t nearest = 0;
    double best = 0.0;
    for (size_t i = 0; i < xs.size(); ++i) {
        ys[i] = -0.5 * a * (xs[i] - f) * (xs[i] - f) + c;
        const double d = std::fabs(ys[i] - c);
        if (i == 0 || d < best) best = d, nearest = i;
    }
    for (size_t i = nearest + 1; i < xs.size(); ++i) ys[i] = 0.5 * a * (xs[i] - f) * (xs[i] - f) + c;
    std::sort(ys.begin(), ys.end());
    return ys;
}

// Clamped cubic knot vector of the reference generator (knots.mojo:192-270):
// parabolic knots outside mean +- 4.5 std, 13 Chebyshev nodes inside, rounded to
// 1e-8 and de-duplicated at 1e-9, boundaries repeated p times.
std::vector<double> reference_knots(int n, double center, double b0, double b1, double mean,
                                    double sd) {
    const int p = 3;
    const double span = b1 - b0;
    const double lo_n = (mean - 4.5 * sd - b0) / span, hi_n = (mean + 4.5 * sd - b0) / span;
    std::vector<double> pool;
    // Chebyshev nodes of the first kind on [mean - 4.5 sd, mean + 4.5 sd] plus ends
    {
        const double a = mean - 4.5 * sd, b = mean + 4.5 * sd;
        const int m = 13;
        std::vector<double> cheb;
        cheb.push_back(a);
        for (int i = 0; i < m; ++i)
            cheb.push_back((b - a) / 2.0 * std::cos((2.0 * double(i) + 1.0) * (M_PI / (2.0 * double(m)))) +
                           (a + b) / 2.0);
        cheb.push_back(b);
        for (double v : cheb) pool.push_back((v - b0) / span);
    }
    std::vector<double> ends = unit_scale({b0, b1});
    for (double v : parabolic_knots(n - 2 * p, ends[0], ends[1], center)) {
        if (std::fabs(v - lo_n) < 1e-10 || v < lo_n) pool.push_back(v);
        if (std::fabs(v - hi_n) < 1e-10 || v > hi_n) pool.push_back(v);
    }
    std::sort(pool.begin(), pool.end());
    std::vector<double> inner;
    for (double v : pool) {
        const double rv = std::round(v * 1e8) / 1e8;
        if (inner.empty() || rv > inner.back() + 1e-9) inner.push_back(rv);
    }
    std::vector<double> knots(p, ends[0]);
    knots.insert(knots.end(), inner.begin(), inner.end());
    knots.insert(knots.end(), p, ends[1]);
    return knots;
}

// Reference insert grid (src/engines/fpe/domain.mojo:20-76).
std::vector<double> insert_grid(double c, double sd, double lo, double hi, int ni, bool near_zero) {
    const int left = std::max(1, int(double(ni) * 0.2));
    const int mid = std::max(3, int(double(ni) * 0.3));
    const int right = std::max(1, int(double(ni) * 0.5));
    std::vector<double> g;
    append_linspace(g, lo, c - 5.0 * sd, left);
    append_linspace(g, c - 5.0 * sd, c - sd / 5.0, std::max(1, mid / 3));
    append_linspace(g, c - sd / 5.0, c + sd / 5.0, std::max(2, mid / 3));
    g.push_back(c);
    append_linspace(g, c + sd / 5.0, c + 5.0 * sd, std::max(1, mid / 3));
    append_linspace(g, c + 5.0 * sd, hi, right);
    append_linspace(g, hi - 0.1, hi, std::max(2, int(double(ni) * 0.1)));
    if (near_zero) append_linspace(g, 0.0, 0.01, std::max(2, int(double(ni) * 0.2)));
    std::sort(g.begin(), g.end());
    g.erase(std::unique(g.begin(), g.end()), g.end());
    return g;
}

// Quadrature points of consecutive intervals. with_left (reference
// discretisation): each interval's left end with weight 0 followed by the
// reference's 4-point rule (src/engines/fpe/domain.mojo:79-111); otherwise the
// 4-point Gauss-Legendre rule.
void interval_quadrature(const std::vector<double>& g, bool with_left, std::vector<double>& x,
                         std::vector<double>& w) {
    static const GL4 ref_rule(false), gauss_rule(true);
    const GL4& gl = with_left ? ref_rule : gauss_rule;
    x.clear();
    w.clear();
    for (size_t k = 0; k + 1 < g.size(); ++k) {
        const double a = g[k], b = g[k + 1];
        if (!(b > a)) continue;
        const double mid = (a + b) * 0.5, half = (b - a) * 0.5;
        if (with_left) x.push_back(a), w.push_back(0.0);
        for (int j = 0; j < 4; ++j) x.push_back(mid + half * gl.x[j]), w.push_back(half * gl.w[j]);
    }
}

// Span-mode breakpoints: n points on [0,1], x = c + d sinh(a1 + xi (a2 - a1)).
std::vector<double> stretched_breakpoints(int n, double c, double d) {
    c = std::min(1.0, std::max(0.0, c));
    const double a1 = std::asinh(-c / d), a2 = std::asinh((1.0 - c) / d);
    std::vector<double> x(n);
    for (int i = 0; i < n; ++i) x[i] = c + d * std::sinh(a1 + (double(i) / double(n - 1)) * (a2 - a1));
    x.front() = 0.0;
    x.back() = 1.0;
    return x;
}

}  // namespace

bool is_in_option(int t) { return t == 0 || t == 1 || t == 4 || t == 5; }

bool option_is_valid(const OptionSpec& o, std::string* why) {
    auto fail = [&](const char* m) {
        if (why) *why = m;
        return false;
    };
    if (!(o.kappa > 0)) return fail("kappa must be positive");
    if (!(o.theta > 0)) return fail("theta must be positive");
    if (!(o.eta > 0)) return fail("sigma must be positive");
    if (!(o.T > 0)) return fail("T must be positive");
    if (!(o.S0 > 0)) return fail("S0 must be positive");
    if (!(o.V0 >= 0)) return fail("V0 must be non-negative");
    if (!(o.rho >= -1 && o.rho <= 1)) return fail("rho must be in [-1, 1]");
    if (!(o.s_min >= 0)) return fail("S_min must be >= 0");
    if (!(o.s_max > o.s_min)) return fail("S_max must be > S_min");
    if (!(o.v_max > 0)) return fail("V_max must be > 0");
    if (o.option_type < 0 || o.option_type > 9) return fail("option_type must be 0-9");
    if (!(o.strike > 0)) return fail("strikes must be positive");
    if (o.option_type <= 3) {
        if (o.barrier <= 0 || o.barrier >= o.S0) return fail("down barrier must satisfy 0 < B < S0");
    } else if (o.option_type <= 7) {
        if (o.barrier <= 0 || o.barrier < o.S0) return fail("up barrier must satisfy B >= S0");
    } else if (o.barrier != 0.0) {
        return fail("vanilla options take barrier = 0");
    }
    if (o.n_s < 4 || o.n_s > 256 || o.n_v < 4 || o.n_v > 256) return fail("n_s, n_v must be in [4, 256]");
    if (o.complex_mode < 0 || o.complex_mode > 1) return fail("complex_mode must be 0 (GMRES) or 1 (direct)");
    if (o.precision < 0 || o.precision > 1) return fail("precision must be 0 (FP64) or 1 (mixed)");
    return true;
}

SolveSpec solve_for(const OptionSpec& o, int t) {
    SolveSpec s{};
    s.kappa = o.kappa, s.theta = o.theta, s.eta = o.eta, s.rho = o.rho, s.r = o.r, s.T = o.T;
    s.S0 = o.S0, s.V0 = o.V0;
    s.s_lo = o.s_min, s.s_hi = o.s_max, s.v_hi = o.v_max;
    const bool down = t <= 3, up = t >= 4 && t <= 7;
    if (down) s.s_lo = o.barrier;
    if (up) s.s_hi = o.barrier;
    s.s_left_neumann = up;
    s.s_right_neumann = !up;
    s.n_s = o.n_s, s.n_v = o.n_v, s.disc_mode = o.disc_mode, s.num_insert = o.num_insert;
    s.ic_sigma_s = o.ic_sigma_s, s.ic_sigma_v = o.ic_sigma_v, s.rtol = o.rtol, s.atol = o.atol;
    s.complex_mode = o.complex_mode, s.precision = o.precision, s.exact_constants = o.exact_constants;
    s.knot_width_s = o.knot_width_s, s.knot_width_v = o.knot_width_v;
    s.knot_center_v = o.knot_center_v;
    s.lu_variant = o.lu_variant, s.max_steps = o.max_steps, s.refine_max = o.refine_max;
    s.ic_variant = o.ic_variant;
    return s;
}

Domain build_domain(const SolveSpec& s) {
    Domain d;
    const double jac = s.s_hi - s.s_lo;
    const double cs = (s.S0 - s.s_lo) / jac;
    if (s.disc_mode == 0) {
        d.s_knots = reference_knots(s.n_s, cs, s.s_lo, s.s_hi, s.S0, 0.1);
        d.v_knots = reference_knots(s.n_v, s.V0, 0.0, s.v_hi, s.V0, 0.001);
        interval_quadrature(unit_scale(insert_grid(s.S0, 0.1, s.s_lo, s.s_hi, s.num_insert, false)), true,
                            d.s_x, d.s_w);
        interval_quadrature(insert_grid(s.V0, 0.001, 0.0, s.v_hi, s.num_insert, true), true, d.v_x, d.v_w);
    } else {
        const double ws = s.knot_width_s > 0 ? s.knot_width_s : 0.1;
        const double wv = s.knot_width_v > 0 ? s.knot_width_v : 0.1;
        std::vector<double> bs = stretched_breakpoints(s.n_s, cs, ws);
        const double cv = s.knot_center_v > 0 ? s.knot_center_v : s.V0;
        std::vector<double> bv = stretched_breakpoints(s.n_v, cv / s.v_hi, wv);
        for (double& x : bv) x *= s.v_hi;
        bv.back() = s.v_hi;
        d.s_knots.assign(3, 0.0);
        d.s_knots.insert(d.s_knots.end(), bs.begin(), bs.end());
        d.s_knots.insert(d.s_knots.end(), 3, 1.0);
        d.v_knots.assign(3, 0.0);
        d.v_knots.insert(d.v_knots.end(), bv.begin(), bv.end());
        d.v_knots.insert(d.v_knots.end(), 3, s.v_hi);
        interval_quadrature(bs, false, d.s_x, d.s_w);
        interval_quadrature(bv, false, d.v_x, d.v_w);
    }
    d.s_phys.resize(d.s_x.size());
    for (size_t i = 0; i < d.s_x.size(); ++i) {
        d.s_w[i] = d.s_w[i] * jac;
        d.s_phys[i] = s.s_lo + d.s_x[i] * (s.s_hi - s.s_lo);
    }
    // clamped cubic: basis count = knots - 4, minus 2 for the recombination
    d.n_sb = int(d.s_knots.size()) - 4 - 2;
    d.n_vb = int(d.v_knots.size()) - 4 - 2;
    return d;
}

double payoff_value(int t, double B, double K, double S) {
    bool active = true;
    switch (t) {
        case 0: case 1: active = S <= B; break;
        case 2: case 3: active = S > B; break;
        case 4: case 5: active = S >= B; break;
        case 6: case 7: active = S < B; break;
        default: break;
    }
    if (!active) return 0.0;
    const double v = (t % 2 == 0) ? S - K : K - S;
    return v > 0.0 ? v : 0.0;
}

}  // namespace fpe
