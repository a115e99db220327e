/* FP64 CPU oracle for the Heston Fokker-Planck pricing path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is what the GPU path is compared
 * against. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; nothing on the product path links or calls it.
 *
 * It restates, single-threaded and in plain FP64, the reference engine's
 * algorithm (elevenwang-creator/FPE_option, Mojo source):
 *   knots / quadrature      src/numerics/bspline/knots.mojo:19-270,
 *                           src/engines/fpe/domain.mojo:20-237
 *   B-splines               src/numerics/bspline/basis.mojo:17-213,
 *                           src/numerics/bspline/recombination.mojo:14-86
 *   Galerkin M, K           src/engines/fpe/galerkin.mojo:16-125
 *   initial condition       src/engines/fpe/initial_cond.mojo:25-131,
 *                           src/numerics/optim/osqp.mojo:52-159
 *   D-scaling + Radau IIA   src/engines/fpe/solver.mojo:57-105,
 *                           src/numerics/ode/radau.mojo:47-1151
 *   PDF / payoff / parity   src/engines/fpe/pdf.mojo:28-32,
 *                           src/server/pricer.mojo:27-155,
 *                           src/server/payoffs.mojo:19-50,
 *                           src/server/option_types.mojo:13-70
 *   Greeks                  src/server/greeks.mojo:21-53
 * The reference's sparse LU (src/numerics/utils/sparse_lu.mojo) is restated as
 * a banded LU with partial pivoting: same linear solve, pivot order differs
 * only in roundoff.
 */
#ifndef FPE_ORACLE_H
#define FPE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Discretisation modes. */
#define ORC_DISC_REFERENCE 0 /* knot generator + insert-grid quadrature (reference) */
#define ORC_DISC_SPAN 1      /* n_s x n_v dofs, sinh-clustered knots, 4-pt GL per span */

/* Complex stage-system solve modes. */
#define ORC_CPLX_GMRES 0  /* reference: GMRES(10) preconditioned by the real LU */
#define ORC_CPLX_DIRECT 1 /* direct LU of (ALPH+i*BETA)*M + h*K */

typedef struct OrcSpec {
    double kappa, theta, sigma, rho, r, T, S0, V0;
    double s_min, s_max; /* physical s-domain before barrier truncation */
    double v_max;        /* reference: 1.0 */
    int32_t n_s, n_v;    /* reference: knot-generator input; span: dof counts */
    int32_t option_type; /* 0..9 as in the reference */
    double barrier;
    int32_t disc_mode;   /* ORC_DISC_* */
    int32_t num_insert;  /* reference mode quadrature density */
    double ic_sigma_s;   /* <=0: reference (sigma0 = 0.1) */
    double ic_sigma_v;   /* <=0: reference (sigma0 / (S_max - S_min)) */
    double rtol, atol;
    int32_t complex_mode;    /* ORC_CPLX_* */
    int32_t exact_constants; /* 0: reference U1/ALPH/BETA, 1: exact */
    double knot_width_s;     /* span mode: sinh clustering width (normalised) */
    double knot_width_v;
} OrcSpec;

typedef struct OrcStats {
    int32_t n_accepted, n_rejected, n_steps_total;
    int32_t n_lu, n_newton, n_err_solves;
    int32_t n_gmres_iters;
    int32_t admm_iters;
    int32_t success; /* 1 = reached T */
    int32_t n_dof, n_sb, n_vb, n_qs, n_qv;
} OrcStats;

/* Fill an OrcSpec with the reference defaults for the given option. */
void orc_spec_default(OrcSpec* s);

/* Price strikes for one option (in-options via in-out parity, two solves).
 * Returns 0 on success, nonzero on invalid parameters (prices zeroed). */
int orc_price(const OrcSpec* spec, const double* K, int n_K, double* prices, OrcStats* stats);

/* One-shot price + finite-difference Greeks (reference PricingEngine):
 * h_s = rel_s*S0, h_v = rel_v*V0; bumped solves regenerate the domain. */
int orc_price_greeks(const OrcSpec* spec, const double* K, int n_K, double rel_s, double rel_v,
                     double* price, double* delta, double* gamma, double* vega);

/* ---- stepwise pipeline (one solve on the barrier-truncated domain) ---- */
typedef struct OrcPipe OrcPipe;
OrcPipe* orc_pipe_create(const OrcSpec* spec); /* NULL on invalid params */
void orc_pipe_destroy(OrcPipe* p);
/* Sizes: n_sk, n_vk knots; n_qs, n_qv quad points; n_sb, n_vb basis dofs. */
void orc_pipe_dims(const OrcPipe* p, int32_t* dims8);
void orc_pipe_knots(const OrcPipe* p, double* s_knots, double* v_knots);
void orc_pipe_grid(const OrcPipe* p, double* s_phys, double* v_phys, double* sw, double* vw,
                   double* s_norm);
/* Basis at quadrature points as CSR (B, dB) for axis 0 = s, 1 = v.
 * Call with data == NULL to get nnz in *nnz. */
void orc_pipe_basis(const OrcPipe* p, int axis, int deriv, int32_t* indptr, int32_t* indices,
                    double* data, int32_t* nnz);
/* Dense 1D factor matrix by name: "Ns","S1","S2","S3","S4","Nv","V2","V3","V4","V5","V6". */
int orc_pipe_factor(const OrcPipe* p, const char* name, double* out);
/* M and K (unscaled) as dense stencils: row (I,i) -> 7x7 entries (J-I+3)*7+(j-i+3). */
void orc_pipe_stencils(const OrcPipe* p, double* M49, double* K49);
int orc_pipe_initial_condition(OrcPipe* p, double* q0);
/* Run the solve; returns number of stored states (n_accepted + 1). */
int orc_pipe_solve(OrcPipe* p, OrcStats* stats);
/* Copy stored state k (0..n-1), D-unscaled, and its time. */
void orc_pipe_state(const OrcPipe* p, int k, double* q, double* t);
void orc_pipe_pdf(OrcPipe* p, double* pdf); /* n_qs x n_qv row-major, terminal */
/* Discounted price of the solved (non-in) option at the given strikes. */
void orc_pipe_price_at(OrcPipe* p, const double* K, int n_K, double* prices);

/* ---- building blocks for known-answer tests ---- */
int orc_gauss_legendre(int order, double* nodes, double* weights);
/* Reference knot generator (non-uniform). Returns count; out may be NULL. */
int orc_generate_knots(int n, int degree, double center, double b0, double b1, double mean,
                       double std, double* out);
/* Cox-de Boor value of basis i (degree p) at x on the given knot vector. */
double orc_de_boor(const double* knots, int n_knots, int degree, int i, double x);
/* Radau IIA on M y' = -K y with dense M, K (n x n, row-major).
 * Writes y(t1) into y_out; returns number of accepted steps or -1. */
int orc_radau_dense(int n, const double* M, const double* K, const double* y0, double t1,
                    double rtol, double atol, double max_step, double first_step,
                    int complex_mode, int exact_constants, double* y_out, OrcStats* stats);
/* NNLS by the reference ADMM on a dense n x n matrix A: min |Ac-b|^2, c>=0. */
int orc_admm_nnls_dense(int n, const double* A, const double* b, double* c);
/* Payoff value (undiscounted) of option type at S for strike K. */
double orc_payoff(int option_type, double barrier, double K, double S);
/* Heston semi-analytic European call (Heston 1993 characteristic function,
 * Gauss-Legendre on [0, umax]); a closed-form check, not a reference restatement. */
double orc_heston_call(double kappa, double theta, double sigma, double rho, double r, double T,
                       double S0, double V0, double K);

#ifdef __cplusplus
}
#endif
#endif
