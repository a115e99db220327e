/* fpe_engine.h - C ABI of the MI355X Heston Fokker-Planck pricing engine.
 *
 * Drop-in replacement for the reference engine's C ABI
 * (elevenwang-creator/FPE_option cpp/include/fpe_engine.h:17-198, implemented in
 * src/bindings/c_abi.mojo:109-494): same symbols, same struct layouts, same
 * ownership rules (the engine allocates every result array; the caller frees it
 * with the matching fpe_compute_free_* call) and the same failure convention
 * (NULL handle / {NULL, 0} results). Numerics run in HIP kernels on the GPU.
 *
 * Additive extensions (no counterpart in the reference): batched pricing of
 * many independent options in one call, per-option status codes, solver
 * options (tolerances, discretisation, complex stage-solve mode, precision)
 * and a last-error query.
 */
#ifndef FPE_ENGINE_H
#define FPE_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/** Opaque handle to an FPE compute pipeline. */
typedef struct FpeCompute FpeCompute;

/** 1-D array result. Replaces fpe_engine.h:21-24. */
struct FpeVecResult {
    double* data;
    int32_t len;
};

/** Two 1-D arrays (s and v). Replaces fpe_engine.h:27-32. */
struct FpeVec2Result {
    double* s_data;
    int32_t s_len;
    double* v_data;
    int32_t v_len;
};

/** Grid points and weights in s and v. Replaces fpe_engine.h:35-42. */
struct FpeGridPtsResult {
    double* s_data;
    int32_t s_len;
    double* v_data;
    int32_t v_len;
    double* sw_data;
    double* vw_data;
};

/** Row-major 2-D matrix. Replaces fpe_engine.h:45-49. */
struct FpeMatResult {
    double* data;
    int32_t n_rows;
    int32_t n_cols;
};

/** Delta, gamma, vega per strike. Replaces fpe_engine.h:52-57. */
struct FpeGreeksResult {
    double* delta;
    double* gamma;
    double* vega;
    int32_t len;
};

/** Price + Greeks per strike. Replaces fpe_engine.h:60-66. */
struct FpeOneshotResult {
    double* price;
    double* delta;
    double* gamma;
    double* vega;
    int32_t len;
};

/** Create a pipeline; NULL on invalid parameters.
 *  Replaces fpe_engine.h:87-94 (c_abi.mojo:109-160): s_min < 0 -> 0,
 *  s_max <= 0 -> 3*S0, v in [0, 1], validity of the type/barrier pair. */
FpeCompute* fpe_compute_create(double kappa, double theta, double sigma, double rho, double r,
                               double T, double S0, double V0, int32_t n_s, int32_t n_v,
                               double barrier, int32_t option_type, int32_t num_insert,
                               double s_min, double s_max);

/** Replaces fpe_engine.h:97 (c_abi.mojo:163-170). */
void fpe_compute_destroy(FpeCompute* ptr);

/** Knot vectors. Replaces fpe_engine.h:103 (c_abi.mojo:173-197). */
void fpe_compute_knots(FpeCompute* ptr, struct FpeVec2Result* result);

/** Quadrature points and weights. Replaces fpe_engine.h:109 (c_abi.mojo:200-230). */
void fpe_compute_grid_points(FpeCompute* ptr, struct FpeGridPtsResult* result);

/** Initial-condition coefficients. Replaces fpe_engine.h:115 (c_abi.mojo:233-247). */
void fpe_compute_initial_condition(FpeCompute* ptr, struct FpeVecResult* result);

/** All accepted Radau states (rows) x dofs. Replaces fpe_engine.h:121 (c_abi.mojo:250-268). */
void fpe_compute_solve(FpeCompute* ptr, struct FpeMatResult* result);

/** Terminal PDF on the quadrature grid. Replaces fpe_engine.h:127 (c_abi.mojo:271-289). */
void fpe_compute_pdf(FpeCompute* ptr, struct FpeMatResult* result);

/** Discounted prices at strikes. Replaces fpe_engine.h:135-140 (c_abi.mojo:292-311). */
void fpe_compute_price(FpeCompute* ptr, const double* K, int32_t n_K, struct FpeVecResult* result);

/** Finite-difference Greeks. Replaces fpe_engine.h:150-157 (c_abi.mojo:314-353).
 *  As in the reference, rel_s / rel_v are used as given (0 is not replaced). */
void fpe_compute_greeks(FpeCompute* ptr, const double* K, int32_t n_K, double rel_s, double rel_v,
                        struct FpeGreeksResult* result);

/** One-shot price + Greeks. Replaces fpe_engine.h:164-173 (c_abi.mojo:356-440). */
void fpe_price_oneshot(double kappa, double theta, double sigma, double rho, double r, double T,
                       double S0, double V0, const double* K, int32_t n_K, double barrier,
                       int32_t option_type, int32_t n_s, int32_t n_v, int32_t num_insert,
                       double s_min, double s_max, struct FpeOneshotResult* result);

/** Free functions. Replace fpe_engine.h:176-186 (c_abi.mojo:443-494). */
void fpe_compute_free_vec(struct FpeVecResult* r);
void fpe_compute_free_vec2(struct FpeVec2Result* r);
void fpe_compute_free_grid_pts(struct FpeGridPtsResult* r);
void fpe_compute_free_mat(struct FpeMatResult* r);
void fpe_compute_free_greeks(struct FpeGreeksResult* r);
void fpe_compute_free_oneshot(struct FpeOneshotResult* r);

/* ------------------------------------------------------------------ extensions */

/* Discretisation modes. */
#define FPE_DISC_REFERENCE 0 /* knot generator + insert-grid quadrature (reference) */
#define FPE_DISC_SPAN 1      /* n_s x n_v dofs, sinh-clustered knots, 4-pt GL per span */

/* Complex Radau stage-system solve. */
#define FPE_CPLX_GMRES 0  /* GMRES(10) preconditioned by the real LU (reference) */
#define FPE_CPLX_DIRECT 1 /* complex block-banded LU of (ALPH + i BETA) M + h K */

/* Factorisation precision. */
#define FPE_PREC_FP64 0
#define FPE_PREC_MIXED 1 /* FP32 block LU + FP64 iterative refinement */

/* Per-option status codes. */
#define FPE_OK 0
#define FPE_INVALID 1      /* parameters rejected by the reference validity rules */
#define FPE_MAX_STEPS 2    /* Radau exceeded 100000 steps */
#define FPE_H_UNDERFLOW 3  /* step size fell below 1e-14 */
#define FPE_NONFINITE 4    /* NaN/Inf in the solution */
#define FPE_TOO_LARGE 5    /* grid exceeds the kernels' limits (n_vb > 64, n_sb > 300); the rest of the batch runs */
#define FPE_SINGULAR 6     /* zero or non-finite pivot in a diagonal block of a stage matrix */
#define FPE_REFINE_FAIL 7  /* FPE_PREC_MIXED: a stage solve's FP64 refinement stopped above 1e-12 (price kept) */

/** One option of a batch: parameters, one strike, type and solver options. */
typedef struct FpeBatchOption {
    double kappa, theta, sigma, rho, r, T, S0, V0;
    double s_min, s_max;  /* s_min < 0 -> 0, s_max <= 0 -> 3*S0 */
    double v_max;         /* <= 0 -> 1 (reference) */
    double barrier;
    double strike;
    int32_t option_type;  /* 0..9 */
    int32_t n_s, n_v;     /* reference: knot-generator input; span: dof counts */
    int32_t disc_mode;    /* FPE_DISC_* */
    int32_t num_insert;   /* reference mode */
    int32_t complex_mode; /* FPE_CPLX_* */
    int32_t precision;    /* FPE_PREC_* */
    int32_t exact_constants;
    double ic_sigma_s, ic_sigma_v; /* <= 0 -> reference */
    double rtol, atol;             /* <= 0 -> 1e-4 / 1e-6 */
    double knot_width_s, knot_width_v; /* span mode, <= 0 -> 0.1 */
    double knot_center_v;              /* span mode: v breakpoints cluster here; <= 0 -> V0 */
    int32_t lu_variant; /* factorisation code path: 0 tuned per grid (tuning_gfx950.txt), 1 generic loops,
                           2 MFMA loops compiled for the block width's k-steps */
    int32_t max_steps;  /* Radau step limit; <= 0 -> 100000 (the reference's) */
    int32_t refine_max; /* FPE_PREC_MIXED refinement rounds per stage solve; 0 -> 10, < 0 -> none */
    int32_t ic_variant; /* initial-condition ADMM launch form: 0 tuned per grid (tuning_gfx950.txt, else the
                           most resident workgroups per CU), 1 8-wave workgroups with up to 159 KB of LDS,
                           2 / 3 4-wave workgroups in a half / third of a CU's LDS; a form the grid does not
                           fit is replaced by 0's choice. Same results in every form up to reduction order. */
} FpeBatchOption;

/** Per-batch counters (summed over the solves of the batch). */
typedef struct FpeBatchStats {
    int64_t n_solves, n_accepted, n_rejected, n_lu, n_newton, n_err_solves, n_inner_iters,
        n_admm_iters;
    double ms_setup, ms_assemble, ms_ic, ms_radau, ms_payoff, ms_total;
    /* Algorithmic work of the Radau kernel, summed over solves: block-banded LU
     * 24.67 n_sb n_vb^3 flops per real factorisation (x4 complex), banded solve
     * 14 n_sb n_vb^2 (x4 complex), 49-point SpMV 98 N; bytes = factor panels read
     * by every solve (7 n_sb n_vb^2 doubles real, x2 complex) plus factor writes. */
    double alg_flops, alg_bytes;
    /* Solver-workgroup seconds summed over solves, by phase: stage-matrix
     * assembly, block factorisation, block solves, SpMV + Newton update, rest. */
    double wg_s_assemble, wg_s_lu, wg_s_solve, wg_s_spmv, wg_s_other, wg_s_total;
    /* Factorisation seconds split: pivot-block inverse, operand staging, L multiply, Schur update. */
    double wg_s_lu_inverse, wg_s_lu_stage, wg_s_lu_lmul, wg_s_lu_schur;
    /* FP64 refinement rounds of the mixed-precision stage solves (FPE_PREC_MIXED). */
    int64_t n_refine;
    /* Radau workgroups launched (two per compute unit, at most one per solve);
     * wg_s_total / (n_radau_slots * ms_radau / 1e3) is their busy fraction. */
    int64_t n_radau_slots;
} FpeBatchStats;

/** CSR matrix (row-major, sorted columns). */
struct FpeCsrResult {
    double* data;
    int32_t* indices;
    int32_t* indptr;
    int32_t n_rows, n_cols, nnz;
};

/** Recombined cubic B-spline values (deriv = 0) or first derivatives (deriv = 1)
 *  at the quadrature points of axis 0 = s or 1 = v, evaluated on the GPU.
 *  Serves the reference's Python-only Compute.basis_1d (_py_pipeline.mojo:46-57,
 *  compute_pipeline.mojo:63-71). Entries that are exactly zero are omitted. */
void fpe_compute_basis_1d(FpeCompute* ptr, int32_t axis, int32_t deriv, struct FpeCsrResult* result);
void fpe_compute_free_csr(struct FpeCsrResult* r);

/** Fill an FpeBatchOption with the reference defaults. */
void fpe_batch_option_default(FpeBatchOption* o);

/** Price n independent options on the GPU (in-options via in-out parity).
 *  prices[i] is the discounted price; status[i] an FPE_* code (an option whose
 *  grid does not fit gets FPE_TOO_LARGE and price 0; the others still run).
 *  Returns 0 on success, nonzero if the batch could not run (see fpe_last_error). */
int32_t fpe_batch_price(const FpeBatchOption* opts, int32_t n, double* prices, int32_t* status,
                        FpeBatchStats* stats);

/** Bytes of the fixed device scratch array that fpe_batch_price_noalloc works in. */
#define FPE_NOALLOC_SCRATCH_BYTES (256ll << 20)

/** fpe_batch_price without any device allocation. No counterpart in the
 *  reference: for callers that cannot allocate device memory (code under HIP
 *  graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *  Every device buffer of the call (instance data, operators, solver
 *  workspace) is a slice of a fixed FPE_NOALLOC_SCRATCH_BYTES array that the
 *  library holds on each GPU it runs on until the process ends; the Radau
 *  solver runs as many concurrent solves as the rest of the scratch holds.
 *  Same arguments, prices and statuses as fpe_batch_price. Returns 3 (message
 *  in fpe_last_error) when the batch does not fit the scratch even with one
 *  solver slot, 0 on success. Calls share the one array per GPU: the library
 *  serialises its own calls, and calls that overlap on one GPU (from other
 *  threads bypassing that lock, or other loaders of the library) would race. */
int32_t fpe_batch_price_noalloc(const FpeBatchOption* opts, int32_t n, double* prices, int32_t* status,
                                FpeBatchStats* stats);

/** Solver record of one option, summed over its solves (two for in-options). */
typedef struct FpeSolveInfo {
    int32_t status, n_solves;
    int32_t n_accepted, n_rejected, n_lu, n_newton, n_err_solves, n_inner_iters, n_admm_iters, n_refine;
    int32_t n_sb, n_vb; /* dofs per axis of the option's first solve */
} FpeSolveInfo;

/** fpe_batch_price with each option's solver record in info[i] (may be NULL)
 *  and, when qT is not NULL, the terminal state q(T) of every vanilla or out
 *  option (one solve) in row i of qT (leading dimension ldq, n_sb * n_vb
 *  values, dof r = I n_vb + i; rows of in-options and of options that did not
 *  run are left untouched). */
int32_t fpe_batch_price_detail(const FpeBatchOption* opts, int32_t n, double* prices, int32_t* status,
                               FpeSolveInfo* info, double* qT, int64_t ldq, FpeBatchStats* stats);

/** Terminal densities q(T) (recombined B-spline coefficients, dof r = I n_vb + i)
 *  of n vanilla (types 8, 9: full domain) or out options (2, 3, 6, 7: the
 *  barrier-truncated domain); the strike does not enter q(T). Row i of qT
 *  (leading dimension ldq) receives n_dof[i] values; status[i] as in
 *  fpe_batch_price (FPE_INVALID for in-types, which need two solves, or when
 *  n_dof exceeds ldq). Returns 0 when the batch ran. */
int32_t fpe_batch_terminal_states(const FpeBatchOption* opts, int32_t n, double* qT, int64_t ldq, int32_t* n_dof,
                                  int32_t* status, FpeBatchStats* stats);

/** Mass and stiffness operators M, K of one vanilla or out option's domain as
 *  assembled on the GPU (before the Radau D-scaling), as 49-point stencils:
 *  row r = I n_vb + i, entry (J - I + 3) * 7 + (j - i + 3) holds the coupling of
 *  dof r to dof J n_vb + j (0 outside the grid). M49, K49: cap_dof x 49 doubles.
 *  Returns 0, FPE_TOO_LARGE for grids that do not fit, else nonzero. */
int32_t fpe_batch_operators(const FpeBatchOption* opt, int32_t cap_dof, double* M49, double* K49, int32_t* n_sb,
                            int32_t* n_vb);

/** K x and M x for one vanilla or out option's domain (unscaled operators as in
 *  fpe_batch_operators) by the sum-factorised Kronecker products from the 1D
 *  factors (src/sparse/kron_spmv.mojo:13-146): x, Kx, Mx hold n_dof = n_sb n_vb
 *  values, dof r = I n_vb + i. Returns 0, FPE_TOO_LARGE, or nonzero on error. */
int32_t fpe_batch_apply_operators(const FpeBatchOption* opt, int32_t n_dof, const double* x, double* Kx,
                                  double* Mx);

/** Pricing operator of one European domain: the s quadrature points s_phys
 *  (n_qs) and the marginal operator marg (n_qs x n_dof, row-major) with
 *  price = exp(-r T) sum_i payoff(s_phys[i]) (marg q(T))_i, i.e.
 *  marg[i][I n_vb + l] = w_s,i Bs[i][I] sum_j Bv[j][l] w_v,j. Capacities:
 *  cap_qs points, cap_dof dofs. Returns 0 on success. */
int32_t fpe_european_pricing_operator(const FpeBatchOption* opt, int32_t cap_qs, int32_t cap_dof, double* s_phys,
                                      double* marg, int32_t* n_qs, int32_t* n_dof);

/* ---- surrogate kernels (no counterpart in the reference). Device pointers
 * (row-major), work queued on `stream` (a hipStream_t; NULL = the null stream).
 * Return 0, or nonzero on a launch error (see fpe_last_error). */

/** Z = X W^T + b (M x N) and A = silu(Z) (act = 1) or A = Z (act = 0); X is M x K
 *  FP32 (rounded to BF16 as an MFMA operand), W is N x K BF16, accumulation FP32.
 *  Z or A may be NULL. */
int32_t fpe_mlp_forward(const float* X, int32_t M, int32_t K, const uint16_t* Wb, const float* b, int32_t N,
                        int32_t act, float* Z, float* A, void* stream);

/** Backward of fpe_mlp_forward given dZ (M x N): dW = dZ^T X (N x K), db = column
 *  sums of dZ, dX = dZ W (M x K) times silu'(Zp) when Zp is not NULL. dX may be
 *  NULL. BF16 operands, FP32 accumulation. */
int32_t fpe_mlp_backward(const float* dZ, const float* X, const uint16_t* Wb, int32_t M, int32_t N, int32_t K,
                         float* dW, float* db, float* dX, const float* Zp, void* stream);

/** fpe_mlp_forward / fpe_mlp_backward with FP32 weights W (N x K) and FP32
 *  operands on v_mfma_f32_16x16x4f32 (no BF16 rounding), FP32 accumulation. */
int32_t fpe_mlp_forward_f32(const float* X, int32_t M, int32_t K, const float* W, const float* b, int32_t N,
                            int32_t act, float* Z, float* A, void* stream);
int32_t fpe_mlp_backward_f32(const float* dZ, const float* X, const float* W, int32_t M, int32_t N, int32_t K,
                             float* dW, float* db, float* dX, const float* Zp, void* stream);

/** Adam step (bias-corrected, step t >= 1) on FP32 master weights W (n), then
 *  Wb = BF16(W) when Wb is not NULL. */
int32_t fpe_mlp_adam(float* W, uint16_t* Wb, const float* G, float* m, float* v, int64_t n, float lr, float beta1,
                     float beta2, float eps, int32_t t, void* stream);

/** Weighted mean-square loss over M x N: dY = 2 wcol[c] (Y - T) / (M N) (wcol
 *  may be NULL = 1); partial[0..1023] receives per-block sums of
 *  wcol[c] (Y - T)^2 / (M N), all 1024 entries written. */
int32_t fpe_mlp_mse(const float* Y, const float* T, const float* wcol, int32_t M, int32_t N, float* dY,
                    float* partial, void* stream);

/** dst[i][:] = src[idx[i]][:] for n rows of `width` floats (mini-batch gather). */
int32_t fpe_mlp_gather_rows(const float* src, const int32_t* idx, int32_t n, int32_t width, float* dst,
                            void* stream);

/** Device memory for callers without a tensor library: allocation (NULL on
 *  failure), release, synchronous copies, device-wide synchronisation. */
void* fpe_device_alloc(int64_t bytes);
void fpe_device_free(void* p);
int32_t fpe_device_upload(void* dst, const void* src, int64_t bytes);
int32_t fpe_device_download(void* dst, const void* src, int64_t bytes);
int32_t fpe_device_synchronize(void);

/** BF16 rounding of n FP32 values. */
int32_t fpe_mlp_to_bf16(const float* W, uint16_t* Wb, int64_t n, void* stream);

/** Surrogate prices from normalised POD coefficients C (n x ncoef, ncoef <= 64,
 *  a multiple of 4): c = C * cscale + cshift; marginal w = w0 + P c with
 *  Pt = P^T (ncoef x nq); price = exp(-r T) sum_i payoff(type, K, sq[i]) w_i
 *  for European types 8 (call) and 9 (put). */
int32_t fpe_surrogate_price(const float* C, int32_t n, int32_t ncoef, const float* cscale, const float* cshift,
                            const float* Pt, const float* w0, const float* sq, int32_t nq, const double* strike,
                            const int32_t* type, const double* r, const double* T, double* prices, void* stream);

/** POD of q(T) snapshots in the pricing inner product (no counterpart in the
 *  reference: the surrogate's basis). Device pointers: Q (n_snap x n_dof,
 *  row-major, 1 <= n_snap <= 16384), G the s-marginal operator (n_q x n_dof).
 *  Outputs: mean (n_dof), Phi (n_dof x k), P (n_q x k, orthonormal columns),
 *  w0 = G mean (n_q), sv (n_q) the singular values of (Q - mean) G^T,
 *  descending. Mode signs make each P column's largest-magnitude entry
 *  positive. n_sweeps (host, may be NULL) receives the Jacobi sweep count.
 *  Returns 0, 1 (invalid arguments), 2 (HIP error) or 3 (no convergence). */
int32_t fpe_pod_fit(const double* Q, int32_t n_snap, int32_t n_dof, const double* G, int32_t n_q, int32_t k,
                    double* mean, double* Phi, double* P, double* w0, double* sv, int32_t* n_sweeps, void* stream);

/** Singular values and right singular vectors of A (n_rows x n_cols,
 *  column-major with leading dimension lda, device memory, n_rows <= 16384)
 *  by one-sided Jacobi; A is overwritten. sv (n_cols) descending; V
 *  (n_cols x n_cols, row-major) holds the matching vectors as columns, signed
 *  as in fpe_pod_fit. No counterpart in the reference. */
int32_t fpe_svd_jacobi(double* A, int32_t n_rows, int32_t n_cols, int32_t lda, double* sv, double* V,
                       int32_t* n_sweeps, void* stream);

/** Human-readable description of the last failure in this thread (empty if none). */
const char* fpe_last_error(void);

/** Number of usable GPUs (0 if the HIP runtime finds none). */
int32_t fpe_device_count(void);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
static_assert(sizeof(struct FpeVecResult) == 16, "FpeVecResult layout mismatch");
static_assert(sizeof(struct FpeVec2Result) == 32, "FpeVec2Result layout mismatch");
static_assert(sizeof(struct FpeGridPtsResult) == 48, "FpeGridPtsResult layout mismatch");
static_assert(sizeof(struct FpeMatResult) == 16, "FpeMatResult layout mismatch");
static_assert(sizeof(struct FpeGreeksResult) == 32, "FpeGreeksResult layout mismatch");
static_assert(sizeof(struct FpeOneshotResult) == 40, "FpeOneshotResult layout mismatch");
#endif

#endif
