// Device-side data layout and workgroup helpers shared by all kernels.
//
// Every kernel runs one 256-thread workgroup (4 wave64s) per unit of work
// (an instance, an axis of an instance, or a persistent solver slot). A "row"
// r of an instance is the tensor dof (I, i) with r = I * n_vb + i: the v index
// is the fast one, so M and K are block-banded with n_vb x n_vb blocks and block
// half-bandwidth 3 (cubic B-splines).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fpe {

constexpr int NT = 256;        // threads per workgroup
constexpr int NWAVE = NT / 64;
constexpr int MAX_BLOCK = 64;  // largest n_vb handled by the block-banded LU

// Global-memory pointer type. Functions that are not inlined into the kernel
// receive generic pointers, which compile to FLAT accesses (counted against the
// LDS wait counter too); casting to this type gives global_load/store.
#define GAS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GAS T* gas(T* p) { return (GAS T*)p; }
// A wave-uniform int argument of a non-inlined function, made visibly uniform
// so that conditions on it compile to scalar branches.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
// A wave-uniform global pointer argument, moved to scalar registers: loads then
// take a scalar base and a 32-bit vector offset.
template <class T>
__device__ __forceinline__ GAS T* uniform(GAS T* p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return (GAS T*)(((unsigned long long)hi << 32) | lo);
}

typedef double d4 __attribute__((ext_vector_type(4)));

// Per-instance descriptor (one Fokker-Planck solve). Offsets are element
// offsets into the engine's device arrays.
struct Inst {
    int nsb, nvb, nqs, nqv;   // dofs per axis, quadrature points per axis
    int nsk, nvk;             // knot counts
    int s_lneu, s_rneu;       // s-boundary: 1 = Neumann (merge), 0 = Dirichlet (drop)
    int cmode, prec, exact, hist_cap;
    int lu_var;               // factorisation code path (LU_GENERIC / LU_COMPILED_K)
    int max_steps, refine_max;  // Radau step limit; mixed-precision refinement rounds per stage solve
    long long o_sk, o_vk;     // knots (array KN)
    long long o_sq, o_vq;     // s: x, w, phys (3*nqs); v: x, w (2*nqv)   (array QD)
    long long o_bs, o_bv;     // basis per point: 8 doubles (BV) and first column (BC)
    long long o_fs, o_fv;     // 1D factors: 5*nsb*7 and 6*nvb*7 (array FC)
    long long o_st;           // stencils Ms, Ks: 98 planes of N (array ST)
    long long o_vec;          // D, Dinv, q0, qT, bt, mw, scratch... (array VC, 12*N)
    long long o_eig;          // Qv, QvT, lv, Av (array EG)
    long long o_eig_s;        // Qs, ls, As (array EG; shared by instances with the same s basis)
    long long o_hist;         // accepted-state history (array HS), hist_cap * N
    double kappa, theta, eta, rho, r, T, S0, V0, jac, s_lo;
    double ic_sig_s, ic_sig_v, rtol, atol;
};

// Per-solve status codes (the FPE_* codes of the C ABI).
enum {
    ST_OK = 0, ST_INVALID = 1, ST_MAX_STEPS = 2, ST_H_UNDERFLOW = 3, ST_NONFINITE = 4, ST_TOO_LARGE = 5,
    ST_SINGULAR = 6, ST_REFINE_FAIL = 7
};

// Per-instance result counters.
struct InstOut {
    int status, n_acc, n_rej, n_lu, n_newt, n_err, n_inner, n_admm;
    int n_hist, n_refine;  // n_refine: FP64 refinement rounds of the mixed-precision stage solves
    double t_end, h_last;
    // wall-clock ticks (100 MHz) spent per phase by the solver workgroup
    unsigned long long tk_asm, tk_lu, tk_solve, tk_mv, tk_other, tk_total;
    unsigned long long tk_lu_part[4];  // factorisation: pivot inverse, staging, L multiply, Schur update
};

// VC vector slots (each N doubles).
enum { V_D = 0, V_DINV, V_Q0, V_QT, V_BT, V_MW, V_ADMM_X, V_ADMM_Z, V_ADMM_U, V_ADMM_Q, V_ADMM_R, V_ADMM_T,
       V_COUNT };

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Workgroup sum over NW waves with a fixed reduction tree (deterministic). red: LDS[NW].
template <int NW>
__device__ __forceinline__ double block_sum_w(double v, double* red) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w];
    return s;
}

__device__ __forceinline__ double block_sum(double v, double* red) { return block_sum_w<NWAVE>(v, red); }

// K independent workgroup sums at once over NW waves. red: LDS[K*NW].
template <int K, int NW = NWAVE>
__device__ __forceinline__ void block_sum_n(double* v, double* red) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[k * NW + (threadIdx.x >> 6)] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += red[k * NW + w];
        v[k] = s;
    }
}

template <int NW>
__device__ __forceinline__ double block_max_w(double v, double* red) {
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) s = fmax(s, red[w]);
    return s;
}

__device__ __forceinline__ double block_max(double v, double* red) { return block_max_w<NWAVE>(v, red); }

}  // namespace fpe
