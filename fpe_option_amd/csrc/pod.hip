// POD of solver snapshots on the GPU, for the surrogate's basis.
//
// Snapshots are q(T) rows Q (n_snap x n_dof, row-major, FP64). The basis is
// orthonormal in the pricing inner product <q1, q2> = (G q1) . (G q2), G the
// s-marginal operator (n_q x n_dof): with X = Q - mean and W = X G^T
// (n_snap x n_q), W = U S V^T gives the pricing-space modes P = V (n_q x k),
// the q-space modes Phi = X^T U_k S_k^-1 = X^T W V_k S_k^-2 (n_dof x k) and
// w0 = G mean.
//
//   k_colmean, k_rowdot   mean snapshot and w0 = G mean (fixed-order sums)
//   k_gemm                W = (Q - mean) G^T and Phi = (Q - mean)^T Z on
//                         v_mfma_f64_16x16x4f64 (64x64 tiles, strided operands)
//   k_jacobi_round        one round of one-sided (Hestenes) Jacobi on W's
//                         columns: disjoint column pairs, one workgroup each
//   k_colnorm, k_modes    singular values; sign-fixed P columns and Z = W V / S^2
//
// One-sided Jacobi works on W itself, never on the Gram matrix W^T W: the
// snapshot singular values fall to 1e-8 of the largest within 64 modes, and a
// Gram matrix would square that spread to 1e-16, the FP64 rounding level.
// Each pair (p, q) is rotated so that its columns become orthogonal; a sweep
// runs every pair once in n-1 rounds of the round-robin ("circle") schedule,
// and sweeps repeat until no pair is rotated. All sums run in a fixed order,
// so results are reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/fpe_engine.h"
#include "engine.hpp"

typedef double pod_d4 __attribute__((ext_vector_type(4)));

namespace fpe_pod {

constexpr int GT = 256;  // GEMM workgroup: 4 waves, 64x64 output tile, 32x32 per wave
constexpr int GM = 64, GN = 64, GK = 16;
constexpr int JT = 512;  // Jacobi workgroup
constexpr int JR = 32;   // rows held per thread: at most JT * JR rows
constexpr int MAX_ROWS = JT * JR;
constexpr int MAX_SWEEPS = 60;

// C(m, n) = sum_k (A(m, k) - s(m, k)) B(n, k), every operand by strides.
struct GemmArgs {
    int M, N, K;
    const double* A;
    long long sam, sak;
    const double* shift;  // null: no shift; else s(m, k) = shift[m * ssm + k * ssk]
    long long ssm, ssk;
    const double* B;
    long long sbn, sbk;
    double* C;
    long long scm, scn;
};

// Lane layout of v_mfma_f64_16x16x4f64: a = A[row0 + lane%16][k0 + lane/16],
// b = B[k0 + lane/16][col0 + lane%16], acc[j] = C[row0 + lane/16 + 4j][col0 + lane%16].
__global__ void __launch_bounds__(GT) k_gemm(GemmArgs g) {
    __shared__ double As[GK][GM + 1];
    __shared__ double Bs[GK][GN + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int m0 = blockIdx.x * GM, n0 = blockIdx.y * GN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const bool a_kfast = g.sak == 1, b_kfast = g.sbk == 1;  // consecutive threads along the contiguous index
    pod_d4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = pod_d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < g.K; k0 += GK) {
#pragma unroll
        for (int r = 0; r < GM * GK / GT; ++r) {
            const int e = tid + r * GT;
            const int mm = a_kfast ? e / GK : e % GM, kk = a_kfast ? e % GK : e / GM;
            const int m = m0 + mm, k = k0 + kk;
            double v = 0.0;
            if (m < g.M && k < g.K) {
                v = g.A[m * g.sam + k * g.sak];
                if (g.shift) v -= g.shift[m * g.ssm + k * g.ssk];
            }
            As[kk][mm] = v;
        }
#pragma unroll
        for (int r = 0; r < GN * GK / GT; ++r) {
            const int e = tid + r * GT;
            const int nn = b_kfast ? e / GK : e % GN, kk = b_kfast ? e % GK : e / GN;
            const int n = n0 + nn, k = k0 + kk;
            Bs[kk][nn] = (n < g.N && k < g.K) ? g.B[n * g.sbn + k * g.sbk] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4) {
            double a[2], b[2];
            for (int i = 0; i < 2; ++i) a[i] = As[ks + l4][wm + 16 * i + l16];
            for (int j = 0; j < 2; ++j) b[j] = Bs[ks + l4][wn + 16 * j + l16];
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int t = 0; t < 4; ++t) {
                const int m = m0 + wm + 16 * i + l4 + 4 * t, n = n0 + wn + 16 * j + l16;
                if (m < g.M && n < g.N) g.C[m * g.scm + n * g.scn] = acc[i][j][t];
            }
}

// Sum over the workgroup of NV values per thread; every thread gets the sums.
// Fixed order: wave butterflies, then the waves' partials in wave order.
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        for (int i = 0; i < NV; ++i) v[i] += __shfl_xor(v[i], o, 64);
    if (lane == 0)
        for (int i = 0; i < NV; ++i) red[w * NV + i] = v[i];
    __syncthreads();
    for (int i = 0; i < NV; ++i) {
        double s = 0.0;
        for (int x = 0; x < NT / 64; ++x) s += red[x * NV + i];
        v[i] = s;
    }
    __syncthreads();
}

// mean[j] = sum_i Q[i][j] / n_rows, rows summed in order.
__global__ void k_colmean(const double* __restrict__ Q, int n_rows, int n_cols, double* __restrict__ mean) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_cols) return;
    double s = 0.0;
    for (int i = 0; i < n_rows; ++i) s += Q[(long long)i * n_cols + j];
    mean[j] = s / n_rows;
}

// y[r] = G[r, :] . x, one workgroup per row.
__global__ void __launch_bounds__(256) k_rowdot(const double* __restrict__ G, int n_cols, const double* __restrict__ x,
                                                double* __restrict__ y) {
    __shared__ double red[4];
    const double* g = G + (long long)blockIdx.x * n_cols;
    double v[1] = {0.0};
    for (int j = threadIdx.x; j < n_cols; j += 256) v[0] += g[j] * x[j];
    block_sum<1, 256>(v, red);
    if (threadIdx.x == 0) y[blockIdx.x] = v[0];
}

// Round `round` of the circle schedule over npad (even) columns: pair 0 is
// (npad-1, round); pair x > 0 is ((round + x) mod (npad-1), (round - x) mod (npad-1)).
// Columns >= nc are the padding of an odd nc and are never touched. A is
// column-major (column c at A + c * lda, nr rows), V column-major (nc x nc).
__global__ void __launch_bounds__(JT) k_jacobi_round(double* __restrict__ A, int nr, long long lda,
                                                     double* __restrict__ V, int nc, int npad, int round, double tol,
                                                     int* __restrict__ nrot) {
    __shared__ double red[3 * JT / 64];
    const int x = blockIdx.x, n1 = npad - 1;
    const int p = x == 0 ? n1 : (round + x) % n1;
    const int q = x == 0 ? round : (round - x + n1) % n1;
    if (p >= nc || q >= nc) return;  // whole workgroup leaves together
    double* ap = A + p * lda;
    double* aq = A + q * lda;
    double xp[JR], xq[JR];
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < JR; ++r) {
        const int i = threadIdx.x + r * JT;
        xp[r] = i < nr ? ap[i] : 0.0;
        xq[r] = i < nr ? aq[i] : 0.0;
        s[0] += xp[r] * xp[r];
        s[1] += xq[r] * xq[r];
        s[2] += xp[r] * xq[r];
    }
    block_sum<3, JT>(s, red);
    const double alpha = s[0], beta = s[1], gamma = s[2];
    const double scale = sqrt(alpha) * sqrt(beta);
    if (!(scale > 0.0) || !(fabs(gamma) > tol * scale)) return;  // already orthogonal (or a zero column)
    // a_p' = c a_p - s a_q, a_q' = s a_p + c a_q with t = s / c the smaller root of
    // t^2 + 2 zeta t - 1 = 0, zeta = (beta - alpha) / (2 gamma): makes a_p' . a_q' = 0.
    const double zeta = (beta - alpha) / (2.0 * gamma);
    const double t = copysign(1.0, zeta) / (fabs(zeta) + hypot(1.0, zeta));
    const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
#pragma unroll
    for (int r = 0; r < JR; ++r) {
        const int i = threadIdx.x + r * JT;
        if (i < nr) {
            ap[i] = c * xp[r] - sn * xq[r];
            aq[i] = sn * xp[r] + c * xq[r];
        }
    }
    double* vp = V + (long long)p * nc;
    double* vq = V + (long long)q * nc;
    for (int i = threadIdx.x; i < nc; i += JT) {
        const double a = vp[i], b = vq[i];
        vp[i] = c * a - sn * b;
        vq[i] = sn * a + c * b;
    }
    if (threadIdx.x == 0) atomicAdd(nrot, 1);
}

__global__ void k_identity(double* __restrict__ V, int n) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (long long)n * n) V[e] = (e / n == e % n) ? 1.0 : 0.0;
}

// sv[c] = |A column c|, one workgroup per column.
__global__ void __launch_bounds__(256) k_colnorm(const double* __restrict__ A, int nr, long long lda,
                                                 double* __restrict__ sv) {
    __shared__ double red[4];
    const double* a = A + blockIdx.x * lda;
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < nr; i += 256) v[0] += a[i] * a[i];
    block_sum<1, 256>(v, red);
    if (threadIdx.x == 0) sv[blockIdx.x] = sqrt(v[0]);
}

// Mode c (< k) is Jacobi column order[c]. Its sign makes the largest-magnitude
// entry of the V column positive (first such entry on ties). Writes
// P[r][c] = sign V[r][col] (P row-major n_q x k) and, when Z is given,
// Z[c][i] = sign A[i][col] / sv[col]^2 (zero for a zero column).
__global__ void __launch_bounds__(256) k_modes(const double* __restrict__ A, int nr, long long lda,
                                               const double* __restrict__ V, int nc, const int* __restrict__ order,
                                               const double* __restrict__ sv, int k, double* __restrict__ P,
                                               double* __restrict__ Z) {
    __shared__ double bv[256];
    __shared__ int bi[256];
    const int c = blockIdx.x, col = order[c];
    const double* v = V + (long long)col * nc;
    double best = -1.0;
    int arg = 0;
    for (int i = threadIdx.x; i < nc; i += 256) {
        const double a = fabs(v[i]);
        if (a > best) { best = a; arg = i; }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = arg;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if (threadIdx.x < o) {
            const double b2 = bv[threadIdx.x + o];
            const int i2 = bi[threadIdx.x + o];
            if (b2 > bv[threadIdx.x] || (b2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
                bv[threadIdx.x] = b2;
                bi[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    const double sg = v[bi[0]] < 0.0 ? -1.0 : 1.0;
    for (int r = threadIdx.x; r < nc; r += 256) P[(long long)r * k + c] = sg * v[r];
    if (Z) {
        const double s2 = sv[col] * sv[col];
        const double f = s2 > 0.0 ? sg / s2 : 0.0;
        const double* a = A + col * lda;
        for (int i = threadIdx.x; i < nr; i += 256) Z[(long long)c * nr + i] = f * a[i];
    }
}

struct Fail {
    int code;
    std::string msg;
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Fail{2, std::string(what) + ": " + hipGetErrorString(e)};
}

template <class T>
struct DevArr {
    T* p = nullptr;
    explicit DevArr(size_t n) { check(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc"); }
    ~DevArr() {
        if (p) (void)hipFree(p);
    }
    DevArr(const DevArr&) = delete;
    DevArr& operator=(const DevArr&) = delete;
};

void gemm(const GemmArgs& g, hipStream_t st) {
    if (g.M == 0 || g.N == 0) return;
    dim3 grid((g.M + GM - 1) / GM, (g.N + GN - 1) / GN);
    hipLaunchKernelGGL(k_gemm, grid, dim3(GT), 0, st, g);
    check(hipGetLastError(), "k_gemm launch");
}

// One-sided Jacobi on A (nr x nc, column-major, leading dimension lda >= nr),
// V = I accumulated. On return A's columns are orthogonal (A_in V = A_out),
// sv[c] = |A_out column c| and order lists the columns by descending sv.
int jacobi(double* A, int nr, long long lda, int nc, double* V, double* sv, std::vector<int>& order, hipStream_t st) {
    DevArr<int> nrot(1);
    hipLaunchKernelGGL(k_identity, dim3((unsigned)(((long long)nc * nc + 255) / 256)), dim3(256), 0, st, V, nc);
    check(hipGetLastError(), "k_identity launch");
    const int npad = nc + (nc & 1);
    const double tol = std::sqrt((double)std::max(nr, 1)) * DBL_EPSILON;
    int sweeps = 0;
    if (nc > 1) {
        for (;;) {
            if (sweeps == MAX_SWEEPS) throw Fail{3, "one-sided Jacobi did not converge in 60 sweeps"};
            ++sweeps;
            check(hipMemsetAsync(nrot.p, 0, sizeof(int), st), "hipMemsetAsync");
            for (int r = 0; r < npad - 1; ++r) {
                hipLaunchKernelGGL(k_jacobi_round, dim3(npad / 2), dim3(JT), 0, st, A, nr, lda, V, nc, npad, r, tol,
                                   nrot.p);
                check(hipGetLastError(), "k_jacobi_round launch");
            }
            int h = 0;
            check(hipMemcpyAsync(&h, nrot.p, sizeof(int), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
            check(hipStreamSynchronize(st), "hipStreamSynchronize");
            if (h == 0) break;
        }
    }
    hipLaunchKernelGGL(k_colnorm, dim3(nc), dim3(256), 0, st, A, nr, lda, sv);
    check(hipGetLastError(), "k_colnorm launch");
    std::vector<double> h(nc);
    check(hipMemcpyAsync(h.data(), sv, nc * sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    order.resize(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h[a] > h[b]; });
    return sweeps;
}

int run(const char* what, int (*body)(void*), void* ctx) {
    try {
        return body(ctx);
    } catch (const Fail& f) {
        fpe::set_last_error(std::string(what) + ": " + f.msg);
        return f.code;
    }
}

}  // namespace fpe_pod

using namespace fpe_pod;

extern "C" {

int32_t fpe_svd_jacobi(double* A, int32_t n_rows, int32_t n_cols, int32_t lda, double* sv, double* V,
                       int32_t* n_sweeps, void* stream) {
    if (n_rows < 0 || n_cols < 1 || lda < n_rows || n_rows > MAX_ROWS || !A || !sv || !V) {
        fpe::set_last_error("fpe_svd_jacobi: invalid arguments (n_rows <= 16384, n_cols >= 1, lda >= n_rows)");
        return 1;
    }
    struct Ctx {
        double *A, *sv, *V;
        int nr, nc;
        long long lda;
        int32_t* sweeps;
        hipStream_t st;
    } c{A, sv, V, n_rows, n_cols, lda, n_sweeps, (hipStream_t)stream};
    return run("fpe_svd_jacobi", [](void* p) -> int {
        Ctx& c = *(Ctx*)p;
        std::vector<int> order;
        DevArr<double> raw(c.nc), Vj((size_t)c.nc * c.nc);
        const int sweeps = jacobi(c.A, c.nr, c.lda, c.nc, Vj.p, raw.p, order, c.st);
        if (c.sweeps) *c.sweeps = sweeps;
        std::vector<double> h(c.nc), hs(c.nc);
        check(hipMemcpy(h.data(), raw.p, c.nc * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        for (int i = 0; i < c.nc; ++i) hs[i] = h[order[i]];
        check(hipMemcpyAsync(c.sv, hs.data(), c.nc * sizeof(double), hipMemcpyHostToDevice, c.st), "hipMemcpyAsync");
        DevArr<int> ord(c.nc);
        check(hipMemcpyAsync(ord.p, order.data(), c.nc * sizeof(int), hipMemcpyHostToDevice, c.st), "hipMemcpyAsync");
        hipLaunchKernelGGL(k_modes, dim3(c.nc), dim3(256), 0, c.st, c.A, c.nr, c.lda, Vj.p, c.nc, ord.p, raw.p, c.nc,
                           c.V, nullptr);
        check(hipGetLastError(), "k_modes launch");
        check(hipStreamSynchronize(c.st), "hipStreamSynchronize");
        return 0;
    }, &c);
}

int32_t fpe_pod_fit(const double* Q, int32_t n_snap, int32_t n_dof, const double* marg, int32_t n_q, int32_t k,
                    double* mean, double* Phi, double* P, double* w0, double* sv, int32_t* n_sweeps, void* stream) {
    if (n_snap < 1 || n_snap > MAX_ROWS || n_dof < 1 || n_q < 1 || k < 1 || k > n_q || !Q || !marg || !mean ||
        !Phi || !P || !w0 || !sv) {
        fpe::set_last_error("fpe_pod_fit: invalid arguments (1 <= n_snap <= 16384, 1 <= k <= n_q)");
        return 1;
    }
    struct Ctx {
        const double *Q, *G;
        int ns, nd, nq, k;
        double *mean, *Phi, *P, *w0, *sv;
        int32_t* sweeps;
        hipStream_t st;
    } c{Q, marg, n_snap, n_dof, n_q, k, mean, Phi, P, w0, sv, n_sweeps, (hipStream_t)stream};
    return run("fpe_pod_fit", [](void* p) -> int {
        Ctx& c = *(Ctx*)p;
        hipLaunchKernelGGL(k_colmean, dim3((c.nd + 255) / 256), dim3(256), 0, c.st, c.Q, c.ns, c.nd, c.mean);
        check(hipGetLastError(), "k_colmean launch");
        hipLaunchKernelGGL(k_rowdot, dim3(c.nq), dim3(256), 0, c.st, c.G, c.nd, c.mean, c.w0);
        check(hipGetLastError(), "k_rowdot launch");
        // W (n_snap x n_q, column-major): W[i][q] = sum_j (Q[i][j] - mean[j]) G[q][j]
        DevArr<double> W((size_t)c.ns * c.nq), V((size_t)c.nq * c.nq), raw(c.nq);
        GemmArgs g1{c.ns, c.nq, c.nd, c.Q, c.nd, 1, c.mean, 0, 1, c.G, c.nd, 1, W.p, 1, c.ns};
        gemm(g1, c.st);
        std::vector<int> order;
        const int sweeps = jacobi(W.p, c.ns, c.ns, c.nq, V.p, raw.p, order, c.st);
        if (c.sweeps) *c.sweeps = sweeps;
        std::vector<double> h(c.nq), hs(c.nq);
        check(hipMemcpy(h.data(), raw.p, c.nq * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        for (int i = 0; i < c.nq; ++i) hs[i] = h[order[i]];
        check(hipMemcpyAsync(c.sv, hs.data(), c.nq * sizeof(double), hipMemcpyHostToDevice, c.st), "hipMemcpyAsync");
        DevArr<int> ord(c.nq);
        DevArr<double> Z((size_t)c.k * c.ns);
        check(hipMemcpyAsync(ord.p, order.data(), c.nq * sizeof(int), hipMemcpyHostToDevice, c.st), "hipMemcpyAsync");
        hipLaunchKernelGGL(k_modes, dim3(c.k), dim3(256), 0, c.st, W.p, c.ns, (long long)c.ns, V.p, c.nq, ord.p, raw.p,
                           c.k, c.P, Z.p);
        check(hipGetLastError(), "k_modes launch");
        // Phi[j][m] = sum_i (Q[i][j] - mean[j]) Z[m][i]
        GemmArgs g2{c.nd, c.k, c.ns, c.Q, 1, c.nd, c.mean, 1, 0, Z.p, c.ns, 1, c.Phi, c.k, 1};
        gemm(g2, c.st);
        check(hipStreamSynchronize(c.st), "hipStreamSynchronize");
        return 0;
    }, &c);
}

}  // extern "C"
