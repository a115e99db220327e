// Initial condition: Galerkin projection of a truncated Gaussian at (S0, V0)
// followed by the reference's ADMM non-negative least-squares fit
// (src/engines/fpe/initial_cond.mojo:25-131, src/numerics/optim/osqp.mojo:52-159).
//
// The ADMM x-update solves (A^T A + alpha I) x = rhs with A = D M D, where
// M = Ns (x) Nv and D = diag(M)^(-1/2) = Ds (x) Dv, so A = As (x) Av with
// As = Ds Ns Ds, Av = Dv Nv Dv symmetric. With As = Qs Ls Qs^T and
// Av = Qv Lv Qv^T the solve is a fast diagonalisation:
//     x = (Qs (x) Qv) [ (Qs^T R Qv) ./ (ls^2 lv^2 + alpha) ] (Qs (x) Qv)^T,
// i.e. four small dense products per iteration instead of a banded
// factor of bandwidth 2(3 n_vb + 3) streamed from memory every iteration.
#pragma once
#include "device_common.hpp"

namespace fpe {

// EG layout: the s part Qs (ns^2), ls (ns), As (ns^2) at o_eig_s, shared by
// every instance of the batch with the same s basis and quadrature (the s mass
// factor, hence As = Ds Ns Ds, depends on nothing else); the v part Qv (nv^2),
// QvT (nv^2), lv (nv), Av (nv^2) at o_eig, one per instance.
struct EigView {
    double *Qs, *Qv, *QvT, *ls, *lv, *As, *Av;
    __device__ EigView(double* EG, const Inst& in) {
        const long long ns = in.nsb, nv = in.nvb;
        Qs = EG + in.o_eig_s;
        ls = Qs + ns * ns;
        As = ls + ns;
        Qv = EG + in.o_eig;
        QvT = Qv + nv * nv;
        lv = QvT + nv * nv;
        Av = lv + nv;
    }
};

// Right-hand side b~ = D (Bs (x) Bv)^T W delta and the basis integrals
// m = (Bs (x) Bv)^T W 1 (separable):
//     b[k][l] = sum_p Bs[p][k] T_l(p),  T_l(p) = sum_q Bv[q][l] (w_p w_q delta(p, q)),
// p ascending over the s points in the support of k, q ascending over the v
// points in the support of l. One workgroup per (instance, l): the threads first
// compute T_l(p) for a block of IC_PBLK s points into LDS (one p each), then one
// thread per k adds its points of the block in ascending order.
// grid: (n_inst, max n_vb).
constexpr int IC_PBLK = 2048;  // s points per LDS block; thread t sums k = t and t + NT (n_sb <= 2 NT)
__global__ void __launch_bounds__(NT) k_ic_rhs(const Inst* __restrict__ inst, const int* __restrict__ order, const double* __restrict__ QD,
                                               const double* __restrict__ BV, const int* __restrict__ BC,
                                               double* __restrict__ VC) {
    const Inst& in = inst[order[blockIdx.x]];  // largest instances first
    const int ns = in.nsb, nv = in.nvb, N = ns * nv, nqs = in.nqs, nqv = in.nqv;
    const int l = blockIdx.y;
    if (l >= nv) return;  // workgroup-uniform
    const double* sw = QD + in.o_sq + nqs;
    const double* sph = QD + in.o_sq + 2 * nqs;
    const double* vx = QD + in.o_vq;
    const double* vw = vx + nqv;
    const double* Ts = BV + in.o_bs;
    const double* Tv = BV + in.o_bv;
    const int* cs = BC + in.o_bs / 8;
    const int* cv = BC + in.o_bv / 8;
    double* b = VC + in.o_vec + (long long)V_BT * N;
    double* mw = VC + in.o_vec + (long long)V_MW * N;
    const double* Di = VC + in.o_vec + (long long)V_DINV * N;
    __shared__ double Tl[IC_PBLK];
    __shared__ double om_v;
    const int q0 = lower_c0(cv, nqv, l - 3), q1 = lower_c0(cv, nqv, l + 1);
    // omega_v[l] = sum_q Bv[q,l] vw_q ; omega_s[k] = sum_p Bs[p,k] sw_p
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int q = q0; q < q1; ++q) a += Tv[(long long)q * 8 + (l - cv[q])] * vw[q];
        om_v = a;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ns; k += NT) {
        double a = 0.0;
        const int p0 = lower_c0(cs, nqs, k - 3), p1 = lower_c0(cs, nqs, k + 1);
        for (int p = p0; p < p1; ++p) a += Ts[(long long)p * 8 + (k - cs[p])] * sw[p];
        mw[(long long)k * nv + l] = a * om_v;
    }
    // truncated product Gaussian (rho = 0), initial_cond.mojo:25-67
    const double sig_s = in.ic_sig_s, sig_v = in.ic_sig_v;
    const double norm = 1.0 / (2.0 * M_PI * sig_s * sig_v * 1.0);
    int kp0[2], kp1[2];
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k = threadIdx.x + h * NT;
        kp0[h] = k < ns ? lower_c0(cs, nqs, k - 3) : 0;
        kp1[h] = k < ns ? lower_c0(cs, nqs, k + 1) : 0;
    }
    for (int P0 = 0; P0 < nqs; P0 += IC_PBLK) {
        const int P1 = nqs - P0 < IC_PBLK ? nqs : P0 + IC_PBLK;
        __syncthreads();
        for (int p = P0 + threadIdx.x; p < P1; p += NT) {
            const double ds = (sph[p] - in.S0) / sig_s, ds2 = ds * ds;
            double T = 0.0;
            if (!(ds2 > 50.0)) {
                for (int q = q0; q < q1; ++q) {
                    const double dv = (vx[q] - in.V0) / sig_v;
                    const double z = ds2 - 2.0 * 0.0 * ds * dv + dv * dv;
                    double dl = 0.0;
                    if (!(z * 0.5 > 50.0)) {
                        dl = norm * exp(-z * 0.5);
                        if (dl < 1e-6) dl = 0.0;
                    }
                    T += Tv[(long long)q * 8 + (l - cv[q])] * (sw[p] * vw[q] * dl);
                }
            }
            Tl[p - P0] = T;
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = threadIdx.x + h * NT;
            const int pa = kp0[h] > P0 ? kp0[h] : P0, pb = kp1[h] < P1 ? kp1[h] : P1;
            for (int p = pa; p < pb; ++p) {
                const double ds = (sph[p] - in.S0) / sig_s, ds2 = ds * ds;
                if (ds2 > 50.0) continue;  // the points the sum skips
                acc[h] += Ts[(long long)p * 8 + (k - cs[p])] * Tl[p - P0];
            }
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k = threadIdx.x + h * NT;
        if (k < ns) {
            const long long r = (long long)k * nv + l;
            b[r] = Di[r] * acc[h];
        }
    }
}

// Jacobi eigendecomposition of As (axis 0) or Av (axis 1) with parallel
// round-robin ordering. One launch per axis: grid n_inst, dynamic LDS
// lds_bytes. A (row stride n + 1) lives in LDS when it fits, else in the
// instance's As/Av scratch; the eigenvector matrix is kept transposed
// (Qt[p][row] = Q[row][p]) so that the column rotations Q <- Q J walk
// contiguous rows: in LDS when A and Qt both fit, else in the Qs region
// (transposed in place at the end). The arithmetic is the same wherever the
// operands live.
constexpr int ENT = 512;  // k_eig workgroup: 8 waves
constexpr int ENW = ENT / 64;
constexpr long long EIG_LDS_S = 150 * 1024;  // the largest dynamic LDS of a k_eig launch
// Dynamic LDS of k_eig for an n x n problem with A and Qt both in LDS.
__host__ __device__ inline long long eig_lds_bytes(int n) { return 16LL * n * (n + 1); }

__global__ void __launch_bounds__(ENT) k_eig(const Inst* __restrict__ inst, const int* __restrict__ order, int ax,
                                             long long lds_bytes, const double* __restrict__ FC,
                                             double* __restrict__ EG) {
    extern __shared__ double dyn[];
    const Inst& in = inst[order[blockIdx.x]];  // largest instances first
    const int ns = in.nsb, nv = in.nvb;
    EigView E(EG, in);
    const int n = ax == 0 ? ns : nv;
    const double* F = FC + (ax == 0 ? in.o_fs : in.o_fv);  // factor 0 = mass
    const bool a_lds = 8LL * n * (n + 1) <= lds_bytes;
    const bool q_lds = a_lds && 16LL * n * (n + 1) <= lds_bytes;
    const int lda = a_lds ? n + 1 : n, ldq = q_lds ? n + 1 : n;
    double* A = a_lds ? dyn : (ax == 0 ? E.As : E.Av);
    double* Qt = q_lds ? dyn + (long long)n * (n + 1) : (ax == 0 ? E.Qs : E.QvT);
    double* lam = ax == 0 ? E.ls : E.lv;
    __shared__ double cc[160], ss[160];
    __shared__ int pp[160], qq[160];
    __shared__ double red2[2 * ENW];
    for (int e = threadIdx.x; e < n * n; e += ENT) {
        const int i = e / n, j = e - i * n;
        double v = 0.0;
        if (j - i >= -3 && j - i <= 3) {
            const double mii = F[(long long)i * 7 + 3], mjj = F[(long long)j * 7 + 3];
            v = F[(long long)i * 7 + (j - i + 3)] / sqrt(mii * mjj);
        }
        A[(long long)i * lda + j] = v;
        Qt[(long long)i * ldq + j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int np = n + (n & 1);
    const int npair = np / 2;
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0, dg = 0.0;
        for (int e = threadIdx.x; e < n * n; e += ENT) {
            const int i = e / n, j = e - i * n;
            const double a = A[(long long)i * lda + j];
            const double v = a * a;
            if (i == j) dg += v; else off += v;
        }
        double v2[2] = {off, dg};
        block_sum_n<2, ENW>(v2, red2);
        if (v2[0] <= 1e-32 * v2[1]) break;
        for (int step = 0; step < np - 1; ++step) {
            for (int k = threadIdx.x; k < npair; k += ENT) {
                auto idx = [&](int a) { return a == 0 ? 0 : 1 + (a - 1 + step) % (np - 1); };
                int p = idx(k), q = idx(np - 1 - k);
                if (p > q) { int t = p; p = q; q = t; }
                double c = 1.0, sn = 0.0;
                if (q < n) {
                    const double apq = A[(long long)p * lda + q];
                    const double app = A[(long long)p * lda + p], aqq = A[(long long)q * lda + q];
                    if (fabs(apq) > 1e-300 && fabs(apq) > 1e-18 * sqrt(fabs(app * aqq))) {
                        const double tau = (aqq - app) / (2.0 * apq);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        sn = t * c;
                    }
                } else {
                    q = -1;
                }
                cc[k] = c, ss[k] = sn, pp[k] = p, qq[k] = q;
            }
            __syncthreads();
            // rows: A <- J^T A, and Qt <- J^T Qt (= (Q J)^T)
            for (int e = threadIdx.x; e < npair * n; e += ENT) {
                const int k = e / n, col = e - k * n;
                const int p = pp[k], q = qq[k];
                if (q < 0 || ss[k] == 0.0) continue;
                const double c = cc[k], sn = ss[k];
                const double ap = A[(long long)p * lda + col], aq = A[(long long)q * lda + col];
                A[(long long)p * lda + col] = c * ap - sn * aq;
                A[(long long)q * lda + col] = sn * ap + c * aq;
                const double bp = Qt[(long long)p * ldq + col], bq = Qt[(long long)q * ldq + col];
                Qt[(long long)p * ldq + col] = c * bp - sn * bq;
                Qt[(long long)q * ldq + col] = sn * bp + c * bq;
            }
            __syncthreads();
            // columns: A <- A J
            for (int e = threadIdx.x; e < npair * n; e += ENT) {
                const int k = e / n, row = e - k * n;
                const int p = pp[k], q = qq[k];
                if (q < 0 || ss[k] == 0.0) continue;
                const double c = cc[k], sn = ss[k];
                const double ap = A[(long long)row * lda + p], aq = A[(long long)row * lda + q];
                A[(long long)row * lda + p] = c * ap - sn * aq;
                A[(long long)row * lda + q] = sn * ap + c * aq;
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += ENT) lam[i] = A[(long long)i * lda + i];
    if (ax == 0) {
        if (q_lds) {
            for (int e = threadIdx.x; e < n * n; e += ENT) {
                const int i = e / n, j = e - i * n;
                E.Qs[e] = Qt[(long long)j * ldq + i];
            }
        } else {  // Qt occupies Qs: transpose in place, one thread per pair
            for (int e = threadIdx.x; e < n * n; e += ENT) {
                const int i = e / n, j = e - i * n;
                if (i < j) {
                    const double a = Qt[(long long)i * n + j], b = Qt[(long long)j * n + i];
                    Qt[(long long)i * n + j] = b;
                    Qt[(long long)j * n + i] = a;
                }
            }
        }
    } else {
        for (int e = threadIdx.x; e < n * n; e += ENT) {
            const int i = e / n, j = e - i * n;
            E.Qv[e] = Qt[(long long)j * ldq + i];
            if (q_lds) E.QvT[e] = Qt[(long long)i * ldq + j];
        }
    }
}

// C = A B (M <= 16 NW MT rows, N <= 16 TN columns) with a K-loop on
// v_mfma_f64_16x16x4f64; operands typically in LDS. A wave owns the 16-row
// strips wave, wave + NW, ... (at most MT of them) across all TN column tiles:
// each k-step feeds one A fragment to TN independent accumulator chains.
// Operand reads are branch-free: indices past the matrix are clamped to its
// last row or column and the value is replaced by zero, so a k-step's TN
// B reads issue together. Padded k-steps add exact zeros (0 * 0).
// AP > 0 (an A operand read from memory: the s eigenvectors when they do not
// fit in LDS): A fragments are loaded a chunk of AP k-steps ahead of their use,
// so the operand waits once per chunk, behind the previous chunk's MFMAs. The
// k order is the same either way, so results are identical.
// T(it, r0, acc) receives strip it (rows r0 .. r0 + 15) with
// acc[u][j] = C[r0 + lane/16 + 4 j][16 u + lane%16].
template <int NW, int MT, int TN, int AP = 0, class FA, class FB, class FT>
__device__ __forceinline__ void mfma_gemm_t(int M, int N, int K, FA A, FB B, FT T) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tm = (M + 15) >> 4;
    const int l16 = lane & 15, l4 = lane >> 4;
#pragma unroll
    for (int it = 0; it < MT; ++it) {
        const int t = wave + it * NW;
        if (t >= tm) break;  // wave-uniform
        const int r0 = t << 4;
        const bool va = r0 + l16 < M;
        const int ra = va ? r0 + l16 : M - 1;
        d4 acc[TN];
#pragma unroll
        for (int u = 0; u < TN; ++u) acc[u] = d4{0.0, 0.0, 0.0, 0.0};
        auto aload = [&](int k) {
            const bool vk = k < K;
            const double x = A(ra, vk ? k : K - 1);
            return (va && vk) ? x : 0.0;
        };
        auto kstep = [&](double a, int k) {
            const bool vk = k < K;
            const int kc = vk ? k : K - 1;
            double b[TN];
#pragma unroll
            for (int u = 0; u < TN; ++u) {
                const int cb = (u << 4) + l16;
                const double x = B(kc, cb < N ? cb : N - 1);
                b[u] = (vk && cb < N) ? x : 0.0;
            }
#pragma unroll
            for (int u = 0; u < TN; ++u) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[u], acc[u], 0, 0, 0);
        };
        int k0 = 0;
        if constexpr (AP > 0) {
            // whole chunks: every k of the chunk is in range
            const int kfull = (K / (4 * AP)) * (4 * AP);
            double a[AP];
            if (kfull > 0) {
#pragma unroll
                for (int j = 0; j < AP; ++j) a[j] = aload(4 * j + l4);
            }
#pragma unroll 1
            for (; k0 < kfull; k0 += 4 * AP) {
                double an[AP];
                const bool more = k0 + 4 * AP < kfull;
#pragma unroll
                for (int j = 0; j < AP; ++j) an[j] = aload(more ? k0 + 4 * AP + 4 * j + l4 : 4 * j + l4);
#pragma unroll
                for (int j = 0; j < AP; ++j) kstep(a[j], k0 + 4 * j + l4);
#pragma unroll
                for (int j = 0; j < AP; ++j) a[j] = an[j];
            }
        }
#pragma unroll 2
        for (; k0 < K; k0 += 4) kstep(aload(k0 + l4), k0 + l4);
        T(it, r0, acc);
    }
}

// Bytes of dynamic LDS k_admm needs for an ns x nv instance, with (qs_in_lds)
// or without the ns x ns eigenvector matrix.
__host__ __device__ inline long long admm_lds_bytes(int ns, int nv, bool qs_in_lds) {
    return 8LL * ((long long)nv * (nv + 1) + 2LL * ns * (nv + 1) + ns + nv +
                  (qs_in_lds ? (long long)ns * (ns + 1) : 0));
}
constexpr long long ADMM_LDS_MAX = 159 * 1024;  // with the 8-wave kernel's 320 B of static LDS: under 160 KB

// Dynamic LDS budgets of the launch groups of k_eig and k_admm. A CU holds
// 160 KB of LDS; an instance whose dynamic LDS need fits a half, third or
// quarter of it (less the kernel's static LDS: 4 KB in k_eig, under 0.5 KB in
// k_admm) runs in a launch whose workgroups take that much, so that many of
// them share a CU.
constexpr long long IC_LDS_HALF = 78 * 1024, IC_LDS_THIRD = 52 * 1024;  // k_admm
constexpr long long EIG_LDS_HALF = 76 * 1024, EIG_LDS_THIRD = 49 * 1024, EIG_LDS_QUARTER = 36 * 1024;

// Strip and column-tile counts of an ns x nv ADMM instance in a workgroup of
// nw waves: MT = ceil(ceil(ns / 16) / nw), TN = ceil(nv / 16). The ADMM state
// stays in registers when MT * TN <= ADMM_REG_TILES.
constexpr int ADMM_REG_TILES = 8;
__host__ __device__ inline int admm_mt(int ns, int nw) { return (((ns + 15) >> 4) + nw - 1) / nw; }
__host__ __device__ inline int admm_tn(int nv) { return (nv + 15) >> 4; }

// ADMM NNLS (osqp.mojo:93-157): rho 0.1, sigma 1e-6, lambda 1e-6, eps 1e-8,
// at most 50000 iterations; then q0 = D z, clip at 0, normalise by m.q0
// (initial_cond.mojo:119-129). The four products per iteration run on MFMA
// with Qv, the two ns x nv work matrices and the squared eigenvalues in LDS
// (Qs too when it fits in lds_budget, the launch's dynamic LDS size).
// The iteration's state is one value per entry, w = x + u before the
// projection: z = max(w, 0) and u = w - z = min(w, 0) exactly, so
//   w' = x' + min(w, 0),  z' = max(w', 0),  u' = min(w', 0)
// reproduce the z and u updates bit for bit. Each thread keeps w and q for the
// entries the last product's epilogue gives it in registers (REG, when the
// wave's MT x TN tiles allow) or in the instance's z and q vectors.
// grid: the instances order[0 .. grid - 1]; ANW waves per workgroup; every
// instance of a launch has admm_mt(ns, ANW) <= MT and admm_tn(nv) == TN.
// Launches are grouped by LDS need: instances whose work matrices fit a third
// (or half) of the CU's LDS run as 4-wave workgroups, three (two) per CU, so
// small grids, whose iterations are latency chains, overlap each other; the
// largest run as 8-wave workgroups with up to ADMM_LDS_MAX bytes, one per CU.
// (The host picks, per instance, the form with the most resident workgroups per
// CU by the occupancy the runtime reports for each kernel.)
// Every product is computed in the same order in every form; only the
// residual norms' and m.q0's partial sums are split over threads differently.
template <int ANW, int MT, int TN, bool REG>
__global__ void __launch_bounds__(ANW * 64) k_admm(const Inst* __restrict__ inst, const int* __restrict__ order,
                                                   const double* __restrict__ ST, const double* __restrict__ EG,
                                                   double* __restrict__ VC, InstOut* __restrict__ out,
                                                   long long lds_budget) {
    constexpr int ANT = ANW * 64;
    extern __shared__ double dyn[];
    const Inst& in = inst[order[blockIdx.x]];
    const int ns = in.nsb, nv = in.nvb, N = ns * nv;
    EigView E(const_cast<double*>(EG), in);
    double* base = VC + in.o_vec;
    const double* Di = base + (long long)V_DINV * N;
    const double* bt = base + (long long)V_BT * N;
    const double* mw = base + (long long)V_MW * N;
    double* wg = base + (long long)V_ADMM_Z * N;  // w when not in registers
    double* qg = base + (long long)V_ADMM_Q * N;  // q when not in registers
    double* q0 = base + (long long)V_Q0 * N;
    const double* Ms = ST + in.o_st;
    const double rho = 0.1, sigma = 1e-6, lam = 1e-6, eps_abs = 1e-8, eps_rel = 1e-8;
    const double alpha = lam + sigma + rho;
    __shared__ double red[5 * ANW];
    // LDS: Qv (nv x nv+1), T and R (ns x nv+1), ls^2 (ns), lv^2 (nv), Qs (ns x ns+1) when it fits
    const int lv = nv + 1, lsd = ns + 1;
    double* sQv = dyn;
    double* sT = sQv + nv * lv;
    double* sR = sT + ns * lv;
    double* sL2 = sR + ns * lv;
    double* sV2 = sL2 + ns;
    const bool qs_lds = admm_lds_bytes(ns, nv, true) <= lds_budget;
    double* sQs = sV2 + nv;
    for (int e = threadIdx.x; e < nv * nv; e += ANT) sQv[(e / nv) * lv + e % nv] = E.Qv[e];
    if (qs_lds)
        for (int e = threadIdx.x; e < ns * ns; e += ANT) sQs[(e / ns) * lsd + e % ns] = E.Qs[e];
    for (int a = threadIdx.x; a < ns; a += ANT) sL2[a] = E.ls[a] * E.ls[a];
    for (int b = threadIdx.x; b < nv; b += ANT) sV2[b] = E.lv[b] * E.lv[b];
    auto Qs = [&](int I, int a) -> double { return qs_lds ? sQs[I * lsd + a] : E.Qs[(long long)I * ns + a]; };
    // q = A^T b~ (A symmetric); it is also the first right-hand side (x = z = u = 0)
    for (int r = threadIdx.x; r < N; r += ANT) {
        const int I = r / nv, i = r - I * nv;
        double acc = 0.0;
        for (int dJ = -3; dJ <= 3; ++dJ) {
            const int J = I + dJ;
            if (J < 0 || J >= ns) continue;
            for (int dj = -3; dj <= 3; ++dj) {
                const int jj = i + dj;
                if (jj < 0 || jj >= nv) continue;
                acc += Ms[(long long)((dJ + 3) * 7 + dj + 3) * N + r] * bt[J * nv + jj];
            }
        }
        if (!REG) {
            qg[r] = acc;
            wg[r] = 0.0;
        }
        sR[I * lv + i] = acc;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, l4 = lane >> 4;
    // this thread's entries of the last product: strip it, tile u, row j
    auto ent_row = [&](int it, int j) { return ((wave + it * ANW) << 4) + l4 + 4 * j; };
    auto ent_col = [&](int u) { return (u << 4) + l16; };
    double wr[REG ? MT : 1][REG ? TN : 1][4], qr[REG ? MT : 1][REG ? TN : 1][4];
    if constexpr (REG) {
#pragma unroll
        for (int it = 0; it < MT; ++it)
#pragma unroll
            for (int u = 0; u < TN; ++u)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int I = ent_row(it, j), i = ent_col(u);
                    qr[it][u][j] = (I < ns && i < nv) ? sR[I * lv + i] : 0.0;
                    wr[it][u][j] = 0.0;
                }
    }
    int iters = 0;
    const int max_iter = 50000;
    for (; iters < max_iter;) {
        // T = Qs^T R
        mfma_gemm_t<ANW, MT, TN, 4>(ns, nv, ns, [&](int a, int I) { return Qs(I, a); },
                                    [&](int I, int i) { return sR[I * lv + i]; },
                                    [&](int, int r0, d4* acc) {
#pragma unroll
                                        for (int u = 0; u < TN; ++u)
#pragma unroll
                                            for (int j = 0; j < 4; ++j) {
                                                const int a = r0 + l4 + 4 * j, i = ent_col(u);
                                                if (a < ns && i < nv) sT[a * lv + i] = acc[u][j];
                                            }
                                    });
        __syncthreads();
        // R = (T Qv) ./ (ls^2 lv^2 + alpha)
        mfma_gemm_t<ANW, MT, TN>(ns, nv, nv, [&](int a, int i) { return sT[a * lv + i]; },
                                 [&](int i, int b) { return sQv[i * lv + b]; },
                                 [&](int, int r0, d4* acc) {
#pragma unroll
                                     for (int u = 0; u < TN; ++u)
#pragma unroll
                                         for (int j = 0; j < 4; ++j) {
                                             const int a = r0 + l4 + 4 * j, b = ent_col(u);
                                             if (a < ns && b < nv) sR[a * lv + b] = acc[u][j] / (sL2[a] * sV2[b] + alpha);
                                         }
                                 });
        __syncthreads();
        // T = Qs R
        mfma_gemm_t<ANW, MT, TN, 4>(ns, nv, ns, [&](int I, int a) { return Qs(I, a); },
                                    [&](int a, int b) { return sR[a * lv + b]; },
                                    [&](int, int r0, d4* acc) {
#pragma unroll
                                        for (int u = 0; u < TN; ++u)
#pragma unroll
                                            for (int j = 0; j < 4; ++j) {
                                                const int I = r0 + l4 + 4 * j, b = ent_col(u);
                                                if (I < ns && b < nv) sT[I * lv + b] = acc[u][j];
                                            }
                                    });
        __syncthreads();
        // x = T Qv^T, then the w update, the five norms and the next right-hand
        // side q + rho (z - u) + sigma x (R is free: GEMM 3 read it before the
        // last barrier; x itself is never stored)
        double nr[5] = {0, 0, 0, 0, 0};
        mfma_gemm_t<ANW, MT, TN>(ns, nv, nv, [&](int I, int b) { return sT[I * lv + b]; },
                                 [&](int b, int i) { return sQv[i * lv + b]; },
                                 [&](int it, int r0, d4* acc) {
                                     double wo[TN][4], qe[TN][4];
#pragma unroll
                                     for (int u = 0; u < TN; ++u)
#pragma unroll
                                         for (int j = 0; j < 4; ++j) {
                                             if constexpr (REG) {
                                                 wo[u][j] = wr[it][u][j], qe[u][j] = qr[it][u][j];
                                             } else {  // every load of the strip before any store
                                                 const int I = r0 + l4 + 4 * j, i = ent_col(u);
                                                 const int e = (I < ns ? I : ns - 1) * nv + (i < nv ? i : nv - 1);
                                                 wo[u][j] = wg[e], qe[u][j] = qg[e];
                                             }
                                         }
#pragma unroll
                                     for (int u = 0; u < TN; ++u)
#pragma unroll
                                         for (int j = 0; j < 4; ++j) {
                                             const int I = r0 + l4 + 4 * j, i = ent_col(u);
                                             if (I < ns && i < nv) {
                                                 const double xv = acc[u][j], w = wo[u][j];
                                                 const double zo = w < 0.0 ? 0.0 : w, uo = w < 0.0 ? w : 0.0;
                                                 const double wn = xv + uo;
                                                 const double zn = wn < 0.0 ? 0.0 : wn, un = wn < 0.0 ? wn : 0.0;
                                                 if constexpr (REG) wr[it][u][j] = wn;
                                                 else wg[I * nv + i] = wn;
                                                 sR[I * lv + i] = qe[u][j] + rho * (zn - un) + sigma * xv;
                                                 const double pv = xv - zn, dv = rho * (zn - zo);
                                                 nr[0] += pv * pv;
                                                 nr[1] += dv * dv;
                                                 nr[2] += zn * zn;
                                                 nr[3] += xv * xv;
                                                 nr[4] += un * un;
                                             }
                                         }
                                 });
        block_sum_n<5, ANW>(nr, red);  // its first barrier also orders the R writes before GEMM 1
        ++iters;
        const double ep = eps_abs + eps_rel * sqrt(fmax(nr[2], nr[3]));
        const double ed = eps_abs + eps_rel * rho * sqrt(nr[4]);
        if (sqrt(nr[0]) < ep && sqrt(nr[1]) < ed) break;
    }
    __syncthreads();
    // q0 = max(D z, 0) / m.q0 with z = max(w, 0)
    double md = 0.0;
    if constexpr (REG) {
#pragma unroll
        for (int it = 0; it < MT; ++it)
#pragma unroll
            for (int u = 0; u < TN; ++u)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int I = ent_row(it, j), i = ent_col(u);
                    if (I < ns && i < nv) {
                        const int e = I * nv + i;
                        const double w = wr[it][u][j];
                        double v = Di[e] * (w < 0.0 ? 0.0 : w);
                        if (v < 0.0) v = 0.0;
                        wr[it][u][j] = v;
                        md += mw[e] * v;
                    }
                }
        md = block_sum_w<ANW>(md, red);
#pragma unroll
        for (int it = 0; it < MT; ++it)
#pragma unroll
            for (int u = 0; u < TN; ++u)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int I = ent_row(it, j), i = ent_col(u);
                    if (I < ns && i < nv) q0[I * nv + i] = md > 0.0 ? wr[it][u][j] / md : wr[it][u][j];
                }
    } else {
        for (int r = threadIdx.x; r < N; r += ANT) {
            const double w = wg[r];
            double v = Di[r] * (w < 0.0 ? 0.0 : w);
            if (v < 0.0) v = 0.0;
            q0[r] = v;
            md += mw[r] * v;
        }
        md = block_sum_w<ANW>(md, red);
        if (md > 0.0)
            for (int r = threadIdx.x; r < N; r += ANT) q0[r] = q0[r] / md;
    }
    if (threadIdx.x == 0) out[order[blockIdx.x]].n_admm = iters;
}

}  // namespace fpe
