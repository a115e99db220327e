// Galerkin assembly kernels: B-spline tables at quadrature points, the banded
// 1D factor matrices, and the Kronecker-sum expansion into D-scaled M and K
// stencils.  Reference maths: src/numerics/bspline/basis.mojo:38-213,
// src/numerics/bspline/recombination.mojo:14-86, src/engines/fpe/galerkin.mojo:16-125,
// src/engines/fpe/solver.mojo:65-87.
#pragma once
#include "device_common.hpp"

namespace fpe {

// Cubic B-spline values and first derivatives of the 4 functions that are
// nonzero on the knot span of x (span algorithm, The NURBS Book A2.2/A2.3).
// Returns the span index k (raw functions k-3..k).
__device__ int bspline4(const double* __restrict__ t, int nk, double x, double* b, double* db) {
    // last span with t[k] < t[k+1] and t[k] <= x; clamped: p = 3
    const int p = 3;
    const int nb = nk - p - 1;
    int lo = p, hi = nb;  // spans p .. nb-1
    if (x >= t[nb]) {
        lo = nb - 1;
    } else {
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (t[mid] <= x) lo = mid; else hi = mid;
        }
    }
    const int k = lo;
    double left[4], right[4], N[4];
    // degree-2 values first (for derivatives), then degree 3
    N[0] = 1.0;
    double N2[3];
    for (int j = 1; j <= p; ++j) {
        left[j] = x - t[k + 1 - j];
        right[j] = t[k + j] - x;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            double tmp = N[r] / (right[r + 1] + left[j - r]);
            N[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        N[j] = saved;
        if (j == 2) N2[0] = N[0], N2[1] = N[1], N2[2] = N[2];
    }
    for (int q = 0; q < 4; ++q) b[q] = N[q];
    // d/dx B_{i,3} = 3/(t[i+3]-t[i]) B_{i,2} - 3/(t[i+4]-t[i+1]) B_{i+1,2}
    for (int q = 0; q < 4; ++q) {
        const int i = k - 3 + q;
        double lo2 = (q >= 1) ? N2[q - 1] : 0.0;  // B_{i,2} nonzero for i in k-2..k
        double hi2 = (q <= 2) ? N2[q] : 0.0;      // B_{i+1,2}
        double d1 = t[i + 3] - t[i], d2 = t[i + 4] - t[i + 1];
        double a = (d1 != 0.0) ? 3.0 / d1 * lo2 : 0.0;
        double c = (d2 != 0.0) ? 3.0 / d2 * hi2 : 0.0;
        db[q] = a - c;
    }
    return k;
}

// Tables per quadrature point: first recombined column c0 and 4 values / 4
// derivatives for columns c0..c0+3. Raw entries with |v| <= 1e-12 are dropped
// as in the reference; the recombination drops (Dirichlet) or merges (Neumann)
// the two end functions.  grid: (n_inst, 2 axes).
__global__ void __launch_bounds__(NT) k_basis(const Inst* __restrict__ inst, const double* __restrict__ KN,
                                              const double* __restrict__ QD, double* __restrict__ BV,
                                              int* __restrict__ BC) {
    const Inst& in = inst[blockIdx.x];
    const int ax = blockIdx.y;
    const double* t = KN + (ax == 0 ? in.o_sk : in.o_vk);
    const int nk = ax == 0 ? in.nsk : in.nvk;
    const int nq = ax == 0 ? in.nqs : in.nqv;
    const double* xq = QD + (ax == 0 ? in.o_sq : in.o_vq);
    const long long ob = ax == 0 ? in.o_bs : in.o_bv;
    const bool lneu = ax == 0 ? in.s_lneu : true, rneu = ax == 0 ? in.s_rneu : true;
    const int nb = nk - 4, nr = nb - 2;
    for (int p = threadIdx.x; p < nq; p += NT) {
        double b[4], db[4];
        int k = bspline4(t, nk, xq[p], b, db);
        int c0 = k - 4 < 0 ? 0 : k - 4;
        double vb[4] = {0, 0, 0, 0}, vd[4] = {0, 0, 0, 0};
        for (int q = 0; q < 4; ++q) {
            const int raw = k - 3 + q;
            int tgt;
            if (raw == 0) tgt = lneu ? 0 : -1;
            else if (raw == nb - 1) tgt = rneu ? nr - 1 : -1;
            else tgt = raw - 1;
            if (tgt < 0) continue;
            if (b[q] > 1e-12 || b[q] < -1e-12) vb[tgt - c0] += b[q];
            if (db[q] > 1e-12 || db[q] < -1e-12) vd[tgt - c0] += db[q];
        }
        double* o = BV + ob + (long long)p * 8;
        for (int q = 0; q < 4; ++q) o[q] = vb[q], o[4 + q] = vd[q];
        BC[ob / 8 + p] = c0;
    }
}

// First point index p with c0[p] >= key (c0 is nondecreasing in p).
__device__ __forceinline__ int lower_c0(const int* c0, int n, int key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (c0[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Banded 1D factors F[f][i][d] (j = i + d - 3), F = sum_p a_i(p) * (w_f(p) * b_j(p)),
// accumulated in ascending p (galerkin.mojo:44-125):
//   s: Ns = B'W B, S1 = dB'(k1 s W)B, S2 = dB'(k2 s W)B, S3 = dB'(k4 s W)B, S4 = dB'(k3 s^2 W)dB
//   v: Nv = B'W B, V2 = B'(v W)B, V3 = B'(v W)dB, V4 = dB'W B, V5 = dB'(v W)B, V6 = dB'(v W)dB
__global__ void __launch_bounds__(NT) k_factors(const Inst* __restrict__ inst, const double* __restrict__ QD,
                                                const double* __restrict__ BV, const int* __restrict__ BC,
                                                double* __restrict__ FC) {
    const Inst& in = inst[blockIdx.x];
    const int ax = blockIdx.y;
    const int nb = ax == 0 ? in.nsb : in.nvb;
    const int nq = ax == 0 ? in.nqs : in.nqv;
    const long long ob = ax == 0 ? in.o_bs : in.o_bv;
    const double* tab = BV + ob;
    const int* c0 = BC + ob / 8;
    const double* xq = QD + (ax == 0 ? in.o_sq : in.o_vq);
    const double* wq = xq + nq;
    const double* phys = ax == 0 ? xq + 2 * nq : xq;
    const double j = in.jac;
    const double k1 = (-in.r + 0.5 * in.rho * in.eta) / j, k2 = 1.0 / j, k3 = 0.5 / (j * j),
                 k4 = 0.5 * in.rho * in.eta / j;
    double* out = FC + (ax == 0 ? in.o_fs : in.o_fv);
    for (int e = threadIdx.x; e < nb * 7; e += NT) {
        const int i = e / 7, d = e - 7 * i, jj = i + d - 3;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        if (jj >= 0 && jj < nb) {
            const int hiC = i < jj ? i : jj, loC = (i > jj ? i : jj) - 3;
            const int p0 = lower_c0(c0, nq, loC), p1 = lower_c0(c0, nq, hiC + 1);
            for (int p = p0; p < p1; ++p) {
                const double* tp = tab + (long long)p * 8;
                const int oi = i - c0[p], oj = jj - c0[p];
                const double bi = tp[oi], bj = tp[oj], di = tp[4 + oi], dj = tp[4 + oj];
                const double w = wq[p], x = phys[p];
                if (ax == 0) {
                    const double sw = x * w;
                    acc[0] += bi * (w * bj);
                    acc[1] += di * ((k1 * sw) * bj);
                    acc[2] += di * ((k2 * sw) * bj);
                    acc[3] += di * ((k4 * sw) * bj);
                    acc[4] += di * ((k3 * (x * x * w)) * dj);
                } else {
                    const double vw = x * w;
                    acc[0] += bi * (w * bj);
                    acc[1] += bi * (vw * bj);
                    acc[2] += bi * (vw * dj);
                    acc[3] += di * (w * bj);
                    acc[4] += di * (vw * bj);
                    acc[5] += di * (vw * dj);
                }
            }
        }
        const int nf = ax == 0 ? 5 : 6;
        for (int f = 0; f < nf; ++f) out[((long long)f * nb + i) * 7 + d] = acc[f];
    }
}

__device__ __forceinline__ double band(const double* F, int nb, int f, int i, int j) {
    return F[((long long)f * nb + i) * 7 + (j - i + 3)];
}

// Kronecker-sum expansion of M = Ns (x) Nv and
// K = S1(x)Nv + S2(x)V2 + Ns(x)(k5 V4 + k6 V5) + S4(x)V2 + k8 Ns(x)V6 + (S3(x)V3)^T + S3(x)V3
// into 49-point stencils, scaled as D^-1 M D^-1 with D = sqrt(diag M).
// Stencil entry e = (J-I+3)*7 + (j-i+3) is stored plane-major: ST[o + e*N + r]
// (M planes 0..48, K planes 49..97). A workgroup takes 256 consecutive rows at a
// time; D^-1 of those rows and of the 3 n_vb + 3 rows on either side (every
// row a stencil of the block reaches) is computed once into LDS.
// grid: (n_inst, row chunks).
constexpr int STENCIL_HALO = 3 * 64 + 3;  // 3 n_vb + 3 rows, n_vb <= 64
__global__ void __launch_bounds__(NT) k_stencil(const Inst* __restrict__ inst, const double* __restrict__ FC,
                                                double* __restrict__ ST, double* __restrict__ VC) {
    const Inst& in = inst[blockIdx.x];
    const int ns = in.nsb, nv = in.nvb, N = ns * nv;
    const double* Fs = FC + in.o_fs;
    const double* Fv = FC + in.o_fv;
    const double eta = in.eta, rho = in.rho, kappa = in.kappa, theta = in.theta;
    const double k5 = 0.5 * eta * eta - kappa * theta, k6 = kappa + 0.5 * rho * eta, k8 = 0.5 * eta * eta;
    double* Ms = ST + in.o_st;
    double* Ks = Ms + 49LL * N;
    double* D = VC + in.o_vec + (long long)V_D * N;
    double* Di = VC + in.o_vec + (long long)V_DINV * N;
    __shared__ double dinv[NT + 2 * STENCIL_HALO];
    const int H = 3 * nv + 3;
    for (int r0 = blockIdx.y * NT; r0 < N; r0 += gridDim.y * NT) {
        __syncthreads();
        for (int k = threadIdx.x; k < NT + 2 * H; k += NT) {
            const int rr = r0 - H + k;
            if (rr >= 0 && rr < N) {
                const int J = rr / nv, jj = rr - J * nv;
                const double m = band(Fs, ns, 0, J, J) * band(Fv, nv, 0, jj, jj);
                dinv[k] = m > 1e-30 ? 1.0 / sqrt(m) : 1.0;
            }
        }
        __syncthreads();
        const int r = r0 + threadIdx.x;
        if (r >= N) continue;
        const int I = r / nv, i = r - I * nv;
        const double mrr = band(Fs, ns, 0, I, I) * band(Fv, nv, 0, i, i);
        const double dr = mrr > 1e-30 ? sqrt(mrr) : 1.0;
        const double dir = mrr > 1e-30 ? 1.0 / dr : 1.0;
        D[r] = dr;
        Di[r] = dir;
        const double* dn = dinv + (threadIdx.x + H);  // D^-1 of row r + o at dn[o], |o| <= H
        // the row's v-factor entries (i, i + dj) and the transposed V3 (i + dj, i), loaded
        // once and used by all seven block offsets dJ (each s-factor entry then by seven dj)
        double v0[7], v1[7], v2[7], v2t[7], v3[7], v4[7], v5[7];
        bool vin[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const int jj = i + q - 3;
            vin[q] = jj >= 0 && jj < nv;
            const int jc = vin[q] ? jj : i;
            v0[q] = band(Fv, nv, 0, i, jc);
            v1[q] = band(Fv, nv, 1, i, jc);
            v2[q] = band(Fv, nv, 2, i, jc);
            v2t[q] = band(Fv, nv, 2, jc, i);
            v3[q] = band(Fv, nv, 3, i, jc);
            v4[q] = band(Fv, nv, 4, i, jc);
            v5[q] = band(Fv, nv, 5, i, jc);
        }
        for (int dJ = -3; dJ <= 3; ++dJ) {
            const int J = I + dJ;
            const bool Jin = J >= 0 && J < ns;
            const int Jc = Jin ? J : I;
            const double s0 = band(Fs, ns, 0, I, Jc), s1 = band(Fs, ns, 1, I, Jc), s2 = band(Fs, ns, 2, I, Jc);
            const double s3 = band(Fs, ns, 3, I, Jc), s3t = band(Fs, ns, 3, Jc, I), s4 = band(Fs, ns, 4, I, Jc);
#pragma unroll
            for (int q = 0; q < 7; ++q) {
                const int e = (dJ + 3) * 7 + q;
                double mv = 0.0, kv = 0.0;
                if (Jin && vin[q]) {
                    const double ns_ = s0;
                    const double nv_ = v0[q];
                    mv = ns_ * nv_;
                    const double a = s1 * nv_;
                    const double b = s2 * v1[q];
                    const double c = ns_ * (k5 * v3[q] + k6 * v4[q]);
                    const double d = s4 * v1[q];
                    const double f = (k8 * ns_) * v5[q];
                    const double g = s3t * v2t[q];
                    const double h = s3 * v2[q];
                    kv = (((((a + b) + c) + d) + f) + g) + h;
                    const double sc = dir * dn[dJ * nv + q - 3];
                    mv = sc * mv;
                    kv = sc * kv;
                }
                Ms[(long long)e * N + r] = mv;
                Ks[(long long)e * N + r] = kv;
            }
        }
    }
}

}  // namespace fpe
