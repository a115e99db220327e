// Sum-factorised operator products (src/sparse/kron_spmv.mojo:13-146 applies
// kron(A, B) x as vec(A X B^T); src/engines/fpe/galerkin.mojo:89-125 is the
// Kronecker sum). With the banded 1D factors of the instance,
//   M = Ns (x) Nv
//   K = S1 (x) Nv + (S2 + S4) (x) V2 + Ns (x) (k5 V4 + k6 V5 + k8 V6) + S3 (x) V3 + S3^T (x) V3^T,
// so for X = reshape(D^-1 x, n_sb, n_vb) (dof r = I n_vb + i):
//   pass 1: Y_t = X B_t^T for B_t = Nv, V2, Vc, V3, V3^T   (7-point rows, v direction)
//   pass 2: D^-1 (S1 Y_Nv + S24 Y_V2 + Ns Y_Vc + S3 Y_V3 + S3^T Y_V3T) = K_s x,
//           D^-1 (Ns Y_Nv) = M_s x                            (7-point columns, s direction)
// where K_s, M_s are the D-scaled operators of the stencils. A row of pass 1 reads
// 7 entries of x and 35 factor entries; a row of pass 2 reads 35 Y entries and 35
// factor entries, against 49 M and 49 K stencil entries of the stencil product.
#pragma once
#include "device_common.hpp"

namespace fpe {

struct KronFactors {
    const double* Fs;  // 5 s factors (Ns, S1, S2, S3, S4), n_sb rows of 7 band entries each
    const double* Fv;  // 6 v factors (Nv, V2, V3, V4, V5, V6), n_vb rows of 7
    double k5, k6, k8;
    int ns, nv;
};

__device__ __forceinline__ KronFactors kron_factors(const Inst& in, const double* FC) {
    KronFactors f;
    f.Fs = FC + in.o_fs, f.Fv = FC + in.o_fv, f.ns = in.nsb, f.nv = in.nvb;
    f.k5 = 0.5 * in.eta * in.eta - in.kappa * in.theta;
    f.k6 = in.kappa + 0.5 * in.rho * in.eta;
    f.k8 = 0.5 * in.eta * in.eta;
    return f;
}

// Band entry (i, j) of factor f (0 when |i - j| > 3 or j outside [0, nb)).
__device__ __forceinline__ double kband(const double* F, int nb, int f, int i, int j) {
    return (j >= 0 && j < nb && j - i <= 3 && i - j <= 3) ? F[((long long)f * nb + i) * 7 + (j - i + 3)] : 0.0;
}

// Pass 1 for NX vectors: Y[5 q + t] = X_q B_t^T (each Y slot N doubles), with
// X_q = D^-1 x_q (Dinv may be null: no scaling). Ends with a barrier.
template <int NX>
__device__ void kron_pass1(const KronFactors& f, const double* Dinv, const double* const* x, double* Y, int nthr,
                           int tid) {
    const int ns = f.ns, nv = f.nv, N = ns * nv;
    for (int r = tid; r < N; r += nthr) {
        const int J = r / nv, i = r - J * nv;
        double b[5][7];
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const int j = i + q - 3;
            b[0][q] = kband(f.Fv, nv, 0, i, j);
            b[1][q] = kband(f.Fv, nv, 1, i, j);
            b[2][q] = f.k5 * kband(f.Fv, nv, 3, i, j) + f.k6 * kband(f.Fv, nv, 4, i, j) + f.k8 * kband(f.Fv, nv, 5, i, j);
            b[3][q] = kband(f.Fv, nv, 2, i, j);
            b[4][q] = (j >= 0 && j < nv) ? kband(f.Fv, nv, 2, j, i) : 0.0;  // V3^T
        }
#pragma unroll
        for (int v = 0; v < NX; ++v) {
            double xv[7];
#pragma unroll
            for (int q = 0; q < 7; ++q) {
                const int j = i + q - 3;
                const int c = J * nv + min(max(j, 0), nv - 1);
                xv[q] = (j >= 0 && j < nv) ? x[v][c] * (Dinv ? Dinv[c] : 1.0) : 0.0;
            }
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                double a = 0.0;
#pragma unroll
                for (int q = 0; q < 7; ++q) a += b[t][q] * xv[q];
                Y[(long long)(5 * v + t) * N + r] = a;
            }
        }
    }
    __syncthreads();
}

// Pass 2 at row r for vector v: (K_s x)_r and (M_s x)_r.
__device__ __forceinline__ void kron_pass2_row(const KronFactors& f, const double* Dinv, const double* Y, int v, int r,
                                               double& kx, double& mx) {
    const int ns = f.ns, nv = f.nv, N = ns * nv;
    const int I = r / nv, i = r - I * nv;
    double k = 0.0, m = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const int J = I + q - 3;
        if (J < 0 || J >= ns) continue;
        const long long c = (long long)J * nv + i;
        const double y0 = Y[(long long)(5 * v + 0) * N + c], y1 = Y[(long long)(5 * v + 1) * N + c];
        const double y2 = Y[(long long)(5 * v + 2) * N + c], y3 = Y[(long long)(5 * v + 3) * N + c];
        const double y4 = Y[(long long)(5 * v + 4) * N + c];
        const double nsij = kband(f.Fs, ns, 0, I, J);
        k += kband(f.Fs, ns, 1, I, J) * y0 + (kband(f.Fs, ns, 2, I, J) + kband(f.Fs, ns, 4, I, J)) * y1 + nsij * y2 +
             kband(f.Fs, ns, 3, I, J) * y3 + kband(f.Fs, ns, 3, J, I) * y4;
        m += nsij * y0;
    }
    const double d = Dinv ? Dinv[r] : 1.0;
    kx = d * k;
    mx = d * m;
}

// Standalone products of one instance's unscaled K and M with x (operator export
// and its test): grid (1), block NT.
__global__ void __launch_bounds__(NT) k_kron_apply(const Inst* __restrict__ inst, int id, const double* __restrict__ FC,
                                                   const double* __restrict__ x, double* __restrict__ Y,
                                                   double* __restrict__ Kx, double* __restrict__ Mx) {
    const KronFactors f = kron_factors(inst[id], FC);
    const double* xs[1] = {x};
    kron_pass1<1>(f, nullptr, xs, Y, NT, threadIdx.x);
    for (int r = threadIdx.x; r < f.ns * f.nv; r += NT) kron_pass2_row(f, nullptr, Y, 0, r, Kx[r], Mx[r]);
}

}  // namespace fpe
