// Block-banded factorisation of the Radau stage matrices E = a M + h K (a real
// or complex) and the matching block solves.
//
// E is block-banded with m x m blocks (m = n_vb) and block half-bandwidth 3.
// Block row I is stored as an m x 7m row-major panel holding (I, I-3..I+3).
// The factorisation is block Gaussian elimination E = L U with unit block-lower
// L and block-upper U, where the diagonal blocks are inverted explicitly:
//   for I = 0..n_sb-1:
//     Ainv_I = inverse of the current diagonal block        (Gauss-Jordan, row pivoting)
//     L_JI   = A_JI Ainv_I               for J = I+1..I+3    (MFMA)
//     A_JK  -= L_JI A_IK                 for J, K = I+1..I+3 (MFMA)
// After it, panel I holds L_IJ (d = 0..2), Ainv_I (d = 3) and U_IK = A_IK
// (d = 4..6). Solves: y_I = b_I - sum_J L_IJ y_J; x_I = Ainv_I (y_I - sum_K U_IK x_K).
// Pivoting is row pivoting inside each diagonal block (the Gauss-Jordan
// inverse); E = a M + h K with D-scaled mass and stiffness is block diagonally
// dominant for the step sizes Radau takes, which the parity tests exercise.
#pragma once
#include "device_common.hpp"

namespace fpe {

constexpr int LS = MAX_BLOCK + 1;  // LDS row stride of an m x m block

struct LuLds {
    double Ar[MAX_BLOCK * LS];  // diagonal block -> its inverse
    double Ai[MAX_BLOCK * LS];
    double Sr[MAX_BLOCK * LS];  // staged operand block
    double Si[MAX_BLOCK * LS];
    int piv[MAX_BLOCK];
    int singular;
};

// Lane layout of v_mfma_f64_16x16x4f64 (checked on gfx950 with an in-bounds
// probe): a = A[row0 + lane%16][k0 + lane/16], b = B[k0 + lane/16][col0 + lane%16],
// acc[j] = C[row0 + lane/16 + 4*j][col0 + lane%16]. (The f32 16x16x4 form places
// acc[j] at row 4*(lane/16) + j instead.)
//
// C (op)= A B, M x N x K, K <= MAX_BLOCK. Each wave owns 16-row strips; the A
// fragments of a strip are loaded once into registers and reused for every
// 16-column tile of that strip.
template <bool SUB, bool CPLX, class FA, class FB, class FC>
__device__ __forceinline__ void mfma_gemm(int M, int N, int K, FA A, FB B, FC C) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tm = (M + 15) >> 4, tn = (N + 15) >> 4, tk = (K + 3) >> 2;
    const int l16 = lane & 15, l4 = lane >> 4;
    for (int t = wave; t < tm; t += NWAVE) {
        const int r0 = t << 4, ra = r0 + l16;
        double ar[MAX_BLOCK / 4], ai[MAX_BLOCK / 4];
#pragma unroll
        for (int s = 0; s < MAX_BLOCK / 4; ++s) {
            const int k = 4 * s + l4;
            const bool ok = s < tk && ra < M && k < K;
            double re = 0.0, im = 0.0;
            if (ok) A(ra, k, re, im);
            ar[s] = re;
            ai[s] = im;
        }
        for (int u = 0; u < tn; ++u) {
            const int c0 = u << 4, cb = c0 + l16;
            d4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < MAX_BLOCK / 4; ++s) {
                if (s < tk) {
                    const int k = 4 * s + l4;
                    double br = 0.0, bi = 0.0;
                    if (cb < N && k < K) B(k, cb, br, bi);
                    accr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[s], br, accr, 0, 0, 0);
                    if (CPLX) {
                        accr = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[s], bi, accr, 0, 0, 0);
                        acci = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[s], bi, acci, 0, 0, 0);
                        acci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[s], br, acci, 0, 0, 0);
                    }
                }
            }
            if (cb < N) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int r = r0 + l4 + 4 * j;
                    if (r < M) C(r, cb, accr[j], acci[j]);
                }
            }
        }
    }
}

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// In-place Gauss-Jordan inverse of the m x m block in (Ar, Ai) with row
// pivoting, run by wave 0 alone (lane i owns row i; LDS ordering within a
// wave needs no workgroup barrier). Sets singular on a zero pivot.
template <bool CPLX>
__device__ void gj_inverse(double* Ar, double* Ai, int* piv, int m, int* singular) {
    const int lane = threadIdx.x & 63;
    for (int k = 0; k < m; ++k) {
        double v = -1.0;
        int idx = lane;
        if (lane >= k && lane < m) v = CPLX ? fabs(Ar[lane * LS + k]) + fabs(Ai[lane * LS + k]) : fabs(Ar[lane * LS + k]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(v, o, 64);
            const int oi = __shfl_xor(idx, o, 64);
            if (ov > v || (ov == v && oi < idx)) v = ov, idx = oi;
        }
        const int p = idx;
        if (lane == 0) {
            piv[k] = p;
            if (!(v > 0.0)) *singular = 1;
        }
        if (p != k && lane < m) {
            double t = Ar[k * LS + lane];
            Ar[k * LS + lane] = Ar[p * LS + lane];
            Ar[p * LS + lane] = t;
            if (CPLX) {
                t = Ai[k * LS + lane];
                Ai[k * LS + lane] = Ai[p * LS + lane];
                Ai[p * LS + lane] = t;
            }
        }
        wave_fence();
        // pivot row: A[k][k] <- 1/piv, A[k][j] <- A[k][j]/piv
        double dr = Ar[k * LS + k], di = CPLX ? Ai[k * LS + k] : 0.0;
        double ir, ii;
        if (CPLX) {
            const double den = dr * dr + di * di;
            ir = den > 0.0 ? dr / den : 0.0;
            ii = den > 0.0 ? -di / den : 0.0;
        } else {
            ir = dr != 0.0 ? 1.0 / dr : 0.0;
            ii = 0.0;
        }
        wave_fence();
        if (lane < m) {
            const double xr = (lane == k) ? 1.0 : Ar[k * LS + lane];
            const double xi = (lane == k) ? 0.0 : (CPLX ? Ai[k * LS + lane] : 0.0);
            Ar[k * LS + lane] = xr * ir - xi * ii;
            if (CPLX) Ai[k * LS + lane] = xr * ii + xi * ir;
        }
        wave_fence();
        // eliminate column k from every other row
        if (lane < m && lane != k) {
            const double fr = Ar[lane * LS + k], fi = CPLX ? Ai[lane * LS + k] : 0.0;
            Ar[lane * LS + k] = 0.0;
            if (CPLX) Ai[lane * LS + k] = 0.0;
            for (int j = 0; j < m; ++j) {
                const double pr = Ar[k * LS + j];
                if (CPLX) {
                    const double pi_ = Ai[k * LS + j];
                    Ar[lane * LS + j] -= fr * pr - fi * pi_;
                    Ai[lane * LS + j] -= fr * pi_ + fi * pr;
                } else {
                    Ar[lane * LS + j] -= fr * pr;
                }
            }
        }
        wave_fence();
    }
    // undo the row interchanges as column interchanges, in reverse order
    for (int k = m - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k && lane < m) {
            double t = Ar[lane * LS + k];
            Ar[lane * LS + k] = Ar[lane * LS + p];
            Ar[lane * LS + p] = t;
            if (CPLX) {
                t = Ai[lane * LS + k];
                Ai[lane * LS + k] = Ai[lane * LS + p];
                Ai[lane * LS + p] = t;
            }
        }
        wave_fence();
    }
}

// Write E = a M_s + h K_s (a = ar + i ai) into the panels from the 49-point stencils.
template <bool CPLX>
__device__ void assemble_E(const double* __restrict__ Ms, const double* __restrict__ Ks, int ns, int m,
                           double ar, double ai, double h, double* __restrict__ Pr, double* __restrict__ Pi) {
    const int N = ns * m;
    const long long tot = (long long)ns * 7 * m * m;
    for (long long e = threadIdx.x; e < tot; e += NT) {
        const int j = int(e % m);
        long long t = e / m;
        const int d = int(t % 7);
        t /= 7;
        const int i = int(t % m);
        const int I = int(t / m);
        const int J = I + d - 3, dj = j - i;
        double re = 0.0, im = 0.0;
        if (J >= 0 && J < ns && dj >= -3 && dj <= 3) {
            const long long r = (long long)I * m + i;
            const int st = d * 7 + dj + 3;
            const double mv = Ms[(long long)st * N + r], kv = Ks[(long long)st * N + r];
            re = ar * mv + h * kv;
            im = ai * mv;
        }
        Pr[e] = re;
        if (CPLX) Pi[e] = im;
    }
}

// Factor the panels in place. Returns nonzero if a diagonal block was singular.
template <bool CPLX>
__device__ int block_lu(double* __restrict__ Pr, double* __restrict__ Pi, int ns, int m, LuLds& L) {
    const int ld = 7 * m;
    const long long ps = (long long)m * ld;
    const int tid = threadIdx.x, wave = tid >> 6;
    if (tid == 0) L.singular = 0;
    for (int I = 0; I < ns; ++I) {
        double* PIr = Pr + I * ps;
        double* PIi = Pi + I * ps;
        for (int e = tid; e < m * m; e += NT) {
            const int i = e / m, j = e - i * m;
            L.Ar[i * LS + j] = PIr[(long long)i * ld + 3 * m + j];
            if (CPLX) L.Ai[i * LS + j] = PIi[(long long)i * ld + 3 * m + j];
        }
        __syncthreads();
        if (wave == 0) gj_inverse<CPLX>(L.Ar, L.Ai, L.piv, m, &L.singular);
        __syncthreads();
        for (int e = tid; e < m * m; e += NT) {
            const int i = e / m, j = e - i * m;
            PIr[(long long)i * ld + 3 * m + j] = L.Ar[i * LS + j];
            if (CPLX) PIi[(long long)i * ld + 3 * m + j] = L.Ai[i * LS + j];
        }
        const int nU = (ns - 1 - I) < 3 ? (ns - 1 - I) : 3;
        // L_JI = A_JI Ainv_I, one block at a time through the staging buffer
        for (int dJ = 1; dJ <= nU; ++dJ) {
            double* Jr = Pr + (I + dJ) * ps + (3 - dJ) * m;
            double* Ji = Pi + (I + dJ) * ps + (3 - dJ) * m;
            for (int e = tid; e < m * m; e += NT) {
                const int i = e / m, j = e - i * m;
                L.Sr[i * LS + j] = Jr[(long long)i * ld + j];
                if (CPLX) L.Si[i * LS + j] = Ji[(long long)i * ld + j];
            }
            __syncthreads();
            mfma_gemm<false, CPLX>(
                m, m, m,
                [&](int i, int k, double& re, double& im) {
                    re = L.Sr[i * LS + k];
                    im = CPLX ? L.Si[i * LS + k] : 0.0;
                },
                [&](int k, int j, double& re, double& im) {
                    re = L.Ar[k * LS + j];
                    im = CPLX ? L.Ai[k * LS + j] : 0.0;
                },
                [&](int i, int j, double re, double im) {
                    Jr[(long long)i * ld + j] = re;
                    if (CPLX) Ji[(long long)i * ld + j] = im;
                });
            __syncthreads();
        }
        // A_JK -= L_JI A_IK: stage A_IK (rows of panel I), L_JI from global
        for (int dK = 1; dK <= nU; ++dK) {
            for (int e = tid; e < m * m; e += NT) {
                const int i = e / m, j = e - i * m;
                L.Sr[i * LS + j] = PIr[(long long)i * ld + (3 + dK) * m + j];
                if (CPLX) L.Si[i * LS + j] = PIi[(long long)i * ld + (3 + dK) * m + j];
            }
            __syncthreads();
            mfma_gemm<true, CPLX>(
                nU * m, m, m,
                [&](int rr, int k, double& re, double& im) {
                    const int dJ = 1 + rr / m, i = rr - (dJ - 1) * m;
                    const long long o = (I + dJ) * ps + (long long)i * ld + (3 - dJ) * m + k;
                    re = Pr[o];
                    im = CPLX ? Pi[o] : 0.0;
                },
                [&](int k, int j, double& re, double& im) {
                    re = L.Sr[k * LS + j];
                    im = CPLX ? L.Si[k * LS + j] : 0.0;
                },
                [&](int rr, int j, double re, double im) {
                    const int dJ = 1 + rr / m, i = rr - (dJ - 1) * m;
                    const long long o = (I + dJ) * ps + (long long)i * ld + (dK - dJ + 3) * m + j;
                    Pr[o] -= re;
                    if (CPLX) Pi[o] -= im;
                });
            __syncthreads();
        }
    }
    return L.singular;
}

// Solve E x = b in place with the block factors. Real: NR right-hand sides
// x[0..NR-1]; complex: one right-hand side (xr[0], xi[0]).
// Thread (i, q): row i = tid / 4 of the current block, part q = tid % 4.
template <bool CPLX, int NR>
__device__ void block_solve(const double* __restrict__ Pr, const double* __restrict__ Pi, int ns, int m,
                            double* const* xr, double* const* xi, double* sT /* LDS [2 * NR * MAX_BLOCK] */) {
    const int ld = 7 * m;
    const long long ps = (long long)m * ld;
    const int tid = threadIdx.x, i = tid >> 2, q = tid & 3;
    const bool row = i < m;
    double* sTr = sT;
    double* sTi = sT + NR * MAX_BLOCK;
    // forward: y_I = b_I - sum_{J=I-3}^{I-1} L_IJ y_J
    for (int I = 0; I < ns; ++I) {
        const double* PIr = Pr + I * ps;
        const double* PIi = Pi + I * ps;
        double ar[NR], ai[NR];
#pragma unroll
        for (int v = 0; v < NR; ++v) ar[v] = 0.0, ai[v] = 0.0;
        if (row) {
            const int d0 = I >= 3 ? 0 : 3 - I;
            for (int c = d0 * m + q; c < 3 * m; c += 4) {
                const int J = I - 3 + c / m, k = c - (c / m) * m;
                const double lr = PIr[(long long)i * ld + c];
                if (CPLX) {
                    const double li = PIi[(long long)i * ld + c];
                    const double yr = xr[0][J * m + k], yi = xi[0][J * m + k];
                    ar[0] += lr * yr - li * yi;
                    ai[0] += lr * yi + li * yr;
                } else {
#pragma unroll
                    for (int v = 0; v < NR; ++v) ar[v] += lr * xr[v][J * m + k];
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NR; ++v) {
            ar[v] += __shfl_xor(ar[v], 1, 64);
            ar[v] += __shfl_xor(ar[v], 2, 64);
            if (CPLX) {
                ai[v] += __shfl_xor(ai[v], 1, 64);
                ai[v] += __shfl_xor(ai[v], 2, 64);
            }
        }
        if (row && q == 0) {
#pragma unroll
            for (int v = 0; v < NR; ++v) {
                xr[v][I * m + i] -= ar[v];
                if (CPLX) xi[v][I * m + i] -= ai[v];
            }
        }
        __syncthreads();
    }
    // backward: x_I = Ainv_I (y_I - sum_{K=I+1}^{I+3} U_IK x_K)
    for (int I = ns - 1; I >= 0; --I) {
        const double* PIr = Pr + I * ps;
        const double* PIi = Pi + I * ps;
        double ar[NR], ai[NR];
#pragma unroll
        for (int v = 0; v < NR; ++v) ar[v] = 0.0, ai[v] = 0.0;
        if (row) {
            const int cmax = (ns - 1 - I < 3 ? ns - 1 - I : 3) * m;
            for (int c = q; c < cmax; c += 4) {
                const int K = I + 1 + c / m, k = c - (c / m) * m;
                const double ur = PIr[(long long)i * ld + 4 * m + c];
                if (CPLX) {
                    const double ui = PIi[(long long)i * ld + 4 * m + c];
                    const double yr = xr[0][K * m + k], yi = xi[0][K * m + k];
                    ar[0] += ur * yr - ui * yi;
                    ai[0] += ur * yi + ui * yr;
                } else {
#pragma unroll
                    for (int v = 0; v < NR; ++v) ar[v] += ur * xr[v][K * m + k];
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NR; ++v) {
            ar[v] += __shfl_xor(ar[v], 1, 64);
            ar[v] += __shfl_xor(ar[v], 2, 64);
            if (CPLX) {
                ai[v] += __shfl_xor(ai[v], 1, 64);
                ai[v] += __shfl_xor(ai[v], 2, 64);
            }
        }
        if (row && q == 0) {
#pragma unroll
            for (int v = 0; v < NR; ++v) {
                sTr[v * MAX_BLOCK + i] = xr[v][I * m + i] - ar[v];
                if (CPLX) sTi[v * MAX_BLOCK + i] = xi[v][I * m + i] - ai[v];
            }
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < NR; ++v) ar[v] = 0.0, ai[v] = 0.0;
        if (row) {
            for (int k = q; k < m; k += 4) {
                const double gr = PIr[(long long)i * ld + 3 * m + k];
                if (CPLX) {
                    const double gi = PIi[(long long)i * ld + 3 * m + k];
                    const double tr = sTr[k], ti = sTi[k];
                    ar[0] += gr * tr - gi * ti;
                    ai[0] += gr * ti + gi * tr;
                } else {
#pragma unroll
                    for (int v = 0; v < NR; ++v) ar[v] += gr * sTr[v * MAX_BLOCK + k];
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NR; ++v) {
            ar[v] += __shfl_xor(ar[v], 1, 64);
            ar[v] += __shfl_xor(ar[v], 2, 64);
            if (CPLX) {
                ai[v] += __shfl_xor(ai[v], 1, 64);
                ai[v] += __shfl_xor(ai[v], 2, 64);
            }
        }
        if (row && q == 0) {
#pragma unroll
            for (int v = 0; v < NR; ++v) {
                xr[v][I * m + i] = ar[v];
                if (CPLX) xi[v][I * m + i] = ai[v];
            }
        }
        __syncthreads();
    }
}

}  // namespace fpe
