// Block-banded LU of the Radau stage matrices E = a M + h K (a real or
// complex) and the matching block triangular solves.
//
// Storage: block row I is an m x 7m row-major panel holding the blocks
// (I, I-3) .. (I, I+3) (m = n_vb).  After factorisation the panel holds
//   d = 0..2 : L_{I,J} multipliers (unpermuted; the solve permutes after subtracting),
//   d = 3    : packed explicit inverses: strict lower = L_II^-1 (unit), upper = U_II^-1,
//   d = 4..6 : U_{I,K} = L_II^-1 P_I A_IK.
// Pivoting is partial pivoting inside the diagonal block of each block row.
// The off-diagonal work (two triangular multiplies by explicit inverses and
// the 3m x m x 3m Schur update) runs on v_mfma_f64_16x16x4f64.
#pragma once
#include "device_common.hpp"

namespace fpe {

struct LuLds {
    double Ar[MAX_BLOCK * (MAX_BLOCK + 1)];
    double Ai[MAX_BLOCK * (MAX_BLOCK + 1)];
    double Xr[MAX_BLOCK * (MAX_BLOCK + 1)];
    double Xi[MAX_BLOCK * (MAX_BLOCK + 1)];
    double lr[MAX_BLOCK], li[MAX_BLOCK];
    int perm[MAX_BLOCK];
    int pivk, singular;
};

// C (op)= A B over an M x N x K product with element accessors; each wave
// computes 16x16 output tiles with v_mfma_f64_16x16x4f64.
// Lane layout of v_mfma_f64_16x16x4f64 (checked on gfx950 with an in-bounds
// probe): a = A[row0 + lane%16][k0 + lane/16], b = B[k0 + lane/16][col0 + lane%16],
// acc[j] = C[row0 + lane/16 + 4*j][col0 + lane%16]. (The f32 16x16x4 form places
// acc[j] at row 4*(lane/16) + j instead.)
template <bool SUB, class FA, class FB, class FC>
__device__ __forceinline__ void mfma_gemm(int M, int N, int K, FA A, FB B, FC C) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tm = (M + 15) >> 4, tn = (N + 15) >> 4;
    const int l16 = lane & 15, l4 = lane >> 4;
    for (int t = wave; t < tm * tn; t += NWAVE) {
        const int r0 = (t / tn) << 4, c0 = (t % tn) << 4;
        const int ra = r0 + l16, cb = c0 + l16;
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int k0 = 0; k0 < K; k0 += 4) {
            const int k = k0 + l4;
            const double a = (ra < M && k < K) ? A(ra, k) : 0.0;
            const double b = (cb < N && k < K) ? B(k, cb) : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
        if (cb < N) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = r0 + l4 + 4 * j;
                if (r < M) {
                    if (SUB) C(r, cb) -= acc[j];
                    else C(r, cb) = acc[j];
                }
            }
        }
    }
}

// Complex version with planar operands: (Cr + i Ci) (op)= (Ar + i Ai)(Br + i Bi).
template <bool SUB, class FAr, class FAi, class FBr, class FBi, class FCr, class FCi>
__device__ __forceinline__ void mfma_gemm_c(int M, int N, int K, FAr Ar, FAi Ai, FBr Br, FBi Bi, FCr Cr,
                                            FCi Ci) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tm = (M + 15) >> 4, tn = (N + 15) >> 4;
    const int l16 = lane & 15, l4 = lane >> 4;
    for (int t = wave; t < tm * tn; t += NWAVE) {
        const int r0 = (t / tn) << 4, c0 = (t % tn) << 4;
        const int ra = r0 + l16, cb = c0 + l16;
        d4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
        for (int k0 = 0; k0 < K; k0 += 4) {
            const int k = k0 + l4;
            const bool ok_a = ra < M && k < K, ok_b = cb < N && k < K;
            const double ar = ok_a ? Ar(ra, k) : 0.0, ai = ok_a ? Ai(ra, k) : 0.0;
            const double br = ok_b ? Br(k, cb) : 0.0, bi = ok_b ? Bi(k, cb) : 0.0;
            accr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, accr, 0, 0, 0);
            accr = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, accr, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acci, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acci, 0, 0, 0);
        }
        if (cb < N) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = r0 + l4 + 4 * j;
                if (r < M) {
                    if (SUB) {
                        Cr(r, cb) -= accr[j];
                        Ci(r, cb) -= acci[j];
                    } else {
                        Cr(r, cb) = accr[j];
                        Ci(r, cb) = acci[j];
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ double cmag(double re, double im) { return fabs(re) + fabs(im); }

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

// Factor the panels in place. Returns nonzero if a zero pivot was met.
template <bool CPLX>
__device__ int block_lu(double* __restrict__ Pr, double* __restrict__ Pi, int* __restrict__ piv, int ns, int m,
                        double* __restrict__ Tr, double* __restrict__ Ti, LuLds& L) {
    const int ld = 7 * m;
    const long long pstride = (long long)m * ld;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int LS = MAX_BLOCK + 1;
    int bad = 0;
    if (tid == 0) L.singular = 0;
    for (int I = 0; I < ns; ++I) {
        double* PIr = Pr + I * pstride;
        double* PIi = Pi + I * pstride;
        for (int e = tid; e < m * m; e += NT) {
            const int i = e / m, j = e - i * m;
            L.Ar[i * LS + j] = PIr[(long long)i * ld + 3 * m + j];
            if (CPLX) L.Ai[i * LS + j] = PIi[(long long)i * ld + 3 * m + j];
        }
        for (int i = tid; i < m; i += NT) L.perm[i] = i;
        __syncthreads();
        // ---- diagonal block: LU with partial pivoting in LDS
        for (int k = 0; k < m; ++k) {
            if (wave == 0) {
                double v = -1.0;
                int idx = lane;
                if (lane >= k && lane < m)
                    v = CPLX ? cmag(L.Ar[lane * LS + k], L.Ai[lane * LS + k]) : fabs(L.Ar[lane * LS + k]);
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    const double ov = __shfl_xor(v, o, 64);
                    const int oi = __shfl_xor(idx, o, 64);
                    if (ov > v || (ov == v && oi < idx)) v = ov, idx = oi;
                }
                if (lane == 0) {
                    L.pivk = idx;
                    if (!(v > 0.0)) L.singular = 1;
                }
            }
            __syncthreads();
            const int p = L.pivk;
            if (p != k) {
                for (int j = tid; j < m; j += NT) {
                    double t = L.Ar[k * LS + j];
                    L.Ar[k * LS + j] = L.Ar[p * LS + j];
                    L.Ar[p * LS + j] = t;
                    if (CPLX) {
                        t = L.Ai[k * LS + j];
                        L.Ai[k * LS + j] = L.Ai[p * LS + j];
                        L.Ai[p * LS + j] = t;
                    }
                }
                if (tid == 0) {
                    const int t = L.perm[k];
                    L.perm[k] = L.perm[p];
                    L.perm[p] = t;
                }
            }
            __syncthreads();
            for (int i = k + 1 + tid; i < m; i += NT) {
                if (CPLX) {
                    const double dr = L.Ar[k * LS + k], di = L.Ai[k * LS + k];
                    const double nr = L.Ar[i * LS + k], ni = L.Ai[i * LS + k];
                    const double den = dr * dr + di * di;
                    const double inv = den > 0.0 ? 1.0 / den : 0.0;
                    L.lr[i] = (nr * dr + ni * di) * inv;
                    L.li[i] = (ni * dr - nr * di) * inv;
                } else {
                    const double dk = L.Ar[k * LS + k];
                    L.lr[i] = dk != 0.0 ? L.Ar[i * LS + k] / dk : 0.0;
                }
            }
            __syncthreads();
            const int w = m - k;
            for (int e = tid; e < (m - k - 1) * w; e += NT) {
                const int i = k + 1 + e / w, j = k + e % w;
                if (j == k) {
                    L.Ar[i * LS + k] = L.lr[i];
                    if (CPLX) L.Ai[i * LS + k] = L.li[i];
                } else {
                    if (CPLX) {
                        const double ur = L.Ar[k * LS + j], ui = L.Ai[k * LS + j];
                        L.Ar[i * LS + j] -= L.lr[i] * ur - L.li[i] * ui;
                        L.Ai[i * LS + j] -= L.lr[i] * ui + L.li[i] * ur;
                    } else {
                        L.Ar[i * LS + j] -= L.lr[i] * L.Ar[k * LS + j];
                    }
                }
            }
            __syncthreads();
        }
        // ---- explicit inverses: threads 0..m-1 -> U^-1 columns, 64..64+m-1 -> L^-1 columns
        if (tid < m) {
            const int c = tid;
            if (CPLX) {
                const double dr = L.Ar[c * LS + c], di = L.Ai[c * LS + c];
                const double den = dr * dr + di * di;
                double xr = den > 0.0 ? dr / den : 0.0, xi = den > 0.0 ? -di / den : 0.0;
                L.Xr[c * LS + c] = xr, L.Xi[c * LS + c] = xi;
                for (int i = c - 1; i >= 0; --i) {
                    double sr = 0.0, si = 0.0;
                    for (int k = i + 1; k <= c; ++k) {
                        const double ur = L.Ar[i * LS + k], ui = L.Ai[i * LS + k];
                        const double vr = L.Xr[k * LS + c], vi = L.Xi[k * LS + c];
                        sr += ur * vr - ui * vi;
                        si += ur * vi + ui * vr;
                    }
                    const double ur = L.Ar[i * LS + i], ui = L.Ai[i * LS + i];
                    const double dd = ur * ur + ui * ui;
                    // -(s / u)
                    L.Xr[i * LS + c] = dd > 0.0 ? -(sr * ur + si * ui) / dd : 0.0;
                    L.Xi[i * LS + c] = dd > 0.0 ? -(si * ur - sr * ui) / dd : 0.0;
                }
            } else {
                const double dcc = L.Ar[c * LS + c];
                L.Xr[c * LS + c] = dcc != 0.0 ? 1.0 / dcc : 0.0;
                for (int i = c - 1; i >= 0; --i) {
                    double s = 0.0;
                    for (int k = i + 1; k <= c; ++k) s += L.Ar[i * LS + k] * L.Xr[k * LS + c];
                    const double uii = L.Ar[i * LS + i];
                    L.Xr[i * LS + c] = uii != 0.0 ? -s / uii : 0.0;
                }
            }
        } else if (tid >= 64 && tid < 64 + m) {
            const int c = tid - 64;
            for (int i = c + 1; i < m; ++i) {
                double sr = CPLX ? 0.0 : 0.0, si = 0.0;
                // Y[i][c] = -(L[i][c] + sum_{k=c+1}^{i-1} L[i][k] Y[k][c])
                sr = L.Ar[i * LS + c];
                if (CPLX) si = L.Ai[i * LS + c];
                for (int k = c + 1; k < i; ++k) {
                    if (CPLX) {
                        const double lr_ = L.Ar[i * LS + k], li_ = L.Ai[i * LS + k];
                        const double yr = L.Xr[k * LS + c], yi = L.Xi[k * LS + c];
                        sr += lr_ * yr - li_ * yi;
                        si += lr_ * yi + li_ * yr;
                    } else {
                        sr += L.Ar[i * LS + k] * L.Xr[k * LS + c];
                    }
                }
                L.Xr[i * LS + c] = -sr;
                if (CPLX) L.Xi[i * LS + c] = -si;
            }
        }
        __syncthreads();
        for (int e = tid; e < m * m; e += NT) {
            const int i = e / m, j = e - i * m;
            PIr[(long long)i * ld + 3 * m + j] = L.Xr[i * LS + j];
            if (CPLX) PIi[(long long)i * ld + 3 * m + j] = L.Xi[i * LS + j];
        }
        for (int i = tid; i < m; i += NT) piv[(long long)I * m + i] = L.perm[i];
        const int nU = (ns - 1 - I) < 3 ? (ns - 1 - I) : 3;
        if (nU > 0) {
            // gather P_I A_{I,I+1..I+3} into T (m x 3m) and the L-candidates into T + 3m^2 (3m x m)
            double* TUr = Tr;
            double* TUi = Ti;
            double* TLr = Tr + 3 * m * m;
            double* TLi = Ti + 3 * m * m;
            const int w3 = nU * m;
            for (int e = tid; e < m * w3; e += NT) {
                const int i = e / w3, c = e - i * w3;
                TUr[i * 3 * m + c] = PIr[(long long)L.perm[i] * ld + 4 * m + c];
                if (CPLX) TUi[i * 3 * m + c] = PIi[(long long)L.perm[i] * ld + 4 * m + c];
            }
            for (int e = tid; e < w3 * m; e += NT) {
                const int rr = e / m, k = e - rr * m;
                const int dJ = 1 + rr / m, i = rr % m;
                const long long src = (I + dJ) * pstride + (long long)i * ld + (3 - dJ) * m + k;
                TLr[rr * m + k] = Pr[src];
                if (CPLX) TLi[rr * m + k] = Pi[src];
            }
            __syncthreads();
            // U_IK = L_II^-1 (P_I A_IK)      (unit lower from the packed block)
            auto Lr_ = [=] __device__(int i, int k) -> double {
                return k < i ? PIr[(long long)i * ld + 3 * m + k] : (k == i ? 1.0 : 0.0);
            };
            auto Li_ = [=] __device__(int i, int k) -> double {
                return k < i ? PIi[(long long)i * ld + 3 * m + k] : 0.0;
            };
            auto TUr_ = [=] __device__(int k, int c) -> double { return TUr[k * 3 * m + c]; };
            auto TUi_ = [=] __device__(int k, int c) -> double { return TUi[k * 3 * m + c]; };
            auto CUr = [=] __device__(int i, int c) -> double& { return PIr[(long long)i * ld + 4 * m + c]; };
            auto CUi = [=] __device__(int i, int c) -> double& { return PIi[(long long)i * ld + 4 * m + c]; };
            // L_JI = A_JI U_II^-1            (upper from the packed block)
            auto TLr_ = [=] __device__(int rr, int k) -> double { return TLr[rr * m + k]; };
            auto TLi_ = [=] __device__(int rr, int k) -> double { return TLi[rr * m + k]; };
            auto Ur_ = [=] __device__(int k, int j) -> double {
                return k <= j ? PIr[(long long)k * ld + 3 * m + j] : 0.0;
            };
            auto Ui_ = [=] __device__(int k, int j) -> double {
                return k <= j ? PIi[(long long)k * ld + 3 * m + j] : 0.0;
            };
            auto CLr = [=] __device__(int rr, int j) -> double& {
                const int dJ = 1 + rr / m, i = rr % m;
                return Pr[(I + dJ) * pstride + (long long)i * ld + (3 - dJ) * m + j];
            };
            auto CLi = [=] __device__(int rr, int j) -> double& {
                const int dJ = 1 + rr / m, i = rr % m;
                return Pi[(I + dJ) * pstride + (long long)i * ld + (3 - dJ) * m + j];
            };
            if (CPLX) {
                mfma_gemm_c<false>(m, w3, m, Lr_, Li_, TUr_, TUi_, CUr, CUi);
                mfma_gemm_c<false>(w3, m, m, TLr_, TLi_, Ur_, Ui_, CLr, CLi);
            } else {
                mfma_gemm<false>(m, w3, m, Lr_, TUr_, CUr);
                mfma_gemm<false>(w3, m, m, TLr_, Ur_, CLr);
            }
            __syncthreads();
            // Schur: A_JK -= L_JI U_IK for J, K in I+1..I+nU
            auto SAr = [=] __device__(int rr, int k) -> double {
                const int dJ = 1 + rr / m, i = rr % m;
                return Pr[(I + dJ) * pstride + (long long)i * ld + (3 - dJ) * m + k];
            };
            auto SAi = [=] __device__(int rr, int k) -> double {
                const int dJ = 1 + rr / m, i = rr % m;
                return Pi[(I + dJ) * pstride + (long long)i * ld + (3 - dJ) * m + k];
            };
            auto SBr = [=] __device__(int k, int c) -> double { return PIr[(long long)k * ld + 4 * m + c]; };
            auto SBi = [=] __device__(int k, int c) -> double { return PIi[(long long)k * ld + 4 * m + c]; };
            auto SCr = [=] __device__(int rr, int c) -> double& {
                const int dJ = 1 + rr / m, i = rr % m, dK = 1 + c / m, j = c % m;
                return Pr[(I + dJ) * pstride + (long long)i * ld + (dK - dJ + 3) * m + j];
            };
            auto SCi = [=] __device__(int rr, int c) -> double& {
                const int dJ = 1 + rr / m, i = rr % m, dK = 1 + c / m, j = c % m;
                return Pi[(I + dJ) * pstride + (long long)i * ld + (dK - dJ + 3) * m + j];
            };
            if (CPLX) mfma_gemm_c<true>(w3, w3, m, SAr, SAi, SBr, SBi, SCr, SCi);
            else mfma_gemm<true>(w3, w3, m, SAr, SBr, SCr);
        }
        __syncthreads();
        bad |= L.singular;
    }
    return bad;
}

// Solve E x = b in place with the block factors. Real: NR right-hand sides
// x[0..NR-1]; complex: one right-hand side (xr[0], xi[0]).
// Threads: row i = tid / 4 of the current block, part q = tid % 4.
template <bool CPLX, int NR>
__device__ void block_solve(const double* __restrict__ Pr, const double* __restrict__ Pi,
                            const int* __restrict__ piv, int ns, int m, double* const* xr, double* const* xi,
                            double* sT /* LDS [2*NR*64] */, int* sP /* LDS [64] */) {
    const int ld = 7 * m;
    const long long pstride = (long long)m * ld;
    const int tid = threadIdx.x, i = tid >> 2, q = tid & 3;
    const bool row = i < m;
    double* sTr = sT;
    double* sTi = sT + NR * MAX_BLOCK;
    // forward: y_I = L_II^-1 P_I (b_I - sum_J L_IJ y_J)
    for (int I = 0; I < ns; ++I) {
        const double* PIr = Pr + I * pstride;
        const double* PIi = Pi + I * pstride;
        double ar[NR], ai[NR];
#pragma unroll
        for (int v = 0; v < NR; ++v) ar[v] = 0.0, ai[v] = 0.0;
        if (row) {
            for (int d = 0; d < 3; ++d) {
                const int J = I + d - 3;
                if (J < 0) continue;
                for (int k = q; k < m; k += 4) {
                    const double lr = PIr[(long long)i * ld + d * m + k];
                    if (CPLX) {
                        const double li = PIi[(long long)i * ld + d * m + k];
                        const double yr = xr[0][J * m + k], yi = xi[0][J * m + k];
                        ar[0] += lr * yr - li * yi;
                        ai[0] += lr * yi + li * yr;
                    } else {
#pragma unroll
                        for (int v = 0; v < NR; ++v) ar[v] += lr * xr[v][J * m + k];
                    }
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
            sP[i] = piv[(long long)I * m + i];
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < NR; ++v) ar[v] = 0.0, ai[v] = 0.0;
        if (row) {
            for (int k = q; k < i; k += 4) {
                const int pk = sP[k];
                const double lr = PIr[(long long)i * ld + 3 * m + k];
                if (CPLX) {
                    const double li = PIi[(long long)i * ld + 3 * m + k];
                    const double tr = sTr[pk], ti = sTi[pk];
                    ar[0] += lr * tr - li * ti;
                    ai[0] += lr * ti + li * tr;
                } else {
#pragma unroll
                    for (int v = 0; v < NR; ++v) ar[v] += lr * sTr[v * MAX_BLOCK + pk];
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
            const int pi_ = sP[i];
#pragma unroll
            for (int v = 0; v < NR; ++v) {
                xr[v][I * m + i] = ar[v] + sTr[v * MAX_BLOCK + pi_];
                if (CPLX) xi[v][I * m + i] = ai[v] + sTi[v * MAX_BLOCK + pi_];
            }
        }
        __syncthreads();
    }
    // backward: x_I = U_II^-1 (y_I - sum_K U_IK x_K)
    for (int I = ns - 1; I >= 0; --I) {
        const double* PIr = Pr + I * pstride;
        const double* PIi = Pi + I * pstride;
        double ar[NR], ai[NR];
#pragma unroll
        for (int v = 0; v < NR; ++v) ar[v] = 0.0, ai[v] = 0.0;
        if (row) {
            for (int d = 4; d < 7; ++d) {
                const int K = I + d - 3;
                if (K >= ns) continue;
                for (int k = q; k < m; k += 4) {
                    const double ur = PIr[(long long)i * ld + d * m + k];
                    if (CPLX) {
                        const double ui = PIi[(long long)i * ld + d * m + k];
                        const double yr = xr[0][K * m + k], yi = xi[0][K * m + k];
                        ar[0] += ur * yr - ui * yi;
                        ai[0] += ur * yi + ui * yr;
                    } else {
#pragma unroll
                        for (int v = 0; v < NR; ++v) ar[v] += ur * xr[v][K * m + k];
                    }
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
            for (int k = i + q; k < m; k += 4) {
                const double ur = PIr[(long long)i * ld + 3 * m + k];
                if (CPLX) {
                    const double ui = PIi[(long long)i * ld + 3 * m + k];
                    const double tr = sTr[k], ti = sTi[k];
                    ar[0] += ur * tr - ui * ti;
                    ai[0] += ur * ti + ui * tr;
                } else {
#pragma unroll
                    for (int v = 0; v < NR; ++v) ar[v] += ur * sTr[v * MAX_BLOCK + k];
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
