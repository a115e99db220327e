// Block-banded factorisation of the Radau stage matrices E = a M + h K (a real
// or complex) and the matching block solves.
//
// E is block-banded with m x m blocks (m = n_vb) and block half-bandwidth 3.
// Block row I is stored as an m x 7m row-major panel holding (I, I-3..I+3).
// The factorisation is block Gaussian elimination E = L U with unit block-lower
// L and block-upper U, where the diagonal blocks are inverted explicitly:
//   for I = 0..n_sb-1:
//     Ainv_I = inverse of the current diagonal block        (Gauss-Jordan, partial pivoting)
//     L_JI   = A_JI Ainv_I               for J = I+1..I+3    (MFMA)
//     A_JK  -= L_JI A_IK                 for J, K = I+1..I+3 (MFMA)
// After it, panel I holds L_IJ (d = 0..2), Ainv_I (d = 3) and U_IK = A_IK
// (d = 4..6). Solves: y_I = b_I - sum_J L_IJ y_J; x_I = Ainv_I (y_I - sum_K U_IK x_K).
// Pivoting is partial pivoting inside each diagonal block (the Gauss-Jordan
// inverse); E = a M + h K with D-scaled mass and stiffness is block diagonally
// dominant for the step sizes Radau takes, which the parity tests exercise.
#pragma once
#include <type_traits>

#include "device_common.hpp"

namespace fpe {

// Radau launch shape: 4-wave workgroups, two per CU (about 74 KB of LDS
// each), so two independent solves share every SIMD and one's memory and
// barrier waits overlap the other's work.
constexpr int RNT = 256;
constexpr int RNW = RNT / 64;
constexpr int RADAU_WG_PER_CU = 2;
constexpr int LS = MAX_BLOCK + 1;  // LDS row stride of an m x m block (column MAX_BLOCK is padding)
constexpr int GJ_PW = 4;           // pivot steps per Gauss-Jordan panel (a multiple of 4)
constexpr int GJ_US = 65;          // LDS stride of a panel column (odd: the MFMA operand reads of
                                   // four lane groups fall in different banks)

struct LuLds {
    // two m x m staging blocks: the pivot block and its inverse (re, im), then
    // the U blocks of the Schur update; idle during solves (solution ring)
    double B[2][MAX_BLOCK * LS];
    int piv[MAX_BLOCK];   // pivot row of each elimination step
    int qinv[MAX_BLOCK];  // step whose pivot row is a given row
    // Gauss-Jordan panel of GJ_PW pivot steps: column s of the panel right after
    // step s (re, im; at [s * GJ_US + row]), 1 / pivot, pivot rows, rows used so far
    double U[2][GJ_PW * GJ_US];
    // U block (I, I+3) as its 7 diagonals (re, im; column j at [7 j + k - j + 3],
    // zeros outside the block), staged for the last Schur pass of block row I
    double UB[2][7 * MAX_BLOCK];
    double ip[2][GJ_PW];
    int pv[GJ_PW];
    unsigned long long umask;
    int singular;
    // wall-clock ticks by sub-phase (thread 0): pivot-block inverse, staging, L multiply, Schur update
    unsigned long long tk[4];
};

// k-step count the MFMA loops of the factorisation are compiled for: the
// smallest of 5, 8, 10, 15, 16 that covers ceil(m / 4) (column tiles:
// ceil(TK / 4) of 16). Operand rows m .. 4 TK - 1 of every staged LDS block are
// kept zero (stage_in, gj_inverse), so padded k-steps add exact zeros and the
// LDS operand reads need no predicate (rows < 4 TK <= 64, columns < 64: inside
// the 64 x 65 block).
__host__ __device__ constexpr int tk_class(int m) {
    return m <= 20 ? 5 : m <= 32 ? 8 : m <= 40 ? 10 : m <= 60 ? 15 : 16;
}

// Lane layout of v_mfma_f64_16x16x4f64 (checked on gfx950 with an in-bounds
// probe): a = A[row0 + lane%16][k0 + lane/16], b = B[k0 + lane/16][col0 + lane%16],
// acc[j] = C[row0 + lane/16 + 4*j][col0 + lane%16]. (The f32 16x16x4 form places
// acc[j] at row 4*(lane/16) + j instead.)
typedef float f4 __attribute__((ext_vector_type(4)));

// Scalar-type traits of the factorisation: FP64 uses v_mfma_f64_16x16x4f64,
// the mixed-precision path FP32 panels with v_mfma_f32_16x16x4f32.
template <class T> struct Mf;
template <> struct Mf<double> {
    typedef d4 acc;
    static __device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int j) { return (lane >> 4) + 4 * j; }
};
template <> struct Mf<float> {
    typedef f4 acc;
    static __device__ __forceinline__ f4 mfma(float a, float b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int j) { return 4 * (lane >> 4) + j; }
};

template <class T, class A>
__device__ __forceinline__ void mfma_c(A& accr, A& acci, T ar, T ai, T br, T bi, bool cplx) {
    accr = Mf<T>::mfma(ar, br, accr);
    if (cplx) {
        accr = Mf<T>::mfma(-ai, bi, accr);
        acci = Mf<T>::mfma(ar, bi, acci);
        acci = Mf<T>::mfma(ai, br, acci);
    }
}

__device__ __forceinline__ double readlane_d(double x, int l) {
    const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane(int(unsigned(v)), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane(int(unsigned(v >> 32)), l);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

__device__ __forceinline__ float readlane_d(float x, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}

// LDS staging block k viewed as elements of T (row stride LS).
template <class T>
__device__ __forceinline__ T* lds_block(LuLds& L, int k) { return reinterpret_cast<T*>(L.B[k]); }
template <class T>
__device__ __forceinline__ const T* lds_block(const LuLds& L, int k) { return reinterpret_cast<const T*>(L.B[k]); }

// Pivot search key of a row: |value| rounded to FP32 with its 7 lowest
// significand bits replaced by a valid flag (64) and 63 - lane, so one unsigned
// 32-bit max yields the pivot row (ties within 2^-16 relative go to the lower
// row; non-finite magnitudes order above finite ones). Used rows: key 0.
__device__ __forceinline__ unsigned pivot_key32(double mag, int lane) {
    const unsigned b = __builtin_bit_cast(unsigned, float(mag));
    return (b & ~127u) | 64u | unsigned(63 - lane);
}
__device__ __forceinline__ unsigned pivot_key32(float mag, int lane) {
    const unsigned b = __builtin_bit_cast(unsigned, mag);
    return (b & ~127u) | 64u | unsigned(63 - lane);
}
// Wave-uniform maximum of an unsigned key: DPP row rotations (8, 4, 2, 1) leave
// each 16-lane row's maximum in all its lanes; the four rows combine in SGPRs.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xF, 0xF, false));
    const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)v, 0), b = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
    const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)v, 32), d = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
    return max(max(a, b), max(c, d));
}
// 1 / d for a nonzero finite pivot: hardware reciprocal and two Newton steps
// (within an ulp or so of the correctly rounded quotient).
__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}
__device__ __forceinline__ float recip(float d) {
    float r = __builtin_amdgcn_rcpf(d);
    return fmaf(r, fmaf(-d, r, 1.0f), r);
}

// Steps k0..k0+nb-1 of the blocked Gauss-Jordan inverse on this wave's 16
// columns c = 16 w + lane % 16. Pivot rows follow the step-by-step recurrence
// (the four lanes of a column each own two of its pivot rows):
//     y_s = x_{p_s}, x_{p_s} <- y_s / piv_s, x_{p_t} <- x_{p_t} + U[p_t][s] y_s (t != s),
// where U[.][s] is panel column s right after step s (-f / piv_s off the pivot
// row); every other row takes x_i <- x_i + sum_s U[i][s] y_s as a rank-nb MFMA
// update per 16-row tile, and the pivot rows are then overwritten. All LDS
// accesses stay inside the 64 x 65 staging block: rows and columns past m are
// padding whose contents never reach rows and columns below m.
template <bool CPLX, class T>
__device__ __forceinline__ void apply_panel(LuLds& L, int m, int k0, int nb) {
    typedef typename Mf<T>::acc A4;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int c = (w << 4) + l16;
    if ((w << 4) >= m) return;  // wave-uniform: no columns here
    T* B0 = lds_block<T>(L, 0);
    T* B1 = lds_block<T>(L, 1);
    const T* Ur = reinterpret_cast<const T*>(L.U[0]);
    const T* Ui = reinterpret_cast<const T*>(L.U[1]);
    // lane group g = lane / 16 owns pivot rows t = g + 4 h (h < GJ_PW / 4) of its
    // column's recurrence; y_s comes from the owner group of s by a lane shuffle
    constexpr int H = GJ_PW / 4;
    int po[H];
    T zr[H], zi[H], br[H], bi[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int t = 4 * h + l4;
        int pt = 0;
#pragma unroll
        for (int g = 0; g < 4; ++g)
            if (g == l4 && 4 * h + g < nb) pt = L.pv[4 * h + g];
        po[h] = t < nb ? pt : 0;
        zr[h] = B0[po[h] * LS + c];
        zi[h] = CPLX ? B1[po[h] * LS + c] : T(0);
        br[h] = bi[h] = T(0);
    }
#pragma unroll
    for (int s = 0; s < GJ_PW; ++s) {
        if (s < nb) {
            const int og = s & 3, oh = s >> 2;  // owner lane group and slot of row p_s
            const int src = l16 + 16 * og;
            const T yr = __shfl(zr[oh], src, 64);
            const T yi = CPLX ? __shfl(zi[oh], src, 64) : T(0);
            if (l4 == og) {  // owner: x_{p_s} <- y_s / piv_s; y_s is also its MFMA B operand
                const T rr = T(L.ip[0][s]), ri = CPLX ? T(L.ip[1][s]) : T(0);
                br[oh] = yr, bi[oh] = yi;
                zr[oh] = CPLX ? yr * rr - yi * ri : yr * rr;
                if (CPLX) zi[oh] = yr * ri + yi * rr;
            }
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int t = 4 * h + l4;
                if (t != s && t < nb) {
                    const T ur = Ur[s * GJ_US + po[h]], ui = CPLX ? Ui[s * GJ_US + po[h]] : T(0);
                    zr[h] += CPLX ? ur * yr - ui * yi : ur * yr;
                    if (CPLX) zi[h] += ur * yi + ui * yr;
                }
            }
        }
    }
    const int tm = (m + 15) >> 4;
    for (int tt = 0; tt < tm; ++tt) {
        const int ra = (tt << 4) + l16;
        A4 accr, acci;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = (tt << 4) + Mf<T>::row(lane, j);
            accr[j] = B0[r * LS + c];
            acci[j] = CPLX ? B1[r * LS + c] : T(0);
        }
#pragma unroll
        for (int kk = 0; kk < GJ_PW / 4; ++kk) {
            if (4 * kk < nb) {
                const T ar = Ur[(4 * kk + l4) * GJ_US + ra], ai = CPLX ? Ui[(4 * kk + l4) * GJ_US + ra] : T(0);
                mfma_c(accr, acci, ar, ai, br[kk], bi[kk], CPLX);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = (tt << 4) + Mf<T>::row(lane, j);
            B0[r * LS + c] = accr[j];
            if (CPLX) B1[r * LS + c] = acci[j];
        }
    }
    // pivot rows last (lane group l4 writes its rows t = l4 + 4 h; t >= nb to padding)
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const int o = (4 * h + l4 < nb) ? po[h] * LS + c : MAX_BLOCK;  // row 0, padding column
        B0[o] = zr[h];
        if (CPLX) B1[o] = zi[h];
    }
}

// Gauss-Jordan inverse of the m x m block in (B[0], B[1]), blocked by panels
// (a call of its own: its pivot-step chain keeps its values in registers instead
// of sharing the register budget of the factorisation's unrolled MFMA loops)
// of GJ_PW pivot steps. Partial pivoting is implicit: step k takes the
// not-yet-used row p_k with the largest |A[.][k]| (|re| + |im| in complex
// arithmetic, compared in FP32 to 2^-16 relative, ties to the lower row) and no
// rows move; diagonal-pivot in-place Gauss-Jordan on P A leaves
// X[p_k][j] = inv(A)[k][p_j], which the last pass unscrambles into natural
// order with p_{q_b} = b.
// Per panel (columns k0..k0+GJ_PW-1):
//  1. the wave owning the panel's 16-column tile runs the steps on the panel
//     columns alone, in registers (lane = row), and publishes each column right
//     after its step, 1 / pivot and the pivot rows;
//  2. every wave applies the steps to its own 16 columns (apply_panel);
//  3. the panel wave stores the panel columns last.
// A wave only reads and writes its own columns in 2-3, so two barriers per
// panel order everything. (One barrier with the panel data double-buffered
// measured the same: the panel wave's own steps and apply are the chain.)
// Sets singular on a zero or non-finite pivot. Ends with a barrier.
template <bool CPLX, class T>
__device__ __noinline__ void gj_inverse(LuLds& L, int m) {
    m = uniform(m);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool lrow = lane < m;
    T* B0 = lds_block<T>(L, 0);
    T* B1 = lds_block<T>(L, 1);
    T* Ur = reinterpret_cast<T*>(L.U[0]);
    T* Ui = reinterpret_cast<T*>(L.U[1]);
    if (threadIdx.x < MAX_BLOCK) L.qinv[threadIdx.x] = 0;  // stays in range if a bad pivot search repeats a row
    if (threadIdx.x == 0) L.umask = 0;
    for (int k0 = 0; k0 < m; k0 += GJ_PW) {
        const int nb = m - k0 < GJ_PW ? m - k0 : GJ_PW;
        const int wp = k0 >> 4;
        T cr[GJ_PW], ci[GJ_PW];  // panel columns (panel wave, lane = row)
        __syncthreads();
        if (w == wp) {
            T ir[GJ_PW], ii[GJ_PW];
            int pv[GJ_PW];
#pragma unroll
            for (int s = 0; s < GJ_PW; ++s) {
                const int col = s < nb ? k0 + s : MAX_BLOCK;  // padding column past the block
                cr[s] = lrow ? B0[lane * LS + col] : T(0);
                ci[s] = (CPLX && lrow) ? B1[lane * LS + col] : T(0);
                ir[s] = ii[s] = T(0);
                pv[s] = 0;
            }
            const unsigned long long um = L.umask;
            bool used = !lrow || ((um >> lane) & 1ull);
            bool bad = false;
#pragma unroll
            for (int s = 0; s < GJ_PW; ++s) {
                if (s < nb) {
                    const T mag = CPLX ? fabs(cr[s]) + fabs(ci[s]) : fabs(cr[s]);
                    const unsigned key = wave_max_u32(used ? 0u : pivot_key32(mag, lane));
                    const int p = 63 - int(key & 63u);
                    const T dr = readlane_d(cr[s], p), di = CPLX ? readlane_d(ci[s], p) : T(0);
                    const T mp = CPLX ? fabs(dr) + fabs(di) : fabs(dr);
                    const bool ok = key != 0u && mp > T(0) && mp < T(INFINITY);
                    bad = bad || !ok;
                    T rr, ri;
                    if (CPLX) {
                        const T rd = recip(dr * dr + di * di);
                        rr = ok ? dr * rd : T(0);
                        ri = ok ? -di * rd : T(0);
                    } else {
                        rr = ok ? recip(dr) : T(0);
                        ri = T(0);
                    }
                    const bool isp = lane == p;
                    if (isp) used = true;
                    const T gr = cr[s], gi = ci[s];
                    ir[s] = rr, ii[s] = ri, pv[s] = p;
#pragma unroll
                    for (int t = s + 1; t < GJ_PW; ++t) {
                        if (t < nb) {
                            const T pr = readlane_d(cr[t], p);
                            const T pi_ = CPLX ? readlane_d(ci[t], p) : T(0);
                            const T sr = CPLX ? pr * rr - pi_ * ri : pr * rr;
                            const T si = CPLX ? pr * ri + pi_ * rr : T(0);
                            const T nr = cr[t] - (CPLX ? gr * sr - gi * si : gr * sr);
                            const T ni = CPLX ? ci[t] - (gr * si + gi * sr) : T(0);
                            cr[t] = isp ? sr : nr;
                            if (CPLX) ci[t] = isp ? si : ni;
                        }
                    }
#pragma unroll
                    for (int t = 0; t < s; ++t) {
                        const T pr = readlane_d(cr[t], p);
                        const T pi_ = CPLX ? readlane_d(ci[t], p) : T(0);
                        const T sr = CPLX ? pr * rr - pi_ * ri : pr * rr;
                        const T si = CPLX ? pr * ri + pi_ * rr : T(0);
                        const T nr = cr[t] - (CPLX ? gr * sr - gi * si : gr * sr);
                        const T ni = CPLX ? ci[t] - (gr * si + gi * sr) : T(0);
                        cr[t] = isp ? sr : nr;
                        if (CPLX) ci[t] = isp ? si : ni;
                    }
                    const T nr = -(CPLX ? gr * rr - gi * ri : gr * rr);
                    const T ni = CPLX ? -(gr * ri + gi * rr) : T(0);
                    cr[s] = isp ? rr : nr;
                    if (CPLX) ci[s] = isp ? ri : ni;
                    Ur[s * GJ_US + lane] = cr[s];
                    if (CPLX) Ui[s * GJ_US + lane] = ci[s];
                } else {
                    Ur[s * GJ_US + lane] = T(0);
                    if (CPLX) Ui[s * GJ_US + lane] = T(0);
                }
            }
            if (lane == 0) {
                unsigned long long nm = um;
#pragma unroll
                for (int s = 0; s < GJ_PW; ++s) {
                    if (s < nb) {
                        L.ip[0][s] = ir[s], L.ip[1][s] = ii[s];
                        L.pv[s] = pv[s];
                        L.piv[k0 + s] = pv[s];
                        nm |= 1ull << pv[s];
                    }
                }
                L.umask = nm;
                if (bad) L.singular = 1;
            }
        }
        __syncthreads();
        apply_panel<CPLX, T>(L, m, k0, nb);
        if (w == wp) {
#pragma unroll
            for (int s = 0; s < GJ_PW; ++s) {
                if (s < nb) {
                    B0[lane * LS + k0 + s] = cr[s];  // rows past m: padding
                    if (CPLX) B1[lane * LS + k0 + s] = ci[s];
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < m) L.qinv[L.piv[threadIdx.x]] = threadIdx.x;
    __syncthreads();
    // unscramble in place through registers: gather, barrier, scatter
    constexpr int GE = MAX_BLOCK * MAX_BLOCK / RNT;
    T gr[GE], gi[GE];
#pragma unroll
    for (int g = 0; g < GE; ++g) {
        const int e = threadIdx.x + g * RNT;
        if (e < m * m) {
            const int a = e / m, c = e - a * m;
            const int src = L.piv[a] * LS + L.qinv[c];
            gr[g] = B0[src];
            gi[g] = CPLX ? B1[src] : T(0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < GE; ++g) {
        const int e = threadIdx.x + g * RNT;
        if (e < m * m) {
            const int a = e / m, c = e - a * m;
            B0[a * LS + c] = gr[g];
            if (CPLX) B1[a * LS + c] = gi[g];
        }
    }
    __syncthreads();
}

// Stacked rows below block row I: row rr = (dJ - 1) m + i is row i of block row
// I + dJ. Offset of (row rr, column 0 of block column I + dJ + d - 3) in the panels.
__device__ __forceinline__ long long stacked_row(long long ps, int ld, int I, int m, int rr, int& dJ) {
    dJ = 1 + rr / m;
    const int i = rr - (dJ - 1) * m;
    return (long long)(I + dJ) * ps + (long long)i * ld;
}

// Generic forms (any m, runtime k-step count; the operand reads of padded
// k-steps hit zero rows, column reads stay inside the block).
// L_JI = A_JI Ainv_I for J = I+1..I+nU, in place. A rows come straight from the
// panels (a wave takes 16-row strips; their fragments stay in registers), the
// inverse from LDS (B[0], B[1]).
template <bool CPLX, class T>
__device__ void lower_multiply_gen(GAS T* __restrict__ Pr, GAS T* __restrict__ Pi, long long ps, int ld, int I, int m, int nU,
                               const LuLds& L) {
    typedef typename Mf<T>::acc A4;
    constexpr int TG = CPLX ? 2 : 4;  // column tiles per accumulator group
    const T* B0 = lds_block<T>(L, 0);
    const T* B1 = lds_block<T>(L, 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int M = nU * m, tm = (M + 15) >> 4, tk = (m + 3) >> 2, tn = (m + 15) >> 4;
    for (int t = wave; t < tm; t += RNW) {
        const int r0 = t << 4, ra = r0 + l16;
        const bool va = ra < M;
        int dJa;
        const long long oa = va ? stacked_row(ps, ld, I, m, ra, dJa) + (3 - dJa) * m : 0;
        T ar[MAX_BLOCK / 4], ai[MAX_BLOCK / 4];
#pragma unroll
        for (int s = 0; s < MAX_BLOCK / 4; ++s) {
            const int k = 4 * s + l4;
            const bool ok = s < tk && va && k < m;
            ar[s] = ok ? Pr[oa + k] : T(0);
            ai[s] = (CPLX && ok) ? Pi[oa + k] : T(0);
        }
        long long oc[4];
        bool vc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = r0 + Mf<T>::row(lane, j);
            vc[j] = r < M;
            int dJ;
            oc[j] = vc[j] ? stacked_row(ps, ld, I, m, r, dJ) + (3 - dJ) * m : 0;
        }
        if constexpr (CPLX) {
            // complex products with three real ones per k-step (re = p1 - p2,
            // im = p3 - p1 - p2): three independent chains per column tile
            for (int u = 0; u < tn; ++u) {
                const int c = (u << 4) + l16;
                A4 p1 = A4{0, 0, 0, 0}, p2 = A4{0, 0, 0, 0}, p3 = A4{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < MAX_BLOCK / 4; ++s) {
                    if (s < tk) {
                        const int k = 4 * s + l4;
                        const T br = B0[k * LS + c];
                        const T bi = B1[k * LS + c];
                        p1 = Mf<T>::mfma(ar[s], br, p1);
                        p2 = Mf<T>::mfma(ai[s], bi, p2);
                        p3 = Mf<T>::mfma(ar[s] + ai[s], br + bi, p3);
                    }
                }
                if (c < m) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (vc[j]) {
                            Pr[oc[j] + c] = p1[j] - p2[j];
                            Pi[oc[j] + c] = p3[j] - p1[j] - p2[j];
                        }
                }
            }
            continue;
        }
        // TG column tiles of the strip at a time: independent accumulator chains
        for (int ug = 0; ug < tn; ug += TG) {
            A4 accr[TG], acci[TG];
#pragma unroll
            for (int u = 0; u < TG; ++u) accr[u] = A4{0, 0, 0, 0}, acci[u] = A4{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < MAX_BLOCK / 4; ++s) {
                if (s < tk) {
                    const int k = 4 * s + l4;
#pragma unroll
                    for (int u = 0; u < TG; ++u) {
                        if (ug + u < tn) {
                            const int c = ((ug + u) << 4) + l16;
                            const T br = B0[k * LS + c];
                            const T bi = CPLX ? B1[k * LS + c] : T(0);
                            mfma_c(accr[u], acci[u], ar[s], ai[s], br, bi, CPLX);
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                const int c = ((ug + u) << 4) + l16;
                if (ug + u < tn && c < m) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (vc[j]) {
                            Pr[oc[j] + c] = accr[u][j];
                            if (CPLX) Pi[oc[j] + c] = acci[u][j];
                        }
                }
            }
        }
    }
}

// A_JK -= L_JI U_IK for J = I+1..I+nU and K = I+1+kb0 .. I+kb0+nkb, with U_IK
// staged in LDS (B[0] re, B[1] im in complex arithmetic, one block per pass;
// B[kb] in real).
// L strips come from the panels into registers once per strip; the column
// tiles of a K block run in groups (C tiles loaded up front, one independent
// accumulator chain per tile).
template <bool CPLX, class T>
__device__ void schur_pass_gen(GAS T* __restrict__ Pr, GAS T* __restrict__ Pi, long long ps, int ld, int I, int m, int nU,
                           int kb0, int nkb, const LuLds& L) {
    typedef typename Mf<T>::acc A4;
    constexpr int TG = CPLX ? 2 : 4;  // column tiles per accumulator group
    constexpr bool PF = !CPLX;        // next-strip prefetch (complex: no registers to spare)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int M = nU * m, tm = (M + 15) >> 4, tk = (m + 3) >> 2, tn = (m + 15) >> 4;
    // L strip of 16 rows (t = strip) into registers; the next strip of this wave
    // is fetched while the current one's MFMA chains run
    auto fetch = [&](int t, T* fr, T* fi) {
        const int ra = (t << 4) + l16;
        const bool va = ra < M;
        int dJa;
        const long long oa = va ? stacked_row(ps, ld, I, m, ra, dJa) + (3 - dJa) * m : 0;
#pragma unroll
        for (int s = 0; s < MAX_BLOCK / 4; ++s) {
            const int k = 4 * s + l4;
            const bool ok = s < tk && va && k < m;
            fr[s] = ok ? Pr[oa + k] : T(0);
            fi[s] = (CPLX && ok) ? Pi[oa + k] : T(0);
        }
    };
    T ar[MAX_BLOCK / 4], ai[MAX_BLOCK / 4], nr[MAX_BLOCK / 4], ni[MAX_BLOCK / 4];
#pragma unroll
    for (int s = 0; s < MAX_BLOCK / 4; ++s) ar[s] = ai[s] = nr[s] = ni[s] = T(0);
    if (wave < tm) fetch(wave, ar, ai);
    for (int t = wave; t < tm; t += RNW) {
        const int r0 = t << 4;
        if (PF && t + RNW < tm) fetch(t + RNW, nr, ni);
        long long oc[4];
        int dj[4], fi[4];
        bool vc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = r0 + Mf<T>::row(lane, j);
            vc[j] = r < M;
            dj[j] = 1;
            oc[j] = vc[j] ? stacked_row(ps, ld, I, m, r, dj[j]) : 0;
            fi[j] = r - (dj[j] - 1) * m;  // row within block row I + dj
        }
        for (int kb = 0; kb < nkb; ++kb) {
            const T* Br = lds_block<T>(L, CPLX ? 2 * kb : kb);
            const T* Bi = lds_block<T>(L, CPLX ? 2 * kb + 1 : kb);
            const int kk = kb0 + kb;  // K = I + 1 + kk; column block d = kk + 4 - dJ in panel I + dJ
            // A block's first update (step I = max(J, K) - 3, or 0) finds it as
            // assembled: zero except within 3 of its diagonal, and only those
            // entries are written beforehand (the first pass writes the whole
            // block). Known zeros are not loaded.
            int band[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) band[j] = (I == 0 || dj[j] == 3 || kk == 2) ? fi[j] : -(1 << 20);
            // TG column tiles of the K block at a time: C tiles loaded up front,
            // independent accumulator chains per tile
            for (int ug = 0; ug < tn; ug += TG) {
                T cr[TG][4], ci[TG][4];
                A4 accr[TG], acci[TG];
#pragma unroll
                for (int u = 0; u < TG; ++u) {
                    const int c = ((ug + u) << 4) + l16;
                    accr[u] = A4{0, 0, 0, 0}, acci[u] = A4{0, 0, 0, 0};
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const bool ok = ug + u < tn && vc[j] && c < m &&
                                        (band[j] < 0 || (c - band[j] <= 3 && band[j] - c <= 3));
                        const long long o = oc[j] + (long long)(kk + 4 - dj[j]) * m + c;
                        cr[u][j] = ok ? Pr[o] : T(0);
                        ci[u][j] = (CPLX && ok) ? Pi[o] : T(0);
                    }
                }
#pragma unroll
                for (int s = 0; s < MAX_BLOCK / 4; ++s) {
                    if (s < tk) {
                        const int k = 4 * s + l4;
#pragma unroll
                        for (int u = 0; u < TG; ++u) {
                            if (ug + u < tn) {
                                const int c = ((ug + u) << 4) + l16;
                                const T br = Br[k * LS + c];
                                const T bi = CPLX ? Bi[k * LS + c] : T(0);
                                mfma_c(accr[u], acci[u], ar[s], ai[s], br, bi, CPLX);
                            }
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < TG; ++u) {
                    const int c = ((ug + u) << 4) + l16;
                    if (ug + u < tn && c < m) {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (vc[j]) {
                                const long long o = oc[j] + (long long)(kk + 4 - dj[j]) * m + c;
                                Pr[o] = cr[u][j] - accr[u][j];
                                if (CPLX) Pi[o] = ci[u][j] - acci[u][j];
                            }
                    }
                }
            }
        }
        if (PF) {
#pragma unroll
            for (int s2 = 0; s2 < MAX_BLOCK / 4; ++s2) ar[s2] = nr[s2], ai[s2] = ni[s2];
        } else if (t + RNW < tm) {
            fetch(t + RNW, ar, ai);
        }
    }
}

// Write E = a M_s + h K_s (a = ar + i ai) into the panels from the 49-point
// stencils, for block_lu: a coalesced zero fill of the blocks it reads before
// writing them, then the (at most) 49 stencil entries of every row (stencil
// planes are row-contiguous). Panels are 16-byte aligned.
// PAIR: also the complex E2 = (cr + i ci) M_s + h K_s into (Qr, Qi) from the
// same stencil loads (the real E into Pr, with ar; ai unused).
template <bool CPLX, class T, bool PAIR = false>
__device__ void assemble_E(const double* __restrict__ Ms, const double* __restrict__ Ks, int ns, int m,
                           double ar, double ai, double h, T* __restrict__ Pr, T* __restrict__ Pi,
                           double cr = 0.0, double ci = 0.0, T* __restrict__ Qr = nullptr, T* __restrict__ Qi = nullptr) {
    static_assert(!(PAIR && CPLX), "the pair form writes one real and one complex matrix");
    const int N = ns * m, ld = 7 * m;
    // Zeros go only where block_lu reads entries it has not written: block rows
    // 0..3 whole (row 0 and column 0 are never updated), and the outermost
    // blocks (d = 0 and 6) of the other rows. Every other in-band block is
    // updated by the Schur pass, whose first update of a block loads only the
    // entries within 3 of its diagonal and then writes the whole block.
    const int head = min(4 * m, N) * ld;  // block rows 0..3, contiguous
    constexpr int V = int(16 / sizeof(T));  // elements per 16-byte store
    const double2 z = {0.0, 0.0};
    for (int e = threadIdx.x; e < head / V; e += RNT) {
        reinterpret_cast<double2*>(Pr)[e] = z;
        if (CPLX) reinterpret_cast<double2*>(Pi)[e] = z;
        if (PAIR) {
            reinterpret_cast<double2*>(Qr)[e] = z;
            reinterpret_cast<double2*>(Qi)[e] = z;
        }
    }
    for (int e = head / V * V + threadIdx.x; e < head; e += RNT) {
        Pr[e] = T(0);
        if (CPLX) Pi[e] = T(0);
        if (PAIR) Qr[e] = Qi[e] = T(0);
    }
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        for (int row = 4 * m + w; row < N; row += RNT / 64) {
            const int I = row / m;
            const long long o = (long long)row * ld;
            if (lane < m) {
                Pr[o + lane] = T(0);  // d = 0 (block row I >= 4 > 3)
                if (CPLX) Pi[o + lane] = T(0);
                if (PAIR) Qr[o + lane] = Qi[o + lane] = T(0);
                if (I + 3 < ns) {
                    Pr[o + 6 * m + lane] = T(0);
                    if (CPLX) Pi[o + 6 * m + lane] = T(0);
                    if (PAIR) Qr[o + 6 * m + lane] = Qi[o + 6 * m + lane] = T(0);
                }
            }
        }
    }
    __syncthreads();
    // Stencil entries: a wave takes 8 rows at a time, lane = (row ri, entry q)
    // for 56 of its lanes, so a load reads 8 consecutive rows of 7 planes and a
    // store writes 8 runs of 7 consecutive entries. Two such groups go together:
    // all 28 loads are issued before the stores.
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int ri = lane / 7, q = lane - 7 * (lane / 7);
        constexpr int G = 8 * (RNT / 64);  // rows per pass of all waves
        for (int r0 = 8 * w; r0 < N; r0 += 2 * G) {
            int rc[2];
            bool act[2];
            double mv[2][7], kv[2][7];
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                const int row = r0 + g * G + ri;
                act[g] = lane < 56 && row < N;
                rc[g] = act[g] ? row : 0;  // every plane has an entry per row (zero for missing neighbours)
#pragma unroll
                for (int d = 0; d < 7; ++d) {
                    const long long o = (long long)(d * 7 + q) * N + rc[g];
                    mv[g][d] = Ms[o];
                    kv[g][d] = Ks[o];
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                const int I = rc[g] / m, i = rc[g] - I * m, j = i + q - 3;
                if (act[g] && j >= 0 && j < m) {
                    const long long o = (long long)rc[g] * ld + j;
#pragma unroll
                    for (int d = 0; d < 7; ++d) {
                        const int J = I + d - 3;
                        if (J >= 0 && J < ns) {
                            Pr[o + d * m] = T(ar * mv[g][d] + h * kv[g][d]);
                            if (CPLX) Pi[o + d * m] = T(ai * mv[g][d]);
                            if (PAIR) {
                                Qr[o + d * m] = T(cr * mv[g][d] + h * kv[g][d]);
                                Qi[o + d * m] = T(ci * mv[g][d]);
                            }
                        }
                    }
                }
            }
        }
    }
}

// Copy the m x m block at (sr, si) (row stride ld) into the LDS staging block
// (dr, di) (row stride LS): wave w takes rows w, w + RNW, ... with lane = column,
// so each load instruction reads one contiguous row segment, and the loads of
// half the rows are issued before their stores (two memory latencies per
// block). Padding columns of the staged rows are written with zeros.
// H = 1 issues the loads of all rows before any store (one memory latency,
// twice the registers).
// Rows m .. mz - 1 of the LDS block are written with zeros (the padded operand
// rows of the TK-compiled MFMA loops; mz = 4 tk_class(m)).
template <bool CPLX, class T, int H = 2>
__device__ __forceinline__ void stage_in(T* dr, T* di, GAS const T* sr, GAS const T* si, int ld, int m) {
    constexpr int R = MAX_BLOCK / RNW / H;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int mz = max(m, 4 * tk_class(m));
#pragma unroll
    for (int h = 0; h < H; ++h) {
        T vr[R], vi[R];
#pragma unroll
        for (int u = 0; u < R; ++u) vr[u] = vi[u] = T(0);
        if (lane < m) {
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int r = w + RNW * (u + R * h);
                if (r < m) {
                    vr[u] = sr[r * ld + lane];
                    if (CPLX) vi[u] = si[r * ld + lane];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int r = w + RNW * (u + R * h);
            if (r < mz) {  // rows m..mz-1: vr, vi hold zeros
                dr[r * LS + lane] = vr[u];
                if (CPLX) di[r * LS + lane] = vi[u];
            }
        }
    }
}
// The reverse copy: LDS block (dr, di) to (sr, si), half the rows at a time.
template <bool CPLX, class T>
__device__ __forceinline__ void stage_out(const T* dr, const T* di, GAS T* sr, GAS T* si, int ld, int m) {
    constexpr int R = MAX_BLOCK / RNW / 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        T vr[R], vi[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int r = w + RNW * (u + R * h);
            vr[u] = r < m ? dr[r * LS + lane] : T(0);
            vi[u] = (CPLX && r < m) ? di[r * LS + lane] : T(0);
        }
        if (lane < m) {
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int r = w + RNW * (u + R * h);
                if (r < m) {
                    sr[r * ld + lane] = vr[u];
                    if (CPLX) si[r * ld + lane] = vi[u];
                }
            }
        }
    }
}
// The pivot block's copies in and out of LDS as calls of their own, so that
// their registers do not add to the factorisation loop's (the copy in issues
// every row's loads at once).
template <bool CPLX, class T>
__device__ __noinline__ void pivot_block_in(LuLds& L, const T* sr, const T* si, int ld, int m) {
    m = uniform(m), ld = uniform(ld);
    stage_in<CPLX, T, 1>(lds_block<T>(L, 0), lds_block<T>(L, 1), uniform(gas(sr)), CPLX ? uniform(gas(si)) : nullptr,
                         ld, m);
}
// U block (I, K) of a Schur pass into staging blocks k (re) and k + 1 (im) as a
// call of its own (same reason as the pivot block's copies).
template <bool CPLX, class T>
__device__ __noinline__ void u_block_in(LuLds& L, int k, const T* sr, const T* si, int ld, int m) {
    m = uniform(m), ld = uniform(ld), k = uniform(k);
    stage_in<CPLX, T, 1>(lds_block<T>(L, k), CPLX ? lds_block<T>(L, k + 1) : nullptr, uniform(gas(sr)),
                         CPLX ? uniform(gas(si)) : nullptr, ld, m);
}
// The band of U block (I, I+3) into UB (sr, si: its first column in panel I).
template <bool CPLX, class T>
__device__ __noinline__ void u_band_in(LuLds& L, const T* sr_, const T* si_, int ld, int m) {
    m = uniform(m), ld = uniform(ld);
    GAS const T* const sr = uniform(gas(sr_));
    GAS const T* const si = CPLX ? uniform(gas(si_)) : nullptr;
    T* ur = reinterpret_cast<T*>(L.UB[0]);
    T* ui = reinterpret_cast<T*>(L.UB[1]);
    for (int e = threadIdx.x; e < 7 * MAX_BLOCK; e += RNT) {
        const int j = e / 7, k = j + e - 7 * j - 3;
        const bool ok = j < m && k >= 0 && k < m;
        const long long o = ok ? (long long)k * ld + j : 0;
        const T xr = sr[o];
        const T xi = CPLX ? si[o] : T(0);
        ur[e] = ok ? xr : T(0);
        if (CPLX) ui[e] = ok ? xi : T(0);
    }
}
template <bool CPLX, class T>
__device__ __noinline__ void pivot_block_out(LuLds& L, T* sr, T* si, int ld, int m) {
    m = uniform(m), ld = uniform(ld);
    stage_out<CPLX, T>(lds_block<T>(L, 0), lds_block<T>(L, 1), uniform(gas(sr)), CPLX ? uniform(gas(si)) : nullptr,
                       ld, m);
}

// LDS pointer whose value the compiler cannot relate to other addresses: every
// operand read through it becomes one ds_read with an immediate offset from
// this one register, instead of a precomputed address per k-step (which the
// register allocator then spilled to scratch inside the MFMA loops).
#define LAS __attribute__((address_space(3)))
template <class T>
__device__ __forceinline__ LAS const T* lds_base(const T* p) {
    LAS const T* q = (LAS const T*)p;
    asm volatile("" : "+v"(q));
    return q;
}

// L_JI = A_JI Ainv_I for J = I+1..I+nU, in place, compiled for TK k-steps
// (tk_class(m)) and ceil(TK / 4) column tiles: branch-free MFMA chains. A rows
// come straight from the panels (a wave takes 16-row strips; clamped loads,
// zeros selected past m), the inverse from LDS (B[0], B[1]) through one opaque
// base per column-tile group with the k-step operand reads at immediate
// offsets, loaded one k-step ahead. Complex products with three real ones per
// k-step (re = p1 - p2, im = p3 - p1 - p2): three independent chains per tile.
template <bool CPLX, class T, int TK>
__device__ __forceinline__ void lower_multiply(GAS T* __restrict__ Pr, GAS T* __restrict__ Pi, long long ps, int ld,
                                                int I, int m, int nU, const LuLds& L) {
    typedef typename Mf<T>::acc A4;
    constexpr int TN = (TK + 3) / 4;
    constexpr int TG = CPLX ? 1 : 4;
    const T* B0 = lds_block<T>(L, 0);
    const T* B1 = lds_block<T>(L, 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int M = nU * m, tm = (M + 15) >> 4;
    for (int t = wave; t < tm; t += RNW) {
        const int r0 = t << 4;
        const int ra = min(r0 + l16, M - 1);
        int dJa;
        const long long oa = stacked_row(ps, ld, I, m, ra, dJa) + (3 - dJa) * m;
        T ar[TK], ai[TK];
#pragma unroll
        for (int s = 0; s < TK; ++s) {
            const int k = 4 * s + l4;
            const int kc = min(k, m - 1);
            const T vr = Pr[oa + kc];
            const T vi = CPLX ? Pi[oa + kc] : T(0);
            ar[s] = k < m ? vr : T(0);
            ai[s] = (CPLX && k < m) ? vi : T(0);
        }
        long long oc[4];
        bool vc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = r0 + Mf<T>::row(lane, j);
            vc[j] = r < M;
            int dJ;
            oc[j] = vc[j] ? stacked_row(ps, ld, I, m, r, dJ) + (3 - dJ) * m : 0;
        }
#pragma unroll
        for (int ug = 0; ug < TN; ug += TG) {
            LAS const T* b0 = lds_base(B0 + l4 * LS + (ug << 4) + l16);
            LAS const T* b1 = lds_base(B1 + l4 * LS + (ug << 4) + l16);
            A4 p1[TG], p2[TG], p3[TG];
#pragma unroll
            for (int u = 0; u < TG; ++u) p1[u] = p2[u] = p3[u] = A4{0, 0, 0, 0};
            T br[TG], bi[TG];
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                br[u] = ug + u < TN ? b0[16 * u] : T(0);
                bi[u] = (CPLX && ug + u < TN) ? b1[16 * u] : T(0);
            }
#pragma unroll
            for (int s = 0; s < TK; ++s) {
                T nbr[TG], nbi[TG];
#pragma unroll
                for (int u = 0; u < TG; ++u) {
                    const bool ldn = s + 1 < TK && ug + u < TN;
                    nbr[u] = ldn ? b0[4 * (s + 1) * LS + 16 * u] : T(0);
                    nbi[u] = (CPLX && ldn) ? b1[4 * (s + 1) * LS + 16 * u] : T(0);
                }
#pragma unroll
                for (int u = 0; u < TG; ++u) {
                    if (ug + u < TN) {
                        p1[u] = Mf<T>::mfma(ar[s], br[u], p1[u]);
                        if (CPLX) {
                            p2[u] = Mf<T>::mfma(ai[s], bi[u], p2[u]);
                            p3[u] = Mf<T>::mfma(ar[s] + ai[s], br[u] + bi[u], p3[u]);
                        }
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < TG; ++u) br[u] = nbr[u], bi[u] = nbi[u];
            }
#pragma unroll
            for (int u = 0; u < TG; ++u) {
                const int c = ((ug + u) << 4) + l16;
                if (ug + u < TN && c < m) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (vc[j]) {
                            if (CPLX) {
                                Pr[oc[j] + c] = p1[u][j] - p2[u][j];
                                Pi[oc[j] + c] = p3[u][j] - p1[u][j] - p2[u][j];
                            } else {
                                Pr[oc[j] + c] = p1[u][j];
                            }
                        }
                }
            }
        }
    }
}

// The U block (I, I+3) still holds only its assembled band (|k - j| <= 3) when
// step I uses it: block (I, I+3) is updated by the steps p with I <= p <= I - 1,
// i.e. by none. It is staged compactly (UB: the 7 diagonals of each column) and
// its update rides on the last dense pass of the block row, reusing that pass's
// L strips; column tile u needs only the k-steps 4 u - 1 .. 4 u + 4, which meet
// rows 16 u - 3 .. 16 u + 18. The skipped k-steps' products are exact zeros.
__device__ __forceinline__ constexpr bool kstep_in_band(int u, int s) { return s >= 4 * u - 1 && s <= 4 * u + 4; }

// A_JK -= L_JI U_IK for J = I+1..I+nU and K = I+1+kb0 .. I+kb0+nkb, with U_IK
// staged in LDS (B[0] re, B[1] im in complex arithmetic, one block per pass;
// B[kb] in real), compiled for TK k-steps like lower_multiply; with BANDK also
// K = I+3 from its staged band (UB). An L strip is loaded into registers once
// per strip and pass; C tiles are loaded before their MFMA chains, which start
// from zero, and written once. Complex tiles one at a time with three real
// product chains, real tiles four at a time (four independent chains).
// DENSE = false: only K = I+3, with U_{I,I+3} staged as a full block pair in
// B[0], B[1] (the k-steps outside its band skipped).
template <bool CPLX, class T, int TK, bool BANDK = false, bool DENSE = true>
__device__ __forceinline__ void schur_pass(GAS T* __restrict__ Pr, GAS T* __restrict__ Pi, long long ps, int ld,
                                           int I, int m, int nU, int kb0, int nkb, const LuLds& L) {
    typedef typename Mf<T>::acc A4;
    constexpr int TN = (TK + 3) / 4;
    // complex: two tiles (six chains) at a time where the registers allow it (FP32, or
    // FP64 up to 8 k-steps); one tile otherwise (two spilled in the surrounding code)
    constexpr int TG = CPLX ? ((sizeof(T) == 4 || TK <= 8) ? 2 : 1) : 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int M = nU * m, tm = (M + 15) >> 4;
    const T* ubr = reinterpret_cast<const T*>(L.UB[0]);
    const T* ubi = reinterpret_cast<const T*>(L.UB[1]);
    for (int t = wave; t < tm; t += RNW) {
        const int r0 = t << 4;
        T ar[TK], ai[TK];
        {
            const int ra = min(r0 + l16, M - 1);
            int dJa;
            const long long oa = stacked_row(ps, ld, I, m, ra, dJa) + (3 - dJa) * m;
#pragma unroll
            for (int s = 0; s < TK; ++s) {
                const int k = 4 * s + l4;
                const int kc = min(k, m - 1);
                const T vr = Pr[oa + kc];
                const T vi = CPLX ? Pi[oa + kc] : T(0);
                ar[s] = k < m ? vr : T(0);
                ai[s] = (CPLX && k < m) ? vi : T(0);
            }
        }
        long long oc[4];
        int dj[4], fi[4];
        bool vc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = r0 + Mf<T>::row(lane, j);
            vc[j] = r < M;
            dj[j] = 1;
            oc[j] = stacked_row(ps, ld, I, m, min(r, M - 1), dj[j]);
            fi[j] = min(r, M - 1) - (dj[j] - 1) * m;
        }
        // one target block column K = I + 1 + kk. Operand mode: 0 dense U from the LDS
        // block pair (Br, Bi); 1 U_{I,I+3} from UB (its band); 2 U_{I,I+3} as a full
        // block in (Br, Bi), only the k-steps that meet its band
        auto block = [&](int kk, const T* Br, const T* Bi, auto mode_tag) {
            constexpr int MODE = decltype(mode_tag)::value;
            constexpr bool BND = MODE != 0;
            // first update of a block: only the entries within 3 of its diagonal are loaded
            int cb[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) cb[j] = (I == 0 || dj[j] == 3 || kk == 2) ? fi[j] : -(1 << 20);
#pragma unroll
            for (int ug = 0; ug < TN; ug += TG) {
                T cr[TG][4], ci[TG][4];
#pragma unroll
                for (int u = 0; u < TG; ++u) {
                    const int c = ((ug + u) << 4) + l16;
                    const int cc = min(c, m - 1);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const bool ok = ug + u < TN && vc[j] && c < m &&
                                        (cb[j] < 0 || (c - cb[j] <= 3 && cb[j] - c <= 3));
                        const long long o = oc[j] + (long long)(kk + 4 - dj[j]) * m + cc;
                        cr[u][j] = ok ? Pr[o] : T(0);
                        ci[u][j] = (CPLX && ok) ? Pi[o] : T(0);
                    }
                }
                LAS const T* b0 = lds_base(Br + l4 * LS + (ug << 4) + l16);
                LAS const T* b1 = lds_base(Bi + l4 * LS + (ug << 4) + l16);
                // operand of k-step s for tile ug + u (zero outside the band / past the tiles)
                auto opnd = [&](int u, int s, T& vr, T& vi) {
                    if (!(ug + u < TN) || (BND && !kstep_in_band(ug + u, s))) {
                        vr = vi = T(0);
                    } else if constexpr (MODE == 1) {
                        const int c = ((ug + u) << 4) + l16, dd = 4 * s + l4 - c + 3;
                        const int e = c * 7 + min(max(dd, 0), 6);
                        const bool in = dd >= 0 && dd <= 6;
                        const T xr = ubr[e];
                        const T xi = CPLX ? ubi[e] : T(0);
                        vr = in ? xr : T(0);
                        vi = (CPLX && in) ? xi : T(0);
                    } else {
                        vr = b0[4 * s * LS + 16 * u];
                        vi = CPLX ? b1[4 * s * LS + 16 * u] : T(0);
                    }
                };
                A4 p1[TG], p2[TG], p3[TG];
#pragma unroll
                for (int u = 0; u < TG; ++u) p1[u] = p2[u] = p3[u] = A4{0, 0, 0, 0};
                T br[TG], bi[TG];
#pragma unroll
                for (int u = 0; u < TG; ++u) opnd(u, 0, br[u], bi[u]);
#pragma unroll
                for (int s = 0; s < TK; ++s) {
                    T nbr[TG], nbi[TG];
#pragma unroll
                    for (int u = 0; u < TG; ++u) {
                        if (s + 1 < TK) opnd(u, s + 1, nbr[u], nbi[u]);
                        else nbr[u] = nbi[u] = T(0);
                    }
#pragma unroll
                    for (int u = 0; u < TG; ++u) {
                        if (ug + u < TN && (!BND || kstep_in_band(ug + u, s))) {
                            p1[u] = Mf<T>::mfma(ar[s], br[u], p1[u]);
                            if (CPLX) {
                                p2[u] = Mf<T>::mfma(ai[s], bi[u], p2[u]);
                                p3[u] = Mf<T>::mfma(ar[s] + ai[s], br[u] + bi[u], p3[u]);
                            }
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < TG; ++u) br[u] = nbr[u], bi[u] = nbi[u];
                }
#pragma unroll
                for (int u = 0; u < TG; ++u) {
                    const int c = ((ug + u) << 4) + l16;
                    if (ug + u < TN && c < m) {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (vc[j]) {
                                const long long o = oc[j] + (long long)(kk + 4 - dj[j]) * m + c;
                                if (CPLX) {
                                    Pr[o] = cr[u][j] - (p1[u][j] - p2[u][j]);
                                    Pi[o] = ci[u][j] - (p3[u][j] - p1[u][j] - p2[u][j]);
                                } else {
                                    Pr[o] = cr[u][j] - p1[u][j];
                                }
                            }
                    }
                }
            }
        };
        if constexpr (DENSE) {
            for (int kb = 0; kb < nkb; ++kb)
                block(kb0 + kb, lds_block<T>(L, CPLX ? 2 * kb : kb), lds_block<T>(L, CPLX ? 2 * kb + 1 : kb),
                      std::integral_constant<int, 0>());
            if constexpr (BANDK) block(2, lds_block<T>(L, 0), lds_block<T>(L, 0), std::integral_constant<int, 1>());
        } else {
            // U_{I,I+3} staged as a full block pair in B[0], B[1]
            block(2, lds_block<T>(L, 0), lds_block<T>(L, CPLX ? 1 : 0), std::integral_constant<int, 2>());
        }
    }
}

// Factor the panels in place (generic loops). Returns nonzero if a diagonal block was singular.
template <bool CPLX, class T>
__device__ __noinline__ int block_lu_gen(T* __restrict__ Pr_, T* __restrict__ Pi_, int ns, int m, LuLds& L) {
    GAS T* const Pr = uniform(gas(Pr_));
    GAS T* const Pi = uniform(gas(Pi_));
    ns = uniform(ns), m = uniform(m);
    const int ld = 7 * m;
    const long long ps = (long long)m * ld;
    const int tid = threadIdx.x;
    if (tid == 0) L.singular = 0;
    unsigned long long mark = wall_clock64();
    auto lap = [&](int k) {
        if (tid == 0) {
            const unsigned long long now = wall_clock64();
            L.tk[k] += now - mark;
            mark = now;
        }
    };
    // U blocks per Schur pass: 2 fit the staging buffers in real arithmetic, 1 in complex
    constexpr int KPASS = CPLX ? 1 : 2;
    T* B0 = lds_block<T>(L, 0);
    T* B1 = lds_block<T>(L, 1);
    for (int I = 0; I < ns; ++I) {
        GAS T* PIr = Pr + I * ps;
        GAS T* PIi = Pi + I * ps;
        pivot_block_in<CPLX, T>(L, (const T*)(PIr + 3 * m), CPLX ? (const T*)(PIi + 3 * m) : nullptr, ld, m);
        __syncthreads();
        lap(1);
        gj_inverse<CPLX, T>(L, m);
        lap(0);
        pivot_block_out<CPLX, T>(L, (T*)(PIr + 3 * m), CPLX ? (T*)(PIi + 3 * m) : nullptr, ld, m);
        const int nU = (ns - 1 - I) < 3 ? (ns - 1 - I) : 3;
        if (nU == 0) break;
        lower_multiply_gen<CPLX, T>(Pr, Pi, ps, ld, I, m, nU, L);
        __syncthreads();
        lap(2);
        for (int kb0 = 0; kb0 < nU; kb0 += KPASS) {
            const int nkb = (nU - kb0) < KPASS ? (nU - kb0) : KPASS;
            for (int kb = 0; kb < nkb; ++kb) {
                T* Br = lds_block<T>(L, CPLX ? 2 * kb : kb);
                T* Bi = CPLX ? lds_block<T>(L, 2 * kb + 1) : nullptr;
                const int col = (4 + kb0 + kb) * m;
                stage_in<CPLX, T, 1>(Br, Bi, PIr + col, CPLX ? PIi + col : nullptr, ld, m);
            }
            __syncthreads();
            lap(1);
            schur_pass_gen<CPLX, T>(Pr, Pi, ps, ld, I, m, nU, kb0, nkb, L);
            __syncthreads();
            lap(3);
        }
    }
    __syncthreads();
    return L.singular;
}

// Factor the panels in place. Returns nonzero if a diagonal block was singular.
// TK = tk_class(m): the k-steps the L multiply and Schur update are compiled for.
template <bool CPLX, class T, int TK>
__device__ __noinline__ int block_lu_tk(T* __restrict__ Pr_, T* __restrict__ Pi_, int ns, int m, LuLds& L) {
    GAS T* const Pr = uniform(gas(Pr_));
    GAS T* const Pi = uniform(gas(Pi_));
    ns = uniform(ns), m = uniform(m);
    const int ld = 7 * m;
    const long long ps = (long long)m * ld;
    const int tid = threadIdx.x;
    if (tid == 0) L.singular = 0;
    unsigned long long mark = wall_clock64();
    auto lap = [&](int k) {
        if (tid == 0) {
            const unsigned long long now = wall_clock64();
            L.tk[k] += now - mark;
            mark = now;
        }
    };
    // U blocks per Schur pass: 2 fit the staging buffers in real arithmetic, 1 in complex
    constexpr int KPASS = CPLX ? 1 : 2;
    for (int I = 0; I < ns; ++I) {
        GAS T* PIr = Pr + I * ps;
        GAS T* PIi = Pi + I * ps;
        pivot_block_in<CPLX, T>(L, (const T*)(PIr + 3 * m), CPLX ? (const T*)(PIi + 3 * m) : nullptr, ld, m);
        __syncthreads();
        lap(1);
        gj_inverse<CPLX, T>(L, m);
        lap(0);
        pivot_block_out<CPLX, T>(L, (T*)(PIr + 3 * m), CPLX ? (T*)(PIi + 3 * m) : nullptr, ld, m);
        const int nU = (ns - 1 - I) < 3 ? (ns - 1 - I) : 3;
        if (nU == 0) break;
        lower_multiply<CPLX, T, TK>(Pr, Pi, ps, ld, I, m, nU, L);
        __syncthreads();
        lap(2);
        // U blocks I+1, I+2 through LDS, KPASS per pass; U_{I,I+3} (still its assembled
        // band) from UB, riding on the last dense pass -- except in FP64 complex
        // arithmetic at TK >= 15, where the merged pass's registers spill and the
        // band update runs as a pass of its own
        constexpr bool MERGE = !(CPLX && sizeof(T) == 8 && TK >= 15);
        const int nD = nU < 3 ? nU : 2;
        for (int kb0 = 0; kb0 < nD; kb0 += KPASS) {
            const int nkb = (nD - kb0) < KPASS ? (nD - kb0) : KPASS;
            const bool band = MERGE && kb0 + nkb == nD && nU == 3;
            for (int kb = 0; kb < nkb; ++kb) {
                const int col = (4 + kb0 + kb) * m;
                u_block_in<CPLX, T>(L, CPLX ? 2 * kb : kb, (const T*)(PIr + col), CPLX ? (const T*)(PIi + col) : nullptr,
                                    ld, m);
            }
            if (band) u_band_in<CPLX, T>(L, (const T*)(PIr + 6 * m), CPLX ? (const T*)(PIi + 6 * m) : nullptr, ld, m);
            __syncthreads();
            lap(1);
            if (MERGE && band) schur_pass<CPLX, T, TK, MERGE>(Pr, Pi, ps, ld, I, m, nU, kb0, nkb, L);
            else schur_pass<CPLX, T, TK>(Pr, Pi, ps, ld, I, m, nU, kb0, nkb, L);
            __syncthreads();
            lap(3);
        }
        if (!MERGE && nU == 3) {
            u_block_in<CPLX, T>(L, 0, (const T*)(PIr + 6 * m), CPLX ? (const T*)(PIi + 6 * m) : nullptr, ld, m);
            __syncthreads();
            lap(1);
            schur_pass<CPLX, T, TK, false, false>(Pr, Pi, ps, ld, I, m, nU, 2, 0, L);
            __syncthreads();
            lap(3);
        }
    }
    __syncthreads();
    return L.singular;
}

// Factorisation code paths: LU_GENERIC (runtime k-step loops, one code path for
// every m) or LU_COMPILED_K (MFMA loops compiled for tk_class(m) k-steps). The
// per-grid choice comes from the tuning table (fpe_option_amd/tuning_gfx950.json).
enum { LU_GENERIC = 0, LU_COMPILED_K = 1 };

// One call site per use in the solver kernel (the code path is chosen inside).
template <bool CPLX, class T>
__device__ __noinline__ int block_lu(T* __restrict__ Pr, T* __restrict__ Pi, int ns, int m, int variant, LuLds& L) {
    m = uniform(m), variant = uniform(variant);
    if (variant == LU_GENERIC) return block_lu_gen<CPLX, T>(Pr, Pi, ns, m, L);
    switch (tk_class(m)) {
        case 5: return block_lu_tk<CPLX, T, 5>(Pr, Pi, ns, m, L);
        case 8: return block_lu_tk<CPLX, T, 8>(Pr, Pi, ns, m, L);
        case 10: return block_lu_tk<CPLX, T, 10>(Pr, Pi, ns, m, L);
        case 15: return block_lu_tk<CPLX, T, 15>(Pr, Pi, ns, m, L);
        default: return block_lu_tk<CPLX, T, 16>(Pr, Pi, ns, m, L);
    }
}

// Single right-hand side solve with the block factors, latency-shaped: each
// thread (row i = tid / 4, part q = tid % 4) loads a block-row's factor
// entries block by block (16 per thread, unrolled and predicated), issuing the
// next block's loads before contracting the current one, so a block-row step
// waits for about one memory latency; the last solution blocks live in an LDS
// ring (4 x 64, re and im) instead of being re-read from memory. ring: LDS of
// at least 10 * 64 doubles. Factors of type T (FP64, or FP32 for the
// mixed-precision path) are widened to FP64; all arithmetic is FP64.
// NF as in block_solve_multi_w: m / 4 known at compile time (NF > 0) lets the
// full column groups go unpredicated; NF = 0 takes any m.
template <bool CPLX, class T, int NF>
__device__ __noinline__ void block_solve1_w(const T* __restrict__ Pr_, const T* __restrict__ Pi_, int ns, int m,
                                            double* __restrict__ xr_, double* __restrict__ xi_, double* ring) {
    GAS const T* const Pr = uniform(gas(Pr_));
    GAS const T* const Pi = uniform(gas(Pi_));
    GAS double* const xr = uniform(gas(xr_));
    GAS double* const xi = uniform(gas(xi_));
    ns = uniform(ns), m = uniform(m);
    const int ld = 7 * m;
    const long long ps = (long long)m * ld;
    constexpr int TPR = RNT / MAX_BLOCK;  // threads per row
    constexpr int U = MAX_BLOCK / TPR;    // columns per thread per block
    const int tid = threadIdx.x, i = tid / TPR, q = tid % TPR;
    const bool row = i < m;
    double* rr = ring;
    double* ri = ring + 4 * 64;
    double* tr = ring + 8 * 64;
    double* ti = ring + 9 * 64;
    constexpr int UA = NF > 0 ? NF + 1 : U;  // column groups held per block
    const int ic = NF > 0 ? min(i, m - 1) : i;  // NF > 0: rows past m read row m - 1, unstored
    double a0[UA], b0[UA], a1[UA], b1[UA];
    GAS const T* pr = nullptr;
    GAS const T* pi = nullptr;
    auto fill = [&](GAS const T* qr, GAS const T* qi, bool ok, double* a, double* b) {
        if constexpr (NF > 0) {
            if (ok) {
#pragma unroll
                for (int u = 0; u < NF; ++u) {
                    a[u] = double(qr[q + TPR * u]);
                    b[u] = CPLX ? double(qi[q + TPR * u]) : 0.0;
                }
                const int k = q + TPR * NF;
                a[NF] = k < m ? double(qr[k]) : 0.0;
                b[NF] = (CPLX && k < m) ? double(qi[k]) : 0.0;
            } else {
#pragma unroll
                for (int u = 0; u < UA; ++u) a[u] = 0.0, b[u] = 0.0;
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = q + TPR * u;
                const bool v = ok && row && k < m;
                a[u] = v ? double(qr[k]) : 0.0;
                b[u] = (CPLX && v) ? double(qi[k]) : 0.0;
            }
        }
    };
    // factor entries (row i, columns col + q + 4u) of block row I
    auto load_row = [&](int I, int col, bool ok, double* a, double* b) {
        fill(Pr + I * ps + (long long)ic * ld + col, CPLX ? Pi + I * ps + (long long)ic * ld + col : nullptr, ok, a,
             b);
    };
    // ... of the current block row
    auto load = [&](int col, bool ok, double* a, double* b) { fill(pr + col, CPLX ? pi + col : nullptr, ok, a, b); };
    // (ar, ai) += block row segment (a, b) times the vector block (yr, yi)
    auto contract = [&](const double* a, const double* b, const double* yr, const double* yi, double& ar,
                        double& ai) {
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            const int k = q + TPR * u;
            if ((NF > 0 && u < NF) || k < m) {
                if (CPLX) {
                    ar += a[u] * yr[k] - b[u] * yi[k];
                    ai += a[u] * yi[k] + b[u] * yr[k];
                } else {
                    ar += a[u] * yr[k];
                }
            }
        }
    };
    auto reduce4 = [&](double& ar, double& ai) {
#pragma unroll
        for (int o = 1; o < TPR; o <<= 1) {
            ar += __shfl_xor(ar, o, 64);
            if (CPLX) ai += __shfl_xor(ai, o, 64);
        }
    };
    // The first block rows contract zero factor entries with ring slots of blocks
    // left of the matrix, which this solve never writes: start from zeros (the
    // staging LDS the ring reuses holds arbitrary values).
    __syncthreads();
    for (int e = tid; e < 8 * 64; e += RNT) ring[e] = 0.0;
    __syncthreads();
    // forward: y_I = b_I - sum_{d=0..2} L_{I, I-3+d} y_{I-3+d} (ring slot of block J: J & 3);
    // the first block of row I + 1 is loaded while row I finishes
    auto d0_of = [&](int I) { return I >= 3 ? 0 : 3 - I; };  // blocks d < d0 lie left of the matrix
    if (ns > 0) load_row(0, 0, d0_of(0) <= 0, a0, b0);
    for (int I = 0; I < ns; ++I) {
        pr = Pr + I * ps + (long long)ic * ld;
        pi = CPLX ? Pi + I * ps + (long long)ic * ld : nullptr;
        const int d0 = d0_of(I);
        double ar = 0.0, ai = 0.0;
        load(m, d0 <= 1, a1, b1);
        contract(a0, b0, rr + ((I + 1) & 3) * 64, ri + ((I + 1) & 3) * 64, ar, ai);
        load(2 * m, d0 <= 2, a0, b0);
        contract(a1, b1, rr + ((I + 2) & 3) * 64, ri + ((I + 2) & 3) * 64, ar, ai);
        if (I + 1 < ns) load_row(I + 1, 0, d0_of(I + 1) <= 0, a1, b1);
        contract(a0, b0, rr + ((I + 3) & 3) * 64, ri + ((I + 3) & 3) * 64, ar, ai);
#pragma unroll
        for (int u = 0; u < UA; ++u) a0[u] = a1[u], b0[u] = b1[u];
        reduce4(ar, ai);
        if (row && q == 0) {
            const double vr = xr[I * m + i] - ar;
            rr[(I & 3) * 64 + i] = vr;
            xr[I * m + i] = vr;
            if (CPLX) {
                const double vi = xi[I * m + i] - ai;
                ri[(I & 3) * 64 + i] = vi;
                xi[I * m + i] = vi;
            }
        }
        __syncthreads();
    }
    // backward: x_I = Ainv_I (y_I - sum_{d=0..2} U_{I, I+1+d} x_{I+1+d}); the first
    // block of row I - 1 is loaded while row I finishes
    auto nK_of = [&](int I) { return (ns - 1 - I) < 3 ? (ns - 1 - I) : 3; };  // blocks right of the matrix are skipped
    if (ns > 0) load_row(ns - 1, 4 * m, nK_of(ns - 1) > 0, a0, b0);
    for (int I = ns - 1; I >= 0; --I) {
        pr = Pr + I * ps + (long long)ic * ld;
        pi = CPLX ? Pi + I * ps + (long long)ic * ld : nullptr;
        const int nK = nK_of(I);
        double ar = 0.0, ai = 0.0;
        load(5 * m, nK > 1, a1, b1);
        contract(a0, b0, rr + ((I + 1) & 3) * 64, ri + ((I + 1) & 3) * 64, ar, ai);
        load(6 * m, nK > 2, a0, b0);
        contract(a1, b1, rr + ((I + 2) & 3) * 64, ri + ((I + 2) & 3) * 64, ar, ai);
        load(3 * m, true, a1, b1);  // Ainv_I row, used after the barrier
        contract(a0, b0, rr + ((I + 3) & 3) * 64, ri + ((I + 3) & 3) * 64, ar, ai);
        if (I > 0) load_row(I - 1, 4 * m, nK_of(I - 1) > 0, a0, b0);
        reduce4(ar, ai);
        if (row && q == 0) {
            tr[i] = xr[I * m + i] - ar;
            if (CPLX) ti[i] = xi[I * m + i] - ai;
        }
        __syncthreads();
        ar = 0.0, ai = 0.0;
        contract(a1, b1, tr, ti, ar, ai);
        reduce4(ar, ai);
        if (row && q == 0) {
            xr[I * m + i] = ar;
            rr[(I & 3) * 64 + i] = ar;
            if (CPLX) {
                xi[I * m + i] = ai;
                ri[(I & 3) * 64 + i] = ai;
            }
        }
        __syncthreads();
    }
}

// NR (<= 3) right-hand sides x0..x{NR-1} with the FP64 real factors (the
// GMRES preconditioner), shaped like block_solve1: a block row's factor
// entries are loaded once, one block ahead, and contracted with every
// right-hand side; each right-hand side keeps its last solution blocks in its
// own LDS ring. ring: LDS of at least 5 * NR * 64 doubles.
// NF > 0: a block width m with m / 4 == NF known at compile time, so the NF
// full column groups (every lane's column inside the block) are loaded and
// contracted without per-lane predicates, and only group NF (m % 4 columns)
// is predicated. Rows i >= m then read row m - 1 and compute values nobody
// stores. NF = 0: any m, every entry predicated.
template <int NR, int NF>
__device__ __noinline__ void block_solve_multi_w(const double* __restrict__ P_, int ns, int m, double* x0_,
                                                 double* x1_, double* x2_, double* ring) {
    static_assert(NR >= 1 && NR <= 3, "one to three right-hand sides");
    GAS const double* const P = uniform(gas(P_));
    GAS double* x[3] = {uniform(gas(x0_)), NR > 1 ? uniform(gas(x1_)) : nullptr,
                        NR > 2 ? uniform(gas(x2_)) : nullptr};
    ns = uniform(ns), m = uniform(m);
    const int ld = 7 * m;
    const long long ps = (long long)m * ld;
    constexpr int TPR = RNT / MAX_BLOCK;  // threads per row
    constexpr int U = MAX_BLOCK / TPR;    // columns per thread per block
    const int tid = threadIdx.x, i = tid / TPR, q = tid % TPR;
    const bool row = i < m;
    double* rg = ring;                // ring of right-hand side v: rg + v * 256, slot J & 3
    double* tv = ring + NR * 4 * 64;  // temporaries: tv + v * 64
    constexpr int UA = NF > 0 ? NF + 1 : U;  // column groups held per block
    const int ic = NF > 0 ? min(i, m - 1) : i;
    double a0[UA], a1[UA];
    GAS const double* pr = nullptr;
    auto fill = [&](GAS const double* qr, bool ok, double* a) {
        if constexpr (NF > 0) {
            if (ok) {
#pragma unroll
                for (int u = 0; u < NF; ++u) a[u] = qr[q + TPR * u];
                const int k = q + TPR * NF;
                a[NF] = k < m ? qr[k] : 0.0;
            } else {
#pragma unroll
                for (int u = 0; u < UA; ++u) a[u] = 0.0;
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = q + TPR * u;
                a[u] = (ok && row && k < m) ? qr[k] : 0.0;
            }
        }
    };
    auto load = [&](int col, bool ok, double* a) { fill(pr + col, ok, a); };
    auto load_row = [&](int I, int col, bool ok, double* a) { fill(P + I * ps + (long long)ic * ld + col, ok, a); };
    auto contract = [&](const double* a, int slot, double* acc) {
#pragma unroll
        for (int v = 0; v < NR; ++v) {
            const double* y = rg + v * 256 + slot * 64;
#pragma unroll
            for (int u = 0; u < UA; ++u) {
                const int k = q + TPR * u;
                if (NF > 0 && u < NF)
                    acc[v] += a[u] * y[k];
                else if (k < m)
                    acc[v] += a[u] * y[k];
            }
        }
    };
    auto reduce4 = [&](double* acc) {
#pragma unroll
        for (int v = 0; v < NR; ++v)
#pragma unroll
            for (int o = 1; o < TPR; o <<= 1) acc[v] += __shfl_xor(acc[v], o, 64);
    };
    // ring slots of blocks left of the matrix are contracted with zero factor
    // entries before this solve writes them: start from zeros
    __syncthreads();
    for (int e = tid; e < NR * 4 * 64; e += RNT) ring[e] = 0.0;
    __syncthreads();
    // forward: y_I = b_I - sum_{d=0..2} L_{I, I-3+d} y_{I-3+d}; the first block of
    // row I + 1 is loaded while row I finishes
    auto d0_of = [&](int I) { return I >= 3 ? 0 : 3 - I; };
    if (ns > 0) load_row(0, 0, d0_of(0) <= 0, a0);
    for (int I = 0; I < ns; ++I) {
        pr = P + I * ps + (long long)ic * ld;
        const int d0 = d0_of(I);
        double acc[NR];
#pragma unroll
        for (int v = 0; v < NR; ++v) acc[v] = 0.0;
        load(m, d0 <= 1, a1);
        contract(a0, (I + 1) & 3, acc);
        load(2 * m, d0 <= 2, a0);
        contract(a1, (I + 2) & 3, acc);
        if (I + 1 < ns) load_row(I + 1, 0, d0_of(I + 1) <= 0, a1);
        contract(a0, (I + 3) & 3, acc);
#pragma unroll
        for (int u = 0; u < UA; ++u) a0[u] = a1[u];
        reduce4(acc);
        if (row && q == 0) {
#pragma unroll
            for (int v = 0; v < NR; ++v) {
                const double y = x[v][I * m + i] - acc[v];
                rg[v * 256 + (I & 3) * 64 + i] = y;
                x[v][I * m + i] = y;
            }
        }
        __syncthreads();
    }
    // backward: x_I = Ainv_I (y_I - sum_{d=0..2} U_{I, I+1+d} x_{I+1+d}); the first
    // block of row I - 1 is loaded while row I finishes
    auto nK_of = [&](int I) { return (ns - 1 - I) < 3 ? (ns - 1 - I) : 3; };
    if (ns > 0) load_row(ns - 1, 4 * m, nK_of(ns - 1) > 0, a0);
    for (int I = ns - 1; I >= 0; --I) {
        pr = P + I * ps + (long long)ic * ld;
        const int nK = nK_of(I);
        double acc[NR];
#pragma unroll
        for (int v = 0; v < NR; ++v) acc[v] = 0.0;
        load(5 * m, nK > 1, a1);
        contract(a0, (I + 1) & 3, acc);
        load(6 * m, nK > 2, a0);
        contract(a1, (I + 2) & 3, acc);
        load(3 * m, true, a1);  // Ainv_I row, used after the barrier
        contract(a0, (I + 3) & 3, acc);
        if (I > 0) load_row(I - 1, 4 * m, nK_of(I - 1) > 0, a0);
        reduce4(acc);
        if (row && q == 0) {
#pragma unroll
            for (int v = 0; v < NR; ++v) tv[v * 64 + i] = x[v][I * m + i] - acc[v];
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < NR; ++v) {
            acc[v] = 0.0;
#pragma unroll
            for (int u = 0; u < UA; ++u) {
                const int k = q + TPR * u;
                if (NF > 0 && u < NF)
                    acc[v] += a1[u] * tv[v * 64 + k];
                else if (k < m)
                    acc[v] += a1[u] * tv[v * 64 + k];
            }
        }
        reduce4(acc);
        if (row && q == 0) {
#pragma unroll
            for (int v = 0; v < NR; ++v) {
                x[v][I * m + i] = acc[v];
                rg[v * 256 + (I & 3) * 64 + i] = acc[v];
            }
        }
        __syncthreads();
    }
}


// Single right-hand side block solve, compiled for the common block widths.
template <bool CPLX, class T>
__device__ __forceinline__ void block_solve1(const T* Pr, const T* Pi, int ns, int m, double* xr, double* xi,
                                             double* ring) {
    switch (m >> 2) {
        case 5: block_solve1_w<CPLX, T, 5>(Pr, Pi, ns, m, xr, xi, ring); break;
        case 7: block_solve1_w<CPLX, T, 7>(Pr, Pi, ns, m, xr, xi, ring); break;
        case 10: block_solve1_w<CPLX, T, 10>(Pr, Pi, ns, m, xr, xi, ring); break;
        case 15: block_solve1_w<CPLX, T, 15>(Pr, Pi, ns, m, xr, xi, ring); break;
        default: block_solve1_w<CPLX, T, 0>(Pr, Pi, ns, m, xr, xi, ring); break;
    }
}

// Block solve with NR right-hand sides, compiled for the common block widths.
// The real solve of x with the real factors P and the complex solve of (zr, zi)
// with the complex factors (Cr, Ci) in one sweep per direction (the two stage
// systems of a direct-mode Newton iteration): each block-row step issues both
// solves' factor loads before either contracts, so the two recurrences share one
// memory latency and one barrier per block row. The per-entry arithmetic is that
// of block_solve1_w for each solve, so the results are the same. NF > 0 only
// (m / 4 known; the factor rows of both solves in registers); ring: LDS of at
// least 15 * 64 doubles. Factors of type T (FP64, or FP32 for the mixed-precision
// path) are widened to FP64; all arithmetic is FP64.
template <class T, int NF>
__device__ __noinline__ void block_solve_pair_w(const T* __restrict__ P_, const T* __restrict__ Cr_,
                                                const T* __restrict__ Ci_, int ns, int m,
                                                double* __restrict__ x_, double* __restrict__ zr_,
                                                double* __restrict__ zi_, double* ring) {
    static_assert(NF > 0, "the paired solve is compiled for known m / 4");
    GAS const T* const P = uniform(gas(P_));
    GAS const T* const Cr = uniform(gas(Cr_));
    GAS const T* const Ci = uniform(gas(Ci_));
    GAS double* const x = uniform(gas(x_));
    GAS double* const zr = uniform(gas(zr_));
    GAS double* const zi = uniform(gas(zi_));
    ns = uniform(ns), m = uniform(m);
    const int ld = 7 * m;
    const long long ps = (long long)m * ld;
    constexpr int TPR = RNT / MAX_BLOCK;  // threads per row
    const int tid = threadIdx.x, i = tid / TPR, q = tid % TPR;
    const bool row = i < m;
    double* rx = ring;             // real solution blocks (4 x 64)
    double* rzr = ring + 4 * 64;   // complex, re
    double* rzi = ring + 8 * 64;   // complex, im
    double* tx = ring + 12 * 64;
    double* tzr = ring + 13 * 64;
    double* tzi = ring + 14 * 64;
    constexpr int UA = NF + 1;   // column groups held per block
    const int ic = min(i, m - 1);  // rows past m read row m - 1, unstored
    double a0[UA], a1[UA], c0[UA], e0[UA], c1[UA], e1[UA];  // real rows a; complex rows (c, e) = (re, im)
    auto fill = [&](GAS const T* qr, bool ok, double* a) {
        if (ok) {
#pragma unroll
            for (int u = 0; u < NF; ++u) a[u] = double(qr[q + TPR * u]);
            const int k = q + TPR * NF;
            a[NF] = k < m ? double(qr[k]) : 0.0;
        } else {
#pragma unroll
            for (int u = 0; u < UA; ++u) a[u] = 0.0;
        }
    };
    // factor entries (row i, columns col + q + 4u) of block row I, both solves
    auto load_row = [&](int I, int col, bool ok, double* a, double* c, double* e) {
        const long long o = I * ps + (long long)ic * ld + col;
        fill(P + o, ok, a);
        fill(Cr + o, ok, c);
        fill(Ci + o, ok, e);
    };
    auto contract_r = [&](const double* a, const double* y, double& ar) {
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            const int k = q + TPR * u;
            if (u < NF || k < m) ar += a[u] * y[k];
        }
    };
    auto contract_c = [&](const double* a, const double* b, const double* yr, const double* yi, double& ar,
                          double& ai) {
#pragma unroll
        for (int u = 0; u < UA; ++u) {
            const int k = q + TPR * u;
            if (u < NF || k < m) {
                ar += a[u] * yr[k] - b[u] * yi[k];
                ai += a[u] * yi[k] + b[u] * yr[k];
            }
        }
    };
    auto reduce = [&](double& v) {
#pragma unroll
        for (int o = 1; o < TPR; o <<= 1) v += __shfl_xor(v, o, 64);
    };
    auto copy = [&](double* dst, const double* src) {
#pragma unroll
        for (int u = 0; u < UA; ++u) dst[u] = src[u];
    };
    // ring slots of blocks left of the matrix are contracted with zero factor
    // entries and never written: start from zeros
    __syncthreads();
    for (int e = tid; e < 12 * 64; e += RNT) ring[e] = 0.0;
    __syncthreads();
    // forward: y_I = b_I - sum_{d=0..2} L_{I, I-3+d} y_{I-3+d} (ring slot of block J: J & 3)
    auto d0_of = [&](int I) { return I >= 3 ? 0 : 3 - I; };
    if (ns > 0) load_row(0, 0, d0_of(0) <= 0, a0, c0, e0);
    for (int I = 0; I < ns; ++I) {
        const int d0 = d0_of(I);
        const int s1 = ((I + 1) & 3) * 64, s2 = ((I + 2) & 3) * 64, s3 = ((I + 3) & 3) * 64;
        double ax = 0.0, azr = 0.0, azi = 0.0;
        load_row(I, m, d0 <= 1, a1, c1, e1);
        contract_r(a0, rx + s1, ax);
        contract_c(c0, e0, rzr + s1, rzi + s1, azr, azi);
        load_row(I, 2 * m, d0 <= 2, a0, c0, e0);
        contract_r(a1, rx + s2, ax);
        contract_c(c1, e1, rzr + s2, rzi + s2, azr, azi);
        if (I + 1 < ns) load_row(I + 1, 0, d0_of(I + 1) <= 0, a1, c1, e1);
        contract_r(a0, rx + s3, ax);
        contract_c(c0, e0, rzr + s3, rzi + s3, azr, azi);
        copy(a0, a1), copy(c0, c1), copy(e0, e1);
        reduce(ax), reduce(azr), reduce(azi);
        if (row && q == 0) {
            const int sl = (I & 3) * 64 + i;
            const double vx = x[I * m + i] - ax, vr = zr[I * m + i] - azr, vi = zi[I * m + i] - azi;
            rx[sl] = vx, rzr[sl] = vr, rzi[sl] = vi;
            x[I * m + i] = vx, zr[I * m + i] = vr, zi[I * m + i] = vi;
        }
        __syncthreads();
    }
    // backward: x_I = Ainv_I (y_I - sum_{d=0..2} U_{I, I+1+d} x_{I+1+d})
    auto nK_of = [&](int I) { return (ns - 1 - I) < 3 ? (ns - 1 - I) : 3; };
    if (ns > 0) load_row(ns - 1, 4 * m, nK_of(ns - 1) > 0, a0, c0, e0);
    for (int I = ns - 1; I >= 0; --I) {
        const int nK = nK_of(I);
        const int s1 = ((I + 1) & 3) * 64, s2 = ((I + 2) & 3) * 64, s3 = ((I + 3) & 3) * 64;
        double ax = 0.0, azr = 0.0, azi = 0.0;
        load_row(I, 5 * m, nK > 1, a1, c1, e1);
        contract_r(a0, rx + s1, ax);
        contract_c(c0, e0, rzr + s1, rzi + s1, azr, azi);
        load_row(I, 6 * m, nK > 2, a0, c0, e0);
        contract_r(a1, rx + s2, ax);
        contract_c(c1, e1, rzr + s2, rzi + s2, azr, azi);
        load_row(I, 3 * m, true, a1, c1, e1);  // Ainv_I rows, used after the barrier
        contract_r(a0, rx + s3, ax);
        contract_c(c0, e0, rzr + s3, rzi + s3, azr, azi);
        if (I > 0) load_row(I - 1, 4 * m, nK_of(I - 1) > 0, a0, c0, e0);
        reduce(ax), reduce(azr), reduce(azi);
        if (row && q == 0) {
            tx[i] = x[I * m + i] - ax;
            tzr[i] = zr[I * m + i] - azr;
            tzi[i] = zi[I * m + i] - azi;
        }
        __syncthreads();
        ax = 0.0, azr = 0.0, azi = 0.0;
        contract_r(a1, tx, ax);
        contract_c(c1, e1, tzr, tzi, azr, azi);
        reduce(ax), reduce(azr), reduce(azi);
        if (row && q == 0) {
            const int sl = (I & 3) * 64 + i;
            x[I * m + i] = ax, rx[sl] = ax;
            zr[I * m + i] = azr, rzr[sl] = azr;
            zi[I * m + i] = azi, rzi[sl] = azi;
        }
        __syncthreads();
    }
}

// The paired real and complex solve where it is compiled (m / 4 in {5, 7, 10});
// returns false for other block widths, which take the two single solves.
template <class T>
__device__ __forceinline__ bool block_solve_pair(const T* P, const T* Cr, const T* Ci, int ns, int m, double* x,
                                                 double* zr, double* zi, double* ring) {
    switch (m >> 2) {
        case 5: block_solve_pair_w<T, 5>(P, Cr, Ci, ns, m, x, zr, zi, ring); return true;
        case 7: block_solve_pair_w<T, 7>(P, Cr, Ci, ns, m, x, zr, zi, ring); return true;
        case 10: block_solve_pair_w<T, 10>(P, Cr, Ci, ns, m, x, zr, zi, ring); return true;
        default: return false;
    }
}

template <int NR>
__device__ __forceinline__ void block_solve_multi(const double* P, int ns, int m, double* x0, double* x1, double* x2,
                                                  double* ring) {
    switch (m >> 2) {
        case 5: block_solve_multi_w<NR, 5>(P, ns, m, x0, x1, x2, ring); break;
        case 7: block_solve_multi_w<NR, 7>(P, ns, m, x0, x1, x2, ring); break;
        case 10: block_solve_multi_w<NR, 10>(P, ns, m, x0, x1, x2, ring); break;
        case 15: block_solve_multi_w<NR, 15>(P, ns, m, x0, x1, x2, ring); break;
        default: block_solve_multi_w<NR, 0>(P, ns, m, x0, x1, x2, ring); break;
    }
}

}  // namespace fpe
