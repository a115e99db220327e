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
#include "device_common.hpp"

namespace fpe {

constexpr int LS = MAX_BLOCK + 1;  // LDS row stride of an m x m block
constexpr int NRG = MAX_BLOCK / NWAVE;  // register columns per lane in the Gauss-Jordan inverse

struct LuLds {
    // four m x m staging blocks: the pivot block and its inverse (0, 1 = re, im),
    // unscramble scratch (2, 3), then the U blocks of the Schur update
    double B[4][MAX_BLOCK * LS];
    int piv[MAX_BLOCK];   // pivot row of each elimination step
    int qinv[MAX_BLOCK];  // step whose pivot row is a given row
    double colr[2][64], coli[2][64];  // pivot column of a step (double-buffered)
    double ipr[2], ipi[2];            // 1 / pivot
    int pk[2];                        // pivot row
    int singular;
    // wall-clock ticks by sub-phase (thread 0): pivot-block inverse, staging, L multiply, Schur update
    unsigned long long tk[4];
};

// Lane layout of v_mfma_f64_16x16x4f64 (checked on gfx950 with an in-bounds
// probe): a = A[row0 + lane%16][k0 + lane/16], b = B[k0 + lane/16][col0 + lane%16],
// acc[j] = C[row0 + lane/16 + 4*j][col0 + lane%16]. (The f32 16x16x4 form places
// acc[j] at row 4*(lane/16) + j instead.)
__device__ __forceinline__ void mfma_c(d4& accr, d4& acci, double ar, double ai, double br, double bi, bool cplx) {
    accr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, accr, 0, 0, 0);
    if (cplx) {
        accr = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, accr, 0, 0, 0);
        acci = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acci, 0, 0, 0);
        acci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, acci, 0, 0, 0);
    }
}

__device__ __forceinline__ double readlane_d(double x, int l) {
    const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane(int(unsigned(v)), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane(int(unsigned(v >> 32)), l);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, int(unsigned(v)), CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, int(unsigned(v >> 32)), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

// Wave-uniform maximum of a double: DPP row rotations (8, 4, 2, 1) leave each
// 16-lane row's maximum in all its lanes, then the four rows are combined.
__device__ __forceinline__ double wave_max_dpp(double v) {
    v = fmax(v, dpp_d<0x128>(v));
    v = fmax(v, dpp_d<0x124>(v));
    v = fmax(v, dpp_d<0x122>(v));
    v = fmax(v, dpp_d<0x121>(v));
    return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}

// Gauss-Jordan inverse of the m x m block in (B[0], B[1]) by all four waves.
// Lane i holds row i and wave w holds the columns j = w (mod 4) in registers,
// so an elimination step is one register update per held entry plus a lane
// broadcast of the pivot row, with one workgroup barrier per step. Partial
// pivoting is implicit: step k takes the not-yet-used row p_k with the largest
// |A[.][k]| (|re| + |im| in complex arithmetic; the search key drops the 6
// lowest mantissa bits to carry the row, ties go to the lower row) and no rows
// move. The owner wave of column k+1 updates that column and searches its
// pivot before updating its other columns, so the search overlaps the other
// waves' updates. Diagonal-pivot in-place Gauss-Jordan on P A leaves
// X[p_k][j] = inv(A)[k][p_j]; the last pass unscrambles inv(A)[a][b] = X[p_a][q_b]
// with p_{q_b} = b. Result in natural order in (B[0], B[1]); B[2], B[3] are
// scratch. Sets singular on a zero or non-finite pivot. Ends with a barrier.
template <bool CPLX>
__device__ void gj_inverse(LuLds& L, int m) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool lrow = lane < m;
    double xr[NRG], xi[NRG];
#pragma unroll
    for (int r = 0; r < NRG; ++r) {
        const int j = NWAVE * r + w;
        xr[r] = (lrow && j < m) ? L.B[0][lane * LS + j] : 0.0;
        xi[r] = (CPLX && lrow && j < m) ? L.B[1][lane * LS + j] : 0.0;
    }
    bool used = !lrow;
    if (threadIdx.x < MAX_BLOCK) L.qinv[threadIdx.x] = 0;  // stays in range if a bad pivot search repeats a row

    // pivot search of column k (owner wave): publishes the column, the pivot row and 1/pivot in buffer b
    auto search = [&](int k, double cr, double ci, int b) {
        const double mag = CPLX ? fabs(cr) + fabs(ci) : fabs(cr);
        const unsigned long long bits = __builtin_bit_cast(unsigned long long, mag);
        double key = used ? -1.0 : __builtin_bit_cast(double, (bits & ~63ull) | (unsigned long long)(63 - lane));
        key = wave_max_dpp(key);
        const int p = __builtin_amdgcn_readfirstlane(63 - int(__builtin_bit_cast(unsigned long long, key) & 63ull));
        const double dr = readlane_d(cr, p), di = CPLX ? readlane_d(ci, p) : 0.0;
        double ir, ii;
        if (CPLX) {
            const double den = dr * dr + di * di;
            ir = den > 0.0 ? dr / den : 0.0;
            ii = den > 0.0 ? -di / den : 0.0;
        } else {
            ir = dr != 0.0 ? 1.0 / dr : 0.0;
            ii = 0.0;
        }
        L.colr[b][lane] = cr;
        if (CPLX) L.coli[b][lane] = ci;
        if (lane == 0) {
            L.pk[b] = p;
            L.ipr[b] = ir;
            L.ipi[b] = ii;
            L.piv[k] = p;
            const double mp = CPLX ? fabs(dr) + fabs(di) : fabs(dr);
            if (!(key > 0.0) || !(mp > 0.0) || !(mp < INFINITY)) L.singular = 1;
        }
    };
    if (w == 0) search(0, xr[0], xi[0], 0);
    for (int k = 0; k < m; ++k) {
        const int b = k & 1;
        __syncthreads();
        const int p = __builtin_amdgcn_readfirstlane(L.pk[b]);
        const double fr = L.colr[b][lane], fi = CPLX ? L.coli[b][lane] : 0.0;
        const double ir = L.ipr[b], ii = CPLX ? L.ipi[b] : 0.0;
        const bool isp = lane == p;
        if (isp) used = true;
        // in-place step on the entry (row lane, column j): pivot row -> row / pivot,
        // column k -> -f / pivot (1 / pivot on the pivot row), rest -> x - f * row_p / pivot
        auto upd = [&](double& vr, double& vi, int j) {
            double sr, si;
            if (j == k) {
                sr = ir, si = ii;
            } else {
                const double pr = readlane_d(vr, p);
                const double pi_ = CPLX ? readlane_d(vi, p) : 0.0;
                sr = CPLX ? pr * ir - pi_ * ii : pr * ir;
                si = CPLX ? pr * ii + pi_ * ir : 0.0;
            }
            if (isp) {
                vr = sr;
                if (CPLX) vi = si;
            } else {
                const double br = (j == k) ? 0.0 : vr, bi = (j == k) ? 0.0 : vi;
                vr = br - (CPLX ? fr * sr - fi * si : fr * sr);
                if (CPLX) vi = bi - (fr * si + fi * sr);
            }
        };
        const int k1 = k + 1;
        if (k1 < m && w == (k1 & (NWAVE - 1))) {
            const int r1 = k1 / NWAVE;
            double cr = 0.0, ci = 0.0;
#pragma unroll
            for (int r = 0; r < NRG; ++r)
                if (r == r1) cr = xr[r], ci = xi[r];
            upd(cr, ci, k1);
            search(k1, cr, ci, b ^ 1);
        }
#pragma unroll
        for (int r = 0; r < NRG; ++r) {
            const int j = NWAVE * r + w;
            if (j < m) upd(xr[r], xi[r], j);
        }
    }
#pragma unroll
    for (int r = 0; r < NRG; ++r) {
        const int j = NWAVE * r + w;
        if (lrow && j < m) {
            L.B[2][lane * LS + j] = xr[r];
            if (CPLX) L.B[3][lane * LS + j] = xi[r];
        }
    }
    __syncthreads();
    if (threadIdx.x < m) L.qinv[L.piv[threadIdx.x]] = threadIdx.x;
    __syncthreads();
    for (int e = threadIdx.x; e < m * m; e += NT) {
        const int a = e / m, c = e - a * m;
        const int src = L.piv[a] * LS + L.qinv[c];
        L.B[0][a * LS + c] = L.B[2][src];
        if (CPLX) L.B[1][a * LS + c] = L.B[3][src];
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

// L_JI = A_JI Ainv_I for J = I+1..I+nU, in place. A rows come straight from the
// panels (a wave takes 16-row strips; their fragments stay in registers), the
// inverse from LDS (B[0], B[1]).
template <bool CPLX>
__device__ void lower_multiply(double* __restrict__ Pr, double* __restrict__ Pi, long long ps, int ld, int I, int m,
                               int nU, const LuLds& L) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int M = nU * m, tm = (M + 15) >> 4, tk = (m + 3) >> 2, tn = (m + 15) >> 4;
    for (int t = wave; t < tm; t += NWAVE) {
        const int r0 = t << 4, ra = r0 + l16;
        const bool va = ra < M;
        int dJa;
        const long long oa = va ? stacked_row(ps, ld, I, m, ra, dJa) + (3 - dJa) * m : 0;
        double ar[MAX_BLOCK / 4], ai[MAX_BLOCK / 4];
#pragma unroll
        for (int s = 0; s < MAX_BLOCK / 4; ++s) {
            const int k = 4 * s + l4;
            const bool ok = s < tk && va && k < m;
            ar[s] = ok ? Pr[oa + k] : 0.0;
            ai[s] = (CPLX && ok) ? Pi[oa + k] : 0.0;
        }
        long long oc[4];
        bool vc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = r0 + l4 + 4 * j;
            vc[j] = r < M;
            int dJ;
            oc[j] = vc[j] ? stacked_row(ps, ld, I, m, r, dJ) + (3 - dJ) * m : 0;
        }
        for (int u = 0; u < tn; ++u) {
            const int c = (u << 4) + l16;
            const bool cv = c < m;
            d4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < MAX_BLOCK / 4; ++s) {
                if (s < tk) {
                    const int k = 4 * s + l4;
                    const bool ok = cv && k < m;
                    const double br = ok ? L.B[0][k * LS + c] : 0.0;
                    const double bi = (CPLX && ok) ? L.B[1][k * LS + c] : 0.0;
                    mfma_c(accr, acci, ar[s], ai[s], br, bi, CPLX);
                }
            }
            if (cv) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (vc[j]) {
                        Pr[oc[j] + c] = accr[j];
                        if (CPLX) Pi[oc[j] + c] = acci[j];
                    }
            }
        }
    }
}

// A_JK -= L_JI U_IK for J = I+1..I+nU and K = I+1+kb0 .. I+kb0+nkb, with U_IK
// staged in LDS (B[2 kb] re, B[2 kb + 1] im in complex arithmetic; B[kb] in real).
// L strips come from the panels into registers once per strip; the C tile of
// the next MFMA chain is loaded before the chain so its latency overlaps it.
template <bool CPLX>
__device__ void schur_pass(double* __restrict__ Pr, double* __restrict__ Pi, long long ps, int ld, int I, int m,
                           int nU, int kb0, int nkb, const LuLds& L) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int M = nU * m, tm = (M + 15) >> 4, tk = (m + 3) >> 2, tn = (m + 15) >> 4;
    for (int t = wave; t < tm; t += NWAVE) {
        const int r0 = t << 4, ra = r0 + l16;
        const bool va = ra < M;
        int dJa;
        const long long oa = va ? stacked_row(ps, ld, I, m, ra, dJa) + (3 - dJa) * m : 0;
        double ar[MAX_BLOCK / 4], ai[MAX_BLOCK / 4];
#pragma unroll
        for (int s = 0; s < MAX_BLOCK / 4; ++s) {
            const int k = 4 * s + l4;
            const bool ok = s < tk && va && k < m;
            ar[s] = ok ? Pr[oa + k] : 0.0;
            ai[s] = (CPLX && ok) ? Pi[oa + k] : 0.0;
        }
        long long oc[4];
        int dj[4];
        bool vc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = r0 + l4 + 4 * j;
            vc[j] = r < M;
            dj[j] = 1;
            oc[j] = vc[j] ? stacked_row(ps, ld, I, m, r, dj[j]) : 0;
        }
        for (int kb = 0; kb < nkb; ++kb) {
            const double* Br = CPLX ? L.B[2 * kb] : L.B[kb];
            const double* Bi = CPLX ? L.B[2 * kb + 1] : L.B[kb];
            const int kk = kb0 + kb;  // K = I + 1 + kk; column block d = kk + 4 - dJ in panel I + dJ
            for (int u = 0; u < tn; ++u) {
                const int c = (u << 4) + l16;
                const bool cv = c < m;
                double cr[4], ci[4];
                long long o[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    o[j] = oc[j] + (long long)(kk + 4 - dj[j]) * m + c;
                    const bool ok = vc[j] && cv;
                    cr[j] = ok ? Pr[o[j]] : 0.0;
                    ci[j] = (CPLX && ok) ? Pi[o[j]] : 0.0;
                }
                d4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s = 0; s < MAX_BLOCK / 4; ++s) {
                    if (s < tk) {
                        const int k = 4 * s + l4;
                        const bool ok = cv && k < m;
                        const double br = ok ? Br[k * LS + c] : 0.0;
                        const double bi = (CPLX && ok) ? Bi[k * LS + c] : 0.0;
                        mfma_c(accr, acci, ar[s], ai[s], br, bi, CPLX);
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (vc[j] && cv) {
                        Pr[o[j]] = cr[j] - accr[j];
                        if (CPLX) Pi[o[j]] = ci[j] - acci[j];
                    }
            }
        }
    }
}

// Write E = a M_s + h K_s (a = ar + i ai) into the panels from the 49-point
// stencils: a coalesced 16-byte zero fill of the panels, then one thread per
// row scatters its (at most) 49 stencil entries, whose loads are coalesced
// across threads (stencil planes are row-contiguous). Panels are 16-byte aligned.
template <bool CPLX>
__device__ void assemble_E(const double* __restrict__ Ms, const double* __restrict__ Ks, int ns, int m,
                           double ar, double ai, double h, double* __restrict__ Pr, double* __restrict__ Pi) {
    const int N = ns * m, ld = 7 * m;
    const long long tot = (long long)N * ld, half = tot >> 1;
    const double2 z = {0.0, 0.0};
    for (long long e = threadIdx.x; e < half; e += NT) {
        reinterpret_cast<double2*>(Pr)[e] = z;
        if (CPLX) reinterpret_cast<double2*>(Pi)[e] = z;
    }
    if ((tot & 1) && threadIdx.x == 0) {
        Pr[tot - 1] = 0.0;
        if (CPLX) Pi[tot - 1] = 0.0;
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += NT) {
        const int I = row / m, i = row - I * m;
        double* pr = Pr + (long long)row * ld;
        double* pi = CPLX ? Pi + (long long)row * ld : nullptr;
        for (int d = 0; d < 7; ++d) {
            const int J = I + d - 3;
            if (J < 0 || J >= ns) continue;
            double mv[7], kv[7];
#pragma unroll
            for (int q = 0; q < 7; ++q) {
                const int j = i + q - 3;
                const bool ok = j >= 0 && j < m;
                const long long o = (long long)(d * 7 + q) * N + row;
                mv[q] = ok ? Ms[o] : 0.0;
                kv[q] = ok ? Ks[o] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 7; ++q) {
                const int j = i + q - 3;
                if (j >= 0 && j < m) {
                    pr[d * m + j] = ar * mv[q] + h * kv[q];
                    if (CPLX) pi[d * m + j] = ai * mv[q];
                }
            }
        }
    }
}

// Factor the panels in place. Returns nonzero if a diagonal block was singular.
template <bool CPLX>
__device__ __noinline__ int block_lu(double* __restrict__ Pr, double* __restrict__ Pi, int ns, int m, LuLds& L) {
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
    // blocks per Schur pass: 3 U blocks fit the staging buffers in real arithmetic, 2 in complex
    constexpr int KPASS = CPLX ? 2 : 3;
    for (int I = 0; I < ns; ++I) {
        double* PIr = Pr + I * ps;
        double* PIi = Pi + I * ps;
        for (int e = tid; e < m * m; e += NT) {
            const int i = e / m, j = e - i * m;
            L.B[0][i * LS + j] = PIr[(long long)i * ld + 3 * m + j];
            if (CPLX) L.B[1][i * LS + j] = PIi[(long long)i * ld + 3 * m + j];
        }
        __syncthreads();
        lap(1);
        gj_inverse<CPLX>(L, m);
        lap(0);
        for (int e = tid; e < m * m; e += NT) {
            const int i = e / m, j = e - i * m;
            PIr[(long long)i * ld + 3 * m + j] = L.B[0][i * LS + j];
            if (CPLX) PIi[(long long)i * ld + 3 * m + j] = L.B[1][i * LS + j];
        }
        const int nU = (ns - 1 - I) < 3 ? (ns - 1 - I) : 3;
        if (nU == 0) break;
        lower_multiply<CPLX>(Pr, Pi, ps, ld, I, m, nU, L);
        __syncthreads();
        lap(2);
        for (int kb0 = 0; kb0 < nU; kb0 += KPASS) {
            const int nkb = (nU - kb0) < KPASS ? (nU - kb0) : KPASS;
            for (int kb = 0; kb < nkb; ++kb) {
                double* Br = CPLX ? L.B[2 * kb] : L.B[kb];
                double* Bi = CPLX ? L.B[2 * kb + 1] : nullptr;
                const int col = (4 + kb0 + kb) * m;
                for (int e = tid; e < m * m; e += NT) {
                    const int i = e / m, j = e - i * m;
                    Br[i * LS + j] = PIr[(long long)i * ld + col + j];
                    if (CPLX) Bi[i * LS + j] = PIi[(long long)i * ld + col + j];
                }
            }
            __syncthreads();
            lap(1);
            schur_pass<CPLX>(Pr, Pi, ps, ld, I, m, nU, kb0, nkb, L);
            __syncthreads();
            lap(3);
        }
    }
    __syncthreads();
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
