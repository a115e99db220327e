// Surrogate kernels: a multilayer perceptron on MFMA, with BF16 operands or
// FP32 operands and FP32 accumulation (forward, backward, Adam, loss), and pricing from POD
// coefficients of the terminal density (coefficients -> s-marginal -> payoff).
// Nothing here has a counterpart in the reference engine; the plain
// Fokker-Planck solve is the surrogate's fallback and its accuracy reference.
//
// Lane layouts (checked on gfx950 with an in-bounds probe):
//   v_mfma_f32_16x16x16_bf16: a[j] = A[l%16][4(l/16)+j], b[j] = B[4(l/16)+j][l%16],
//                             acc[j] = C[4(l/16)+j][l%16]
//   v_mfma_f32_16x16x4_f32:   a = A[l%16][l/16], b = B[l/16][l%16], acc[j] = C[4(l/16)+j][l%16]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/fpe_engine.h"
#include "engine.hpp"

namespace fpe_sur {

constexpr int NT = 256;
constexpr int TILE = 64;  // workgroup output tile (4 waves x 16 rows, 64 columns)
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even; NaN stays NaN.
__device__ __forceinline__ unsigned short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// One wave: acc[t] = C[r0 .. r0+15][c0 + 16t .. c0 + 16t + 15] for t = 0..3,
// C = A B with inner dimension K. A(m, k) and B(k, n) return BF16 bit patterns.
template <class FA, class FB>
__device__ __forceinline__ void wave_gemm(int r0, int c0, int M, int N, int K, FA A, FB B, f4 acc[4]) {
    const int l = threadIdx.x & 63, lr = l & 15, lk = (l >> 4) * 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const int m = r0 + lr;
    for (int k0 = 0; k0 < K; k0 += 16) {
        s4 a;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + lk + j;
            a[j] = (m < M && k < K) ? (short)A(m, k) : (short)0;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int n = c0 + 16 * t + lr;
            s4 b;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + lk + j;
                b[j] = (n < N && k < K) ? (short)B(k, n) : (short)0;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
}

// FP32 operands: acc[t] = C[r0 .. r0+15][c0 + 16t .. c0 + 16t + 15] for t = 0..3,
// C = A B with inner dimension K on v_mfma_f32_16x16x4f32 (A(m, k), B(k, n)
// return floats).
template <class FA, class FB>
__device__ __forceinline__ void wave_gemm_f32(int r0, int c0, int M, int N, int K, FA A, FB B, f4 acc[4]) {
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const int m = r0 + lr;
    for (int k0 = 0; k0 < K; k0 += 4) {
        const int k = k0 + lk;
        const float a = (m < M && k < K) ? A(m, k) : 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int n = c0 + 16 * t + lr;
            const float b = (n < N && k < K) ? B(k, n) : 0.0f;
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
        }
    }
}

// Weight element n, k as a BF16 bit pattern (BF16 path) or a float (FP32 path).
template <bool F32>
struct WeightT;
template <>
struct WeightT<false> {
    typedef unsigned short type;
};
template <>
struct WeightT<true> {
    typedef float type;
};

// C = A B over the BF16 operands of the BF16 path (X and dZ rounded to BF16) or
// the FP32 operands of the FP32 path.
template <bool F32, class FA, class FB>
__device__ __forceinline__ void wave_gemm_p(int r0, int c0, int M, int N, int K, FA A, FB B, f4 acc[4]) {
    if constexpr (F32)
        wave_gemm_f32(r0, c0, M, N, K, A, B, acc);
    else
        wave_gemm(
            r0, c0, M, N, K, [&](int m, int k) { return f2bf(A(m, k)); }, [&](int k, int n) { return f2bf(B(k, n)); },
            acc);
}
template <bool F32>
__device__ __forceinline__ float wval(const typename WeightT<F32>::type* W, long long i) {
    if constexpr (F32)
        return W[i];
    else
        return __uint_as_float(unsigned(W[i]) << 16);  // exact: BF16 -> FP32 -> BF16 round trips
}

// Z = X W^T + b, A = act(Z). grid (ceil(M/64), ceil(N/64)).
template <bool F32>
__global__ void __launch_bounds__(NT) k_lin_fwd(const float* __restrict__ X, int M, int K,
                                                const typename WeightT<F32>::type* __restrict__ W,
                                                const float* __restrict__ bias, int N, int act, float* __restrict__ Z,
                                                float* __restrict__ Aout) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = blockIdx.x * TILE + 16 * w, c0 = blockIdx.y * TILE;
    f4 acc[4];
    wave_gemm_p<F32>(
        r0, c0, M, N, K, [&](int m, int k) { return X[(long long)m * K + k]; },
        [&](int k, int n) { return wval<F32>(W, (long long)n * K + k); }, acc);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = c0 + 16 * t + (l & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = r0 + 4 * (l >> 4) + j;
            if (m < M && n < N) {
                const float z = acc[t][j] + bias[n];
                if (Z) Z[(long long)m * N + n] = z;
                if (Aout) Aout[(long long)m * N + n] = act ? z * sigmoid(z) : z;
            }
        }
    }
}

// dX = dZ W (M x K), times silu'(Zp) when Zp != nullptr. grid (ceil(M/64), ceil(K/64)).
template <bool F32>
__global__ void __launch_bounds__(NT) k_lin_bwd_data(const float* __restrict__ dZ,
                                                     const typename WeightT<F32>::type* __restrict__ W, int M, int N,
                                                     int K, const float* __restrict__ Zp, float* __restrict__ dX) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = blockIdx.x * TILE + 16 * w, c0 = blockIdx.y * TILE;
    f4 acc[4];
    wave_gemm_p<F32>(
        r0, c0, M, K, N, [&](int m, int n) { return dZ[(long long)m * N + n]; },
        [&](int n, int k) { return wval<F32>(W, (long long)n * K + k); }, acc);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int k = c0 + 16 * t + (l & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = r0 + 4 * (l >> 4) + j;
            if (m < M && k < K) {
                float g = acc[t][j];
                if (Zp) {
                    const float z = Zp[(long long)m * K + k], s = sigmoid(z);
                    g *= s * (1.0f + z * (1.0f - s));
                }
                dX[(long long)m * K + k] = g;
            }
        }
    }
}

// dW = dZ^T X (N x K). grid (ceil(N/64), ceil(K/64)).
template <bool F32>
__global__ void __launch_bounds__(NT) k_lin_bwd_weight(const float* __restrict__ dZ, const float* __restrict__ X,
                                                       int M, int N, int K, float* __restrict__ dW) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = blockIdx.x * TILE + 16 * w, c0 = blockIdx.y * TILE;
    f4 acc[4];
    wave_gemm_p<F32>(
        r0, c0, N, K, M, [&](int n, int m) { return dZ[(long long)m * N + n]; },
        [&](int m, int k) { return X[(long long)m * K + k]; }, acc);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int k = c0 + 16 * t + (l & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = r0 + 4 * (l >> 4) + j;
            if (n < N && k < K) dW[(long long)n * K + k] = acc[t][j];
        }
    }
}

// db[n] = sum_m dZ[m][n]. grid ceil(N/256); fixed summation order.
__global__ void __launch_bounds__(NT) k_colsum(const float* __restrict__ dZ, int M, int N, float* __restrict__ db) {
    const int n = blockIdx.x * NT + threadIdx.x;
    if (n >= N) return;
    float a = 0.0f;
    for (int m = 0; m < M; ++m) a += dZ[(long long)m * N + n];
    db[n] = a;
}

__global__ void __launch_bounds__(NT) k_adam(float* __restrict__ W, unsigned short* __restrict__ Wb,
                                             const float* __restrict__ G, float* __restrict__ m1,
                                             float* __restrict__ m2, long long n, float lr, float b1, float b2,
                                             float eps, float c1, float c2) {
    for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT) {
        const float g = G[i];
        const float a = b1 * m1[i] + (1.0f - b1) * g;
        const float b = b2 * m2[i] + (1.0f - b2) * g * g;
        m1[i] = a, m2[i] = b;
        const float wv = W[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
        W[i] = wv;
        if (Wb) Wb[i] = f2bf(wv);
    }
}

__global__ void __launch_bounds__(NT) k_to_bf16(const float* __restrict__ W, unsigned short* __restrict__ Wb,
                                                long long n) {
    for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < n; i += (long long)gridDim.x * NT)
        Wb[i] = f2bf(W[i]);
}

// dY = 2 wcol (Y - T) / (M N); partial[block] = sum of wcol (Y - T)^2 / (M N).
__global__ void __launch_bounds__(NT) k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                            const float* __restrict__ wcol, int M, int N, float* __restrict__ dY,
                                            float* __restrict__ partial) {
    const long long tot = (long long)M * N;
    const float sc = 1.0f / float(tot);
    float a = 0.0f;
    for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < tot; i += (long long)gridDim.x * NT) {
        const float d = Y[i] - T[i], wc = wcol ? wcol[i % N] : 1.0f;
        dY[i] = 2.0f * wc * d * sc;
        a += wc * d * d * sc;
    }
    __shared__ float red[NT];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Prices from normalised coefficients: one workgroup per 64 options, a wave per
// 16. The marginal tile (16 options x 16 s points) comes from FP32 MFMA over the
// coefficients; each lane folds its entries into per-option payoff sums.
__global__ void __launch_bounds__(NT) k_sur_price(const float* __restrict__ C, int n, int nc,
                                                  const float* __restrict__ cscale, const float* __restrict__ cshift,
                                                  const float* __restrict__ Pt, const float* __restrict__ w0,
                                                  const float* __restrict__ sq, int nq,
                                                  const double* __restrict__ strike, const int* __restrict__ type,
                                                  const double* __restrict__ rr, const double* __restrict__ TT,
                                                  double* __restrict__ prices) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, lr = l & 15, lg = l >> 4;
    const int o0 = blockIdx.x * TILE + 16 * w;
    const int ns = nc >> 2;
    float a[16];
    {
        const int o = o0 + lr;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int k = 4 * s + lg;
            a[s] = (s < ns && o < n) ? C[(long long)o * nc + k] * cscale[k] + cshift[k] : 0.0f;
        }
    }
    double Ko[4];
    int to[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int o = o0 + 4 * lg + j;
        Ko[j] = o < n ? strike[o] : 0.0;
        to[j] = o < n ? type[o] : 8;
    }
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (int u = 0; u < nq; u += 16) {
        const int i = u + lr;
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            if (s < ns) {
                const float b = i < nq ? Pt[(long long)(4 * s + lg) * nq + i] : 0.0f;
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], b, acc, 0, 0, 0);
            }
        }
        if (i < nq) {
            const double S = sq[i], wi = w0[i];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double v = (to[j] == 8) ? S - Ko[j] : Ko[j] - S;
                if (v > 0.0) sum[j] += v * (double(acc[j]) + wi);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) sum[j] += __shfl_xor(sum[j], o, 64);
        const int opt = o0 + 4 * lg + j;
        if (lr == 0 && opt < n) prices[opt] = exp(-rr[opt] * TT[opt]) * sum[j];
    }
}

// dst[i][:] = src[idx[i]][:] (float rows of `width`).
__global__ void __launch_bounds__(NT) k_gather(const float* __restrict__ src, const int* __restrict__ idx, int n,
                                               int width, float* __restrict__ dst) {
    const long long tot = (long long)n * width;
    for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < tot; e += (long long)gridDim.x * NT) {
        const int i = int(e / width), c = int(e - (long long)i * width);
        dst[e] = src[(long long)idx[i] * width + c];
    }
}

int grid1(long long n) {
    long long b = (n + NT - 1) / NT;
    return int(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

int32_t check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fpe::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
        return 2;
    }
    return 0;
}

int32_t bad(const std::string& msg) {
    fpe::set_last_error(msg);
    return 1;
}

template <bool F32, class WT>
int32_t mlp_forward(const float* X, int32_t M, int32_t K, const WT* W, const float* b, int32_t N, int32_t act,
                    float* Z, float* A, void* stream, const char* name) {
    if (M < 0 || K <= 0 || N <= 0 || !X || !W || !b) return bad(std::string(name) + ": invalid arguments");
    if (M == 0) return 0;
    dim3 g((M + TILE - 1) / TILE, (N + TILE - 1) / TILE);
    hipLaunchKernelGGL(k_lin_fwd<F32>, g, dim3(NT), 0, (hipStream_t)stream, X, M, K, W, b, N, act, Z, A);
    return check(name);
}

template <bool F32, class WT>
int32_t mlp_backward(const float* dZ, const float* X, const WT* W, int32_t M, int32_t N, int32_t K, float* dW,
                     float* db, float* dX, const float* Zp, void* stream, const char* name) {
    if (M <= 0 || K <= 0 || N <= 0 || !dZ || !X || !W || !dW || !db)
        return bad(std::string(name) + ": invalid arguments");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_lin_bwd_weight<F32>, dim3((N + TILE - 1) / TILE, (K + TILE - 1) / TILE), dim3(NT), 0, st, dZ,
                       X, M, N, K, dW);
    hipLaunchKernelGGL(k_colsum, dim3((N + NT - 1) / NT), dim3(NT), 0, st, dZ, M, N, db);
    if (dX)
        hipLaunchKernelGGL(k_lin_bwd_data<F32>, dim3((M + TILE - 1) / TILE, (K + TILE - 1) / TILE), dim3(NT), 0, st,
                           dZ, W, M, N, K, Zp, dX);
    return check(name);
}
}  // namespace fpe_sur

using namespace fpe_sur;

extern "C" {


int32_t fpe_mlp_forward(const float* X, int32_t M, int32_t K, const uint16_t* Wb, const float* b, int32_t N,
                        int32_t act, float* Z, float* A, void* stream) {
    return mlp_forward<false>(X, M, K, reinterpret_cast<const unsigned short*>(Wb), b, N, act, Z, A, stream,
                              "fpe_mlp_forward");
}

int32_t fpe_mlp_backward(const float* dZ, const float* X, const uint16_t* Wb, int32_t M, int32_t N, int32_t K,
                         float* dW, float* db, float* dX, const float* Zp, void* stream) {
    return mlp_backward<false>(dZ, X, reinterpret_cast<const unsigned short*>(Wb), M, N, K, dW, db, dX, Zp, stream,
                               "fpe_mlp_backward");
}

int32_t fpe_mlp_forward_f32(const float* X, int32_t M, int32_t K, const float* W, const float* b, int32_t N,
                            int32_t act, float* Z, float* A, void* stream) {
    return mlp_forward<true>(X, M, K, W, b, N, act, Z, A, stream, "fpe_mlp_forward_f32");
}

int32_t fpe_mlp_backward_f32(const float* dZ, const float* X, const float* W, int32_t M, int32_t N, int32_t K,
                             float* dW, float* db, float* dX, const float* Zp, void* stream) {
    return mlp_backward<true>(dZ, X, W, M, N, K, dW, db, dX, Zp, stream, "fpe_mlp_backward_f32");
}

int32_t fpe_mlp_adam(float* W, uint16_t* Wb, const float* G, float* m, float* v, int64_t n, float lr, float beta1,
                     float beta2, float eps, int32_t t, void* stream) {
    if (n <= 0 || !W || !G || !m || !v || t < 1) return bad("fpe_mlp_adam: invalid arguments");
    const float c1 = 1.0f / (1.0f - powf(beta1, float(t))), c2 = 1.0f / (1.0f - powf(beta2, float(t)));
    hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(NT), 0, (hipStream_t)stream, W,
                       reinterpret_cast<unsigned short*>(Wb), G, m, v, (long long)n, lr, beta1, beta2, eps, c1, c2);
    return check("k_adam");
}

int32_t fpe_mlp_mse(const float* Y, const float* T, const float* wcol, int32_t M, int32_t N, float* dY,
                    float* partial, void* stream) {
    if (M <= 0 || N <= 0 || !Y || !T || !dY || !partial) return bad("fpe_mlp_mse: invalid arguments");
    // always 1024 blocks, so every partial entry is written
    hipLaunchKernelGGL(k_mse, dim3(1024), dim3(NT), 0, (hipStream_t)stream, Y, T, wcol, M, N, dY, partial);
    return check("k_mse");
}

int32_t fpe_mlp_gather_rows(const float* src, const int32_t* idx, int32_t n, int32_t width, float* dst,
                            void* stream) {
    if (n < 0 || width <= 0 || !src || !idx || !dst) return bad("fpe_mlp_gather_rows: invalid arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_gather, dim3(grid1((long long)n * width)), dim3(NT), 0, (hipStream_t)stream, src, idx, n,
                       width, dst);
    return check("k_gather");
}

void* fpe_device_alloc(int64_t bytes) {
    void* p = nullptr;
    if (bytes <= 0) {
        fpe::set_last_error("fpe_device_alloc: size must be positive");
        return nullptr;
    }
    const hipError_t e = hipMalloc(&p, size_t(bytes));
    if (e != hipSuccess) {
        fpe::set_last_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return nullptr;
    }
    return p;
}

void fpe_device_free(void* p) {
    if (p) (void)hipFree(p);
}

int32_t fpe_device_upload(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return bad("fpe_device_upload: invalid arguments");
    const hipError_t e = hipMemcpy(dst, src, size_t(bytes), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bad(std::string("hipMemcpy: ") + hipGetErrorString(e));
    return 0;
}

int32_t fpe_device_download(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return bad("fpe_device_download: invalid arguments");
    const hipError_t e = hipMemcpy(dst, src, size_t(bytes), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return bad(std::string("hipMemcpy: ") + hipGetErrorString(e));
    return 0;
}

int32_t fpe_device_synchronize(void) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return bad(std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
    return 0;
}

int32_t fpe_mlp_to_bf16(const float* W, uint16_t* Wb, int64_t n, void* stream) {
    if (n <= 0 || !W || !Wb) return bad("fpe_mlp_to_bf16: invalid arguments");
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1(n)), dim3(NT), 0, (hipStream_t)stream, W,
                       reinterpret_cast<unsigned short*>(Wb), (long long)n);
    return check("k_to_bf16");
}

int32_t fpe_surrogate_price(const float* C, int32_t n, int32_t ncoef, const float* cscale, const float* cshift,
                            const float* Pt, const float* w0, const float* sq, int32_t nq, const double* strike,
                            const int32_t* type, const double* r, const double* T, double* prices, void* stream) {
    if (n < 0 || ncoef <= 0 || ncoef > 64 || (ncoef & 3) || nq <= 0 || !C || !cscale || !cshift || !Pt || !w0 ||
        !sq || !strike || !type || !r || !T || !prices)
        return bad("fpe_surrogate_price: invalid arguments (ncoef must be a multiple of 4, at most 64)");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_sur_price, dim3((n + TILE - 1) / TILE), dim3(NT), 0, (hipStream_t)stream, C, n, ncoef,
                       cscale, cshift, Pt, w0, sq, nq, strike, type, r, T, prices);
    return check("k_sur_price");
}

}  // extern "C"
