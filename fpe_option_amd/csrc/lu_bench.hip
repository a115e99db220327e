// Block-banded factorisation microbenchmark (a measurement tool, not part of
// the library): nwg workgroups (two per CU, the Radau kernel's launch shape and
// register budget) each factor their own copy of the same block diagonally
// dominant banded matrix (ns block rows of m x m blocks, half-bandwidth 3), in
// real or complex arithmetic, FP64 or FP32, with a chosen factorisation code
// path, reps times. Prints one JSON line per (arithmetic, path): mean wall time
// per factorisation and its sub-phases (pivot-block inverse, staging, L
// multiply, Schur update) in microseconds, and the largest difference of the
// factors from the first path's.
//
// usage: lu_bench m ns reps nwg path...   (path 0 generic loops, 1 compiled k-steps)
//        LU_BENCH_F32=1 for FP32 panels.
#include "blocklu.inl"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
using namespace fpe;

template <bool CPLX, class T>
__global__ __launch_bounds__(RNT, 2) void k_lu_bench(T* P, T* Pi, const T* P0, const T* Pi0, long long pstride, int ns,
                                                     int m, int var, int reps, unsigned long long* out) {
    __shared__ LuLds L;
    T* pr = P + blockIdx.x * pstride;
    T* pi = CPLX ? Pi + blockIdx.x * pstride : nullptr;
    const long long n = (long long)ns * m * 7 * m;
    if (threadIdx.x < 4) L.tk[threadIdx.x] = 0;
    unsigned long long tot = 0;
    int bad = 0;
    for (int r = 0; r < reps; ++r) {
        for (long long e = threadIdx.x; e < n; e += RNT) {
            pr[e] = P0[e];
            if (CPLX) pi[e] = Pi0[e];
        }
        __syncthreads();
        const unsigned long long t0 = wall_clock64();
        bad |= block_lu<CPLX, T>(pr, pi, ns, m, var, L);
        tot += wall_clock64() - t0;
    }
    if (threadIdx.x == 0) {
        unsigned long long* o = out + blockIdx.x * 6;
        o[0] = tot, o[1] = L.tk[0], o[2] = L.tk[1], o[3] = L.tk[2], o[4] = L.tk[3], o[5] = bad;
    }
}

#define CK(x)                                                                                     \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);                 \
            exit(1);                                                                              \
        }                                                                                         \
    } while (0)

template <class T>
static void run(int m, int ns, int reps, int nwg, const std::vector<int>& vars) {
    const int ld = 7 * m;
    const long long n = (long long)ns * m * ld;
    // panel I, block d = 0..6 holds block (I, I+d-3): diagonal 3 + x, entries within 3 of the diagonal x
    std::vector<T> h0(n, T(0)), h1(n, T(0));
    srand(7);
    for (int I = 0; I < ns; ++I)
        for (int i = 0; i < m; ++i)
            for (int d = 0; d < 7; ++d) {
                const int J = I + d - 3;
                if (J < 0 || J >= ns) continue;
                for (int j = 0; j < m; ++j) {
                    const double x = (rand() / (double)RAND_MAX - 0.5) * 0.02;
                    double v = (std::abs(i - j) <= 3) ? x : 0.0;
                    if (d == 3 && i == j) v = 3.0 + x;
                    const long long o = (long long)I * m * ld + (long long)i * ld + d * m + j;
                    h0[o] = T(v);
                    h1[o] = T((d == 3 && i == j) ? 1.0 : 0.5 * v);
                }
            }
    T *P0, *Pi0, *P, *Pi;
    unsigned long long* out;
    const long long pstride = (n + 31) / 32 * 32;
    CK(hipMalloc(&P0, n * sizeof(T)));
    CK(hipMalloc(&Pi0, n * sizeof(T)));
    CK(hipMemcpy(P0, h0.data(), n * sizeof(T), hipMemcpyHostToDevice));
    CK(hipMemcpy(Pi0, h1.data(), n * sizeof(T), hipMemcpyHostToDevice));
    CK(hipMalloc(&P, pstride * nwg * sizeof(T)));
    CK(hipMalloc(&Pi, pstride * nwg * sizeof(T)));
    CK(hipMalloc(&out, nwg * 6 * 8));
    std::vector<unsigned long long> ho(nwg * 6);
    std::vector<T> rr(n), ri(n), fr(n), fi(n);
    for (int cp = 0; cp < 2; ++cp) {
        for (size_t vi = 0; vi < vars.size(); ++vi) {
            hipEvent_t a, b;
            CK(hipEventCreate(&a));
            CK(hipEventCreate(&b));
            CK(hipEventRecord(a));
            if (cp)
                hipLaunchKernelGGL((k_lu_bench<true, T>), dim3(nwg), dim3(RNT), 0, 0, P, Pi, P0, Pi0, pstride, ns, m,
                                   vars[vi], reps, out);
            else
                hipLaunchKernelGGL((k_lu_bench<false, T>), dim3(nwg), dim3(RNT), 0, 0, P, Pi, P0, Pi0, pstride, ns, m,
                                   vars[vi], reps, out);
            CK(hipEventRecord(b));
            CK(hipEventSynchronize(b));
            float ms = 0;
            CK(hipEventElapsedTime(&ms, a, b));
            CK(hipMemcpy(ho.data(), out, nwg * 6 * 8, hipMemcpyDeviceToHost));
            CK(hipMemcpy(rr.data(), P, n * sizeof(T), hipMemcpyDeviceToHost));
            if (cp) CK(hipMemcpy(ri.data(), Pi, n * sizeof(T), hipMemcpyDeviceToHost));
            double s[6] = {0, 0, 0, 0, 0, 0};
            for (int w = 0; w < nwg; ++w)
                for (int k = 0; k < 6; ++k) s[k] += double(ho[w * 6 + k]);
            const double f = 1.0 / nwg / reps / 100.0;  // 100 MHz ticks -> us per factorisation
            double dmax = 0;
            if (vi == 0) {
                fr = rr, fi = ri;
            } else {
                for (long long e = 0; e < n; ++e)
                    dmax = std::fmax(dmax, std::fabs(double(rr[e]) - double(fr[e])) +
                                               (cp ? std::fabs(double(ri[e]) - double(fi[e])) : 0.0));
            }
            printf("{\"fp32\": %d, \"complex\": %d, \"m\": %d, \"ns\": %d, \"path\": %d, \"ms\": %.3f, "
                   "\"us_per_lu\": %.1f, \"inverse\": %.1f, \"stage\": %.1f, \"lmul\": %.1f, \"schur\": %.1f, "
                   "\"singular\": %d, \"dmax_vs_first\": %.3g}\n",
                   int(sizeof(T) == 4), cp, m, ns, vars[vi], ms, s[0] * f, s[1] * f, s[2] * f, s[3] * f, s[4] * f,
                   int(s[5]), dmax);
            fflush(stdout);
        }
    }
    CK(hipFree(P0));
    CK(hipFree(Pi0));
    CK(hipFree(P));
    CK(hipFree(Pi));
    CK(hipFree(out));
}

int main(int argc, char** argv) {
    const int m = argc > 1 ? atoi(argv[1]) : 60;
    const int ns = argc > 2 ? atoi(argv[2]) : 24;
    const int reps = argc > 3 ? atoi(argv[3]) : 4;
    const int nwg = argc > 4 ? atoi(argv[4]) : 512;
    std::vector<int> vars;
    for (int a = 5; a < argc; ++a) vars.push_back(atoi(argv[a]));
    if (vars.empty()) vars = {0, 1};
    if (m < 4 || m > MAX_BLOCK || ns < 4 || reps < 1 || nwg < 1) {
        printf("need 4 <= m <= %d, ns >= 4, reps >= 1, nwg >= 1\n", MAX_BLOCK);
        return 2;
    }
    const char* f32 = getenv("LU_BENCH_F32");
    if (f32 && atoi(f32)) run<float>(m, ns, reps, nwg, vars);
    else run<double>(m, ns, reps, nwg, vars);
    return 0;
}
