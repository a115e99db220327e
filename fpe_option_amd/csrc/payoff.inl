// Terminal-PDF payoff integration (src/server/pricer.mojo:27-64,153-155) and the
// full PDF grid for the stepwise API (src/engines/fpe/pdf.mojo:28-32).
//
// price = e^{-rT} sum_i payoff(s_i) w_s,i sum_j p(s_i, v_j) w_v,j with
// p = Bs Q Bv^T.  The v-sum is folded before expanding in s:
//   sum_j p_ij w_v,j = sum_k Bs[i,k] (Q omega)_k,  omega_l = sum_j Bv[j,l] w_v,j,
// so pricing never materialises the n_qs x n_qv grid.
#pragma once
#include "device_common.hpp"

namespace fpe {

// WM[o_w + i] = w_s,i * sum_j p_ij w_v,j for every s quadrature point.  grid: n_inst.
__global__ void __launch_bounds__(NT) k_marginal(const Inst* __restrict__ inst, const double* __restrict__ QD,
                                                 const double* __restrict__ BV, const int* __restrict__ BC,
                                                 const double* __restrict__ VC, const long long* __restrict__ o_w,
                                                 double* __restrict__ WM) {
    const Inst& in = inst[blockIdx.x];
    const int ns = in.nsb, nv = in.nvb, N = ns * nv, nqs = in.nqs, nqv = in.nqv;
    const double* sw = QD + in.o_sq + nqs;
    const double* vw = QD + in.o_vq + nqv;
    const double* Ts = BV + in.o_bs;
    const double* Tv = BV + in.o_bv;
    const int* cs = BC + in.o_bs / 8;
    const int* cv = BC + in.o_bv / 8;
    const double* q = VC + in.o_vec + (long long)V_QT * N;
    __shared__ double om[320], qo[320];
    for (int l = threadIdx.x; l < nv; l += NT) {
        double a = 0.0;
        const int q0 = lower_c0(cv, nqv, l - 3), q1 = lower_c0(cv, nqv, l + 1);
        for (int j = q0; j < q1; ++j) a += Tv[(long long)j * 8 + (l - cv[j])] * vw[j];
        om[l] = a;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ns; k += NT) {
        double a = 0.0;
        for (int l = 0; l < nv; ++l) a += q[(long long)k * nv + l] * om[l];
        qo[k] = a;
    }
    __syncthreads();
    double* w = WM + o_w[blockIdx.x];
    for (int i = threadIdx.x; i < nqs; i += NT) {
        double a = 0.0;
        for (int t = 0; t < 4; ++t) {
            const int k = cs[i] + t;
            if (k < ns) a += Ts[(long long)i * 8 + t] * qo[k];
        }
        w[i] = a * sw[i];
    }
}

struct PriceReq {
    int inst;        // solve instance
    int option_type; // payoff type of that solve
    double barrier, strike, sign;
    int out;         // output index (accumulated with sign)
    int pad;
};

__device__ __forceinline__ double payoff_dev(int t, double B, double K, double S) {
    bool active = true;
    if (t == 0 || t == 1) active = S <= B;
    else if (t == 2 || t == 3) active = S > B;
    else if (t == 4 || t == 5) active = S >= B;
    else if (t == 6 || t == 7) active = S < B;
    if (!active) return 0.0;
    const double v = (t % 2 == 0) ? S - K : K - S;
    return v > 0.0 ? v : 0.0;
}

// One workgroup per pricing request; writes the signed discounted price into
// part[req] (in-options combine two requests on the host, in a fixed order).
__global__ void __launch_bounds__(NT) k_price(const Inst* __restrict__ inst, const double* __restrict__ QD,
                                              const long long* __restrict__ o_w, const double* __restrict__ WM,
                                              const PriceReq* __restrict__ req, double* __restrict__ part) {
    const PriceReq R = req[blockIdx.x];
    const Inst& in = inst[R.inst];
    const int nqs = in.nqs;
    const double* sph = QD + in.o_sq + 2 * nqs;
    const double* w = WM + o_w[R.inst];
    __shared__ double red[NWAVE];
    double a = 0.0;
    for (int i = threadIdx.x; i < nqs; i += NT) a += payoff_dev(R.option_type, R.barrier, R.strike, sph[i]) * w[i];
    a = block_sum(a, red);
    if (threadIdx.x == 0) part[blockIdx.x] = exp(-in.r * in.T) * a;
}

// Full PDF grid p = Bs Q Bv^T (n_qs x n_qv row-major) for one instance.
// grid: (row chunks).
__global__ void __launch_bounds__(NT) k_pdf(const Inst* __restrict__ inst, int id, const double* __restrict__ BV,
                                            const int* __restrict__ BC, const double* __restrict__ q,
                                            double* __restrict__ pdf) {
    const Inst& in = inst[id];
    const int nv = in.nvb, nqs = in.nqs, nqv = in.nqv, ns = in.nsb;
    const double* Ts = BV + in.o_bs;
    const double* Tv = BV + in.o_bv;
    const int* cs = BC + in.o_bs / 8;
    const int* cv = BC + in.o_bv / 8;
    const long long tot = (long long)nqs * nqv;
    for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < tot; e += (long long)gridDim.x * NT) {
        const int i = int(e / nqv), j = int(e - (long long)i * nqv);
        double a = 0.0;
        for (int t = 0; t < 4; ++t) {
            const int k = cs[i] + t;
            if (k >= ns) continue;
            double wv = 0.0;
            for (int u = 0; u < 4; ++u) {
                const int l = cv[j] + u;
                if (l < nv) wv += Tv[(long long)j * 8 + u] * q[(long long)k * nv + l];
            }
            a += Ts[(long long)i * 8 + t] * wv;
        }
        pdf[e] = a;
    }
}

}  // namespace fpe
