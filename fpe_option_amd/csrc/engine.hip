// Host orchestration of the batched GPU pricing path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "assemble.inl"
#include "engine.hpp"
#include "ic.inl"
#include "payoff.inl"
#include "radau.inl"

namespace fpe {
namespace {

#define HIP_OK(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                       \
            return false;                                                                  \
        }                                                                                  \
    } while (0)

// Grow-only device buffer.
struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes, std::string& err) {
        if (bytes <= cap) return true;
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max(bytes, size_t(256));
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            err = std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e);
            p = nullptr;
            return false;
        }
        cap = want;
        return true;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct DeviceState {
    bool init = false;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8];
    DBuf inst, out, kn, qd, bv, bc, fc, st, vc, eg, hs, wm, ow, ws, wi, req, part, queue, order, pdf;
};

std::mutex g_mu;
DeviceState g_dev;

bool dev_init(std::string& err) {
    if (g_dev.init) return true;
    int n = 0;
    HIP_OK(hipGetDeviceCount(&n));
    if (n <= 0) {
        err = "no HIP device";
        return false;
    }
    HIP_OK(hipSetDevice(0));
    HIP_OK(hipDeviceGetAttribute(&g_dev.n_cu, hipDeviceAttributeMultiprocessorCount, 0));
    HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_admm), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(ADMM_LDS_MAX)));
    HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_eig), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(EIG_LDS_S)));
    HIP_OK(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
    for (auto& e : g_dev.ev) HIP_OK(hipEventCreate(&e));
    g_dev.init = true;
    return true;
}

template <class T>
bool upload(DBuf& b, const std::vector<T>& h, std::string& err) {
    if (!b.ensure(h.size() * sizeof(T), err)) return false;
    if (!h.empty()) HIP_OK(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, g_dev.stream));
    return true;
}

}  // namespace

bool grid_fits(const Domain& d, std::string* why) {
    if (d.n_vb < 4 || d.n_vb > MAX_BLOCK || d.n_sb < 4 || d.n_sb > MAX_SB) {
        if (why)
            *why = "grid outside the kernels' limits (n_vb in [4, " + std::to_string(MAX_BLOCK) + "], n_sb in [4, " +
                   std::to_string(MAX_SB) + "]): n_sb=" + std::to_string(d.n_sb) + " n_vb=" + std::to_string(d.n_vb);
        return false;
    }
    if (admm_lds_bytes(d.n_sb, d.n_vb, false) > ADMM_LDS_MAX) {
        if (why) *why = "grid too large for the initial-condition kernel's LDS budget";
        return false;
    }
    return true;
}

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

bool run_batch(const std::vector<SolveSpec>& solves, const std::vector<Domain>& doms, const std::vector<Request>& reqs,
               int n_out, const RunOptions& opt, std::vector<double>& out_prices, std::vector<SolveResult>& results,
               BatchTimes& times, std::string& err) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto t_start = std::chrono::steady_clock::now();
    if (!dev_init(err)) return false;
    const int n = int(solves.size());
    results.assign(n, SolveResult());
    out_prices.assign(n_out, 0.0);
    if (n == 0) return true;

    // ---- pack host arrays
    std::vector<Inst> hi(n);
    std::vector<double> kn, qd;
    std::vector<long long> ow(n);
    long long o_bv = 0, o_fc = 0, o_st = 0, o_vc = 0, o_eg = 0, o_hs = 0, o_wm = 0;
    long long max_panel = 0, max_N = 0, max_nm = 0;
    int max_m = 0;
    bool any_direct = false;
    for (int s = 0; s < n; ++s) {
        const SolveSpec& sp = solves[s];
        const Domain& d = doms[s];
        Inst& I = hi[s];
        std::memset(&I, 0, sizeof(I));
        I.nsb = d.n_sb, I.nvb = d.n_vb;
        I.nqs = int(d.s_x.size()), I.nqv = int(d.v_x.size());
        I.nsk = int(d.s_knots.size()), I.nvk = int(d.v_knots.size());
        if (!grid_fits(d, &err)) return false;  // callers screen grids per option first
        if (opt.given_qT && (*opt.given_qT)[s].size() != size_t(I.nsb) * I.nvb) {
            err = "given terminal state has the wrong length";
            return false;
        }
        I.s_lneu = sp.s_left_neumann, I.s_rneu = sp.s_right_neumann;
        I.cmode = sp.complex_mode, I.prec = sp.precision, I.exact = sp.exact_constants;
        I.hist_cap = opt.hist_cap;
        I.o_sk = (long long)kn.size();
        kn.insert(kn.end(), d.s_knots.begin(), d.s_knots.end());
        I.o_vk = (long long)kn.size();
        kn.insert(kn.end(), d.v_knots.begin(), d.v_knots.end());
        I.o_sq = (long long)qd.size();
        qd.insert(qd.end(), d.s_x.begin(), d.s_x.end());
        qd.insert(qd.end(), d.s_w.begin(), d.s_w.end());
        qd.insert(qd.end(), d.s_phys.begin(), d.s_phys.end());
        I.o_vq = (long long)qd.size();
        qd.insert(qd.end(), d.v_x.begin(), d.v_x.end());
        qd.insert(qd.end(), d.v_w.begin(), d.v_w.end());
        I.o_bs = o_bv;
        o_bv += 8LL * I.nqs;
        I.o_bv = o_bv;
        o_bv += 8LL * I.nqv;
        const long long N = (long long)I.nsb * I.nvb;
        I.o_fs = o_fc;
        o_fc += 5LL * I.nsb * 7;
        I.o_fv = o_fc;
        o_fc += 6LL * I.nvb * 7;
        I.o_st = o_st;
        o_st += 98 * N;
        I.o_vec = o_vc;
        o_vc += (long long)V_COUNT * N;
        I.o_eig = o_eg;
        o_eg += 2LL * I.nsb * I.nsb + 3LL * I.nvb * I.nvb + I.nsb + I.nvb;
        I.o_hist = o_hs;
        o_hs += (long long)opt.hist_cap * N;
        ow[s] = o_wm;
        o_wm += I.nqs;
        const double jac = sp.s_hi - sp.s_lo;
        I.kappa = sp.kappa, I.theta = sp.theta, I.eta = sp.eta, I.rho = sp.rho, I.r = sp.r, I.T = sp.T;
        I.S0 = sp.S0, I.V0 = sp.V0, I.jac = jac, I.s_lo = sp.s_lo;
        const double sigma0 = 0.1;
        I.ic_sig_s = sp.ic_sigma_s > 0 ? sp.ic_sigma_s : sigma0;
        I.ic_sig_v = sp.ic_sigma_v > 0 ? sp.ic_sigma_v : (jac > 0 ? sigma0 / jac : sigma0);
        I.rtol = sp.rtol > 0 ? sp.rtol : 1e-4;
        I.atol = sp.atol > 0 ? sp.atol : 1e-6;
        max_panel = std::max(max_panel, 7LL * I.nsb * I.nvb * I.nvb);
        max_N = std::max(max_N, N);
        max_nm = std::max(max_nm, N);
        max_m = std::max(max_m, I.nvb);
        any_direct |= (sp.complex_mode == 1 || sp.precision == 1);
    }
    max_panel = (max_panel + 31) & ~31LL;  // panels start 256-byte aligned (16-byte stores in assembly)
    SlotLayout L{};
    L.o_p = 0;
    L.o_cr = L.o_p + max_panel;
    L.o_ci = L.o_cr + (any_direct ? max_panel : 0);
    L.o_t = L.o_ci + (any_direct ? max_panel : 0);
    L.o_v = L.o_t + 12LL * max_m * max_m;
    L.stride = L.o_v + (long long)S_NVEC * max_N;
    L.stride = (L.stride + 31) & ~31LL;
    L.istride = ((2 * max_nm) + 63) & ~63LL;
    L.has_complex = any_direct;
    const int n_slots = std::min(n, std::max(1, RADAU_WG_PER_CU * g_dev.n_cu));
    times.n_slots = n_slots;
    // Radau work queue order: largest estimated factorisation cost (n_sb n_vb^3) first,
    // so the slots finish together instead of waiting on a late large instance.
    std::vector<int> order(n);
    for (int s = 0; s < n; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const double ca = double(hi[a].nsb) * hi[a].nvb * hi[a].nvb * hi[a].nvb;
        const double cb = double(hi[b].nsb) * hi[b].nvb * hi[b].nvb * hi[b].nvb;
        return ca > cb;
    });

    std::vector<PriceReq> hr(reqs.size());
    for (size_t k = 0; k < reqs.size(); ++k) {
        hr[k].inst = reqs[k].solve, hr[k].option_type = reqs[k].option_type;
        hr[k].barrier = reqs[k].barrier, hr[k].strike = reqs[k].strike, hr[k].sign = reqs[k].sign;
        hr[k].out = reqs[k].out;
    }

    // ---- device buffers
    DeviceState& G = g_dev;
    hipStream_t st = G.stream;
    if (!upload(G.inst, hi, err) || !upload(G.kn, kn, err) || !upload(G.qd, qd, err) || !upload(G.ow, ow, err) ||
        !upload(G.req, hr, err) || !upload(G.order, order, err))
        return false;
    if (!G.out.ensure(sizeof(InstOut) * n, err) || !G.bv.ensure(sizeof(double) * o_bv, err) ||
        !G.bc.ensure(sizeof(int) * (o_bv / 8 + 1), err) || !G.fc.ensure(sizeof(double) * o_fc, err) ||
        !G.st.ensure(sizeof(double) * o_st, err) || !G.vc.ensure(sizeof(double) * o_vc, err) ||
        !G.eg.ensure(sizeof(double) * o_eg, err) || !G.hs.ensure(sizeof(double) * std::max(o_hs, 1LL), err) ||
        !G.wm.ensure(sizeof(double) * o_wm, err) || !G.ws.ensure(sizeof(double) * L.stride * n_slots, err) ||
        !G.wi.ensure(sizeof(int) * L.istride * n_slots, err) ||
        !G.part.ensure(sizeof(double) * std::max<size_t>(reqs.size(), 1), err) || !G.queue.ensure(sizeof(int), err))
        return false;
    HIP_OK(hipMemsetAsync(G.out.p, 0, sizeof(InstOut) * n, st));
    HIP_OK(hipMemsetAsync(G.queue.p, 0, sizeof(int), st));
    const Inst* dI = G.inst.as<Inst>();
    auto t_setup = std::chrono::steady_clock::now();
    times.setup = std::chrono::duration<double, std::milli>(t_setup - t_start).count();

    // ---- kernels
    // given_qT: price (and PDF) from terminal states the caller already holds; only
    // the basis tables are rebuilt, then the states are copied into their slots.
    const bool given = opt.given_qT != nullptr;
    const bool solved = !opt.stop_after_ic;
    HIP_OK(hipEventRecord(G.ev[0], st));
    hipLaunchKernelGGL(k_basis, dim3(n, 2), dim3(NT), 0, st, dI, G.kn.as<double>(), G.qd.as<double>(),
                       G.bv.as<double>(), G.bc.as<int>());
    if (!given) {
        hipLaunchKernelGGL(k_factors, dim3(n, 2), dim3(NT), 0, st, dI, G.qd.as<double>(), G.bv.as<double>(),
                           G.bc.as<int>(), G.fc.as<double>());
        const int chunks = int(std::min<long long>(64, (max_N + NT - 1) / NT));
        hipLaunchKernelGGL(k_stencil, dim3(n, chunks), dim3(NT), 0, st, dI, G.fc.as<double>(), G.st.as<double>(),
                           G.vc.as<double>());
    } else {
        for (int s = 0; s < n; ++s) {
            const long long N = (long long)hi[s].nsb * hi[s].nvb;
            HIP_OK(hipMemcpyAsync(G.vc.as<double>() + hi[s].o_vec + V_QT * N, (*opt.given_qT)[s].data(),
                                  sizeof(double) * N, hipMemcpyHostToDevice, st));
        }
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(G.ev[1], st));
    if (!given) {
        hipLaunchKernelGGL(k_ic_rhs, dim3(n), dim3(NT), 0, st, dI, G.order.as<int>(), G.qd.as<double>(),
                           G.bv.as<double>(), G.bc.as<int>(), G.vc.as<double>());
        hipLaunchKernelGGL(k_eig, dim3(n), dim3(ENT), EIG_LDS_S, st, dI, G.order.as<int>(), 0, EIG_LDS_S,
                           G.fc.as<double>(), G.eg.as<double>());
        hipLaunchKernelGGL(k_eig, dim3(n), dim3(ENT), EIG_LDS_V, st, dI, G.order.as<int>(), 1, EIG_LDS_V,
                           G.fc.as<double>(), G.eg.as<double>());
        hipLaunchKernelGGL(k_admm, dim3(n), dim3(ANT), ADMM_LDS_MAX, st, dI, G.order.as<int>(), G.st.as<double>(),
                           G.eg.as<double>(), G.vc.as<double>(), G.out.as<InstOut>());
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipEventRecord(G.ev[2], st));
    if (solved && !given) {
        hipLaunchKernelGGL(k_radau, dim3(n_slots), dim3(RNT), 0, st, dI, n, G.queue.as<int>(), G.order.as<int>(),
                           G.st.as<double>(), G.vc.as<double>(), G.ws.as<double>(), G.wi.as<int>(), L,
                           G.hs.as<double>(), G.out.as<InstOut>());
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipEventRecord(G.ev[3], st));
    if (solved && !reqs.empty()) {
        hipLaunchKernelGGL(k_marginal, dim3(n), dim3(NT), 0, st, dI, G.qd.as<double>(), G.bv.as<double>(),
                           G.bc.as<int>(), G.vc.as<double>(), G.ow.as<long long>(), G.wm.as<double>());
        hipLaunchKernelGGL(k_price, dim3(int(reqs.size())), dim3(NT), 0, st, dI, G.qd.as<double>(),
                           G.ow.as<long long>(), G.wm.as<double>(), G.req.as<PriceReq>(), G.part.as<double>());
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipEventRecord(G.ev[4], st));
    std::vector<InstOut> ho(n);
    std::vector<double> parts(reqs.size());
    HIP_OK(hipMemcpyAsync(ho.data(), G.out.p, sizeof(InstOut) * n, hipMemcpyDeviceToHost, st));
    if (!reqs.empty() && !opt.stop_after_ic)
        HIP_OK(hipMemcpyAsync(parts.data(), G.part.p, sizeof(double) * reqs.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, G.ev[0], G.ev[1]));
    times.assemble = ms;
    HIP_OK(hipEventElapsedTime(&ms, G.ev[1], G.ev[2]));
    times.ic = ms;
    HIP_OK(hipEventElapsedTime(&ms, G.ev[2], G.ev[3]));
    times.radau = ms;
    HIP_OK(hipEventElapsedTime(&ms, G.ev[3], G.ev[4]));
    times.payoff = ms;
    for (size_t k = 0; k < reqs.size(); ++k) out_prices[reqs[k].out] += reqs[k].sign * parts[k];

    for (int s = 0; s < n; ++s) {
        SolveResult& R = results[s];
        const InstOut& o = ho[s];
        R.status = o.status, R.n_acc = o.n_acc, R.n_rej = o.n_rej, R.n_lu = o.n_lu, R.n_newt = o.n_newt;
        R.n_err = o.n_err, R.n_inner = o.n_inner, R.n_admm = o.n_admm, R.n_refine = o.n_refine;
        R.nsb = hi[s].nsb, R.nvb = hi[s].nvb;
        R.direct = hi[s].cmode == 1 || hi[s].prec == 1;
        const unsigned long long tks[6] = {o.tk_asm, o.tk_lu, o.tk_solve, o.tk_mv, o.tk_other, o.tk_total};
        for (int k = 0; k < 6; ++k) R.tk[k] = double(tks[k]) * 1e-8;
        for (int k = 0; k < 4; ++k) R.tk[6 + k] = double(o.tk_lu_part[k]) * 1e-8;
        const long long N = (long long)hi[s].nsb * hi[s].nvb;
        if (opt.want_basis) {
            for (int ax = 0; ax < 2; ++ax) {
                const int nq = ax == 0 ? hi[s].nqs : hi[s].nqv;
                const long long ob = ax == 0 ? hi[s].o_bs : hi[s].o_bv;
                R.bas[ax].resize(8LL * nq);
                R.bas_c0[ax].resize(nq);
                HIP_OK(hipMemcpy(R.bas[ax].data(), G.bv.as<double>() + ob, sizeof(double) * 8 * nq,
                                 hipMemcpyDeviceToHost));
                HIP_OK(hipMemcpy(R.bas_c0[ax].data(), G.bc.as<int>() + ob / 8, sizeof(int) * nq,
                                 hipMemcpyDeviceToHost));
            }
        }
        if (opt.want_q0) {
            R.q0.resize(N);
            HIP_OK(hipMemcpy(R.q0.data(), G.vc.as<double>() + hi[s].o_vec + V_Q0 * N, sizeof(double) * N,
                             hipMemcpyDeviceToHost));
        }
        if (opt.want_stencils && !given) {
            // unscaled M and K as N x 49 rows (entry (dJ + 3) * 7 + dj + 3), from the
            // D-scaled plane-major stencils: M[r][e] = D[r] Ms[r][e] D[col(r, e)]
            std::vector<double> st98(98 * N), Dh(N);
            HIP_OK(hipMemcpy(st98.data(), G.st.as<double>() + hi[s].o_st, sizeof(double) * 98 * N,
                             hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(Dh.data(), G.vc.as<double>() + hi[s].o_vec + V_D * N, sizeof(double) * N,
                             hipMemcpyDeviceToHost));
            const int nvb = hi[s].nvb;
            R.M49.assign(49 * N, 0.0);
            R.K49.assign(49 * N, 0.0);
            for (long long r = 0; r < N; ++r)
                for (int e = 0; e < 49; ++e) {
                    const long long c = r + (long long)(e / 7 - 3) * nvb + (e % 7 - 3);
                    const int I = int(r / nvb), i = int(r % nvb), J = I + e / 7 - 3, j = i + e % 7 - 3;
                    if (J < 0 || J >= hi[s].nsb || j < 0 || j >= nvb) continue;
                    R.M49[r * 49 + e] = Dh[r] * st98[e * N + r] * Dh[c];
                    R.K49[r * 49 + e] = Dh[r] * st98[(49 + e) * N + r] * Dh[c];
                }
        }
        if (opt.want_qT && !opt.stop_after_ic) {
            R.qT.resize(N);
            HIP_OK(hipMemcpy(R.qT.data(), G.vc.as<double>() + hi[s].o_vec + V_QT * N, sizeof(double) * N,
                             hipMemcpyDeviceToHost));
        }
        if (opt.hist_cap > 0 && !opt.stop_after_ic) {
            const int nh = std::min(o.n_hist, opt.hist_cap);
            R.hist.assign(nh, std::vector<double>(N));
            for (int k = 0; k < nh; ++k)
                HIP_OK(hipMemcpy(R.hist[k].data(), G.hs.as<double>() + hi[s].o_hist + k * N, sizeof(double) * N,
                                 hipMemcpyDeviceToHost));
        }
        if (opt.want_pdf && !opt.stop_after_ic) {
            const long long np = (long long)hi[s].nqs * hi[s].nqv;
            if (!G.pdf.ensure(sizeof(double) * np, err)) return false;
            const int blocks = int(std::min<long long>(1024, (np + NT - 1) / NT));
            hipLaunchKernelGGL(k_pdf, dim3(blocks), dim3(NT), 0, st, dI, s, G.bv.as<double>(), G.bc.as<int>(),
                               G.vc.as<double>() + hi[s].o_vec + V_QT * N, G.pdf.as<double>());
            HIP_OK(hipGetLastError());
            R.pdf.resize(np);
            HIP_OK(hipMemcpyAsync(R.pdf.data(), G.pdf.p, sizeof(double) * np, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
        }
    }
    times.total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return true;
}

bool price_options(const std::vector<OptionSpec>& opts, std::vector<double>& prices, std::vector<int>& status,
                   std::vector<SolveResult>* results_out, BatchTimes* times_out, std::string& err,
                   std::vector<std::vector<int>>* used_out, bool want_qT) {
    const int n = int(opts.size());
    prices.assign(n, 0.0);
    status.assign(n, 0);
    std::vector<SolveSpec> solves;
    std::vector<Domain> doms;
    std::vector<Request> reqs;
    std::vector<std::vector<int>> used(n);
    // an option's solves: (spec, domain, payoff type, sign)
    struct Part {
        SolveSpec s;
        Domain d;
        int type;
        double sign;
    };
    for (int i = 0; i < n; ++i) {
        const OptionSpec& o = opts[i];
        std::string why;
        if (!option_is_valid(o, &why)) {
            status[i] = 1;
            continue;
        }
        const int t = o.option_type;
        std::vector<std::pair<OptionSpec, std::pair<int, double>>> want;
        if (is_in_option(t)) {  // in = vanilla (full domain) - out (truncated)
            const bool call = t % 2 == 0, down = t <= 1;
            OptionSpec van = o;
            van.barrier = 0.0;
            want.push_back({van, {call ? 8 : 9, 1.0}});
            want.push_back({o, {(down ? 2 : 6) + (call ? 0 : 1), -1.0}});
        } else {
            want.push_back({o, {t, 1.0}});
        }
        std::vector<Part> parts;
        bool fits = true;
        for (auto& w : want) {
            Part p{solve_for(w.first, w.second.first), Domain(), w.second.first, w.second.second};
            p.d = build_domain(p.s);
            fits = fits && grid_fits(p.d, nullptr);
            parts.push_back(std::move(p));
        }
        if (!fits) {  // one oversized option does not stop the rest of the batch
            status[i] = 5;
            continue;
        }
        for (auto& p : parts) {
            solves.push_back(p.s);
            doms.push_back(std::move(p.d));
            Request r{int(solves.size()) - 1, p.type, opts[i].barrier, opts[i].strike, p.sign, i};
            if (p.type >= 8) r.barrier = 0.0;
            reqs.push_back(r);
            used[i].push_back(int(solves.size()) - 1);
        }
    }
    std::vector<SolveResult> res;
    BatchTimes bt;
    RunOptions ro;
    ro.want_qT = want_qT;
    if (!run_batch(solves, doms, reqs, n, ro, prices, res, bt, err)) return false;
    for (int i = 0; i < n; ++i) {
        for (int s : used[i])
            if (res[s].status != 0 && status[i] == 0) status[i] = res[s].status;
        if (status[i] == 1 || status[i] == 5) prices[i] = 0.0;
    }
    if (results_out) *results_out = std::move(res);
    if (used_out) *used_out = std::move(used);
    if (times_out) *times_out = bt;
    return true;
}

}  // namespace fpe
