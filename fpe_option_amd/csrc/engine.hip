// Host orchestration of the batched GPU pricing path.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "assemble.inl"
#include "engine.hpp"
#include "ic.inl"
#include "kron.inl"
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

// Fixed device scratch of the allocation-free entry point (fpe_batch_price_noalloc):
// every device buffer of such a call is a slice of it. Each GPU the library runs
// on holds its own copy until the process ends. In host code the name stands for
// a host-side symbol, not a device address: only the address hipGetSymbolAddress
// returns (DeviceState::scratch) is passed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

struct DeviceState {
    bool init = false;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8];
    // streams beside `stream` for independent launch groups (default priority),
    // and the events that fork them from it and join them back
    hipStream_t aux[3] = {};
    hipEvent_t fork_ev = nullptr, join_ev[3] = {};
    unsigned char* scratch = nullptr;  // device address of `scratch` on this device
    DBuf inst, out, kn, qd, bv, bc, fc, st, vc, eg, hs, wm, ow, ws, wi, req, part, queue, order, pdf, kron;
};

// Device pointers of one batch, from the grow-only buffers or from the scratch.
struct DevPtrs {
    Inst* inst = nullptr;
    InstOut* out = nullptr;
    double *kn = nullptr, *qd = nullptr, *bv = nullptr, *fc = nullptr, *st = nullptr, *vc = nullptr, *eg = nullptr,
           *hs = nullptr, *wm = nullptr, *ws = nullptr, *part = nullptr;
    int *bc = nullptr, *wi = nullptr, *queue = nullptr, *order = nullptr;
    long long* ow = nullptr;
    PriceReq* req = nullptr;
};

std::mutex g_mu;
DeviceState g_dev;

// Tuning table: measured factorisation code path per grid shape (n_sb, n_vb)
// and, optionally, the initial-condition ADMM launch form, read once from
// tuning_gfx950.txt next to this library (written by fpe_option_amd/autotune.py
// from GPU timings). Lines: "n_sb n_vb variant [admm_form]" (admm_form: the
// k_admm launch class 0..2; absent or -1: chosen by occupancy).
struct TuneRow {
    int nsb, nvb, variant, ic;
};
std::vector<TuneRow> g_tune;
bool g_tune_loaded = false;

void load_tuning() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    Dl_info info;
    if (!dladdr(reinterpret_cast<void*>(&load_tuning), &info) || !info.dli_fname) return;
    std::string path(info.dli_fname);
    const size_t slash = path.rfind('/');
    path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning_gfx950.txt";
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        TuneRow r;
        if (line[0] == '#') continue;
        r.ic = -1;
        const int got = std::sscanf(line, "%d %d %d %d", &r.nsb, &r.nvb, &r.variant, &r.ic);
        if (got >= 3 && (r.variant == 0 || r.variant == 1)) {
            if (got < 4 || r.ic < -1 || r.ic > 2) r.ic = -1;
            g_tune.push_back(r);
        }
    }
    std::fclose(f);
}

// The tuned code path of a grid: the measured shape with the nearest block
// width at or above n_vb (else the widest), then the nearest block-row count.
int tuned_variant(int nsb, int nvb) {
    load_tuning();
    int best = -1;
    for (size_t k = 0; k < g_tune.size(); ++k) {
        const TuneRow& r = g_tune[k];
        if (best < 0) {
            best = int(k);
            continue;
        }
        const TuneRow& b = g_tune[best];
        const bool r_cov = r.nvb >= nvb, b_cov = b.nvb >= nvb;
        if (r_cov != b_cov) {
            if (r_cov) best = int(k);
            continue;
        }
        if (r.nvb != b.nvb) {
            if (r_cov ? r.nvb < b.nvb : r.nvb > b.nvb) best = int(k);
            continue;
        }
        if (std::abs(r.nsb - nsb) < std::abs(b.nsb - nsb)) best = int(k);
    }
    return best < 0 ? LU_GENERIC : g_tune[best].variant;
}

// Launch groups of the initial-condition kernels (budgets in ic.inl).
struct IcGroup {
    int start, count;  // a slice of the group order
    int key;           // the group's launch key
};

// Sorts the instances into ord by launch key (ascending) and cost (largest
// first) and cuts it into one group per key.
template <class Key, class Cost>
void ic_groups(int n, Key key, Cost cost, std::vector<int>& ord, std::vector<IcGroup>& groups) {
    std::vector<int> k(n);
    for (int s = 0; s < n; ++s) k[s] = key(s);
    ord.resize(n);
    for (int s = 0; s < n; ++s) ord[s] = s;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        if (k[a] != k[b]) return k[a] < k[b];
        return cost(a) > cost(b);
    });
    groups.clear();
    for (int i = 0; i < n;) {
        int j = i;
        while (j < n && k[ord[j]] == k[ord[i]]) ++j;
        groups.push_back(IcGroup{i, j - i, k[ord[i]]});
        i = j;
    }
}

// k_eig launch class: the smallest LDS budget that holds A and Qt, else the largest.
const long long kEigBudgets[4] = {EIG_LDS_S, EIG_LDS_HALF, EIG_LDS_THIRD, EIG_LDS_QUARTER};
int eig_class(int n) {
    int c = 0;
    for (int b = 1; b < 4; ++b)
        if (eig_lds_bytes(n) <= kEigBudgets[b]) c = b;
    return c;
}

using AdmmKernel = void (*)(const Inst*, const int*, const double*, const double*, double*, InstOut*, long long);
template <int NW, int MT, int TN>
constexpr AdmmKernel admm_kernel() {
    return k_admm<NW, MT, TN, (MT * TN <= ADMM_REG_TILES)>;
}
// [MT - 1][TN - 1]: 8-wave workgroups up to 3 strips per wave (n_sb <= 384), 4-wave ones up to 2
const AdmmKernel kAdmm8[3][4] = {
    {admm_kernel<8, 1, 1>(), admm_kernel<8, 1, 2>(), admm_kernel<8, 1, 3>(), admm_kernel<8, 1, 4>()},
    {admm_kernel<8, 2, 1>(), admm_kernel<8, 2, 2>(), admm_kernel<8, 2, 3>(), admm_kernel<8, 2, 4>()},
    {admm_kernel<8, 3, 1>(), admm_kernel<8, 3, 2>(), admm_kernel<8, 3, 3>(), admm_kernel<8, 3, 4>()}};
const AdmmKernel kAdmm4[2][4] = {
    {admm_kernel<4, 1, 1>(), admm_kernel<4, 1, 2>(), admm_kernel<4, 1, 3>(), admm_kernel<4, 1, 4>()},
    {admm_kernel<4, 2, 1>(), admm_kernel<4, 2, 2>(), admm_kernel<4, 2, 3>(), admm_kernel<4, 2, 4>()}};

// k_admm launch key: class (0: 8-wave workgroups with ADMM_LDS_MAX bytes; 1, 2:
// 4-wave workgroups with a half or third of the CU's LDS) and the kernel's strip
// and column-tile counts. Among the forms an instance fits (class 0 always; 1 and
// 2 when its work matrices fit the budget and it has at most 8 strips), the one
// with the most resident workgroups per CU (g_admm_occ, from the runtime's
// occupancy of each kernel at its budget) wins; ties go to the lower class.
// want (0..2): a class asked for (forced by the caller, or measured best in the
// tuning table), taken when the instance fits it.
const long long kAdmmBudgets[3] = {ADMM_LDS_MAX, IC_LDS_HALF, IC_LDS_THIRD};
int g_admm_occ[3][3][4];  // [class][MT - 1][TN - 1] workgroups per CU (class 0: MT up to 3, else 2)
bool admm_fits(int ns, int nv, int c) {
    return c == 0 || (((ns + 15) >> 4) <= 8 && admm_lds_bytes(ns, nv, false) <= kAdmmBudgets[c]);
}
int admm_key(int ns, int nv, int want) {
    const int tn = admm_tn(nv);
    int best = 0;
    if (want >= 0 && want <= 2 && admm_fits(ns, nv, want)) {
        best = want;
    } else {
        int best_occ = g_admm_occ[0][admm_mt(ns, 8) - 1][tn - 1];
        for (int c = 1; c < 3; ++c)
            if (admm_fits(ns, nv, c) && g_admm_occ[c][admm_mt(ns, 4) - 1][tn - 1] > best_occ)
                best = c, best_occ = g_admm_occ[c][admm_mt(ns, 4) - 1][tn - 1];
    }
    const int mt = admm_mt(ns, best == 0 ? 8 : 4);
    return best * 16 + (mt - 1) * 4 + (tn - 1);
}

// The measured ADMM launch class of a grid shape (exact match), else -1.
int tuned_admm_class(int nsb, int nvb) {
    load_tuning();
    for (const TuneRow& r : g_tune)
        if (r.nsb == nsb && r.nvb == nvb) return r.ic;
    return -1;
}

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
    for (const auto& row : kAdmm8)
        for (AdmmKernel k : row)
            HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(ADMM_LDS_MAX)));
    for (const auto& row : kAdmm4)
        for (AdmmKernel k : row)
            HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(IC_LDS_HALF)));
    for (int c = 0; c < 3; ++c)
        for (int mt = 1; mt <= (c == 0 ? 3 : 2); ++mt)
            for (int tn = 1; tn <= 4; ++tn) {
                const AdmmKernel k = c == 0 ? kAdmm8[mt - 1][tn - 1] : kAdmm4[mt - 1][tn - 1];
                int nb = 0;
                HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(k),
                                                                    (c == 0 ? 8 : 4) * 64, size_t(kAdmmBudgets[c])));
                g_admm_occ[c][mt - 1][tn - 1] = nb;
            }
    HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_eig), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(EIG_LDS_S)));
    HIP_OK(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
    for (auto& e : g_dev.ev) HIP_OK(hipEventCreate(&e));
    for (auto& a : g_dev.aux) HIP_OK(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&g_dev.fork_ev, hipEventDisableTiming));
    for (auto& e : g_dev.join_ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // once per device, with the device current: the scratch array's device address
    HIP_OK(hipGetSymbolAddress(reinterpret_cast<void**>(&g_dev.scratch), HIP_SYMBOL(scratch)));
    if (!g_dev.scratch) {
        err = "hipGetSymbolAddress(scratch) returned no address";
        return false;
    }
    g_dev.init = true;
    return true;
}

template <class T>
bool upload(void* dst, const std::vector<T>& h, std::string& err) {
    if (!h.empty()) HIP_OK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, g_dev.stream));
    return true;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }


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
    const int n = int(solves.size());
    results.assign(n, SolveResult());
    out_prices.assign(n_out, 0.0);
    if (n == 0) return true;  // nothing to run (every option was screened out on the host)
    if (!dev_init(err)) return false;

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
        I.lu_var = sp.lu_variant == 1 ? LU_GENERIC : sp.lu_variant == 2 ? LU_COMPILED_K : tuned_variant(I.nsb, I.nvb);
        I.max_steps = sp.max_steps > 0 ? sp.max_steps : 100000;  // radau.mojo:257
        I.refine_max = sp.refine_max == 0 ? 10 : (sp.refine_max < 0 ? 0 : sp.refine_max);
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
        I.o_eig = o_eg;  // the v part; the shared s parts follow all of them (below)
        o_eg += 3LL * I.nvb * I.nvb + I.nvb;
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
    // s eigenproblems: one per distinct s basis and quadrature (the s mass factor,
    // hence Qs, ls and As, depends on nothing else; European options and the
    // vanilla solves of in-options on one grid all share it). Candidates are
    // bucketed by a hash of a few sizes and entries, then compared exactly.
    std::vector<int> s_reps;
    {
        std::unordered_map<unsigned long long, std::vector<int>> buckets;
        auto bits = [](double x) {
            unsigned long long u;
            std::memcpy(&u, &x, sizeof u);
            return u;
        };
        for (int s = 0; s < n; ++s) {
            const Domain& d = doms[s];
            unsigned long long h = 1469598103934665603ull;
            auto mix = [&](unsigned long long v) { h = (h ^ v) * 1099511628211ull; };
            mix(hi[s].nsb), mix(hi[s].nqs), mix(hi[s].nsk), mix(hi[s].s_lneu), mix(hi[s].s_rneu);
            mix(bits(d.s_knots[d.s_knots.size() / 2])), mix(bits(d.s_x[d.s_x.size() / 2]));
            mix(bits(d.s_w[d.s_w.size() / 3]));
            // bitwise equality (equal inputs give bit-identical factors and eigenvectors)
            auto same = [](const std::vector<double>& a, const std::vector<double>& b) {
                return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(double) * a.size()) == 0;
            };
            std::vector<int>& b = buckets[h];
            int rep = -1;
            for (int c : b) {
                const Domain& e = doms[c];
                if (hi[c].nsb == hi[s].nsb && hi[c].s_lneu == hi[s].s_lneu && hi[c].s_rneu == hi[s].s_rneu &&
                    same(e.s_knots, d.s_knots) && same(e.s_x, d.s_x) && same(e.s_w, d.s_w)) {
                    rep = c;
                    break;
                }
            }
            if (rep < 0) {
                b.push_back(s);
                s_reps.push_back(s);
                hi[s].o_eig_s = o_eg;
                o_eg += 2LL * hi[s].nsb * hi[s].nsb + hi[s].nsb;
            } else {
                hi[s].o_eig_s = hi[rep].o_eig_s;
            }
        }
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
    int n_slots = std::min(n, std::max(1, RADAU_WG_PER_CU * g_dev.n_cu));
    // Radau work queue order: largest estimated cost first, so the slots finish
    // together instead of waiting on a late large instance. Cost per instance
    // n_sb n_vb^2 (n_vb + k): the factorisations grow as n_sb n_vb^3, the block
    // solves and operator products as n_sb n_vb^2, and k weighs the second by the
    // mode's measured share of the solver's time at n_vb = 40 (GMRES mode: solves and
    // products three times the factorisation time; mixed precision, with its refined
    // solves: 0.6 times; direct FP64: 0.3 times).
    std::vector<double> rcost(n);
    for (int s = 0; s < n; ++s) {
        const double k = hi[s].prec == 1 ? 24.0 : hi[s].cmode == 1 ? 11.0 : 120.0;
        rcost[s] = double(hi[s].nsb) * hi[s].nvb * hi[s].nvb * (hi[s].nvb + k);
    }
    std::vector<int> order(n);
    for (int s = 0; s < n; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rcost[a] > rcost[b]; });

    // initial-condition launch groups: the s and v eigenproblems (A and Qt in
    // LDS when the class budget holds both), then ADMM
    std::vector<int> ord_e0, ord_e1, ord_a;
    std::vector<IcGroup> grp_e0, grp_e1, grp_a;
    ic_groups(int(s_reps.size()), [&](int k) { return eig_class(hi[s_reps[k]].nsb); },
              [&](int k) { return double(hi[s_reps[k]].nsb) * hi[s_reps[k]].nsb * hi[s_reps[k]].nsb; }, ord_e0,
              grp_e0);
    for (int& k : ord_e0) k = s_reps[k];  // group order over instances (one per distinct s basis)
    ic_groups(n, [&](int s) { return eig_class(hi[s].nvb); },
              [&](int s) { return double(hi[s].nvb) * hi[s].nvb * hi[s].nvb; }, ord_e1, grp_e1);
    std::vector<int> admm_want(n);  // forced form (ic_variant 1..3), else the tuning table's, else -1
    for (int s = 0; s < n; ++s) {
        const int v = solves[s].ic_variant;
        admm_want[s] = (v >= 1 && v <= 3) ? v - 1 : tuned_admm_class(hi[s].nsb, hi[s].nvb);
    }
    ic_groups(n, [&](int s) { return admm_key(hi[s].nsb, hi[s].nvb, admm_want[s]); },
              [&](int s) { return double(hi[s].nsb) * hi[s].nvb * (hi[s].nsb + hi[s].nvb); }, ord_a, grp_a);
    // offsets of the group orders after the Radau order
    const long long oo_e0 = n, oo_e1 = oo_e0 + (long long)ord_e0.size(), oo_a = oo_e1 + n;
    order.insert(order.end(), ord_e0.begin(), ord_e0.end());
    order.insert(order.end(), ord_e1.begin(), ord_e1.end());
    order.insert(order.end(), ord_a.begin(), ord_a.end());

    std::vector<PriceReq> hr(reqs.size());
    for (size_t k = 0; k < reqs.size(); ++k) {
        hr[k].inst = reqs[k].solve, hr[k].option_type = reqs[k].option_type;
        hr[k].barrier = reqs[k].barrier, hr[k].strike = reqs[k].strike, hr[k].sign = reqs[k].sign;
        hr[k].out = reqs[k].out;
    }

    // ---- device buffers: grow-only allocations, or (RunOptions::noalloc) slices of the
    // fixed scratch array, with as many Radau slots as the rest of the scratch holds
    DeviceState& G = g_dev;
    hipStream_t st = G.stream;
    const size_t b_inst = sizeof(Inst) * n, b_out = sizeof(InstOut) * n, b_kn = sizeof(double) * kn.size(),
                 b_qd = sizeof(double) * qd.size(), b_bv = sizeof(double) * o_bv,
                 b_bc = sizeof(int) * (o_bv / 8 + 1), b_fc = sizeof(double) * o_fc, b_st = sizeof(double) * o_st,
                 b_vc = sizeof(double) * o_vc, b_eg = sizeof(double) * o_eg,
                 b_hs = sizeof(double) * std::max(o_hs, 1LL), b_wm = sizeof(double) * o_wm,
                 b_ow = sizeof(long long) * ow.size(), b_req = sizeof(PriceReq) * std::max<size_t>(hr.size(), 1),
                 b_part = sizeof(double) * std::max<size_t>(reqs.size(), 1), b_queue = sizeof(int),
                 b_order = sizeof(int) * order.size();
    const size_t slot_bytes = sizeof(double) * L.stride + sizeof(int) * L.istride;
    DevPtrs D;
    if (opt.noalloc) {
        if (opt.want_pdf || opt.hist_cap > 0) {
            err = "the allocation-free path prices only (no PDF or state history)";
            return false;
        }
        const size_t sizes[] = {b_inst, b_out, b_kn, b_qd, b_bv, b_bc, b_fc, b_st, b_vc,
                                b_eg, b_hs, b_wm, b_ow, b_req, b_part, b_queue, b_order};
        size_t fixed = 0;
        for (size_t b : sizes) fixed += align256(b);
        const size_t cap = NOALLOC_SCRATCH_BYTES;
        if (fixed + slot_bytes > cap) {
            times.scratch_exceeded = true;
            err = "batch needs " + std::to_string(fixed + slot_bytes) + " bytes of device memory with one solver "
                  "slot; the allocation-free scratch holds " + std::to_string(cap) + " bytes";
            return false;
        }
        n_slots = int(std::min<size_t>(size_t(n_slots), (cap - fixed) / slot_bytes));
        unsigned char* p = G.scratch;
        auto take = [&](size_t b) {
            unsigned char* q = p;
            p += align256(b);
            return q;
        };
        D.inst = reinterpret_cast<Inst*>(take(b_inst));
        D.out = reinterpret_cast<InstOut*>(take(b_out));
        D.kn = reinterpret_cast<double*>(take(b_kn));
        D.qd = reinterpret_cast<double*>(take(b_qd));
        D.bv = reinterpret_cast<double*>(take(b_bv));
        D.bc = reinterpret_cast<int*>(take(b_bc));
        D.fc = reinterpret_cast<double*>(take(b_fc));
        D.st = reinterpret_cast<double*>(take(b_st));
        D.vc = reinterpret_cast<double*>(take(b_vc));
        D.eg = reinterpret_cast<double*>(take(b_eg));
        D.hs = reinterpret_cast<double*>(take(b_hs));
        D.wm = reinterpret_cast<double*>(take(b_wm));
        D.ow = reinterpret_cast<long long*>(take(b_ow));
        D.req = reinterpret_cast<PriceReq*>(take(b_req));
        D.part = reinterpret_cast<double*>(take(b_part));
        D.queue = reinterpret_cast<int*>(take(b_queue));
        D.order = reinterpret_cast<int*>(take(b_order));
        D.ws = reinterpret_cast<double*>(take(sizeof(double) * L.stride * n_slots));
        D.wi = reinterpret_cast<int*>(take(sizeof(int) * L.istride * n_slots));
    } else {
        if (!G.inst.ensure(b_inst, err) || !G.out.ensure(b_out, err) || !G.kn.ensure(b_kn, err) ||
            !G.qd.ensure(b_qd, err) || !G.bv.ensure(b_bv, err) || !G.bc.ensure(b_bc, err) ||
            !G.fc.ensure(b_fc, err) || !G.st.ensure(b_st, err) || !G.vc.ensure(b_vc, err) ||
            !G.eg.ensure(b_eg, err) || !G.hs.ensure(b_hs, err) || !G.wm.ensure(b_wm, err) ||
            !G.ow.ensure(b_ow, err) || !G.req.ensure(b_req, err) || !G.part.ensure(b_part, err) ||
            !G.queue.ensure(b_queue, err) || !G.order.ensure(b_order, err) ||
            !G.ws.ensure(sizeof(double) * L.stride * n_slots, err) ||
            !G.wi.ensure(sizeof(int) * L.istride * n_slots, err))
            return false;
        D.inst = G.inst.as<Inst>(), D.out = G.out.as<InstOut>(), D.kn = G.kn.as<double>();
        D.qd = G.qd.as<double>(), D.bv = G.bv.as<double>(), D.bc = G.bc.as<int>(), D.fc = G.fc.as<double>();
        D.st = G.st.as<double>(), D.vc = G.vc.as<double>(), D.eg = G.eg.as<double>(), D.hs = G.hs.as<double>();
        D.wm = G.wm.as<double>(), D.ow = G.ow.as<long long>(), D.req = G.req.as<PriceReq>();
        D.part = G.part.as<double>(), D.queue = G.queue.as<int>(), D.order = G.order.as<int>();
        D.ws = G.ws.as<double>(), D.wi = G.wi.as<int>();
    }
    times.n_slots = n_slots;
    if (!upload(D.inst, hi, err) || !upload(D.kn, kn, err) || !upload(D.qd, qd, err) || !upload(D.ow, ow, err) ||
        !upload(D.req, hr, err) || !upload(D.order, order, err))
        return false;
    // the scratch keeps the previous call's contents: the counters and the queue
    // start from zero here (every other buffer is written before it is read)
    HIP_OK(hipMemsetAsync(D.out, 0, b_out, st));
    HIP_OK(hipMemsetAsync(D.queue, 0, sizeof(int), st));
    const Inst* dI = D.inst;
    auto t_setup = std::chrono::steady_clock::now();
    times.setup = std::chrono::duration<double, std::milli>(t_setup - t_start).count();

    // ---- kernels
    // given_qT: price (and PDF) from terminal states the caller already holds; only
    // the basis tables are rebuilt, then the states are copied into their slots.
    const bool given = opt.given_qT != nullptr;
    const bool solved = !opt.stop_after_ic;
    HIP_OK(hipEventRecord(G.ev[0], st));
    hipLaunchKernelGGL(k_basis, dim3(n, 2), dim3(NT), 0, st, dI, D.kn, D.qd,
                       D.bv, D.bc);
    if (!given) {
        hipLaunchKernelGGL(k_factors, dim3(n, 2), dim3(NT), 0, st, dI, D.qd, D.bv,
                           D.bc, D.fc);
        const int chunks = int(std::min<long long>(64, (max_N + NT - 1) / NT));
        hipLaunchKernelGGL(k_stencil, dim3(n, chunks), dim3(NT), 0, st, dI, D.fc, D.st,
                           D.vc);
    } else {
        for (int s = 0; s < n; ++s) {
            const long long N = (long long)hi[s].nsb * hi[s].nvb;
            HIP_OK(hipMemcpyAsync(D.vc + hi[s].o_vec + V_QT * N, (*opt.given_qT)[s].data(),
                                  sizeof(double) * N, hipMemcpyHostToDevice, st));
        }
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(G.ev[1], st));
    if (!given) {
        // Two stages of mutually independent launches, each spread over the batch
        // stream and the auxiliary ones (fork, launches, join), so that one launch's
        // last workgroups overlap the next launch's first: (1) the right-hand sides
        // and the s and v eigenproblem groups (disjoint outputs in VC and EG); (2) the
        // ADMM groups (disjoint instances; EG and the stencils read only). A group goes
        // to the stream with the least estimated work so far, largest group first.
        constexpr int NS = 1 + int(sizeof(G.aux) / sizeof(G.aux[0]));
        const hipStream_t sts[NS] = {st, G.aux[0], G.aux[1], G.aux[2]};
        struct Job {
            double cost;
            int kind, g;  // kind 0: s eigenproblems, 1: v eigenproblems, 2: ADMM; g: group index
        };
        auto group_cost = [&](const std::vector<int>& ord, const IcGroup& g, auto cost) {
            double c = 0.0;
            for (int i = g.start; i < g.start + g.count; ++i) c += cost(ord[i]);
            return c;
        };
        auto launch = [&](const Job& j, hipStream_t s_) {
            if (j.kind < 2) {
                const IcGroup& g = (j.kind == 0 ? grp_e0 : grp_e1)[j.g];
                hipLaunchKernelGGL(k_eig, dim3(g.count), dim3(ENT), kEigBudgets[g.key], s_, dI,
                                   D.order + (j.kind == 0 ? oo_e0 : oo_e1) + g.start, j.kind, kEigBudgets[g.key],
                                   D.fc, D.eg);
            } else {
                const IcGroup& g = grp_a[j.g];
                const int cls = g.key / 16, mt = (g.key / 4) % 4 + 1, tn = g.key % 4 + 1;
                const long long lds = kAdmmBudgets[cls];
                const AdmmKernel k = cls == 0 ? kAdmm8[mt - 1][tn - 1] : kAdmm4[mt - 1][tn - 1];
                hipLaunchKernelGGL(k, dim3(g.count), dim3((cls == 0 ? 8 : 4) * 64), lds, s_, dI,
                                   D.order + oo_a + g.start, D.st, D.eg, D.vc, D.out, lds);
            }
        };
        auto stage = [&](std::vector<Job> jobs) -> bool {
            std::stable_sort(jobs.begin(), jobs.end(), [](const Job& x, const Job& y) { return x.cost > y.cost; });
            HIP_OK(hipEventRecord(G.fork_ev, st));
            for (int i = 1; i < NS; ++i) HIP_OK(hipStreamWaitEvent(sts[i], G.fork_ev, 0));
            double load[NS] = {};
            for (const Job& j : jobs) {
                int best = 0;
                for (int i = 1; i < NS; ++i)
                    if (load[i] < load[best]) best = i;
                launch(j, sts[best]);
                load[best] += j.cost;
            }
            HIP_OK(hipGetLastError());
            for (int i = 1; i < NS; ++i) {
                HIP_OK(hipEventRecord(G.join_ev[i - 1], sts[i]));
                HIP_OK(hipStreamWaitEvent(st, G.join_ev[i - 1], 0));
            }
            return true;
        };
        static_assert(MAX_SB <= 2 * NT, "k_ic_rhs sums two k per thread");
        hipLaunchKernelGGL(k_ic_rhs, dim3(n, max_m), dim3(NT), 0, st, dI, D.order, D.qd, D.bv, D.bc, D.vc);
        std::vector<Job> jobs;
        auto cube = [](double x) { return x * x * x; };
        for (int i = 0; i < int(grp_e0.size()); ++i)
            jobs.push_back({group_cost(ord_e0, grp_e0[i], [&](int s) { return cube(hi[s].nsb); }), 0, i});
        for (int i = 0; i < int(grp_e1.size()); ++i)
            jobs.push_back({group_cost(ord_e1, grp_e1[i], [&](int s) { return cube(hi[s].nvb); }), 1, i});
        if (!stage(jobs)) return false;
        jobs.clear();
        for (int i = 0; i < int(grp_a.size()); ++i)
            jobs.push_back({group_cost(ord_a, grp_a[i],
                                       [&](int s) { return double(hi[s].nsb) * hi[s].nvb * (hi[s].nsb + hi[s].nvb); }),
                            2, i});
        if (!stage(jobs)) return false;
    }
    HIP_OK(hipEventRecord(G.ev[2], st));
    if (solved && !given) {
        hipLaunchKernelGGL(k_radau, dim3(n_slots), dim3(RNT), 0, st, dI, n, D.queue, D.order,
                           D.st, D.vc, D.ws, D.wi, L,
                           D.hs, D.out);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipEventRecord(G.ev[3], st));
    if (solved && !reqs.empty()) {
        hipLaunchKernelGGL(k_marginal, dim3(n), dim3(NT), 0, st, dI, D.qd, D.bv,
                           D.bc, D.vc, D.ow, D.wm);
        hipLaunchKernelGGL(k_price, dim3(int(reqs.size())), dim3(NT), 0, st, dI, D.qd,
                           D.ow, D.wm, D.req, D.part);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipEventRecord(G.ev[4], st));
    std::vector<InstOut> ho(n);
    std::vector<double> parts(reqs.size());
    HIP_OK(hipMemcpyAsync(ho.data(), D.out, sizeof(InstOut) * n, hipMemcpyDeviceToHost, st));
    if (!reqs.empty() && !opt.stop_after_ic)
        HIP_OK(hipMemcpyAsync(parts.data(), D.part, sizeof(double) * reqs.size(), hipMemcpyDeviceToHost, st));
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
                HIP_OK(hipMemcpy(R.bas[ax].data(), D.bv + ob, sizeof(double) * 8 * nq,
                                 hipMemcpyDeviceToHost));
                HIP_OK(hipMemcpy(R.bas_c0[ax].data(), D.bc + ob / 8, sizeof(int) * nq,
                                 hipMemcpyDeviceToHost));
            }
        }
        if (opt.want_q0) {
            R.q0.resize(N);
            HIP_OK(hipMemcpy(R.q0.data(), D.vc + hi[s].o_vec + V_Q0 * N, sizeof(double) * N,
                             hipMemcpyDeviceToHost));
        }
        if (opt.apply_x && s == 0 && !given) {
            if (opt.apply_x->size() != size_t(N)) {
                err = "apply_x has the wrong length";
                return false;
            }
            if (!G.kron.ensure(sizeof(double) * 8 * N, err)) return false;
            double* kb = G.kron.as<double>();
            HIP_OK(hipMemcpyAsync(kb, opt.apply_x->data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_kron_apply, dim3(1), dim3(NT), 0, st, dI, 0, D.fc, kb, kb + 3 * N, kb + N, kb + 2 * N);
            HIP_OK(hipGetLastError());
            R.Kx.resize(N);
            R.Mx.resize(N);
            HIP_OK(hipMemcpyAsync(R.Kx.data(), kb + N, sizeof(double) * N, hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(R.Mx.data(), kb + 2 * N, sizeof(double) * N, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
        }
        if (opt.want_stencils && !given) {
            // unscaled M and K as N x 49 rows (entry (dJ + 3) * 7 + dj + 3), from the
            // D-scaled plane-major stencils: M[r][e] = D[r] Ms[r][e] D[col(r, e)]
            std::vector<double> st98(98 * N), Dh(N);
            HIP_OK(hipMemcpy(st98.data(), D.st + hi[s].o_st, sizeof(double) * 98 * N,
                             hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(Dh.data(), D.vc + hi[s].o_vec + V_D * N, sizeof(double) * N,
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
            HIP_OK(hipMemcpy(R.qT.data(), D.vc + hi[s].o_vec + V_QT * N, sizeof(double) * N,
                             hipMemcpyDeviceToHost));
        }
        if (opt.hist_cap > 0 && !opt.stop_after_ic) {
            const int nh = std::min(o.n_hist, opt.hist_cap);
            R.hist.assign(nh, std::vector<double>(N));
            for (int k = 0; k < nh; ++k)
                HIP_OK(hipMemcpy(R.hist[k].data(), D.hs + hi[s].o_hist + k * N, sizeof(double) * N,
                                 hipMemcpyDeviceToHost));
        }
        if (opt.want_pdf && !opt.stop_after_ic) {
            const long long np = (long long)hi[s].nqs * hi[s].nqv;
            if (!G.pdf.ensure(sizeof(double) * np, err)) return false;
            const int blocks = int(std::min<long long>(1024, (np + NT - 1) / NT));
            hipLaunchKernelGGL(k_pdf, dim3(blocks), dim3(NT), 0, st, dI, s, D.bv, D.bc,
                               D.vc + hi[s].o_vec + V_QT * N, G.pdf.as<double>());
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
                   std::vector<std::vector<int>>* used_out, bool want_qT, bool noalloc) {
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
    ro.noalloc = noalloc;
    if (!run_batch(solves, doms, reqs, n, ro, prices, res, bt, err)) {
        if (times_out) *times_out = bt;
        return false;
    }
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
