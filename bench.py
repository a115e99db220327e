"""Throughput benchmark: batched Heston Fokker-Planck option pricing on one MI355X.

One step = pricing the seeded 4096-option batch (span discretisation, grids
(40,20)..(120,60), 50 % European / 50 % single-barrier, in-options by parity)
end to end on the GPU: host setup, assembly, initial condition, Radau IIA,
payoff. Prints one JSON line with options/s, the Radau kernel's roofline
figures, and the FP64 CPU oracle timed on a bounded sample on the host cores.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def _oracle_worker(args):
    kw, K = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    t = time.perf_counter()
    p, st, rc = oracle_lib.price([K], **kw)
    return p[0], time.perf_counter() - t


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_start(cols, n_sample, procs):
    """Starts the FP64 CPU oracle (the C++ restatement of the reference
    algorithm, one option per process, one thread each) on options spread
    evenly over the same batch; collect with cpu_baseline_finish."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    oracle_lib.build()
    from fpe_option_amd import workloads
    n = len(cols["strike"])
    idx = np.linspace(0, n - 1, n_sample).astype(int)
    sub = workloads.take(cols, idx)
    jobs = []
    for i in range(n_sample):
        kw = {k: (v[i].item() if isinstance(v, np.ndarray) else v) for k, v in sub.items()
              if k not in ("strike", "precision")}
        jobs.append((kw, float(sub["strike"][i])))
    pool = mp.get_context("spawn").Pool(procs)
    return pool, pool.map_async(_oracle_worker, jobs, chunksize=1), idx, time.perf_counter()


def cpu_baseline_finish(handle, procs):
    pool, fut, idx, t0 = handle
    res = fut.get()
    wall = time.perf_counter() - t0
    pool.close()
    pool.join()
    prices = np.array([r[0] for r in res])
    cpu_s = float(sum(r[1] for r in res))
    # aggregate throughput of `procs` cores kept busy: options per core-second x cores
    return procs * len(res) / cpu_s, len(res) / cpu_s, cpu_s, wall, prices, idx


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_radau_batch4096_allshapes_hbm_counters.md")


def profiled_traffic(key):
    """HBM bytes per k_radau launch from the committed FETCH_SIZE / WRITE_SIZE
    counter runs (copy-calibrated) when they were taken on this same workload
    (their workload_key line), else None."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    vals = {}
    for line in open(TRAFFIC_PROFILE):
        if ":" in line and line.split(":")[0] in ("workload_key", "traffic_bytes_per_launch"):
            vals[line.split(":")[0]] = line.split(":", 1)[1].strip()
    if vals.get("workload_key") != key or "traffic_bytes_per_launch" not in vals:
        return None
    return float(vals["traffic_bytes_per_launch"])


def tuning_table():
    """The per-grid choices the library ran with (fpe_option_amd/tuning_gfx950.txt):
    factorisation code path and initial-condition ADMM launch form per (n_sb, n_vb)."""
    from fpe_option_amd import autotune
    forms = {-1: "auto", 0: "waves8_full", 1: "waves4_half", 2: "waves4_third"}
    rows = autotune.read_table(autotune.TABLE)
    return [[nsb, nvb, "compiled_k" if v == 1 else "generic", forms.get(f, "auto")]
            for (nsb, nvb), (v, f) in sorted(rows.items())]


def surrogate_leg(n=1_000_000, audit_frac=0.01, n_out=256, seed=20261205, tol=1e-4, n_rate=4096):
    """The learned path on n seeded European options from the training box
    (plus n_out rows outside it): the ensemble prices every in-box option with
    an error estimate; those certified at or below tol are served by the
    surrogate, the rest (and the out-of-box rows) by the mixed-precision solver
    on the same 80x40 grid, and a random audit_frac of the in-box options is
    re-solved by the plain FP64 solver. Job time = inference + solver fallback
    + audit; the solver times come from their measured rates on n_rate options
    of the same job (the full fallback would take hours). The ratio is against
    the plain FP64 solver on the same options, and the audit states the error
    of the served prices; the same audit gives the frontier at looser bars."""
    from fpe_option_amd import surrogate as S
    if not os.path.exists(S.MODEL_PATH):
        return None
    sur = S.Surrogate()
    rng = np.random.Generator(np.random.PCG64(seed))
    P = S.sample_params(n, seed)
    out_rows = rng.choice(n, n_out, replace=False)
    P[out_rows, 2] = rng.uniform(0.85, 1.0, n_out)  # sigma above the box
    K = rng.uniform(45.0, 75.0, n)
    ty = np.where(rng.uniform(0.0, 1.0, n) < 0.5, 8, 9).astype(np.int32)
    sur.price(P[:4096], K[:4096], ty[:4096])  # warm-up (weights resident, kernels loaded)
    t0 = time.perf_counter()
    sp, est, ok = sur.price(P, K, ty)
    t_inf = time.perf_counter() - t0
    inbox = np.nonzero(ok)[0]
    pick = np.sort(rng.choice(inbox, size=int(round(audit_frac * len(inbox))), replace=False))
    t1 = time.perf_counter()
    ref, st = S.solver_prices(P[pick], K[pick], ty[pick])
    t_audit = time.perf_counter() - t1
    good = st == 0
    err, e_est = np.abs(sp[pick] - ref)[good], est[pick][good]

    def rate(idx, precision):
        cols = {name: P[idx, j] for j, name in enumerate(S.PARAMS)}
        cols.update(S.GRID)
        cols.update(strike=K[idx], option_type=ty[idx], barrier=0.0, precision=precision)
        from fpe_option_amd import api
        api.price_batch(**{k: (v[:64] if isinstance(v, np.ndarray) else v) for k, v in cols.items()})
        t = time.perf_counter()
        api.price_batch(**cols)
        return len(idx) / (time.perf_counter() - t)

    sample = rng.choice(n, n_rate, replace=False)
    r_plain, r_mixed = rate(sample, 0), rate(sample, 1)
    frontier = []
    for b in (tol, 1e-3, 1e-2, 1e-1, float("inf")):
        cert = ok & (est <= b)
        n_fb = int(n - cert.sum())
        t_job = t_inf + n_fb / r_mixed + t_audit
        sel = e_est <= b
        frontier.append({"bar": b if b != float("inf") else None, "certified_frac": round(float(cert.mean()), 5),
                         "options_per_s": round(n / t_job, 1), "vs_plain_fp64": round(n / t_job / r_plain, 3),
                         "audit_n": int(sel.sum()),
                         "audit_max_abs_dprice": float(err[sel].max()) if sel.any() else None,
                         "audit_p99_abs_dprice": float(np.percentile(err[sel], 99)) if sel.any() else None})
    head = frontier[0]
    return {"options": n, "tol": tol, "value": head["options_per_s"], "unit": "options/s",
            "vs_plain_fp64": head["vs_plain_fp64"], "certified_frac": head["certified_frac"],
            "served_audit_max_abs_dprice": head["audit_max_abs_dprice"],
            "served_audit_p99_abs_dprice": head["audit_p99_abs_dprice"], "served_audit_n": head["audit_n"],
            "plain_fp64_options_per_s_same_options": round(r_plain, 1),
            "mixed_options_per_s_same_options": round(r_mixed, 1),
            "inference_options_per_s": round(int(ok.sum()) / t_inf, 1),
            "seconds": {"inference": round(t_inf, 3), "audit": round(t_audit, 3)},
            "audit_all_inbox": {"n": int(good.sum()), "max_abs_dprice": float(err.max()),
                                "p99_abs_dprice": float(np.percentile(err, 99)), "mean_abs_dprice": float(err.mean()),
                                "estimate_covers_error_frac": round(float(np.mean(err <= e_est)), 5)},
            "frontier": frontier,
            "model": {"members": len(sur.members), "samples": int(sur.info["meta_samples"]), "gamma": sur.gamma,
                      "val_max_abs_dprice": float(sur.info["meta_val_max_abs_dprice"])},
            "workload": f"{n} seeded European calls/puts from the training box (kappa, theta, sigma, rho, V0, r, T), "
                        f"K in [45, 75], S0 = 60, {n_out} of them outside the box; {audit_frac:.0%} audit of the "
                        f"in-box options by the plain FP64 solver on the same 80x40 grid; fallback and plain rates "
                        f"measured on {n_rate} options of the job"}


def mixed_leg(api, cols, prices, status, plain_value, n, precision, path):
    """The same batch with FP32 factors and FP64-refined stage solves
    (precision 1), one timed step, prices compared with the plain FP64 step's."""
    mcols = dict(cols, precision=precision, complex_mode=1)
    t = time.perf_counter()
    mprices, mstatus, mst = api.price_batch(return_stats=True, **mcols)
    mwall = time.perf_counter() - t
    ok = (mstatus == 0) & (status == 0)
    d = np.abs(mprices - prices)[ok]
    return {"value": round(n / mwall, 3), "unit": "options/s", "ms_per_step": round(1e3 * mwall, 1),
            "vs_plain_fp64": round((n / mwall) / plain_value, 3), "precision": precision, "path": path,
            "max_abs_dprice_vs_fp64": float(np.max(d)), "p99_abs_dprice_vs_fp64": float(np.percentile(d, 99)),
            "refine_rounds_per_stage_solve": round(mst["n_refine"] / max(1, 2 * mst["n_newton"] +
                                                                          mst["n_err_solves"]), 3),
            "newton_iterations": int(mst["n_newton"]), "lu_refreshes": int(mst["n_lu"]),
            "phases_ms": {k: round(mst[k], 1) for k in ("ms_ic", "ms_radau")},
            "radau_wg_s": {k[5:]: round(mst[k], 1) for k in (
                "wg_s_assemble", "wg_s_lu", "wg_s_solve", "wg_s_spmv", "wg_s_other", "wg_s_total",
                "wg_s_lu_inverse", "wg_s_lu_stage", "wg_s_lu_lmul", "wg_s_lu_schur")},
            "status_ok": int(np.sum(mstatus == 0))}


def copy_calibration(mbytes):
    """One plain device-to-device copy of a known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import torch
    n = int(mbytes * (1 << 20)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": 4 * n, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4096, help="options per batch")
    ap.add_argument("--mode", default="direct", choices=["direct", "gmres"])
    ap.add_argument("--grids", default="independent", choices=["independent", "paired"],
                    help="n_s, n_v drawn independently (16 shapes) or index-paired (4 shapes)")
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="MIB",
                    help="only run a plain copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        copy_calibration(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; sharding over several GPUs is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from fpe_option_amd import api, workloads
    cols = workloads.batch_columns(args.n, complex_mode=1 if args.mode == "direct" else 0, pairing=args.grids)
    workload_key = f"{args.n}-{'allshapes' if args.grids == 'independent' else 'paired'}-{args.mode}"
    # the host's cores (capped at the 16 a one-GPU share of the machine has)
    procs = max(1, min(16, len(os.sched_getaffinity(0))))
    for _ in range(args.warmup):
        api.price_batch(**cols)
    t0 = time.perf_counter()
    stats_acc = []
    prices = None
    for _ in range(args.steps):
        prices, status, st = api.price_batch(return_stats=True, **cols)
        stats_acc.append(st)
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    value = args.n / (wall / args.steps)
    st = stats_acc[-1]
    radau_s = st["ms_radau"] / 1e3
    achieved_tf = st["alg_flops"] / radau_s / 1e12
    achieved_gbs = st["alg_bytes"] / radau_s / 1e9
    line = {
        "metric": "options/sec on 1 MI355X: plain FP64 vs mixed+surrogate at |dprice|<=1e-4",
        "value": round(value, 3), "unit": "options/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched plain FP64 Heston FPE pricing, {args.n} seeded options (grids "
                               + ("n_s x n_v drawn independently from {40,60,80,120} x {20,30,40,60}"
                                  if args.grids == "independent" else "index-paired 40x20, 60x30, 80x40, 120x60")
                               + ", 50% European / 50% barrier, in-options by parity)",
                   "path": f"fp64 {args.mode} complex stage solve, reference Radau constants, factorisation "
                           "code path per grid shape from fpe_option_amd/tuning_gfx950.txt",
                   "solves_per_step": workloads.n_solves(cols), "rtol": 1e-4, "atol": 1e-6, "seed": workloads.SEED},
        "status_ok": int(np.sum(status == 0)),
        "phases_ms": {k: round(st[k], 2) for k in ("ms_setup", "ms_assemble", "ms_ic", "ms_radau", "ms_payoff",
                                                   "ms_total")},
        "solver": {k: int(st[k]) for k in ("n_solves", "n_accepted", "n_rejected", "n_lu", "n_newton",
                                           "n_err_solves", "n_inner_iters", "n_admm_iters")},
        # Radau workgroups: busy fraction over the kernel's duration, and where their time goes
        "radau_slots": {"slots": int(st["n_radau_slots"]),
                        "busy_frac": round(st["wg_s_total"] / max(1, st["n_radau_slots"]) / radau_s, 4),
                        "wg_s": {k[5:]: round(st[k], 1) for k in (
                            "wg_s_assemble", "wg_s_lu", "wg_s_solve", "wg_s_spmv", "wg_s_other", "wg_s_total",
                            "wg_s_lu_inverse", "wg_s_lu_stage", "wg_s_lu_lmul", "wg_s_lu_schur")}},
        "roofline": {"bound": "mfma", "kernel": "k_radau", "achieved": round(achieved_tf, 4),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved_tf / FP64_PEAK_TFLOPS, 5),
                     "traffic": profiled_traffic(workload_key),
                     "traffic_source": "rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 (copy-calibrated), "
                                       "profiles/k_radau_batch4096_allshapes_hbm_counters.md",
                     "alg_bytes_per_launch": st["alg_bytes"],
                     "alg_flops_per_launch": st["alg_flops"], "hbm_achieved_GBs": round(achieved_gbs, 2),
                     "hbm_frac": round(achieved_gbs / HBM_PEAK_GBS, 5)},
        "cpu_baseline": None,
        "tuning": {"source": "fpe_option_amd/tuning_gfx950.txt",
                   "columns": ["n_sb", "n_vb", "factorisation", "admm_launch"], "shapes": tuning_table()},
    }
    if not args.no_mixed:
        line["mixed"] = mixed_leg(api, cols, prices, status, value, args.n, 1,
                                  "FP32 block factorisation on v_mfma_f32_16x16x4f32, FP64 refinement of every "
                                  "stage solve to |r| <= 1e-12 |b|; FP64 Newton residuals, error control, payoff")
    if not args.no_surrogate:
        line["mixed_surrogate"] = surrogate_leg()
    if not args.no_cpu:
        # after the GPU legs, so that neither timing disturbs the other
        rate, rate1, cpu_s, cwall, cprices, idx = cpu_baseline_finish(
            cpu_baseline_start(cols, args.cpu_sample, procs), procs)
        d = np.abs(prices[idx] - cprices)
        line["cpu_baseline"] = {"value": round(rate, 4), "unit": "options/s", "cores": procs, "kind": "port",
                                "cpu_model": cpu_model(), "host_cores": len(os.sched_getaffinity(0)),
                                "sample": f"{len(idx)} options spread evenly over the same batch, one "
                                          f"single-threaded FP64 oracle process per option on {procs} host cores, "
                                          f"after the GPU timing; value = cores x options / {cpu_s:.1f} "
                                          f"core-seconds ({cwall:.1f} s wall)",
                                "per_core": round(rate1, 5),
                                "max_abs_dprice_vs_gpu": float(np.max(d))}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
