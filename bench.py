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


def cpu_baseline_start(cols, n_sample, procs):
    """Starts the FP64 CPU oracle (the C++ restatement of the reference
    algorithm, one option per process, one thread each) on options spread
    evenly over the same batch. Runs while the GPU is timed; collect with
    cpu_baseline_finish."""
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


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_radau_batch4096_hbm_counters.md")


def profiled_traffic(n, mode):
    """HBM bytes per k_radau launch from the committed FETCH_SIZE / WRITE_SIZE
    counter runs of this same workload (4096 options, direct mode), scaled by
    the copy-kernel calibration recorded there; None for other workloads."""
    if n != 4096 or mode != "direct" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def surrogate_leg(plain_value, n=1_000_000, audit_frac=0.01, n_out=256, seed=20261205):
    """The learned path on n seeded European options from the training box
    (plus n_out rows outside it, which go to the solver): MLP + POD pricing on
    the GPU, a random audit fraction re-solved by the plain FP64 solver on the
    same grid, throughput over the whole job (inference + fallback + audit)."""
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
    prices, ok, n_fb = sur.price_with_fallback(P, K, ty)
    t1 = time.perf_counter()
    rows = np.nonzero(ok)[0]
    pick = np.sort(rng.choice(rows, size=int(round(audit_frac * len(rows))), replace=False))
    ref, st = S.solver_prices(P[pick], K[pick], ty[pick])
    t2 = time.perf_counter()
    d = np.abs(prices[pick] - ref)[st == 0]
    value = n / (t2 - t0)
    return {"options": n, "value": round(value, 1), "unit": "options/s",
            "vs_plain_fp64": round(value / plain_value, 1) if plain_value else None,
            "inference_options_per_s": round(int(ok.sum()) / max(1e-9, t1 - t0 - 0.0), 1),
            "seconds": {"surrogate_and_fallback": round(t1 - t0, 3), "audit": round(t2 - t1, 3)},
            "fallback_solves": int(n_fb),
            "audit": {"n": int(len(pick)), "max_abs_dprice": float(np.max(d)) if len(d) else None,
                      "p99_abs_dprice": float(np.percentile(d, 99)) if len(d) else None,
                      "mean_abs_dprice": float(np.mean(d)) if len(d) else None,
                      "checked_against": "plain FP64 solve, same 80x40 grid"},
            "model": {"samples": int(sur.info["meta_samples"]), "val_max_abs_dprice":
                      float(sur.info["meta_val_max_abs_dprice"])},
            "workload": f"{n} seeded European calls/puts from the training box (kappa, theta, sigma, rho, V0, r, T), "
                        f"K in [45, 75], S0 = 60, {n_out} of them outside the box (solver fallback); "
                        f"{audit_frac:.0%} audit"}


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
    ap.add_argument("--cpu-sample", type=int, default=15)
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
    cols = workloads.batch_columns(args.n, complex_mode=1 if args.mode == "direct" else 0)
    cores = len(os.sched_getaffinity(0))
    procs = max(1, min(15, cores - 1))
    cpu = None if args.no_cpu else cpu_baseline_start(cols, args.cpu_sample, procs)
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
        "config": {"workload": f"batched plain FP64 Heston FPE pricing, {args.n} seeded options "
                               "(grids 40x20..120x60, 50% European / 50% barrier, in-options by parity)",
                   "path": f"fp64 {args.mode} complex stage solve, reference Radau constants",
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
                     "traffic": profiled_traffic(args.n, args.mode),
                     "traffic_source": "rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 (copy-calibrated), "
                                       "profiles/k_radau_batch4096_hbm_counters.md",
                     "alg_bytes_per_launch": st["alg_bytes"],
                     "alg_flops_per_launch": st["alg_flops"], "hbm_achieved_GBs": round(achieved_gbs, 2),
                     "hbm_frac": round(achieved_gbs / HBM_PEAK_GBS, 5)},
        "cpu_baseline": None,
    }
    if not args.no_mixed:
        line["mixed"] = mixed_leg(api, cols, prices, status, value, args.n, 1,
                                  "FP32 block factorisation on v_mfma_f32_16x16x4f32, FP64 refinement of every "
                                  "stage solve to |r| <= 1e-12 |b|; FP64 Newton residuals, error control, payoff")
    if not args.no_surrogate:
        line["mixed_surrogate"] = surrogate_leg(value)
    if cpu is not None:
        rate, rate1, cpu_s, cwall, cprices, idx = cpu_baseline_finish(cpu, procs)
        d = np.abs(prices[idx] - cprices)
        line["cpu_baseline"] = {"value": round(rate, 4), "unit": "options/s", "cores": procs, "kind": "port",
                                "sample": f"{len(idx)} options spread evenly over the same batch (all four grids), "
                                          f"one single-threaded oracle process per option on {procs} host cores, "
                                          f"run beside the GPU timing; value = cores x options / "
                                          f"{cpu_s:.1f} core-seconds ({cwall:.1f} s wall)",
                                "per_core": round(rate1, 5),
                                "max_abs_dprice_vs_gpu": float(np.max(d))}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
