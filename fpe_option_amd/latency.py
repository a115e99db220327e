"""Per-call latency of the drop-in surface on the reference's own benchmark case.

The reference publishes single-call times for one configuration (README.md:620,
benchmarks/bench_vs_python.mojo:31-74, benchmarks/bench_inout_parity.mojo:89-92):
n_s = n_v = 38, num_insert = 251, s in [0, 150], S0 = 60, V0 = 0.1, r = 0.1,
kappa = 1.2, theta = 0.05, eta = 0.35, rho = -0.4, T = 0.6, strikes 65..100 step 5,
rtol 1e-4 / atol 1e-6, its default GMRES stage solve. This module times the same
calls through `Compute` (python/fpe_engine/pricer.py:104-197): a fresh pipeline per
call, so each time includes the handle, the GPU solve(s) and the price.

    python -m fpe_option_amd.latency [--reps 5]

prints one JSON line: the median wall time of each call after one warm-up call.
"""
import argparse
import json
import time

import numpy as np

from .api import Compute

STRIKES = [65.0, 70.0, 75.0, 80.0, 85.0, 90.0, 95.0, 100.0]
CASE = dict(kappa=1.2, theta=0.05, sigma=0.35, rho=-0.4, r=0.1, T=0.6, S0=60.0, V0=0.1, n_s=38, n_v=38,
            num_insert=251, s_max=150.0)
# seconds the reference reports for the same calls (Mojo engine, Apple M1 Pro, one core)
REFERENCE_S = {
    "doc_solve_price": 3.44,       # README.md:633, native; 4.22 through its Python binding (README.md:631)
    "vanilla_solve_price": 3.57,   # README.md:691
    "dic_solve_price": 7.00,       # README.md:693 (two solves, in-out parity)
    "doc_price_greeks": 15.83,     # README.md:635, price + Greeks through its C++ binding (five solves)
}


def _calls():
    def doc():
        return Compute(option_type="down_and_out_call", barrier=50.0, **CASE).payoff_price(STRIKES)

    def vanilla():
        return Compute(option_type="european_call", **CASE).payoff_price(STRIKES)

    def dic():
        return Compute(option_type="down_and_in_call", barrier=50.0, **CASE).payoff_price(STRIKES)

    def doc_greeks():
        c = Compute(option_type="down_and_out_call", barrier=50.0, **CASE)
        return c.payoff_price(STRIKES), c.greeks(STRIKES)

    return {"doc_solve_price": doc, "vanilla_solve_price": vanilla, "dic_solve_price": dic,
            "doc_price_greeks": doc_greeks}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)
    calls = _calls()
    calls["doc_solve_price"]()  # device set-up and first use of each kernel
    out = {}
    for name, fn in calls.items():
        ts = []
        for _ in range(args.reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        med = float(np.median(ts))
        out[name] = {"s": round(med, 4), "min_s": round(min(ts), 4), "reference_s": REFERENCE_S[name],
                     "speedup": round(REFERENCE_S[name] / med, 1)}
    doc = calls["doc_solve_price"]()
    print(json.dumps({"metric": "wall seconds per drop-in call, reference README benchmark case",
                      "config": {"n_s": 38, "n_v": 38, "num_insert": 251, "s_max": 150.0, "strikes": STRIKES,
                                 "stage_solve": "GMRES (the reference's default)"},
                      "calls": out, "doc_prices": [round(float(x), 4) for x in doc]}))


if __name__ == "__main__":
    main()
