"""Loader of the oracle fixtures under tests/golden/ (written by
tests/golden/make_golden.py): per case the solver settings, strikes, prices,
terminal state q(T), solver counters and grid."""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INT_KEYS = {"n_s", "n_v", "option_type", "disc_mode", "num_insert", "complex_mode", "exact_constants"}
COUNTER_NAMES = ("n_accepted", "n_rejected", "n_lu", "n_newton", "n_err_solves", "n_gmres_iters", "success")


def load(group):
    """{case name: dict(kw, strikes, prices, qT, counters, dims)} of one group."""
    z = np.load(os.path.join(HERE, f"{group}.npz"), allow_pickle=False)
    out = {}
    for n in z["names"]:
        n = str(n)
        kw = {}
        for k, v in zip(z[f"{n}.keys"], z[f"{n}.vals"]):
            k = str(k)
            kw[k] = int(v) if k in INT_KEYS else float(v)
        case = dict(kw=kw, dims=tuple(int(x) for x in z[f"{n}.dims"]))
        for f in ("strikes", "prices", "qT", "M", "K"):
            if f"{n}.{f}" in z:
                case[f] = z[f"{n}.{f}"]
        if f"{n}.counters" in z:
            case["counters"] = dict(zip(COUNTER_NAMES, (int(x) for x in z[f"{n}.counters"])))
        out[n] = case
    return out
