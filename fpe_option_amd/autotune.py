"""Per-grid autotuner of the block-banded factorisation (run on the GPU).

For every grid shape (n_s, n_v) in {40, 60, 80, 120} x {20, 30, 40, 60} it
prices the same seeded single-shape batch (the 4096-option parameter box,
span discretisation, direct complex stage solve) once with each factorisation
code path -- the generic loops (one code path for every block width) and the
MFMA loops compiled for the block width's k-step count -- and keeps the one
with the shorter Radau kernel time. The choice is written to
tuning_gfx950.txt next to libfpe_engine.so (read by the library at first use,
one line "n_sb n_vb variant" per shape, variant 0 generic / 1 compiled) and the
measurements to tuning_gfx950.json.

Usage: python -m fpe_option_amd.autotune [--n OPTIONS] [--reps R] [--out DIR]
"""
import argparse
import json
import os
import time

import numpy as np

from . import api, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning_gfx950.txt")
RECORD = os.path.join(HERE, "tuning_gfx950.json")
SHAPES = [(ns, nv) for nv in (20, 30, 40, 60) for ns in (40, 60, 80, 120)]
PATHS = {1: "generic", 2: "compiled_k"}  # FpeBatchOption.lu_variant values


def measure(shape, n, reps):
    cols = workloads.batch_columns(n, grid=shape)
    out = {}
    for lu in PATHS:
        c = dict(cols, lu_variant=lu)
        api.price_batch(**c)  # warm-up
        radau, wg = [], []
        for _ in range(reps):
            prices, status, st, info = api.price_batch(return_stats=True, return_info=True, **c)
            assert np.all(status == 0), f"{shape} {PATHS[lu]}: status {np.unique(status)}"
            radau.append(st["ms_radau"])
            wg.append(st["wg_s_lu"])
        out[PATHS[lu]] = {"ms_radau": float(np.median(radau)), "wg_s_lu": float(np.median(wg)),
                          "prices": prices}
        dims = (info[0]["n_sb"], info[0]["n_vb"])
    # both paths solve the same systems: the prices agree to factorisation roundoff
    d = float(np.max(np.abs(out["generic"].pop("prices") - out["compiled_k"].pop("prices"))))
    return dims, out, d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024, help="options per shape")
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--out", default=HERE, help="directory for the table and the record")
    args = ap.parse_args()
    table, record = os.path.join(args.out, os.path.basename(TABLE)), os.path.join(args.out, os.path.basename(RECORD))
    t0 = time.perf_counter()
    rows, rec = [], []
    for shape in SHAPES:
        dims, res, d = measure(shape, args.n, args.reps)
        best = min(PATHS, key=lambda v: res[PATHS[v]]["ms_radau"])
        gain = res["generic"]["ms_radau"] / res[PATHS[best]]["ms_radau"]
        rows.append((dims[0], dims[1], best - 1))
        rec.append({"n_s": shape[0], "n_v": shape[1], "n_sb": dims[0], "n_vb": dims[1], "chosen": PATHS[best],
                    "speedup_vs_generic": round(gain, 4), "max_abs_dprice_between_paths": d, **res})
        print(json.dumps(rec[-1]), flush=True)
    os.makedirs(args.out, exist_ok=True)
    with open(table, "w") as f:
        f.write("# n_sb n_vb variant (0 generic loops, 1 compiled k-steps); written by fpe_option_amd.autotune\n")
        for r in rows:
            f.write(f"{r[0]} {r[1]} {r[2]}\n")
    with open(record, "w") as f:
        json.dump({"options_per_shape": args.n, "reps": args.reps, "seconds": round(time.perf_counter() - t0, 1),
                   "shapes": rec}, f, indent=1)
    print("wrote", table, "and", record)


if __name__ == "__main__":
    main()
