"""Writes the oracle fixtures under tests/golden/ (run in the build container:
`python tests/golden/make_golden.py [group ...]`).

The FP64 CPU oracle (oracle/fpe_oracle.cpp) is run on cases too slow to run
inside a test, and its outputs are stored as .npz vectors (numpy, no pickle):
prices at several strikes, the terminal state q(T) (recombined B-spline
coefficients, dof r = I * n_vb + i) and the solver counters, so GPU tests can
hold the HIP path to 1e-8 on the same trajectory. Groups:

  config2   single instance, 120 x 60 span grid, rtol 1e-8 (European call on
            [0, 180] and down-and-out call B = 50 on [50, 180]; the
            down-and-in is their difference), direct complex stage solve
  grids     one seeded instance of the 4096-option parameter box per grid
            shape (n_s, n_v) in {40, 60, 80, 120} x {20, 30, 40, 60}, rtol 1e-4,
            direct complex stage solve; GMRES mode for the four index-paired
            shapes
  stencils  unscaled M and K 49-point stencils of two small grids (reference
            and span discretisation, one of them barrier-truncated)
"""
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib  # noqa: E402

HESTON = dict(kappa=1.2, theta=0.05, sigma=0.35, rho=-0.4, r=0.1, T=0.6, S0=60.0, V0=0.1)
SPAN = dict(disc_mode=1, v_max=1.5, s_max=180.0, ic_sigma_s=0.5, ic_sigma_v=0.005)
GRIDS_S, GRIDS_V = (40, 60, 80, 120), (20, 30, 40, 60)
PAIRED = ((40, 20), (60, 30), (80, 40), (120, 60))
NON_IN_TYPES = (8, 9, 2, 3, 6, 7)


def solve_case(args):
    """One non-in solve on the (truncated) domain: counters, q(T), prices."""
    name, kw, strikes = args
    t0 = time.time()
    P = oracle_lib.Pipe(**kw)
    st = P.solve()
    qT, tT = P.state(P.n_states - 1)
    prices = P.price_at(strikes)
    dt = time.time() - t0
    print(f"{name}: {dt:.1f} s, N={P.N} ({P.n_sb}x{P.n_vb}), steps {st['n_accepted']}+{st['n_rejected']}, "
          f"lu {st['n_lu']}, newton {st['n_newton']}, gmres {st['n_gmres_iters']}, prices {prices}", flush=True)
    counters = np.array([st[c] for c in ("n_accepted", "n_rejected", "n_lu", "n_newton", "n_err_solves",
                                         "n_gmres_iters", "success")], dtype=np.int64)
    return name, kw, np.asarray(strikes, dtype=np.float64), prices, qT, counters, (P.n_sb, P.n_vb), tT


def _pack(results, path):
    out = {}
    names = []
    for name, kw, strikes, prices, qT, counters, dims, tT in results:
        names.append(name)
        keys = sorted(kw)
        out[f"{name}.keys"] = np.array(keys)
        out[f"{name}.vals"] = np.array([float(kw[k]) for k in keys])
        out[f"{name}.strikes"] = strikes
        out[f"{name}.prices"] = prices
        out[f"{name}.qT"] = qT
        out[f"{name}.counters"] = counters
        out[f"{name}.dims"] = np.array(dims, dtype=np.int64)
        out[f"{name}.t_end"] = np.array(tT)
    out["names"] = np.array(names)
    np.savez_compressed(path, **out)
    print("wrote", path, flush=True)


def config2_cases():
    base = dict(HESTON, **SPAN, n_s=120, n_v=60, rtol=1e-8, atol=1e-11, complex_mode=1)
    K = [50.0, 55.0, 60.0, 65.0, 70.0, 80.0]
    return [("call", dict(base, option_type=8, barrier=0.0), K),
            ("doc", dict(base, option_type=2, barrier=50.0), K)]


def grid_cases():
    """One instance per grid shape, parameters from the 4096-option box
    (seed 20261204): kappa U[0.5,4], theta U[0.02,0.2], eta U[0.1,0.8],
    rho U[-0.9,0], V0 U[0.02,0.3], r U[0,0.1], T U[0.1,2], S0 = 60,
    K U[45,75], barrier 10-25 % from S0."""
    rng = np.random.default_rng(20261204)
    cases = []
    k = 0
    for ns in GRIDS_S:
        for nv in GRIDS_V:
            t = NON_IN_TYPES[k % len(NON_IN_TYPES)]
            u = rng.uniform(0.10, 0.25)
            p = dict(kappa=rng.uniform(0.5, 4), theta=rng.uniform(0.02, 0.2), sigma=rng.uniform(0.1, 0.8),
                     rho=rng.uniform(-0.9, 0), V0=rng.uniform(0.02, 0.3), r=rng.uniform(0, 0.1),
                     T=rng.uniform(0.1, 2), S0=60.0)
            K0 = rng.uniform(45, 75)
            B = 0.0 if t >= 8 else (60.0 * (1 - u) if t <= 3 else 60.0 * (1 + u))
            kw = dict(p, **SPAN, n_s=ns, n_v=nv, rtol=1e-4, atol=1e-6, option_type=t, barrier=B)
            strikes = [K0, 55.0, 60.0, 65.0]
            cases.append((f"g{ns}x{nv}_direct", dict(kw, complex_mode=1), strikes))
            if (ns, nv) in PAIRED:
                cases.append((f"g{ns}x{nv}_gmres", dict(kw, complex_mode=0), strikes))
            k += 1
    return cases


def stencil_fixture(path):
    out = {}
    cases = {
        "ref16_doc": dict(HESTON, n_s=16, n_v=16, num_insert=50, s_max=180.0, option_type=2, barrier=50.0),
        "span16x8": dict(HESTON, **SPAN, n_s=16, n_v=8, option_type=8),
        "span12x10_uoc": dict(HESTON, **SPAN, n_s=12, n_v=10, option_type=6, barrier=72.0),
    }
    for name, kw in cases.items():
        P = oracle_lib.Pipe(**kw)
        M, K = P.stencils()
        keys = sorted(kw)
        out[f"{name}.keys"] = np.array(keys)
        out[f"{name}.vals"] = np.array([float(kw[k]) for k in keys])
        out[f"{name}.M"] = M
        out[f"{name}.K"] = K
        out[f"{name}.dims"] = np.array([P.n_sb, P.n_vb], dtype=np.int64)
    out["names"] = np.array(list(cases))
    np.savez_compressed(path, **out)
    print("wrote", path, flush=True)


def main(groups):
    oracle_lib.build()
    if "stencils" in groups:
        stencil_fixture(os.path.join(HERE, "stencils.npz"))
    jobs = []
    if "config2" in groups:
        jobs += [("config2", c) for c in config2_cases()]
    if "grids" in groups:
        jobs += [("grids", c) for c in grid_cases()]
    if not jobs:
        return
    # largest first so the pool finishes together
    order = sorted(range(len(jobs)), key=lambda i: -jobs[i][1][1]["n_s"] * jobs[i][1][1]["n_v"] ** 3 *
                   (100 if jobs[i][1][1]["rtol"] < 1e-6 else 1))
    with Pool(int(os.environ.get("GOLDEN_PROCS", "6"))) as pool:
        res = pool.map(solve_case, [jobs[i][1] for i in order], chunksize=1)
    by_group = {}
    for i, r in zip(order, res):
        by_group.setdefault(jobs[i][0], []).append(r)
    for g, rs in by_group.items():
        rs.sort(key=lambda r: r[0])
        _pack(rs, os.path.join(HERE, f"{g}.npz"))


if __name__ == "__main__":
    main(sys.argv[1:] or ["stencils", "grids", "config2"])
