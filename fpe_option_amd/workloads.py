"""Seeded synthetic option batches for the batched benchmarks.

The 4096-option batch: Heston parameters drawn uniformly from
kappa in [0.5, 4], theta in [0.02, 0.2], eta in [0.1, 0.8], rho in [-0.9, 0],
V0 in [0.02, 0.3], r in [0, 0.1], T in [0.1, 2], S0 = 60, K in [45, 75];
grids (n_s, n_v) with n_s and n_v drawn independently from {40, 60, 80, 120}
and {20, 30, 40, 60} (all 16 shapes; pairing="paired" draws one index for
both, the 4 shapes of aspect 2:1); 50 % European calls/puts, 50 % single-barrier out/in options
over {down, up} x {out, in} x {call, put} with B = S0 (1 -+ u), u in [0.10, 0.25].
Span discretisation: cubic B-splines with sinh-clustered knots, 4-point
Gauss-Legendre per span, s in [0, 180], v in [0, 1.5], initial Gaussian
sigma = (0.5, 0.005). Generator: numpy PCG64 with the given seed.
"""
import numpy as np

SEED = 20261203
GRIDS = np.array([[40, 20], [60, 30], [80, 40], [120, 60]])
BARRIER_TYPES = np.array([2, 3, 6, 7, 0, 1, 4, 5])  # DO call/put, UO call/put, DI call/put, UI call/put


def batch_columns(n=4096, seed=SEED, rtol=1e-4, atol=1e-6, complex_mode=1, precision=0, grid=None,
                  exact_constants=0, pairing="independent"):
    rng = np.random.Generator(np.random.PCG64(seed))
    cols = dict(kappa=rng.uniform(0.5, 4.0, n), theta=rng.uniform(0.02, 0.2, n), sigma=rng.uniform(0.1, 0.8, n),
                rho=rng.uniform(-0.9, 0.0, n), V0=rng.uniform(0.02, 0.3, n), r=rng.uniform(0.0, 0.1, n),
                T=rng.uniform(0.1, 2.0, n))
    S0 = 60.0
    cols["S0"] = np.full(n, S0)
    cols["strike"] = rng.uniform(45.0, 75.0, n)
    k = rng.integers(0, 4, n)
    kv = k if pairing == "paired" else rng.integers(0, 4, n)
    cols["n_s"] = GRIDS[k, 0] if grid is None else np.full(n, grid[0])
    cols["n_v"] = GRIDS[kv, 1] if grid is None else np.full(n, grid[1])
    euro = rng.uniform(0.0, 1.0, n) < 0.5
    call = rng.uniform(0.0, 1.0, n) < 0.5
    bt = BARRIER_TYPES[rng.integers(0, 8, n)]
    u = rng.uniform(0.10, 0.25, n)
    types = np.where(euro, np.where(call, 8, 9), bt)
    down = (types <= 3)
    cols["option_type"] = types.astype(np.int32)
    cols["barrier"] = np.where(types >= 8, 0.0, np.where(down, S0 * (1.0 - u), S0 * (1.0 + u)))
    cols.update(disc_mode=1, v_max=1.5, s_min=0.0, s_max=3.0 * S0, ic_sigma_s=0.5, ic_sigma_v=0.005, rtol=rtol,
                atol=atol, complex_mode=complex_mode, precision=precision, exact_constants=exact_constants,
                knot_width_s=0.1, knot_width_v=0.1)
    return cols


def n_solves(cols):
    t = np.asarray(cols["option_type"])
    return int(np.sum(np.where(np.isin(t, [0, 1, 4, 5]), 2, 1)))


def take(cols, idx):
    """Sub-batch with the given indices (scalars are kept)."""
    return {k: (v[idx] if isinstance(v, np.ndarray) else v) for k, v in cols.items()}
