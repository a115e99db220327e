"""GPU parity against the committed oracle fixtures (tests/golden/): cases the
oracle takes minutes for, so they are not recomputed in a test.

* config2: the single-instance 120 x 60 span grid at rtol 1e-8 (7200 dofs),
  European call and down-and-out call (B = 50, domain [50, 180]) at six
  strikes, in FP64 and in mixed precision (FP32 block LU + FP64 refinement).
* grids: one seeded instance of the 4096-option box per grid shape
  {40, 60, 80, 120} x {20, 30, 40, 60}, direct complex solve in FP64 and in
  mixed precision, and GMRES mode for the index-paired shapes.
* stencils: the GPU-assembled mass and stiffness operators of three small
  grids.

Bar: prices within 1e-8 relative and q(T) within 1e-8 of max|q(T)| whenever
the GPU solve takes the oracle's trajectory (the counters are compared per
option); the Newton bound, recorded in the session summary, when it does not.
"""
import numpy as np
import pytest

import golden_data
from conftest import record_fallback

pytestmark = pytest.mark.gpu

RTOL = 1e-8
NEWTON_BOUND = 1e-5
KEYS = ("n_accepted", "n_rejected", "n_lu", "n_newton", "n_err_solves")


def _run(cases, precision=0, lu_variant=0):
    """One GPU batch: every case at every fixture strike."""
    from fpe_option_amd import api
    rows = []
    for name, c in cases.items():
        for j, K in enumerate(c["strikes"]):
            rows.append((name, j, dict(c["kw"], strike=float(K), precision=precision, lu_variant=lu_variant)))
    cols = {k: np.array([r[2][k] for r in rows]) for k in rows[0][2]}
    g, status, info, qT = api.price_batch(return_info=True, return_qT=True, **cols)
    return rows, g, status, info, qT


def _check(rows, g, status, info, qT, cases, label):
    n_same = n_all = 0
    for i, (name, j, _) in enumerate(rows):
        c = cases[name]
        assert status[i] == 0, (name, status[i])
        oc = c["counters"]
        gmres = c["kw"]["complex_mode"] == 0 and rows[i][2]["precision"] == 0
        same = all(info[i][k] == oc[k] for k in KEYS) and (not gmres or info[i]["n_inner_iters"] == oc["n_gmres_iters"])
        o = c["prices"][j]
        err = abs(g[i] - o)
        qerr = np.max(np.abs(qT[i] - c["qT"])) / np.max(np.abs(c["qT"]))
        assert qT[i].shape == c["qT"].shape
        n_all += 1
        if same:
            n_same += 1
            assert err <= RTOL * abs(o) + 1e-13, f"{label} {name} K={c['strikes'][j]}: {g[i]} vs {o}"
            assert qerr <= RTOL, f"{label} {name}: q(T) rel err {qerr:.2e}"
        else:
            record_fallback(f"{label}:{name}:K{j}", info[i], dict(oc), err / max(abs(o), 1e-300))
            assert err <= NEWTON_BOUND * abs(o) + 1e-12, f"{label} {name}: {g[i]} vs {o}"
            assert qerr <= NEWTON_BOUND, f"{label} {name}: q(T) rel err {qerr:.2e}"
    print(f"{label}: {n_same}/{n_all} options on the oracle's trajectory")
    return n_same, n_all


@pytest.mark.parametrize("precision", [0, 1])
def test_config2_against_fixture(engine, precision):
    """120 x 60, rtol 1e-8 (src/engines/fpe/solver.mojo:98-105 settings scaled to
    the configuration), European call and down-and-out call; the down-and-in
    is their difference (in-out parity)."""
    cases = golden_data.load("config2")
    rows, g, status, info, qT = _run(cases, precision)
    _check(rows, g, status, info, qT, cases, f"config2_prec{precision}")


@pytest.mark.parametrize("mode,lu_variant", [("direct", 0), ("mixed", 0), ("gmres", 0), ("direct", 1), ("direct", 2),
                                             ("mixed", 1), ("mixed", 2)])
def test_grid_shapes_against_fixture(engine, mode, lu_variant):
    """lu_variant 0: the tuned factorisation code path of each shape; 1 and 2
    force the generic loops and the loops compiled for the block width."""
    allc = golden_data.load("grids")
    suffix = "_gmres" if mode == "gmres" else "_direct"
    cases = {k: v for k, v in allc.items() if k.endswith(suffix)}
    assert len(cases) == (4 if mode == "gmres" else 16)
    rows, g, status, info, qT = _run(cases, 1 if mode == "mixed" else 0, lu_variant)
    _check(rows, g, status, info, qT, cases, f"grids_{mode}_lu{lu_variant}")


def test_sum_factorised_products_against_fixture(engine):
    """K x and M x by the sum-factorised Kronecker products (the GMRES operator's
    form) against the oracle's assembled operators times x, to 1e-13 of
    max|A| max|x| per entry."""
    import oracle_lib
    from fpe_option_amd import api
    rng = np.random.default_rng(5)
    for name, c in golden_data.load("stencils").items():
        ns, nv = c["dims"]
        x = rng.standard_normal(ns * nv)
        Kx, Mx = api.apply_operators(x, **c["kw"])
        for got, S49, lab in ((Kx, c["K"], "K"), (Mx, c["M"], "M")):
            A = oracle_lib.dense_to_stencil_matrix(S49, ns, nv)
            err = np.max(np.abs(got - A @ x)) / (np.max(np.abs(A)) * np.max(np.abs(x)))
            assert err <= 1e-13, f"{name} {lab}: {err:.2e}"


def test_operators_against_fixture(engine):
    """Mass and stiffness operators as assembled on the GPU (de Boor basis,
    1D Gram factors, Kronecker sums: src/engines/fpe/galerkin.mojo:16-125)
    against the oracle's, entry by entry, to 1e-13 of the largest entry."""
    from fpe_option_amd import api
    for name, c in golden_data.load("stencils").items():
        M, K, dims = api.operators(**c["kw"])
        assert dims == c["dims"], name
        for A, B, lab in ((M, c["M"], "M"), (K, c["K"], "K")):
            scale = np.max(np.abs(B))
            err = np.max(np.abs(A - B)) / scale
            assert err <= 1e-13, f"{name} {lab}: max err {err:.2e} of max entry"
            # the sparsity pattern: entries outside the grid are exact zeros on both sides
            assert np.array_equal(A == 0.0, B == 0.0) or np.max(np.abs(A[B == 0.0])) <= 1e-15 * scale, (name, lab)
