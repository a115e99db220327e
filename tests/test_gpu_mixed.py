"""Mixed precision (FPE_PREC_MIXED): FP32 block factorisation on MFMA with FP64
iterative refinement of every stage solve to a 1e-12 relative residual against
the factored FP64 operator. The refined solves equal the FP64 solves to that
residual, so prices are held to the FP64 parity bars against the oracle (1e-8
relative when the Radau trajectory matches, the Newton bound otherwise)."""
import numpy as np
import pytest

from test_gpu_parity import HESTON, _close, _check_option

pytestmark = pytest.mark.gpu


def _run(K, **kw):
    from fpe_option_amd import api
    prices, status, st, info = api.price_batch(return_stats=True, return_info=True, strike=np.asarray(K, float),
                                               **kw)
    assert np.all(status == 0), status
    return prices, st, info


@pytest.mark.parametrize("grid", [(16, 8), (40, 20)])
def test_mixed_matches_oracle(engine, oracle, grid):
    kw = dict(HESTON, disc_mode=1, n_s=grid[0], n_v=grid[1], v_max=1.5, s_max=180.0, ic_sigma_s=0.5,
              ic_sigma_v=0.005, rtol=1e-6, atol=1e-9, option_type=8, complex_mode=1)
    K = [50.0, 60.0, 70.0]
    g, st, info = _run(K, precision=1, **kw)
    o, ost, rc = oracle.price(K, **kw)
    assert rc == 0
    assert _check_option(f"mixed_{grid[0]}x{grid[1]}", g, o, info[0], ost, gmres=False)
    assert st["n_refine"] > 0
    # rounds per stage solve: FP32 factors of these well-conditioned systems need a few
    n_stage = st["n_newton"] * 2 + st["n_err_solves"]
    assert st["n_refine"] <= 6 * n_stage


def test_mixed_reference_mode_barrier(engine, oracle):
    kw = dict(HESTON, n_s=16, n_v=16, num_insert=50, s_max=180.0, option_type=2, barrier=50.0, complex_mode=1)
    K = [55.0, 60.0, 80.0]
    g, st, info = _run(K, precision=1, **kw)
    f, _, _ = _run(K, precision=0, **kw)
    o, ost, rc = oracle.price(K, **kw)
    _check_option("mixed_reference_doc", g, o, info[0], ost, gmres=False)
    _close(g, f)


def test_mixed_in_option_parity(engine):
    base = dict(HESTON, disc_mode=1, n_s=24, n_v=12, v_max=1.5, s_max=180.0, ic_sigma_s=0.5, ic_sigma_v=0.005,
                complex_mode=1, precision=1)
    K = [55.0, 65.0]
    van, _, _ = _run(K, option_type=8, **base)
    doc, _, _ = _run(K, option_type=2, barrier=50.0, **base)
    dic, _, _ = _run(K, option_type=0, barrier=50.0, **base)
    np.testing.assert_allclose(dic + doc, van, rtol=0, atol=1e-13)


def test_unknown_precision_and_complex_mode_are_rejected(engine):
    from fpe_option_amd import api
    prices, status = api.price_batch(precision=np.array([0, 2, 1, -1], dtype=np.int32), complex_mode=1,
                                     strike=60.0, **HESTON, n_s=16, n_v=16, num_insert=50, s_max=180.0,
                                     option_type=8)
    assert status[0] == 0 and status[2] == 0
    assert status[1] != 0 and status[3] != 0
    _, status = api.price_batch(complex_mode=np.array([1, 2], dtype=np.int32), strike=60.0, **HESTON, n_s=16,
                                n_v=16, num_insert=50, s_max=180.0, option_type=8)
    assert status[0] == 0 and status[1] != 0
