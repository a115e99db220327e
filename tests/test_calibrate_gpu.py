"""Calibration over batched GPU pricing: synthetic markets generated from known
parameters are fitted back from perturbed starting points (the reference's own
check, tests/test_calibrator.mojo:20-66, uses 1 % perturbations and a 10 %
recovery bar; here three problems run in one batch)."""
import numpy as np
import pytest

from fpe_option_amd import calibration as cal

pytestmark = pytest.mark.gpu

BASE = dict(S0=100.0, r=0.02, s_min=0.0, s_max=300.0, disc_mode=1, n_s=24, n_v=12, v_max=1.5, ic_sigma_s=0.5,
            ic_sigma_v=0.005, rtol=1e-6, atol=1e-9, complex_mode=1)


def test_calibration_recovers_parameters(engine):
    truth = np.array([[1.1, 0.06, 0.3, -0.35, 0.05], [2.0, 0.04, 0.5, -0.6, 0.08], [0.8, 0.1, 0.25, -0.2, 0.1]])
    K = np.tile([85.0, 95.0, 100.0, 105.0, 115.0, 90.0, 100.0, 110.0], (3, 1))
    T = np.tile([0.25, 0.25, 0.25, 0.25, 0.25, 0.75, 0.75, 0.75], (3, 1))
    market = cal.model_prices(truth[:, None, :], K, T, BASE)[:, 0]
    init = truth * np.array([1.01, 0.99, 1.01, 0.99, 1.01])
    out = cal.calibrate_batch(market, K, T, init, base=BASE, max_iter=20, tol=1e-10)
    rel = np.abs(out["params"] - truth) / np.abs(truth)
    print("relative parameter errors", rel.max(axis=1), "iterations", out["iterations"])
    assert np.all(rel < 0.10)
    assert np.max(np.abs(out["residuals"])) < 1e-4
    # the fit improved on the starting point
    start = cal.model_prices(init[:, None, :], K, T, BASE)[:, 0] - market
    assert np.all(np.sum(out["residuals"] ** 2, axis=1) < np.sum(start ** 2, axis=1))
