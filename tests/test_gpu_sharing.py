"""Work shared inside a batch does not change any result. Options on the same
s grid (European options and the vanilla legs of in-options) share one s-axis
eigenproblem of the initial-condition solve, and the initial-condition kernels
run in launch groups by grid shape; each option of a mixed batch must price
bit for bit as it does alone in a batch of one."""
import numpy as np
import pytest

from fpe_option_amd import api, workloads

pytestmark = pytest.mark.gpu


def _single(cols, i):
    return api.price_batch(**workloads.take(cols, np.array([i])))


def test_batch_prices_equal_single_option_prices():
    cols = workloads.batch_columns(24, seed=11)
    # three grids, so that several options share each s grid and the launch groups mix shapes
    cols["n_s"] = np.array([40, 60, 40] * 8)
    cols["n_v"] = np.array([20, 30, 20] * 8)
    prices, status = api.price_batch(**cols)
    assert np.all(status == 0)
    for i in range(24):
        p1, s1 = _single(cols, i)
        assert s1[0] == 0
        assert p1[0] == prices[i], (i, p1[0], prices[i])


def test_shared_s_grid_across_types_and_strikes():
    # one grid: every European and in-option vanilla leg shares the s eigenproblem
    n = 12
    cols = dict(kappa=np.linspace(0.8, 3.0, n), theta=0.05, sigma=0.4, rho=-0.5, r=0.03, T=0.5, S0=60.0,
                V0=np.linspace(0.04, 0.2, n), strike=np.linspace(50.0, 70.0, n),
                option_type=np.array([8, 9, 6, 7, 8, 9, 4, 5, 8, 9, 0, 1]),
                barrier=np.array([0, 0, 72, 72, 0, 0, 70, 70, 0, 0, 48, 48], dtype=float),
                n_s=40, n_v=20, disc_mode=1, v_max=1.5, s_min=0.0, s_max=180.0, ic_sigma_s=0.5, ic_sigma_v=0.005,
                complex_mode=1)
    prices, status = api.price_batch(**cols)
    assert np.all(status == 0)
    for i in range(n):
        p1, s1 = _single(cols, i)
        assert s1[0] == 0 and p1[0] == prices[i], (i, p1[0], prices[i])


def test_initial_condition_launch_forms_agree():
    # every ADMM launch form (FpeBatchOption.ic_variant 1..3, and 0: the library's choice)
    # computes the same products; only the residual norms' partial sums are split differently
    cols = workloads.batch_columns(16, seed=5)
    cols["n_s"] = np.array([40, 80, 120, 60] * 4)
    cols["n_v"] = np.array([20, 40, 30, 60] * 4)
    cols["T"] = np.full(16, 0.05)
    ref, status = api.price_batch(**cols, ic_variant=0)
    assert np.all(status == 0)
    for v in (1, 2, 3):
        p, s = api.price_batch(**cols, ic_variant=v)
        assert np.all(s == 0)
        np.testing.assert_allclose(p, ref, rtol=1e-11, atol=1e-13, err_msg=f"ic_variant {v}")


def test_concurrent_launch_groups_are_deterministic():
    # 16 grid shapes: several eigenproblem and ADMM launch groups run concurrently on
    # the library's streams, and the Radau queue order depends on the batch's mix;
    # neither may change a price. Repeat the batch, then price it in two halves.
    cols = workloads.batch_columns(64, seed=3)
    p1, s1 = api.price_batch(**cols)
    assert np.all(s1 == 0)
    p2, _ = api.price_batch(**cols)
    assert np.array_equal(p1, p2)
    idx = np.argsort(np.asarray(cols["n_s"]) * 100 + np.asarray(cols["n_v"]), kind="stable")
    parts = np.empty_like(p1)
    for half in (idx[:32], idx[32:]):
        ph, sh = api.price_batch(**workloads.take(cols, half))
        assert np.all(sh == 0)
        parts[half] = ph
    assert np.array_equal(parts, p1)
