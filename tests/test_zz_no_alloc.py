"""The allocation-free batch entry point (fpe_batch_price_noalloc) against
fpe_batch_price through the C ABI: the same kernels on the same inputs, so the
prices and statuses must be identical, on two calls in a row (the second with
a smaller batch, on the scratch the first one left behind), and a batch larger
than the fixed scratch is refused with a message."""
import ctypes as C

import numpy as np
import pytest

from fpe_option_amd import _lib as L
from fpe_option_amd import api

pytestmark = pytest.mark.gpu

HESTON = dict(kappa=1.2, theta=0.05, sigma=0.35, rho=-0.4, r=0.1, T=0.6, S0=60.0, V0=0.1)
SPAN = dict(disc_mode=1, v_max=1.5, s_max=180.0, ic_sigma_s=0.5, ic_sigma_v=0.005)


def _cols(n, seed):
    rng = np.random.default_rng(seed)
    types = np.arange(n) % 10
    u = rng.uniform(0.1, 0.25, n)
    return dict(kappa=rng.uniform(0.5, 4, n), theta=rng.uniform(0.02, 0.2, n), sigma=rng.uniform(0.1, 0.8, n),
                rho=rng.uniform(-0.9, 0, n), V0=rng.uniform(0.02, 0.3, n), r=rng.uniform(0, 0.1, n),
                T=rng.uniform(0.1, 2, n), S0=60.0, strike=rng.uniform(45, 75, n), option_type=types,
                barrier=np.where(types >= 8, 0.0, np.where(types <= 3, 60.0 * (1 - u), 60.0 * (1 + u))),
                n_s=rng.choice([16, 24, 40], n), n_v=rng.choice([8, 12, 20], n),
                complex_mode=rng.choice([0, 1], n), precision=np.where(np.arange(n) % 3 == 2, 1, 0), **SPAN)


def _noalloc(arr):
    n = len(arr)
    prices = np.zeros(n)
    status = np.zeros(n, dtype=np.int32)
    st = L.FpeBatchStats()
    rc = L.lib().fpe_batch_price_noalloc(arr, n, prices.ctypes.data_as(L._DP),
                                         status.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st))
    return rc, prices, status, st.as_dict()


def _both(cols):
    arr = api._batch_array(cols)
    ref, ref_status, _ = L.price_batch_raw(arr)
    rc, p, status, st = _noalloc(arr)
    assert rc == 0, L.last_error()
    return ref, ref_status, p, status, st


def test_noalloc_matches_batch_price_twice(engine):
    ref, rs, p, s, st = _both(_cols(40, 11))
    assert np.all(rs == 0)
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(p, ref)
    assert st["n_radau_slots"] >= 1
    # a second, smaller call on the scratch the first one left behind
    ref2, rs2, p2, s2, _ = _both(_cols(7, 12))
    assert np.all(rs2 == 0)
    np.testing.assert_array_equal(s2, rs2)
    np.testing.assert_array_equal(p2, ref2)


def test_noalloc_large_grids_share_fewer_slots(engine):
    # 120 x 60 solves need about 72 MB of workspace each: the scratch holds a few
    # at once and the rest queue behind them
    cols = dict(HESTON, **SPAN, n_s=120, n_v=60, strike=np.array([55.0, 60.0, 65.0, 70.0, 75.0, 80.0]),
                option_type=8, complex_mode=1)
    ref, rs, p, s, st = _both(cols)
    assert np.all(rs == 0) and np.all(s == 0)
    np.testing.assert_array_equal(p, ref)
    assert 1 <= st["n_radau_slots"] < 6


def test_noalloc_rejects_a_batch_beyond_the_scratch(engine):
    # 400 solves of 120 x 60: about 2.6 GB of operators alone
    cols = dict(HESTON, **SPAN, n_s=120, n_v=60, strike=np.linspace(50.0, 70.0, 400), option_type=8)
    rc, p, s, _ = _noalloc(api._batch_array(cols))
    assert rc == 3
    assert "scratch" in L.last_error()
