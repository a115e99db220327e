"""The trained surrogate (fpe_option_amd/surrogate_80x40.npz) against its accuracy reference,
the plain FP64 solve on the same 80x40 grid, on fresh seeded parameter sets
that were not in the training data; and what the network adds over the same
POD pricing without it (every option priced at the mean coefficients)."""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from fpe_option_amd import surrogate as S

needs_model = pytest.mark.skipif(not os.path.exists(S.MODEL_PATH), reason="surrogate model not built")


@pytest.fixture(scope="module")
def sur(engine):
    return S.Surrogate()


@pytest.fixture(scope="module")
def fresh():
    n = 96
    P = S.sample_params(n, 777)  # training used seeds 20261204 and 20261205
    rng = np.random.default_rng(777)
    K = rng.uniform(45.0, 75.0, n)
    ty = np.where(np.arange(n) % 2 == 0, 8, 9).astype(np.int32)
    return P, K, ty


@needs_model
def test_surrogate_against_plain_solve(sur, fresh, engine):
    P, K, ty = fresh
    sp, ok = sur.price(P, K, ty)
    assert ok.all()
    ref, st = S.solver_prices(P, K, ty)
    assert np.all(st == 0)
    d = np.abs(sp - ref)
    # the no-network baseline: the same POD pricing at the mean coefficients
    C0 = S.DeviceArray.from_numpy(np.zeros((len(P), sur.k), np.float32))
    out = S.DeviceArray((len(P),), np.float64)
    args = [S.DeviceArray.from_numpy(a) for a in (K.astype(np.float64), ty, P[:, 5].copy(), P[:, 6].copy())]
    assert S.L.lib().fpe_surrogate_price(C0.c, len(P), sur.k, sur.cscale.c, sur.cshift.c, sur.Pt.c, sur.w0.c,
                                         sur.sq.c, sur.nq, args[0].c, args[1].c, args[2].c, args[3].c, out.c,
                                         None) == 0
    d0 = np.abs(out.numpy() - ref)
    print(f"surrogate |dprice| max {d.max():.3e} mean {d.mean():.3e}; without the network max {d0.max():.3e} "
          f"mean {d0.mean():.3e}")
    assert np.mean(d) < 0.05 * np.mean(d0)
    assert np.max(d) < 0.5


@needs_model
def test_out_of_box_rows_fall_back_to_the_solver(sur, fresh, engine):
    P, K, ty = fresh
    P = P[:6].copy()
    P[1, 2] = 0.95   # sigma above the box
    P[4, 6] = 2.5    # T above the box
    prices, ok, n_fb = sur.price_with_fallback(P, K[:6], ty[:6])
    assert n_fb == 2 and not ok[1] and not ok[4]
    ref, st = S.solver_prices(P[[1, 4]], K[[1, 4]], ty[[1, 4]])
    np.testing.assert_array_equal(prices[[1, 4]], ref)


@needs_model
def test_surrogate_throughput_gain(sur, fresh, engine):
    P, K, ty = fresh
    n = 65536
    Pb = S.sample_params(n, 4242)
    Kb = np.full(n, 60.0)
    tb = np.full(n, 8, np.int32)
    sur.price(Pb[:1024], Kb[:1024], tb[:1024])
    t = time.perf_counter()
    sur.price(Pb, Kb, tb)
    ts = (time.perf_counter() - t) / n
    t = time.perf_counter()
    S.solver_prices(P[:32], K[:32], ty[:32])
    tp = (time.perf_counter() - t) / 32
    print(f"per option: surrogate {ts * 1e6:.2f} us, plain solve {tp * 1e3:.1f} ms (batch of 32)")
    assert tp / ts > 1000
