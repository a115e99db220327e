"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/fpe_engine.h declares with the reference's struct layouts, and
the host-side setup (validation, truncation, knots, quadrature) matches the
oracle exactly. No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fpe_option_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HESTON = dict(kappa=1.2, theta=0.05, sigma=0.35, rho=-0.4, r=0.1, T=0.6, S0=60.0, V0=0.1)


def _header_functions():
    text = open(os.path.join(ROOT, "include", "fpe_engine.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fpe_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = L.lib()
    names = _header_functions()
    assert len(names) >= 22
    for n in names:
        assert hasattr(lib, n), n
        assert n in L.SIGNATURES, f"ctypes signature missing for {n}"


def test_reference_symbols_present():
    # the reference's 17 C symbols (cpp/include/fpe_engine.h:87-186)
    ref = ["fpe_compute_create", "fpe_compute_destroy", "fpe_compute_knots", "fpe_compute_grid_points",
           "fpe_compute_initial_condition", "fpe_compute_solve", "fpe_compute_pdf", "fpe_compute_price",
           "fpe_compute_greeks", "fpe_price_oneshot", "fpe_compute_free_vec", "fpe_compute_free_vec2",
           "fpe_compute_free_grid_pts", "fpe_compute_free_mat", "fpe_compute_free_greeks",
           "fpe_compute_free_oneshot"]
    assert set(ref) <= set(_header_functions())


def test_struct_layouts_match_reference():
    # cpp/include/fpe_engine.h:193-198
    assert C.sizeof(L.FpeVecResult) == 16
    assert C.sizeof(L.FpeVec2Result) == 32
    assert C.sizeof(L.FpeGridPtsResult) == 48
    assert C.sizeof(L.FpeMatResult) == 16
    assert C.sizeof(L.FpeGreeksResult) == 32
    assert C.sizeof(L.FpeOneshotResult) == 40


@pytest.mark.parametrize("args", [
    dict(kappa=-1.0), dict(theta=0.0), dict(sigma=0.0), dict(T=0.0), dict(S0=0.0), dict(V0=-0.1),
    dict(rho=1.5), dict(option_type=2, barrier=70.0), dict(option_type=2, barrier=0.0),
    dict(option_type=6, barrier=50.0), dict(option_type=8, barrier=10.0), dict(option_type=11),
])
def test_create_rejects_invalid_parameters(args):
    # src/server/option_types.mojo:13-32, heston_params.mojo:22-36 -> NULL handle
    p = dict(HESTON, n_s=16, n_v=16, barrier=0.0, option_type=8, num_insert=50, s_min=0.0, s_max=180.0)
    p.update(args)
    h = L.lib().fpe_compute_create(p["kappa"], p["theta"], p["sigma"], p["rho"], p["r"], p["T"], p["S0"], p["V0"],
                                   p["n_s"], p["n_v"], p["barrier"], p["option_type"], p["num_insert"], p["s_min"],
                                   p["s_max"])
    assert not h


def test_host_screened_statuses_without_a_gpu():
    """Invalid parameters (FPE_INVALID, 1) and grids wider than the kernels take
    (FPE_TOO_LARGE, 5) are decided on the host: a batch of only such options
    completes without touching the device, price 0 each."""
    from fpe_option_amd import api
    kw = dict(HESTON, kappa=[1.2, -1.0, 1.2])
    prices, status = api.price_batch(**kw, strike=60.0, option_type=[8, 8, 2], n_s=[16, 16, 16],
                                     n_v=[100, 8, 8], barrier=[0.0, 0.0, 70.0], disc_mode=1, v_max=1.5,
                                     s_max=180.0)
    assert list(status) == [5, 1, 1]
    assert np.all(prices == 0.0)


def test_null_handles_give_empty_results():
    lib = L.lib()
    r = L.FpeVecResult()
    lib.fpe_compute_initial_condition(None, C.byref(r))
    assert r.len == 0 and not r.data
    m = L.FpeMatResult()
    lib.fpe_compute_pdf(None, C.byref(m))
    assert m.n_rows == 0
    k = np.array([60.0])
    lib.fpe_compute_price(None, k.ctypes.data_as(L._DP), 1, C.byref(r))
    assert r.len == 0
    lib.fpe_compute_free_vec(C.byref(r))  # freeing an empty result is a no-op


@pytest.mark.parametrize("ns,nv,ot,B,smax,ni", [(38, 38, 8, 0.0, 180.0, 251), (38, 38, 2, 50.0, 150.0, 251),
                                                (16, 16, 8, 0.0, 180.0, 50), (38, 38, 6, 75.0, 180.0, 50),
                                                (22, 48, 1, 45.0, 180.0, 100), (42, 30, 4, 70.0, 200.0, 251)])
def test_host_domain_matches_oracle_exactly(oracle, ns, nv, ot, B, smax, ni):
    lib = L.lib()
    h = lib.fpe_compute_create(1.2, 0.05, 0.35, -0.4, 0.1, 0.6, 60.0, 0.1, ns, nv, B, ot, ni, 0.0, smax)
    assert h
    r2 = L.FpeVec2Result()
    lib.fpe_compute_knots(h, C.byref(r2))
    sk, vk = L.vec(r2.s_data, r2.s_len), L.vec(r2.v_data, r2.v_len)
    lib.fpe_compute_free_vec2(C.byref(r2))
    g = L.FpeGridPtsResult()
    lib.fpe_compute_grid_points(h, C.byref(g))
    s, v = L.vec(g.s_data, g.s_len), L.vec(g.v_data, g.v_len)
    sw, vw = L.vec(g.sw_data, g.s_len), L.vec(g.vw_data, g.v_len)
    lib.fpe_compute_free_grid_pts(C.byref(g))
    lib.fpe_compute_destroy(h)
    P = oracle.Pipe(**HESTON, n_s=ns, n_v=nv, option_type=ot, barrier=B, s_max=smax, num_insert=ni)
    osk, ovk = P.knots()
    og = P.grid()
    for a, b in ((sk, osk), (vk, ovk), (s, og["s"]), (v, og["v"]), (sw, og["sw"]), (vw, og["vw"])):
        assert a.shape == b.shape
        assert np.array_equal(a, b)


def test_python_surface_validation_errors():
    import fpe_engine as fe
    with pytest.raises(ValueError):
        fe.FpeParams(option_type="bogus")
    with pytest.raises(ValueError):
        fe.FpeParams(n_s=3)
    with pytest.raises(ValueError):
        fe.price(option_type="not_a_type")
    with pytest.raises(ValueError):
        fe.price(K=[])
    with pytest.raises(ValueError):
        fe.price(K=list(range(1, 1026)))
    with pytest.raises(ValueError):
        fe.price(option_type="down_and_out_call", barrier=70.0)  # B >= S0 for a down option
    with pytest.raises(ValueError):
        fe.Compute(S0=60.0, option_type="up_and_out_call", barrier=50.0)
    assert fe._OPTION_TYPES["european_put"] == 9


def test_native_module_surface():
    # the names the reference's Python layer imports from its native module
    # (src/bindings/_fpe_native.mojo:40-50) and the methods pricer.Compute calls
    from fpe_option_amd import native
    assert callable(native.price)
    for m in ("knots", "grid_points", "basis_1d", "basis_2d", "initial_condition", "solve", "pdf",
              "payoff_price", "greeks"):
        assert callable(getattr(native.Compute, m)), m
