"""Python surface of the engine, mirroring the reference `fpe_engine` package
(python/fpe_engine/__init__.py:50-103, python/fpe_engine/pricer.py:12-197) over
the C ABI in include/fpe_engine.h. All numerics run on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
from scipy import sparse

from . import _lib as L

_OPTION_TYPES = {
    "down_and_in_call": 0, "down_and_in_put": 1,
    "down_and_out_call": 2, "down_and_out_put": 3,
    "up_and_in_call": 4, "up_and_in_put": 5,
    "up_and_out_call": 6, "up_and_out_put": 7,
    "european_call": 8, "european_put": 9,
}
MAX_STRIKES = 1024  # src/bindings/_params.mojo:5


@dataclass
class PriceResult:
    prices: np.ndarray
    deltas: np.ndarray
    gammas: np.ndarray
    vegas: np.ndarray


@dataclass
class KnotsResult:
    s: np.ndarray
    v: np.ndarray


@dataclass
class GridPointsResult:
    s: np.ndarray
    v: np.ndarray
    s_weights: np.ndarray
    v_weights: np.ndarray


@dataclass
class Basis1DResult:
    Bs: sparse.csr_matrix
    dBs: sparse.csr_matrix
    Bv: sparse.csr_matrix
    dBv: sparse.csr_matrix


@dataclass
class GreeksResult:
    delta: np.ndarray
    gamma: np.ndarray
    vega: np.ndarray


@dataclass
class FpeParams:
    """Parameter container with the reference's eager checks (pricer.py:65-93)."""
    kappa: float = 1.2
    theta: float = 0.05
    sigma: float = 0.35
    rho: float = -0.4
    r: float = 0.05
    T: float = 0.5
    S0: float = 100.0
    V0: float = 0.1
    n_s: int = 38
    n_v: int = 38
    option_type: str | int = "european_call"
    K: float | list | None = None
    barrier: float = 0.0
    s_min: float | None = None
    s_max: float | None = None

    def __post_init__(self):
        if isinstance(self.option_type, str):
            if self.option_type not in _OPTION_TYPES:
                raise ValueError(f"unknown option_type '{self.option_type}'")
        else:
            if not (0 <= self.option_type <= 9):
                raise ValueError(f"option_type must be 0-9, got {self.option_type}")
        if not (4 <= self.n_s <= 256):
            raise ValueError(f"n_s must be in [4, 256], got {self.n_s}")
        if not (4 <= self.n_v <= 256):
            raise ValueError(f"n_v must be in [4, 256], got {self.n_v}")


def is_available() -> bool:
    """True when the HIP library loads and a GPU is present."""
    try:
        return L.lib().fpe_device_count() > 0
    except (OSError, RuntimeError):
        return False


def _option_type_int(ot) -> int:
    # src/bindings/_params.mojo:7-41
    if isinstance(ot, str):
        if ot not in _OPTION_TYPES:
            raise ValueError(f"unknown option_type: {ot}")
        return _OPTION_TYPES[ot]
    if isinstance(ot, (int, np.integer)) and not isinstance(ot, bool):
        if ot < 0 or ot > 9:
            raise ValueError(f"option_type must be 0-9, got {ot}")
        return int(ot)
    raise ValueError("option_type must be str or int")


def _strikes(K) -> list:
    if isinstance(K, np.ndarray):
        K = K.tolist()
    if isinstance(K, (list, tuple)):
        if len(K) > MAX_STRIKES:
            raise ValueError(f"K list too large, max {MAX_STRIKES}")
        ks = [float(k) for k in K]
    elif isinstance(K, (int, float, np.floating, np.integer)) and not isinstance(K, bool):
        ks = [float(K)]
    else:
        raise ValueError("K must be float, int, or list of floats")
    if not ks:
        raise ValueError("K must not be empty")
    return ks


def _parse(kw: dict) -> dict:
    """Keyword parsing and validation of src/bindings/_params.mojo:53-121."""
    p = dict(kappa=float(kw.get("kappa", 1.2)), theta=float(kw.get("theta", 0.05)),
             sigma=float(kw.get("sigma", 0.35)), rho=float(kw.get("rho", -0.4)), r=float(kw.get("r", 0.05)),
             T=float(kw.get("T", 0.5)), S0=float(kw.get("S0", 100.0)), V0=float(kw.get("V0", 0.1)),
             n_s=int(kw.get("n_s", 38)), n_v=int(kw.get("n_v", 38)), barrier=float(kw.get("barrier", 0.0)))
    p["s_min"] = float(kw.get("s_min", 0.0))
    p["s_max"] = float(kw.get("s_max", p["S0"] * 3.0))
    if p["n_s"] < 4 or p["n_s"] > 256:
        raise ValueError(f"n_s must be in [4, 256], got {p['n_s']}")
    if p["n_v"] < 4 or p["n_v"] > 256:
        raise ValueError(f"n_v must be in [4, 256], got {p['n_v']}")
    if "option_type" not in kw:
        raise KeyError("option_type")
    p["option_type"] = _option_type_int(kw["option_type"])
    p["K"] = _strikes(kw.get("K", 100.0))
    ok = (p["kappa"] > 0 and p["theta"] > 0 and p["sigma"] > 0 and p["T"] > 0 and p["S0"] > 0 and p["V0"] >= 0
          and p["s_min"] >= 0 and p["s_max"] > p["s_min"] and -1.0 <= p["rho"] <= 1.0)
    if not ok:
        raise ValueError("invalid Heston parameters")
    t, B = p["option_type"], p["barrier"]
    valid = all(k > 0 for k in p["K"])
    if t <= 3:
        valid = valid and 0 < B < p["S0"]
    elif t <= 7:
        valid = valid and B > 0 and B >= p["S0"]
    else:
        valid = valid and B == 0.0
    if not valid:
        raise ValueError("invalid FPE parameters (check barrier/option_type combo)")
    return p


def price(kappa: float = 1.2, theta: float = 0.05, sigma: float = 0.35, rho: float = -0.4, r: float = 0.1,
          T: float = 0.6, S0: float = 60.0, V0: float = 0.1, K=100.0, barrier: float = 0.0,
          option_type: str | int = "european_call", n_s: int = 38, n_v: int = 38,
          s_min: float | None = None, s_max: float | None = None) -> PriceResult:
    """One-shot price + Greeks (python/fpe_engine/__init__.py:63-103): one base
    and four bumped solves, priced together in one GPU batch (num_insert = 50)."""
    if isinstance(K, (int, float)):
        K = [float(K)]
    else:
        K = [float(k) for k in K]
    if isinstance(option_type, str):
        if option_type not in _OPTION_TYPES:
            raise ValueError(f"unknown option_type '{option_type}'")
        option_type = _OPTION_TYPES[option_type]
    kw = dict(kappa=kappa, theta=theta, sigma=sigma, rho=rho, r=r, T=T, S0=S0, V0=V0, K=K, barrier=barrier,
              option_type=int(option_type), n_s=n_s, n_v=n_v)
    if s_min is not None:
        kw["s_min"] = s_min
    if s_max is not None:
        kw["s_max"] = s_max
    p = _parse(kw)
    ks = np.ascontiguousarray(p["K"], dtype=np.float64)
    res = L.FpeOneshotResult()
    L.lib().fpe_price_oneshot(p["kappa"], p["theta"], p["sigma"], p["rho"], p["r"], p["T"], p["S0"], p["V0"],
                              ks.ctypes.data_as(L._DP), len(ks), p["barrier"], p["option_type"], p["n_s"],
                              p["n_v"], 50, p["s_min"], p["s_max"], C.byref(res))
    if res.len == 0:
        raise RuntimeError(f"pricing failed: {L.last_error()}")
    out = PriceResult(prices=L.vec(res.price, res.len), deltas=L.vec(res.delta, res.len),
                      gammas=L.vec(res.gamma, res.len), vegas=L.vec(res.vega, res.len))
    L.lib().fpe_compute_free_oneshot(C.byref(res))
    return out


class Compute:
    """Stepwise pipeline (python/fpe_engine/pricer.py:104-197): properties are
    cached on the Python side; the handle caches device results natively."""

    def __init__(self, **kwargs):
        if "option_type" not in kwargs:
            kwargs["option_type"] = "european_put"
        kw = {k: v for k, v in kwargs.items() if v is not None}
        num_insert = int(kw.pop("num_insert", 251))
        p = _parse(kw)
        self._p = p
        self._h = L.lib().fpe_compute_create(p["kappa"], p["theta"], p["sigma"], p["rho"], p["r"], p["T"], p["S0"],
                                             p["V0"], p["n_s"], p["n_v"], p["barrier"], p["option_type"],
                                             num_insert, p["s_min"], p["s_max"])
        if not self._h:
            raise ValueError(f"invalid FPE parameters: {L.last_error()}")
        self._knots_cache = None
        self._grid_points_cache = None
        self._basis_1d_cache = None
        self._basis_2d_cache = None
        self._ic_cache = None
        self._solve_cache = None
        self._pdf_cache = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib().fpe_compute_destroy(h)
            self._h = None

    @property
    def knots(self) -> KnotsResult:
        if self._knots_cache is None:
            r = L.FpeVec2Result()
            L.lib().fpe_compute_knots(self._h, C.byref(r))
            self._knots_cache = KnotsResult(s=L.vec(r.s_data, r.s_len), v=L.vec(r.v_data, r.v_len))
            L.lib().fpe_compute_free_vec2(C.byref(r))
        return self._knots_cache

    @property
    def grid_points(self) -> GridPointsResult:
        if self._grid_points_cache is None:
            r = L.FpeGridPtsResult()
            L.lib().fpe_compute_grid_points(self._h, C.byref(r))
            self._grid_points_cache = GridPointsResult(
                s=L.vec(r.s_data, r.s_len), v=L.vec(r.v_data, r.v_len),
                s_weights=L.vec(r.sw_data, r.s_len), v_weights=L.vec(r.vw_data, r.v_len))
            L.lib().fpe_compute_free_grid_pts(C.byref(r))
        return self._grid_points_cache

    def _csr(self, axis, deriv):
        r = L.FpeCsrResult()
        L.lib().fpe_compute_basis_1d(self._h, axis, deriv, C.byref(r))
        if not r.indptr:
            raise RuntimeError(f"basis evaluation failed: {L.last_error()}")
        data = L.vec(r.data, r.nnz)
        idx = np.ctypeslib.as_array(r.indices, shape=(max(r.nnz, 1),))[:r.nnz].copy()
        ptr = np.ctypeslib.as_array(r.indptr, shape=(r.n_rows + 1,)).copy()
        m = sparse.csr_matrix((data, idx, ptr), shape=(r.n_rows, r.n_cols))
        L.lib().fpe_compute_free_csr(C.byref(r))
        return m

    @property
    def basis_1d(self) -> Basis1DResult:
        if self._basis_1d_cache is None:
            self._basis_1d_cache = Basis1DResult(Bs=self._csr(0, 0), dBs=self._csr(0, 1), Bv=self._csr(1, 0),
                                                 dBv=self._csr(1, 1))
        return self._basis_1d_cache

    @property
    def basis_2d(self) -> sparse.csr_matrix:
        if self._basis_2d_cache is None:
            b = self.basis_1d
            self._basis_2d_cache = sparse.kron(b.Bs, b.Bv, format="csr")
        return self._basis_2d_cache

    @property
    def initial_condition(self) -> np.ndarray:
        if self._ic_cache is None:
            r = L.FpeVecResult()
            L.lib().fpe_compute_initial_condition(self._h, C.byref(r))
            if r.len == 0:
                raise RuntimeError(f"initial condition failed: {L.last_error()}")
            self._ic_cache = L.vec(r.data, r.len)
            L.lib().fpe_compute_free_vec(C.byref(r))
        return self._ic_cache

    @property
    def solve(self) -> list:
        if self._solve_cache is None:
            r = L.FpeMatResult()
            L.lib().fpe_compute_solve(self._h, C.byref(r))
            if r.n_rows == 0:
                raise RuntimeError(f"solve failed: {L.last_error()}")
            m = L.vec(r.data, r.n_rows * r.n_cols).reshape(r.n_rows, r.n_cols)
            L.lib().fpe_compute_free_mat(C.byref(r))
            self._solve_cache = [m[i].copy() for i in range(m.shape[0])]
        return self._solve_cache

    @property
    def pdf(self) -> np.ndarray:
        if self._pdf_cache is None:
            r = L.FpeMatResult()
            L.lib().fpe_compute_pdf(self._h, C.byref(r))
            if r.n_rows == 0:
                raise RuntimeError(f"pdf failed: {L.last_error()}")
            self._pdf_cache = L.vec(r.data, r.n_rows * r.n_cols).reshape(r.n_rows, r.n_cols)
            L.lib().fpe_compute_free_mat(C.byref(r))
        return self._pdf_cache

    def payoff_price(self, K) -> np.ndarray:
        ks = np.ascontiguousarray(_strikes(K), dtype=np.float64)
        r = L.FpeVecResult()
        L.lib().fpe_compute_price(self._h, ks.ctypes.data_as(L._DP), len(ks), C.byref(r))
        if r.len != len(ks):
            raise RuntimeError(f"pricing failed: {L.last_error()}")
        out = L.vec(r.data, r.len)
        L.lib().fpe_compute_free_vec(C.byref(r))
        return out

    def greeks(self, K, rel_s: float = 0.01, rel_v: float = 0.1) -> GreeksResult:
        ks = np.ascontiguousarray(_strikes(K), dtype=np.float64)
        r = L.FpeGreeksResult()
        L.lib().fpe_compute_greeks(self._h, ks.ctypes.data_as(L._DP), len(ks), rel_s, rel_v, C.byref(r))
        if r.len != len(ks):
            raise RuntimeError(f"greeks failed: {L.last_error()}")
        out = GreeksResult(delta=L.vec(r.delta, r.len), gamma=L.vec(r.gamma, r.len), vega=L.vec(r.vega, r.len))
        L.lib().fpe_compute_free_greeks(C.byref(r))
        return out


_BATCH_FIELDS = ("kappa", "theta", "sigma", "rho", "r", "T", "S0", "V0", "s_min", "s_max", "v_max", "barrier",
                 "strike", "option_type", "n_s", "n_v", "disc_mode", "num_insert", "complex_mode", "precision",
                 "exact_constants", "ic_sigma_s", "ic_sigma_v", "rtol", "atol", "knot_width_s", "knot_width_v",
                 "knot_center_v", "lu_variant", "max_steps", "refine_max", "ic_variant")


def _batch_array(columns):
    """FpeBatchOption array from scalar or per-option columns, written through a
    numpy view of the ctypes array (one assignment per field)."""
    n = max([np.size(v) for v in columns.values()] + [1])
    arr = L.batch_options(n)
    view = np.ctypeslib.as_array(arr)
    for k, v in columns.items():
        if k not in _BATCH_FIELDS:
            raise KeyError(k)
        if k == "option_type":
            a = np.asarray(v)
            if a.dtype.kind in "iu":
                if a.size and (a.min() < 0 or a.max() > 9):
                    raise ValueError(f"option_type must be 0-9, got {a.min() if a.min() < 0 else a.max()}")
            else:
                v = [_option_type_int(x) for x in np.atleast_1d(v)] if not np.isscalar(v) else _option_type_int(v)
        a = np.asarray(v)
        if view.dtype[k].kind == "i" and a.dtype.kind not in "iub":
            raise TypeError(f"{k} must be an integer, got {a.dtype}")
        if a.dtype.kind not in "iubf":
            raise TypeError(f"{k} must be numeric, got {a.dtype}")
        view[k] = np.broadcast_to(a, (n,))
    return arr


def price_batch(return_stats: bool = False, return_info: bool = False, return_qT: bool = False, **columns):
    """Price many independent options in one GPU call (extension; no
    counterpart in the reference). Each keyword is a scalar or an array of one
    value per option; names follow FpeBatchOption. Returns (prices, status),
    then, when asked, the batch statistics (return_stats), a list of per-option
    solver records (return_info: status, solve counters, grid) and a list of
    terminal states q(T) (return_qT; None for in-options)."""
    arr = _batch_array(columns)
    prices, status, stats, *rest = L.price_batch_raw(arr, want_info=return_info, want_qT=return_qT)
    out = (prices, status)
    if return_stats:
        out += (stats,)
    return out + tuple(rest)


def terminal_states(**columns):
    """q(T) of vanilla or out options (one solve each, on the barrier-truncated
    domain for out-types) as a list of arrays, with the per-option status."""
    arr = _batch_array(columns)
    n = len(arr)
    ldq = L.MAX_DOF
    q = np.zeros((n, ldq))
    nd = np.zeros(n, dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    stats = L.FpeBatchStats()
    rc = L.lib().fpe_batch_terminal_states(arr, n, q.ctypes.data_as(L._DP), ldq,
                                           nd.ctypes.data_as(C.POINTER(C.c_int32)),
                                           st.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(stats))
    if rc != 0:
        raise RuntimeError(f"fpe_batch_terminal_states failed ({rc}): {L.last_error()}")
    return [q[i, :nd[i]].copy() for i in range(n)], st


def apply_operators(x, **columns):
    """K x and M x (unscaled) of one vanilla or out option's domain by the
    sum-factorised Kronecker products on the GPU."""
    arr = _batch_array(columns)
    x = np.ascontiguousarray(x, dtype=np.float64)
    Kx, Mx = np.zeros_like(x), np.zeros_like(x)
    rc = L.lib().fpe_batch_apply_operators(arr, len(x), x.ctypes.data_as(L._DP), Kx.ctypes.data_as(L._DP),
                                           Mx.ctypes.data_as(L._DP))
    if rc != 0:
        raise RuntimeError(f"fpe_batch_apply_operators failed ({rc}): {L.last_error()}")
    return Kx, Mx


def operators(**columns):
    """GPU-assembled M and K (unscaled) of one vanilla or out option as 49-point
    stencils (N x 49 each) and the grid (n_sb, n_vb)."""
    arr = _batch_array(columns)
    cap = L.MAX_DOF
    M = np.zeros(cap * 49)
    K = np.zeros(cap * 49)
    ns, nv = C.c_int32(), C.c_int32()
    rc = L.lib().fpe_batch_operators(arr, cap, M.ctypes.data_as(L._DP), K.ctypes.data_as(L._DP), C.byref(ns),
                                     C.byref(nv))
    if rc != 0:
        raise RuntimeError(f"fpe_batch_operators failed ({rc}): {L.last_error()}")
    N = ns.value * nv.value
    return M[:N * 49].reshape(N, 49), K[:N * 49].reshape(N, 49), (ns.value, nv.value)
