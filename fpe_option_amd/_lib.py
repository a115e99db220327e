"""ctypes binding of libfpe_engine.so (include/fpe_engine.h).

The library is built in-tree by fpe_option_amd.build; loading fails loudly if
it is missing -- there is no CPU fallback on the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfpe_engine.so")

DISC_REFERENCE, DISC_SPAN = 0, 1
CPLX_GMRES, CPLX_DIRECT = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "MAX_STEPS", 3: "H_UNDERFLOW", 4: "NONFINITE", 5: "TOO_LARGE",
                6: "SINGULAR", 7: "REFINE_FAIL"}


class FpeVecResult(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_double)), ("len", C.c_int32)]


class FpeVec2Result(C.Structure):
    _fields_ = [("s_data", C.POINTER(C.c_double)), ("s_len", C.c_int32),
                ("v_data", C.POINTER(C.c_double)), ("v_len", C.c_int32)]


class FpeGridPtsResult(C.Structure):
    _fields_ = [("s_data", C.POINTER(C.c_double)), ("s_len", C.c_int32),
                ("v_data", C.POINTER(C.c_double)), ("v_len", C.c_int32),
                ("sw_data", C.POINTER(C.c_double)), ("vw_data", C.POINTER(C.c_double))]


class FpeMatResult(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_double)), ("n_rows", C.c_int32), ("n_cols", C.c_int32)]


class FpeGreeksResult(C.Structure):
    _fields_ = [("delta", C.POINTER(C.c_double)), ("gamma", C.POINTER(C.c_double)),
                ("vega", C.POINTER(C.c_double)), ("len", C.c_int32)]


class FpeOneshotResult(C.Structure):
    _fields_ = [("price", C.POINTER(C.c_double)), ("delta", C.POINTER(C.c_double)),
                ("gamma", C.POINTER(C.c_double)), ("vega", C.POINTER(C.c_double)), ("len", C.c_int32)]


class FpeCsrResult(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_double)), ("indices", C.POINTER(C.c_int32)),
                ("indptr", C.POINTER(C.c_int32)), ("n_rows", C.c_int32), ("n_cols", C.c_int32),
                ("nnz", C.c_int32)]


class FpeBatchOption(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("kappa", "theta", "sigma", "rho", "r", "T", "S0", "V0",
                                           "s_min", "s_max", "v_max", "barrier", "strike")] + [
        (n, C.c_int32) for n in ("option_type", "n_s", "n_v", "disc_mode", "num_insert",
                                 "complex_mode", "precision", "exact_constants")] + [
        (n, C.c_double) for n in ("ic_sigma_s", "ic_sigma_v", "rtol", "atol", "knot_width_s", "knot_width_v",
                                  "knot_center_v")] + [
        (n, C.c_int32) for n in ("lu_variant", "max_steps", "refine_max", "ic_variant")]


class FpeBatchStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("n_solves", "n_accepted", "n_rejected", "n_lu", "n_newton",
                                          "n_err_solves", "n_inner_iters", "n_admm_iters")] + [
        (n, C.c_double) for n in ("ms_setup", "ms_assemble", "ms_ic", "ms_radau", "ms_payoff", "ms_total",
                                  "alg_flops", "alg_bytes", "wg_s_assemble", "wg_s_lu", "wg_s_solve",
                                  "wg_s_spmv", "wg_s_other", "wg_s_total", "wg_s_lu_inverse", "wg_s_lu_stage",
                                  "wg_s_lu_lmul", "wg_s_lu_schur")] + [("n_refine", C.c_int64), ("n_radau_slots", C.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class FpeSolveInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("status", "n_solves", "n_accepted", "n_rejected", "n_lu", "n_newton",
                                          "n_err_solves", "n_inner_iters", "n_admm_iters", "n_refine", "n_sb",
                                          "n_vb")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# Every symbol the header declares, with its ctypes signature.
_D, _I, _P = C.c_double, C.c_int32, C.c_void_p
_DP = C.POINTER(C.c_double)
SIGNATURES = {
    "fpe_compute_create": ([_D] * 8 + [_I, _I, _D, _I, _I, _D, _D], _P),
    "fpe_compute_destroy": ([_P], None),
    "fpe_compute_knots": ([_P, C.POINTER(FpeVec2Result)], None),
    "fpe_compute_grid_points": ([_P, C.POINTER(FpeGridPtsResult)], None),
    "fpe_compute_initial_condition": ([_P, C.POINTER(FpeVecResult)], None),
    "fpe_compute_solve": ([_P, C.POINTER(FpeMatResult)], None),
    "fpe_compute_pdf": ([_P, C.POINTER(FpeMatResult)], None),
    "fpe_compute_price": ([_P, _DP, _I, C.POINTER(FpeVecResult)], None),
    "fpe_compute_greeks": ([_P, _DP, _I, _D, _D, C.POINTER(FpeGreeksResult)], None),
    "fpe_price_oneshot": ([_D] * 8 + [_DP, _I, _D, _I, _I, _I, _I, _D, _D, C.POINTER(FpeOneshotResult)], None),
    "fpe_compute_free_vec": ([C.POINTER(FpeVecResult)], None),
    "fpe_compute_free_vec2": ([C.POINTER(FpeVec2Result)], None),
    "fpe_compute_free_grid_pts": ([C.POINTER(FpeGridPtsResult)], None),
    "fpe_compute_free_mat": ([C.POINTER(FpeMatResult)], None),
    "fpe_compute_free_greeks": ([C.POINTER(FpeGreeksResult)], None),
    "fpe_compute_free_oneshot": ([C.POINTER(FpeOneshotResult)], None),
    "fpe_compute_basis_1d": ([_P, _I, _I, C.POINTER(FpeCsrResult)], None),
    "fpe_compute_free_csr": ([C.POINTER(FpeCsrResult)], None),
    "fpe_batch_option_default": ([C.POINTER(FpeBatchOption)], None),
    "fpe_batch_price": ([C.POINTER(FpeBatchOption), _I, _DP, C.POINTER(C.c_int32), C.POINTER(FpeBatchStats)], _I),
    "fpe_batch_price_detail": ([C.POINTER(FpeBatchOption), _I, _DP, C.POINTER(C.c_int32), C.POINTER(FpeSolveInfo),
                                _DP, C.c_int64, C.POINTER(FpeBatchStats)], _I),
    "fpe_batch_price_noalloc": ([C.POINTER(FpeBatchOption), _I, _DP, C.POINTER(C.c_int32),
                                 C.POINTER(FpeBatchStats)], _I),
    "fpe_batch_apply_operators": ([C.POINTER(FpeBatchOption), _I, _DP, _DP, _DP], _I),
    "fpe_batch_operators": ([C.POINTER(FpeBatchOption), _I, _DP, _DP, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
                            _I),
    "fpe_batch_terminal_states": ([C.POINTER(FpeBatchOption), _I, _DP, C.c_int64, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(FpeBatchStats)], _I),
    "fpe_european_pricing_operator": ([C.POINTER(FpeBatchOption), _I, _I, _DP, _DP, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32)], _I),
    "fpe_mlp_forward": ([_P, _I, _I, _P, _P, _I, _I, _P, _P, _P], _I),
    "fpe_mlp_backward": ([_P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P], _I),
    "fpe_mlp_forward_f32": ([_P, _I, _I, _P, _P, _I, _I, _P, _P, _P], _I),
    "fpe_mlp_backward_f32": ([_P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P], _I),
    "fpe_mlp_adam": ([_P, _P, _P, _P, _P, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, _I, _P], _I),
    "fpe_mlp_mse": ([_P, _P, _P, _I, _I, _P, _P, _P], _I),
    "fpe_mlp_to_bf16": ([_P, _P, C.c_int64, _P], _I),
    "fpe_mlp_gather_rows": ([_P, _P, _I, _I, _P, _P], _I),
    "fpe_device_alloc": ([C.c_int64], _P),
    "fpe_device_free": ([_P], None),
    "fpe_device_upload": ([_P, _P, C.c_int64], _I),
    "fpe_device_download": ([_P, _P, C.c_int64], _I),
    "fpe_device_synchronize": ([], _I),
    "fpe_surrogate_price": ([_P, _I, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P], _I),
    "fpe_pod_fit": ([_P, _I, _I, _P, _I, _I, _P, _P, _P, _P, _P, C.POINTER(C.c_int32), _P], _I),
    "fpe_svd_jacobi": ([_P, _I, _I, _I, _P, _P, C.POINTER(C.c_int32), _P], _I),
    "fpe_last_error": ([], C.c_char_p),
    "fpe_device_count": ([], _I),
}

_lib = None


def lib():
    """Load libfpe_engine.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not found: build it with `python -m fpe_option_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def last_error():
    return lib().fpe_last_error().decode()


def vec(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 and ptr else np.zeros(0)


def batch_options(n):
    arr = (FpeBatchOption * n)()
    d = (FpeBatchOption * 1)()
    lib().fpe_batch_option_default(C.byref(d[0]))
    if n > 0:
        np.ctypeslib.as_array(arr)[:] = np.ctypeslib.as_array(d)[0]
    return arr


MAX_DOF = 300 * 64  # largest grid the kernels take (n_sb <= 300, n_vb <= 64)


def price_batch_raw(arr, want_info=False, want_qT=False):
    """fpe_batch_price_detail over an FpeBatchOption array: prices, status,
    batch statistics, then the per-option records (want_info) and the list of
    terminal states (want_qT; None for in-options) when asked."""
    n = len(arr)
    prices = np.zeros(n)
    status = np.zeros(n, dtype=np.int32)
    st = FpeBatchStats()
    info = (FpeSolveInfo * max(n, 1))()
    q = np.zeros((n, MAX_DOF)) if want_qT else None
    rc = lib().fpe_batch_price_detail(arr, n, prices.ctypes.data_as(_DP),
                                      status.ctypes.data_as(C.POINTER(C.c_int32)), info,
                                      q.ctypes.data_as(_DP) if want_qT else None, MAX_DOF if want_qT else 0,
                                      C.byref(st))
    if rc != 0:
        raise RuntimeError(f"fpe_batch_price failed ({rc}): {last_error()}")
    out = (prices, status, st.as_dict())
    if not (want_info or want_qT):
        return out
    names = [f for f, _ in FpeSolveInfo._fields_]
    recs = [dict(zip(names, row)) for row in np.ctypeslib.as_array(info)[:n].tolist()]
    if want_info:
        out += (recs,)
    if want_qT:
        out += ([q[i, :r["n_sb"] * r["n_vb"]].copy() if r["n_solves"] == 1 and r["status"] in (0, 7) else None
                 for i, r in enumerate(recs)],)
    return out
