"""ctypes binding of the FP64 CPU oracle (oracle/fpe_oracle.h).

Test infrastructure only: results are compared against the oracle; it is never
the thing measured.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "libfpe_oracle.so")

DISC_REFERENCE, DISC_SPAN = 0, 1
CPLX_GMRES, CPLX_DIRECT = 0, 1


class OrcSpec(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("kappa", "theta", "sigma", "rho", "r", "T", "S0", "V0",
                                           "s_min", "s_max", "v_max")] + [
        ("n_s", C.c_int32), ("n_v", C.c_int32), ("option_type", C.c_int32),
        ("barrier", C.c_double), ("disc_mode", C.c_int32), ("num_insert", C.c_int32),
        ("ic_sigma_s", C.c_double), ("ic_sigma_v", C.c_double),
        ("rtol", C.c_double), ("atol", C.c_double),
        ("complex_mode", C.c_int32), ("exact_constants", C.c_int32),
        ("knot_width_s", C.c_double), ("knot_width_v", C.c_double)]


class OrcStats(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "n_accepted", "n_rejected", "n_steps_total", "n_lu", "n_newton", "n_err_solves",
        "n_gmres_iters", "admm_iters", "success", "n_dof", "n_sb", "n_vb", "n_qs", "n_qv")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.orc_spec_default.argtypes = [C.POINTER(OrcSpec)]
        L.orc_price.argtypes = [C.POINTER(OrcSpec), dp, C.c_int, dp, C.POINTER(OrcStats)]
        L.orc_price_greeks.argtypes = [C.POINTER(OrcSpec), dp, C.c_int, C.c_double, C.c_double,
                                       dp, dp, dp, dp]
        L.orc_pipe_create.argtypes = [C.POINTER(OrcSpec)]
        L.orc_pipe_create.restype = C.c_void_p
        L.orc_pipe_destroy.argtypes = [C.c_void_p]
        L.orc_pipe_dims.argtypes = [C.c_void_p, ip]
        L.orc_pipe_knots.argtypes = [C.c_void_p, dp, dp]
        L.orc_pipe_grid.argtypes = [C.c_void_p, dp, dp, dp, dp, dp]
        L.orc_pipe_basis.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, ip, dp, ip]
        L.orc_pipe_factor.argtypes = [C.c_void_p, C.c_char_p, dp]
        L.orc_pipe_stencils.argtypes = [C.c_void_p, dp, dp]
        L.orc_pipe_initial_condition.argtypes = [C.c_void_p, dp]
        L.orc_pipe_solve.argtypes = [C.c_void_p, C.POINTER(OrcStats)]
        L.orc_pipe_state.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.orc_pipe_pdf.argtypes = [C.c_void_p, dp]
        L.orc_pipe_price_at.argtypes = [C.c_void_p, dp, C.c_int, dp]
        L.orc_gauss_legendre.argtypes = [C.c_int, dp, dp]
        L.orc_generate_knots.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5 + [dp]
        L.orc_de_boor.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_de_boor.restype = C.c_double
        L.orc_radau_dense.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_int, C.c_int, dp,
                                      C.POINTER(OrcStats)]
        L.orc_admm_nnls_dense.argtypes = [C.c_int, dp, dp, dp]
        L.orc_payoff.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_payoff.restype = C.c_double
        L.orc_heston_call.argtypes = [C.c_double] * 9
        L.orc_heston_call.restype = C.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def spec(**kw):
    s = OrcSpec()
    lib().orc_spec_default(C.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


def price(K, **kw):
    s = spec(**kw)
    K = np.ascontiguousarray(np.atleast_1d(np.asarray(K, dtype=np.float64)))
    out = np.zeros(len(K))
    st = OrcStats()
    rc = lib().orc_price(C.byref(s), _dp(K), len(K), _dp(out), C.byref(st))
    return out, st.as_dict(), rc


def price_greeks(K, rel_s=0.01, rel_v=0.1, **kw):
    s = spec(**kw)
    K = np.ascontiguousarray(np.atleast_1d(np.asarray(K, dtype=np.float64)))
    outs = [np.zeros(len(K)) for _ in range(4)]
    rc = lib().orc_price_greeks(C.byref(s), _dp(K), len(K), rel_s, rel_v, *[_dp(o) for o in outs])
    return outs, rc


class Pipe:
    """Stepwise oracle pipeline for one (non-in) solve on the truncated domain."""

    def __init__(self, **kw):
        self.spec = spec(**kw)
        self.h = lib().orc_pipe_create(C.byref(self.spec))
        if not self.h:
            raise ValueError("invalid oracle spec")
        d = np.zeros(8, dtype=np.int32)
        lib().orc_pipe_dims(self.h, _ip(d))
        self.n_sk, self.n_vk, self.n_qs, self.n_qv, self.n_sb, self.n_vb, self.N = map(int, d[:7])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pipe_destroy(self.h)
            self.h = None

    def knots(self):
        s, v = np.zeros(self.n_sk), np.zeros(self.n_vk)
        lib().orc_pipe_knots(self.h, _dp(s), _dp(v))
        return s, v

    def grid(self):
        a = [np.zeros(self.n_qs), np.zeros(self.n_qv), np.zeros(self.n_qs), np.zeros(self.n_qv),
             np.zeros(self.n_qs)]
        lib().orc_pipe_grid(self.h, *[_dp(x) for x in a])
        return dict(s=a[0], v=a[1], sw=a[2], vw=a[3], s_norm=a[4])

    def basis(self, axis, deriv):
        import scipy.sparse as sp
        nnz = C.c_int32()
        lib().orc_pipe_basis(self.h, axis, deriv, None, None, None, C.byref(nnz))
        nr = self.n_qs if axis == 0 else self.n_qv
        nc = self.n_sb if axis == 0 else self.n_vb
        ptr = np.zeros(nr + 1, dtype=np.int32)
        idx = np.zeros(nnz.value, dtype=np.int32)
        val = np.zeros(nnz.value)
        lib().orc_pipe_basis(self.h, axis, deriv, _ip(ptr), _ip(idx), _dp(val), C.byref(nnz))
        return sp.csr_matrix((val, idx, ptr), shape=(nr, nc))

    def factor(self, name):
        n = self.n_sb if name[0] in "NS" and name != "Nv" else self.n_vb
        out = np.zeros(n * n)
        lib().orc_pipe_factor(self.h, name.encode(), _dp(out))
        return out.reshape(n, n)

    def stencils(self):
        M, K = np.zeros(self.N * 49), np.zeros(self.N * 49)
        lib().orc_pipe_stencils(self.h, _dp(M), _dp(K))
        return M.reshape(self.N, 49), K.reshape(self.N, 49)

    def initial_condition(self):
        q = np.zeros(self.N)
        it = lib().orc_pipe_initial_condition(self.h, _dp(q))
        return q, it

    def solve(self):
        st = OrcStats()
        n = lib().orc_pipe_solve(self.h, C.byref(st))
        self.n_states = n
        return st.as_dict()

    def state(self, k):
        q = np.zeros(self.N)
        t = C.c_double()
        lib().orc_pipe_state(self.h, k, _dp(q), C.byref(t))
        return q, t.value

    def pdf(self):
        out = np.zeros(self.n_qs * self.n_qv)
        lib().orc_pipe_pdf(self.h, _dp(out))
        return out.reshape(self.n_qs, self.n_qv)

    def price_at(self, K):
        K = np.ascontiguousarray(np.atleast_1d(np.asarray(K, dtype=np.float64)))
        out = np.zeros(len(K))
        lib().orc_pipe_price_at(self.h, _dp(K), len(K), _dp(out))
        return out


def dense_to_stencil_matrix(S49, n_sb, n_vb):
    """Expand a 49-point stencil (rows (I,i)) into a dense N x N matrix."""
    N = n_sb * n_vb
    A = np.zeros((N, N))
    for I in range(n_sb):
        for i in range(n_vb):
            r = I * n_vb + i
            for dJ in range(-3, 4):
                J = I + dJ
                if J < 0 or J >= n_sb:
                    continue
                for dj in range(-3, 4):
                    j = i + dj
                    if j < 0 or j >= n_vb:
                        continue
                    A[r, J * n_vb + j] = S49[r, (dJ + 3) * 7 + dj + 3]
    return A
