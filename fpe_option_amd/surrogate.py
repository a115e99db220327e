"""POD + MLP surrogate for European prices on a fixed grid, with the plain
Fokker-Planck solve as its fallback and its accuracy reference.

Model: the terminal density q(T) of a European solve on a fixed span-mode grid
(n_s = 80, n_v = 40, s in [0, 180], v in [0, 1.5], S0 = 60, v-knots clustered
at a fixed 0.12) enters prices only through its s-marginal w = G q(T) at the s
quadrature points (price = exp(-rT) sum_i payoff(s_i) w_i). w is approximated
as w0 + Pb c: a POD basis of the marginal snapshots (the pricing inner product
<q1, q2> = (G q1) . (G q2)), rotated into price space -- the singular basis of
the payoff matrix of calls and puts at 61 strikes over the workload's strike
range [45, 75] applied to the POD basis -- so that squared errors of the 64
output coordinates add up to squared price errors over those strikes. An MLP
maps (kappa, theta, sigma, rho, V0, r, T), scaled to [-1, 1] over the training
box, to the normalised coordinates c. Several members (different seeds) form an
ensemble: the served price is their mean and their spread, scaled by a factor
calibrated on held-out solver prices, is the error estimate stated for every
price. A price is certified when its estimate is at or below the tolerance;
the others go to the solver.

Each network is 7 -> 256 -> 256 -> 256 -> 256 -> 64 with SiLU activations and
FP32 accumulation, trained with Adam; its layers run either with BF16 weights
and operands (FP32 master copy) or with FP32 weights and operands (the model
file says which). BF16 rounds the 7 inputs to 2^-8 relative, which alone moves
prices by ~1e-2 (a theta of ~3 per year times 0.4 % of T): a model meant to
certify prices at 1e-4 runs in FP32. Forward, backward, loss and optimiser
steps are this library's kernels (fpe_mlp_forward[_f32] /
fpe_mlp_backward[_f32] / fpe_mlp_mse / fpe_mlp_adam); prices come
from fpe_surrogate_price, which folds the reconstruction into the s-marginal
pricing operator of the grid. Device memory comes from the library's own
allocator (fpe_device_alloc), so no second GPU runtime enters the process.

Training data are solver-generated (fpe_batch_terminal_states); the POD basis
is fitted on the GPU too (fpe_pod_fit: FP64 MFMA projection GEMMs and a
one-sided Jacobi SVD of the marginal snapshots), and so is the price-space
rotation (fpe_svd_jacobi). Inputs outside
the training box, and non-European types, go to the solver. `audit` re-solves
a random fraction with the plain solver.
"""
import ctypes as C
import os
import time

import numpy as np

from . import _lib as L

PARAMS = ("kappa", "theta", "sigma", "rho", "V0", "r", "T")
BOX_LO = np.array([0.5, 0.02, 0.1, -0.9, 0.02, 0.0, 0.1])
BOX_HI = np.array([4.0, 0.2, 0.8, 0.0, 0.3, 0.1, 2.0])
GRID = dict(S0=60.0, disc_mode=1, n_s=80, n_v=40, s_min=0.0, s_max=180.0, v_max=1.5, ic_sigma_s=0.5,
            ic_sigma_v=0.005, rtol=1e-4, atol=1e-6, complex_mode=1, knot_width_s=0.1, knot_width_v=0.1,
            knot_center_v=0.12)
WIDTHS = (7, 256, 256, 256, 256, 64)
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_80x40.npz")


def sample_params(n, seed):
    """Uniform draws over the training box, columns in PARAMS order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return BOX_LO + (BOX_HI - BOX_LO) * rng.uniform(0.0, 1.0, (n, 7))


def in_box(P):
    P = np.asarray(P, dtype=np.float64)
    return np.all((P >= BOX_LO) & (P <= BOX_HI), axis=1)


def _options(P, strike=None, otype=None):
    """FpeBatchOption array on the surrogate's grid, written through a numpy view."""
    P = np.asarray(P, dtype=np.float64)
    n = len(P)
    arr = L.batch_options(n)
    view = np.ctypeslib.as_array(arr) if n else None
    if n:
        for k, v in GRID.items():
            view[k] = v
        for j, name in enumerate(PARAMS):
            view[name] = P[:, j]
        view["strike"] = np.asarray(strike, dtype=np.float64) if strike is not None else 60.0
        view["option_type"] = np.asarray(otype, dtype=np.int32) if otype is not None else 8
        view["barrier"] = 0.0
    return arr


def terminal_states(P, chunk=8192):
    """q(T) rows for parameter rows P on the fixed grid (GPU solves)."""
    lib = L.lib()
    n = len(P)
    ldq = 80 * 40
    out, ndof, stat = None, np.zeros(n, np.int32), np.zeros(n, np.int32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        arr = _options(P[a:b])
        q = np.zeros((b - a, ldq))
        nd = np.zeros(b - a, np.int32)
        st = np.zeros(b - a, np.int32)
        rc = lib.fpe_batch_terminal_states(arr, b - a, q.ctypes.data_as(L._DP), ldq,
                                           nd.ctypes.data_as(C.POINTER(C.c_int32)),
                                           st.ctypes.data_as(C.POINTER(C.c_int32)), None)
        if rc != 0:
            raise RuntimeError(f"fpe_batch_terminal_states failed: {L.last_error()}")
        if out is None:
            out = np.zeros((n, int(nd.max())))
        out[a:b] = q[:, :out.shape[1]]
        ndof[a:b], stat[a:b] = nd, st
    return out, stat


def pricing_operator():
    """s quadrature points and the marginal operator (n_qs x n_dof) of the grid."""
    lib = L.lib()
    arr = _options(np.array([[1.0, 0.05, 0.3, -0.5, 0.1, 0.05, 1.0]]))
    cap_q, cap_d = 4096, 80 * 40
    s = np.zeros(cap_q)
    marg = np.zeros(cap_q * cap_d)
    nq, nd = C.c_int32(), C.c_int32()
    rc = lib.fpe_european_pricing_operator(arr, cap_q, cap_d, s.ctypes.data_as(L._DP), marg.ctypes.data_as(L._DP),
                                           C.byref(nq), C.byref(nd))
    if rc != 0:
        raise RuntimeError(f"fpe_european_pricing_operator failed: {L.last_error()}")
    return s[:nq.value].copy(), marg[:nq.value * nd.value].reshape(nq.value, nd.value).copy()


def solver_prices(P, strike, otype):
    """The plain FP64 solve on the same fixed grid (the surrogate's accuracy reference)."""
    from . import api
    cols = {name: P[:, j] for j, name in enumerate(PARAMS)}
    cols.update(GRID)
    cols.update(strike=np.asarray(strike, float), option_type=np.asarray(otype, np.int32), barrier=0.0)
    prices, status = api.price_batch(**cols)
    return prices, status


# ---------------------------------------------------------------- device MLP
class DeviceArray:
    """Device buffer from the library's allocator (fpe_device_alloc), with a
    numpy shape and dtype; views share the parent's memory."""

    def __init__(self, shape, dtype, _base=None, _ptr=None):
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._base = _base
        if _ptr is not None:
            self.ptr = _ptr
        else:
            self.ptr = L.lib().fpe_device_alloc(max(self.nbytes, 16))
            if not self.ptr:
                raise MemoryError(L.last_error())

    @staticmethod
    def from_numpy(a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = DeviceArray(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if L.lib().fpe_device_upload(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes) != 0:
            raise RuntimeError(L.last_error())

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        L.lib().fpe_device_synchronize()
        if L.lib().fpe_device_download(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes) != 0:
            raise RuntimeError(L.last_error())
        return out

    def rows(self, a, b):
        """View of rows a..b-1 of a 2-D (or 1-D) array."""
        row = self.nbytes // self.shape[0] if self.shape[0] else 0
        return DeviceArray((b - a,) + self.shape[1:], self.dtype, _base=self, _ptr=self.ptr + a * row)

    @property
    def c(self):
        return C.c_void_p(self.ptr)

    def __del__(self):
        if self._base is None and getattr(self, "ptr", None):
            L.lib().fpe_device_free(C.c_void_p(self.ptr))
            self.ptr = None


def _c(x):
    return x.c if x is not None else None


class DeviceMLP:
    """MLP parameters on the GPU: FP32 master weights and their BF16 copies.
    fp32: the layers run with FP32 weights and operands (fpe_mlp_forward_f32 /
    fpe_mlp_backward_f32) instead of BF16 ones."""

    def __init__(self, weights, biases, fp32=False):
        self.lib = L.lib()
        if self.lib.fpe_device_count() < 1:
            raise RuntimeError("the surrogate kernels need a GPU")
        self.fp32 = bool(fp32)
        self.W = [DeviceArray.from_numpy(w, np.float32) for w in weights]
        self.b = [DeviceArray.from_numpy(b, np.float32) for b in biases]
        self.Wb = [DeviceArray(w.shape, np.uint16) for w in self.W]
        self.stream = None  # the null stream: every call is ordered after the previous one
        self._cache = {}
        for w, wb in zip(self.W, self.Wb):
            self._ok(self.lib.fpe_mlp_to_bf16(w.c, wb.c, int(np.prod(w.shape)), self.stream))

    @staticmethod
    def init(seed=0, fp32=False):
        rng = np.random.Generator(np.random.PCG64(seed))
        Ws, bs = [], []
        for a, b in zip(WIDTHS[:-1], WIDTHS[1:]):
            Ws.append(rng.normal(0.0, np.sqrt(1.0 / a), (b, a)))
            bs.append(np.zeros(b))
        return DeviceMLP(Ws, bs, fp32)

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(L.last_error())

    def buf(self, name, shape, dtype=np.float32):
        """Named device buffer reused across calls (device frees synchronise the
        GPU, so steady-state training and inference allocate nothing)."""
        b = self._cache.get(name)
        if b is None or b.shape != tuple(shape) or b.dtype != np.dtype(dtype):
            b = DeviceArray(shape, dtype)
            self._cache[name] = b
        return b

    @property
    def n_layers(self):
        return len(self.W)

    def forward(self, X, keep=False):
        """X: (M, 7) float32 DeviceArray. Returns the output and, with keep, the
        per-layer (input, pre-activation) pairs for backward."""
        M = X.shape[0]
        h, saved = X, []
        for i in range(self.n_layers):
            N, K = self.W[i].shape
            last = i == self.n_layers - 1
            Z = self.buf(f"Z{i}", (M, N)) if (keep and not last) else None
            A = self.buf(f"A{i}", (M, N))
            fwd = self.lib.fpe_mlp_forward_f32 if self.fp32 else self.lib.fpe_mlp_forward
            Wl = self.W[i] if self.fp32 else self.Wb[i]
            self._ok(fwd(h.c, M, K, Wl.c, self.b[i].c, N, 0 if last else 1, _c(Z), A.c, self.stream))
            saved.append((h, Z))
            h = A
        return (h, saved) if keep else h

    def backward(self, dY, saved):
        """Gradients (dW, db) per layer for output gradient dY."""
        M = dY.shape[0]
        grads = [None] * self.n_layers
        g = dY
        for i in reversed(range(self.n_layers)):
            N, K = self.W[i].shape
            X = saved[i][0]
            dW, db = self.buf(f"dW{i}", (N, K)), self.buf(f"db{i}", (N,))
            dX = self.buf(f"dX{i}", (M, K)) if i > 0 else None
            Zp = saved[i - 1][1] if i > 0 else None
            bwd = self.lib.fpe_mlp_backward_f32 if self.fp32 else self.lib.fpe_mlp_backward
            Wl = self.W[i] if self.fp32 else self.Wb[i]
            self._ok(bwd(g.c, X.c, Wl.c, M, N, K, dW.c, db.c, _c(dX), _c(Zp), self.stream))
            grads[i] = (dW, db)
            g = dX
        return grads

    def numpy(self):
        return [w.numpy() for w in self.W], [b.numpy() for b in self.b]


class Adam:
    def __init__(self, net, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.net, self.lr, self.b1, self.b2, self.eps, self.t = net, lr, beta1, beta2, eps, 0
        z = lambda a: DeviceArray.from_numpy(np.zeros(a.shape, np.float32))
        self.m = [(z(w), z(b)) for w, b in zip(net.W, net.b)]
        self.v = [(z(w), z(b)) for w, b in zip(net.W, net.b)]

    def step(self, grads, lr=None):
        self.t += 1
        lr = self.lr if lr is None else lr
        net = self.net
        for i, (dW, db) in enumerate(grads):
            net._ok(net.lib.fpe_mlp_adam(net.W[i].c, net.Wb[i].c, dW.c, self.m[i][0].c, self.v[i][0].c,
                                         int(np.prod(net.W[i].shape)), lr, self.b1, self.b2, self.eps, self.t,
                                         net.stream))
            net._ok(net.lib.fpe_mlp_adam(net.b[i].c, None, db.c, self.m[i][1].c, self.v[i][1].c,
                                         int(np.prod(net.b[i].shape)), lr, self.b1, self.b2, self.eps, self.t,
                                         net.stream))


def normalise(P):
    return (2.0 * (np.asarray(P, np.float64) - BOX_LO) / (BOX_HI - BOX_LO) - 1.0).astype(np.float32)


def mse_step(net, xb, tb, wc, partial):
    """One forward + loss + backward pass; returns the gradients and the loss."""
    out, saved = net.forward(xb, keep=True)
    dY = net.buf("dY", out.shape)
    net._ok(net.lib.fpe_mlp_mse(out.c, tb.c, _c(wc), out.shape[0], out.shape[1], dY.c, partial.c, net.stream))
    grads = net.backward(dY, saved)
    return grads, float(partial.numpy().sum())


def train(P, Y, epochs=60, batch=1024, lr=1e-3, lr_end=1e-5, wcol=None, seed=0, log=print, val=None, fp32=False):
    """Fit the MLP to normalised targets Y (n x k) for inputs P (n x 7).
    Returns the trained DeviceMLP and the per-epoch loss history."""
    net = DeviceMLP.init(seed, fp32)
    opt = Adam(net, lr=lr)
    n, k = Y.shape
    X = DeviceArray.from_numpy(normalise(P))
    T = DeviceArray.from_numpy(np.asarray(Y, np.float32))
    wc = DeviceArray.from_numpy(np.asarray(wcol if wcol is not None else np.ones(k), np.float32))
    partial = DeviceArray((1024,), np.float32)
    idx = DeviceArray((n,), np.int32)
    xb, tb = DeviceArray((batch, 7), np.float32), DeviceArray((batch, k), np.float32)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    steps = epochs * (n // batch)
    hist, step = [], 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        idx.upload(rng.permutation(n).astype(np.int32))
        tot = 0.0
        for a in range(0, n - batch + 1, batch):
            ib = idx.rows(a, a + batch)
            net._ok(net.lib.fpe_mlp_gather_rows(X.c, ib.c, batch, 7, xb.c, net.stream))
            net._ok(net.lib.fpe_mlp_gather_rows(T.c, ib.c, batch, k, tb.c, net.stream))
            grads, loss = mse_step(net, xb, tb, wc, partial)
            frac = step / max(1, steps - 1)
            opt.step(grads, lr=float(lr_end + 0.5 * (lr - lr_end) * (1.0 + np.cos(np.pi * frac))))
            tot += loss
            step += 1
        hist.append(tot / max(1, n // batch))
        if log and (ep % 10 == 0 or ep == epochs - 1):
            msg = f"epoch {ep} loss {hist[-1]:.3e} ({time.perf_counter() - t0:.1f} s)"
            if val is not None:
                msg += f" val {val(net):.3e}"
            log(msg)
    return net, hist


class Surrogate:
    """Trained surrogate (one or more ensemble members) loaded on the GPU."""

    def __init__(self, path=MODEL_PATH):
        z = np.load(path, allow_pickle=False)
        n_layers = int(z["n_layers"])
        self.members = []
        self.fp32 = bool(int(z["fp32"])) if "fp32" in z.files else False
        for j in range(int(z["n_members"])):
            self.members.append(DeviceMLP([z[f"m{j}_W{i}"] for i in range(n_layers)],
                                          [z[f"m{j}_b{i}"] for i in range(n_layers)], self.fp32))
        self.net = self.members[0]
        self.k = int(z["cmean"].shape[0])
        self.cscale = DeviceArray.from_numpy(z["cstd"], np.float32)
        self.cshift = DeviceArray.from_numpy(z["cmean"], np.float32)
        self.Pt = DeviceArray.from_numpy(np.ascontiguousarray(z["Pb"].T), np.float32)
        self.w0 = DeviceArray.from_numpy(z["w0"], np.float32)
        self.sq = DeviceArray.from_numpy(z["s_q"], np.float32)
        self.nq = int(z["s_q"].shape[0])
        self.gamma = float(z["cert_gamma"])  # error estimate = gamma x ensemble spread
        self.info = {k: z[k] for k in z.files if k.startswith("meta_") or k in ("loss", "pod_sv")}

    def coefficients(self, P, member=0):
        """Normalised output coordinates of one member for parameter rows P (DeviceArray, n x k)."""
        net = self.members[member]
        X = net.buf("X", (len(P), 7))
        X.upload(normalise(P))
        return net.forward(X)

    def _member_prices(self, P, strike, otype, rows, member):
        m = len(rows)
        Cn = self.coefficients(P[rows], member)
        buf = self.members[member].buf
        K, ty = buf("K", (m,), np.float64), buf("type", (m,), np.int32)
        r, T, pr = buf("r", (m,), np.float64), buf("T", (m,), np.float64), buf("price", (m,), np.float64)
        K.upload(np.asarray(strike, np.float64)[rows])
        ty.upload(np.asarray(otype, np.int32)[rows])
        r.upload(P[rows, 5])
        T.upload(P[rows, 6])
        net = self.members[member]
        net._ok(net.lib.fpe_surrogate_price(Cn.c, m, self.k, self.cscale.c, self.cshift.c, self.Pt.c, self.w0.c,
                                            self.sq.c, self.nq, K.c, ty.c, r.c, T.c, pr.c, net.stream))
        return pr.numpy()

    def price(self, P, strike, otype, chunk=131072):
        """Ensemble-mean prices of European calls (8) / puts (9) for in-box rows
        and their error estimates; NaN elsewhere. Returns (prices, error
        estimates, mask of rows the surrogate priced)."""
        P = np.asarray(P, np.float64)
        n = len(P)
        otype = np.asarray(otype, np.int32)
        ok = in_box(P) & np.isin(otype, (8, 9))
        out = np.full(n, np.nan)
        est = np.full(n, np.nan)
        idx = np.nonzero(ok)[0]
        for a in range(0, len(idx), chunk):
            rows = idx[a:a + chunk]
            ps = np.stack([self._member_prices(P, strike, otype, rows, j) for j in range(len(self.members))])
            out[rows] = ps.mean(axis=0)
            est[rows] = self.gamma * (ps.max(axis=0) - ps.min(axis=0))
        return out, est, ok

    def price_certified(self, P, strike, otype, tol=1e-4):
        """Surrogate prices where the error estimate is at or below tol; the
        other rows (estimate above tol, outside the box, other types) are left
        NaN for the solver. Returns (prices, estimates, certified mask)."""
        out, est, ok = self.price(P, strike, otype)
        cert = ok & (est <= tol)
        out[~cert] = np.nan
        return out, est, cert

    def price_with_fallback(self, P, strike, otype, tol=None):
        """Surrogate where it applies (and, with tol, where its estimate is at or
        below tol), the plain solve elsewhere. Returns prices, error estimates
        (0 for solver prices), the surrogate mask and the number of solves."""
        if tol is None:
            out, est, ok = self.price(P, strike, otype)
        else:
            out, est, ok = self.price_certified(P, strike, otype, tol)
        rest = np.nonzero(~ok)[0]
        if len(rest):
            out[rest], _ = solver_prices(np.asarray(P)[rest], np.asarray(strike)[rest], np.asarray(otype)[rest])
            est[rest] = 0.0
        return out, est, ok, len(rest)


def audit(sur, P, strike, otype, frac=0.01, seed=1):
    """Re-solves a random fraction of the surrogate-priced rows with the plain
    solver. Returns (rows, surrogate prices, estimates, solver prices, status)."""
    sp, est, ok = sur.price(P, strike, otype)
    rows = np.nonzero(ok)[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    pick = np.sort(rng.choice(rows, size=max(1, int(round(frac * len(rows)))), replace=False))
    ref, st = solver_prices(np.asarray(P)[pick], np.asarray(strike)[pick], np.asarray(otype)[pick])
    return pick, sp[pick], est[pick], ref, st


STRIKES = np.linspace(45.0, 75.0, 61)  # the workload's strike range, for the price-space basis


def payoff_matrix(s_q, strikes=STRIKES):
    """Undiscounted call and put payoffs at the s quadrature points (2 n_K x n_qs)."""
    s = np.asarray(s_q, np.float64)[None]
    K = np.asarray(strikes, np.float64)[:, None]
    return np.concatenate([np.maximum(s - K, 0.0), np.maximum(K - s, 0.0)])


def price_basis(P, s_q, rel_floor=1e-6, svd=None):
    """Rotation of the POD coordinates into price space: with A P = U S V^T (A
    the payoff matrix), coordinates c'' = S' V^T c (S' = S with a floor of
    rel_floor S[0]) have |A P (c1 - c2)| = |c1'' - c2''| where S' = S. Returns
    the basis Pb = P V S'^-1 for w = w0 + Pb c'' and the map R = S' V^T.
    svd(M) -> (singular values, V^T); the GPU's one-sided Jacobi by default."""
    sv, Vt = (svd or svd_jacobi)(payoff_matrix(s_q) @ P)
    sp = np.maximum(sv, rel_floor * sv[0])
    R = sp[:, None] * Vt
    return P @ (Vt.T / sp), R


def fit_pod(Q, marg, k=64):
    """POD of q(T) snapshot rows Q in the pricing inner product, on the GPU
    (fpe_pod_fit: projection GEMMs and one-sided Jacobi SVD on FP64 MFMA).
    Returns mu (n_dof), Phi (n_dof x k), P = marg Phi (n_qs x k, orthonormal
    columns), w0 = marg mu and the singular values."""
    Q = np.ascontiguousarray(Q, np.float64)
    marg = np.ascontiguousarray(marg, np.float64)
    ns, nd = Q.shape
    nq = marg.shape[0]
    k = min(k, nq)
    dQ, dG = DeviceArray.from_numpy(Q), DeviceArray.from_numpy(marg)
    mu, Phi, P = DeviceArray((nd,), np.float64), DeviceArray((nd, k), np.float64), DeviceArray((nq, k), np.float64)
    w0, sv = DeviceArray((nq,), np.float64), DeviceArray((nq,), np.float64)
    sweeps = C.c_int32()
    rc = L.lib().fpe_pod_fit(dQ.c, ns, nd, dG.c, nq, k, mu.c, Phi.c, P.c, w0.c, sv.c, C.byref(sweeps), None)
    if rc != 0:
        raise RuntimeError(f"fpe_pod_fit failed ({rc}): {L.last_error()}")
    return mu.numpy(), Phi.numpy(), P.numpy(), w0.numpy(), sv.numpy()


def svd_jacobi(A):
    """Singular values (descending) and V^T of a dense FP64 matrix on the GPU
    (fpe_svd_jacobi, one-sided Jacobi). Same conventions as
    numpy.linalg.svd(A, full_matrices=False)[1:] for A with rows >= columns;
    each right singular vector's largest-magnitude entry is positive."""
    A = np.asarray(A, np.float64)
    m, n = A.shape
    dA = DeviceArray.from_numpy(np.asfortranarray(A).ravel(order="F"))
    sv, V = DeviceArray((n,), np.float64), DeviceArray((n, n), np.float64)
    sweeps = C.c_int32()
    rc = L.lib().fpe_svd_jacobi(dA.c, m, n, m, sv.c, V.c, C.byref(sweeps), None)
    if rc != 0:
        raise RuntimeError(f"fpe_svd_jacobi failed ({rc}): {L.last_error()}")
    r = min(m, n)
    return sv.numpy()[:r], V.numpy()[:, :r].T.copy()


def generate(n, seed, pod=None, n_pod=8192, k=64, chunk=8192, log=print):
    """Solver-generated training data on the fixed grid: parameter rows and
    their POD coefficients. Without `pod`, the basis is fitted on the first
    n_pod snapshots and returned. Returns (P, coefficients, ok mask, pod dict)."""
    t0 = time.perf_counter()
    P = sample_params(n, seed)
    s_q, marg = pricing_operator()
    C_out = np.zeros((n, k), np.float32)
    ok = np.zeros(n, bool)
    bounds = [0] + list(range(min(n_pod if pod is None else chunk, n), n, chunk)) + [n]
    for a, b in zip(bounds[:-1], bounds[1:]):
        if b <= a:
            continue
        Q, st = terminal_states(P[a:b])
        ok[a:b] = st == 0
        if pod is None:
            mu, Phi, Pm, w0, sv = fit_pod(Q[st == 0], marg, k)
            pod = dict(mu=mu, Phi=Phi.astype(np.float32), P=Pm, w0=w0, s_q=s_q, pod_sv=sv, pod_snapshots=np.array(
                int(np.sum(st == 0))))
            log(f"POD from {int(np.sum(st == 0))} snapshots: sv[{k - 1}]/sv[0] = {sv[k - 1] / sv[0]:.3e}")
        C_out[a:b] = ((Q - pod["mu"]) @ marg.T) @ pod["P"]
        log(f"solved {b}/{n} ({time.perf_counter() - t0:.1f} s)")
    return P, C_out, ok, pod


def load_data(data_dir, seeds):
    """Parameter rows (regenerated from their seeds) and POD coefficients of
    generated data files C_<seed>.npy / ok_<seed>.npy."""
    Ps, Cs = [], []
    for sd in seeds:
        C = np.load(os.path.join(data_dir, f"C_{sd}.npy"))
        ok = np.load(os.path.join(data_dir, f"ok_{sd}.npy"))
        P = sample_params(len(C), sd)
        Ps.append(P[ok])
        Cs.append(C[ok].astype(np.float64))
    return np.concatenate(Ps), np.concatenate(Cs)


def load_pod(data_dir):
    return {k: np.load(os.path.join(data_dir, f"pod_{k}.npy")) for k in ("mu", "P", "w0", "s_q", "pod_sv")}


def _val_prices(pod, Pb, cmean, cstd, Y, Pv, Kv, tv):
    W = (Y * cstd + cmean) @ Pb.T + pod["w0"]
    pay = np.maximum(np.where(tv[:, None] == 8, pod["s_q"][None] - Kv[:, None], Kv[:, None] - pod["s_q"][None]), 0)
    return np.exp(-Pv[:, 5] * Pv[:, 6]) * np.sum(pay * W, axis=1)


def step_seconds(k, batch=1024, fp32=True, steps=20):
    """Time of one training step (forward, loss, backward, Adam) of a fresh network."""
    net = DeviceMLP.init(0, fp32)
    opt = Adam(net)
    rng = np.random.default_rng(0)
    xb = DeviceArray.from_numpy(rng.uniform(-1, 1, (batch, 7)).astype(np.float32))
    tb = DeviceArray.from_numpy(rng.standard_normal((batch, k)).astype(np.float32))
    partial = DeviceArray((1024,), np.float32)
    grads, _ = mse_step(net, xb, tb, None, partial)
    opt.step(grads)
    net.lib.fpe_device_synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        grads, _ = mse_step(net, xb, tb, None, partial)
        opt.step(grads)
    net.lib.fpe_device_synchronize()
    return (time.perf_counter() - t) / steps


def build_model(data_dir, seeds, out=MODEL_PATH, members=2, epochs=200, batch=1024, lr=2e-3, n_val=8192,
                log=print, seed=0, safety=2.0, fp32=True, budget_s=None):
    """Trains `members` networks on generated data (price-space coordinates),
    calibrates the error estimate on n_val held-out samples and saves the model.
    gamma = safety x the largest ratio |error| / spread over the held-out
    samples whose spread is not negligible (1e-9 of the median price)."""
    pod = load_pod(data_dir)
    P, Cpod = load_data(data_dir, seeds)
    Pb, R = price_basis(pod["P"], pod["s_q"])
    Cf = Cpod @ R.T
    cmean, cstd = Cf.mean(axis=0), Cf.std(axis=0) + 1e-30
    Y = (Cf - cmean) / cstd
    wcol = cstd ** 2 / np.mean(cstd ** 2)
    Pv, Yv = P[:n_val], Y[:n_val]
    Ptr, Ytr = P[n_val:], Y[n_val:]
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    Kv = rng.uniform(45.0, 75.0, n_val)
    tv = np.where(rng.uniform(0.0, 1.0, n_val) < 0.5, 8, 9)
    pv = _val_prices(pod, Pb, cmean, cstd, Yv, Pv, Kv, tv)
    Xv = DeviceArray.from_numpy(normalise(Pv))

    def val(net):
        e = np.abs(_val_prices(pod, Pb, cmean, cstd, net.forward(Xv).numpy().astype(np.float64), Pv, Kv, tv) - pv)
        return float(np.max(e))

    if budget_s:  # as many epochs as the time budget holds (all members), at most `epochs`
        ts = step_seconds(Y.shape[1], batch, fp32)
        fit = int(budget_s / (members * ts * (len(Ptr) // batch)))
        log(f"{ts * 1e3:.2f} ms per step: {fit} epochs per member fit in {budget_s} s")
        epochs = max(1, min(epochs, fit))
    log(f"training {members} member(s) on {len(Ptr)} samples, validating on {n_val}, {epochs} epochs, "
        f"{'FP32' if fp32 else 'BF16'} weights")
    arrays, hists, vprices = {}, [], []
    for j in range(members):
        net, hist = train(Ptr, Ytr, epochs=epochs, batch=batch, lr=lr, wcol=wcol, seed=seed + 7919 * j, log=log,
                          val=val, fp32=fp32)
        Ws, bs = net.numpy()
        arrays.update({f"m{j}_W{i}": w for i, w in enumerate(Ws)})
        arrays.update({f"m{j}_b{i}": b for i, b in enumerate(bs)})
        hists.append(hist)
        vprices.append(_val_prices(pod, Pb, cmean, cstd, net.forward(Xv).numpy().astype(np.float64), Pv, Kv, tv))
    vp = np.stack(vprices)
    mean, spread = vp.mean(axis=0), vp.max(axis=0) - vp.min(axis=0)
    err = np.abs(mean - pv)
    floor = 1e-9 * np.median(np.abs(pv))
    ratio = err / np.maximum(spread, floor)
    gamma = safety * float(np.max(ratio)) if members > 1 else float("inf")
    est = gamma * spread
    cert = est <= 1e-4
    log(f"held-out |dprice|: max {err.max():.3e} p99 {np.percentile(err, 99):.3e} mean {err.mean():.3e}; "
        f"gamma {gamma:.3g}; certified at 1e-4: {cert.mean():.3%} (max error there "
        f"{(err[cert].max() if cert.any() else 0.0):.2e})")
    arrays.update(n_layers=np.array(len(WIDTHS) - 1), n_members=np.array(members), fp32=np.array(int(fp32)),
                  cmean=cmean, cstd=cstd,
                  box_lo=BOX_LO, box_hi=BOX_HI, loss=np.array(hists), meta_samples=np.array(len(Ptr)),
                  meta_val=np.array(n_val), meta_val_max_abs_dprice=np.array(float(err.max())),
                  meta_val_p99_abs_dprice=np.array(float(np.percentile(err, 99))),
                  meta_val_mean_abs_dprice=np.array(float(err.mean())),
                  meta_val_certified_frac_1e4=np.array(float(cert.mean())), cert_gamma=np.array(gamma),
                  Pb=Pb, w0=pod["w0"], s_q=pod["s_q"], pod_sv=pod["pod_sv"], meta_seeds=np.array(seeds))
    np.savez(out, **arrays)
    log(f"saved {out}")
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="surrogate data generation and training (needs a GPU)")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate", help="solve n parameter draws and store their POD coefficients")
    g.add_argument("--samples", type=int, default=100000)
    g.add_argument("--seed", type=int, default=20261204)
    g.add_argument("--pod", default=None, help="directory of an existing POD; otherwise fit one into --out")
    g.add_argument("--out", required=True, help="directory for C_<seed>.npy, ok_<seed>.npy (and pod_*.npy)")
    t = sub.add_parser("train", help="train the ensemble on generated data and write the model")
    t.add_argument("--data", required=True)
    t.add_argument("--seeds", type=int, nargs="+", required=True)
    t.add_argument("--members", type=int, default=2)
    t.add_argument("--epochs", type=int, default=200)
    t.add_argument("--batch", type=int, default=1024)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--n-val", type=int, default=8192)
    t.add_argument("--bf16", action="store_true", help="BF16 weights and operands instead of FP32")
    t.add_argument("--budget", type=float, default=None, help="seconds of training for all members")
    t.add_argument("--out", default=MODEL_PATH)
    a = ap.parse_args()
    say = lambda m: print(m, flush=True)
    if a.cmd == "generate":
        os.makedirs(a.out, exist_ok=True)
        pod = load_pod(a.pod) if a.pod else None
        P, Cf, ok, pod2 = generate(a.samples, a.seed, pod=pod, log=say)
        np.save(os.path.join(a.out, f"C_{a.seed}.npy"), Cf.astype(np.float32))
        np.save(os.path.join(a.out, f"ok_{a.seed}.npy"), ok)
        if pod is None:
            for k in ("mu", "P", "w0", "s_q", "pod_sv"):
                np.save(os.path.join(a.out, f"pod_{k}.npy"), pod2[k])
    else:
        build_model(a.data, a.seeds, out=a.out, members=a.members, epochs=a.epochs, batch=a.batch, lr=a.lr,
                    n_val=a.n_val, log=say, fp32=not a.bf16, budget_s=a.budget)
