"""POD + MLP surrogate for European prices on a fixed grid, with the plain
Fokker-Planck solve as its fallback and its accuracy reference.

Model: the terminal density q(T) of a European solve on a fixed span-mode grid
(n_s = 80, n_v = 40, s in [0, 180], v in [0, 1.5], S0 = 60, v-knots clustered
at a fixed 0.12) is approximated as q = mu + Phi c with a POD basis Phi of
64 modes, and an MLP maps (kappa, theta, sigma, rho, V0, r, T), scaled to
[-1, 1] over the training box, to the normalised coefficients c.

The POD uses the pricing inner product <q1, q2> = (G q1) . (G q2), where G is
the grid's s-marginal operator (price = exp(-rT) sum_i payoff(s_i) (G q)_i):
the modes are the leading right singular vectors of the centred marginal
snapshots G (q_j - mu), lifted to q space by the method of snapshots, so that
G Phi has orthonormal columns. European prices depend on q(T) only through
G q (an n_s-dimensional image), which 64 modes resolve to ~1e-8 in price,
whereas 64 Euclidean q-space modes leave O(1) price errors (their energy is
dominated by the density spike near v = 0 that the v-integral removes). The network
is 7 -> 256 -> 256 -> 256 -> 256 -> 64 with SiLU activations, BF16 weights
(FP32 master copy) and FP32 accumulation; it is trained with Adam. Forward,
backward, loss and optimiser steps are this library's kernels
(fpe_mlp_forward / fpe_mlp_backward / fpe_mlp_mse / fpe_mlp_adam); prices come
from fpe_surrogate_price, which folds the reconstruction into the s-marginal
pricing operator of the grid. Device memory comes from the library's own
allocator (fpe_device_alloc), so no second GPU runtime enters the process.

Training data are solver-generated (fpe_batch_terminal_states); the POD basis
is an offline host step (float64 eigendecomposition of the snapshot
covariance). Inputs outside the training box, and non-European types, go to
the solver. `audit` re-solves a random fraction with the plain solver.
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
    n = len(P)
    arr = L.batch_options(n)
    for i in range(n):
        o = arr[i]
        for k, v in GRID.items():
            setattr(o, k, v)
        for j, name in enumerate(PARAMS):
            setattr(o, name, float(P[i, j]))
        o.strike = float(strike[i]) if strike is not None else 60.0
        o.option_type = int(otype[i]) if otype is not None else 8
        o.barrier = 0.0
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


def pod(Q, k=64):
    """Mean and the k leading POD modes of snapshot rows Q (float64 host step)."""
    mu = Q.mean(axis=0)
    X = Q - mu
    cov = X.T @ X
    ev, V = np.linalg.eigh(cov)
    order = np.argsort(ev)[::-1][:k]
    return mu, V[:, order], ev[order]


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
    """MLP parameters on the GPU: FP32 master weights and their BF16 copies."""

    def __init__(self, weights, biases):
        self.lib = L.lib()
        if self.lib.fpe_device_count() < 1:
            raise RuntimeError("the surrogate kernels need a GPU")
        self.W = [DeviceArray.from_numpy(w, np.float32) for w in weights]
        self.b = [DeviceArray.from_numpy(b, np.float32) for b in biases]
        self.Wb = [DeviceArray(w.shape, np.uint16) for w in self.W]
        self.stream = None  # the null stream: every call is ordered after the previous one
        self._cache = {}
        for w, wb in zip(self.W, self.Wb):
            self._ok(self.lib.fpe_mlp_to_bf16(w.c, wb.c, int(np.prod(w.shape)), self.stream))

    @staticmethod
    def init(seed=0):
        rng = np.random.Generator(np.random.PCG64(seed))
        Ws, bs = [], []
        for a, b in zip(WIDTHS[:-1], WIDTHS[1:]):
            Ws.append(rng.normal(0.0, np.sqrt(1.0 / a), (b, a)))
            bs.append(np.zeros(b))
        return DeviceMLP(Ws, bs)

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
            self._ok(self.lib.fpe_mlp_forward(h.c, M, K, self.Wb[i].c, self.b[i].c, N, 0 if last else 1, _c(Z), A.c,
                                              self.stream))
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
            self._ok(self.lib.fpe_mlp_backward(g.c, X.c, self.Wb[i].c, M, N, K, dW.c, db.c, _c(dX), _c(Zp),
                                               self.stream))
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


def train(P, Y, epochs=60, batch=1024, lr=1e-3, lr_end=1e-5, wcol=None, seed=0, log=print, val=None):
    """Fit the MLP to normalised targets Y (n x k) for inputs P (n x 7).
    Returns the trained DeviceMLP and the per-epoch loss history."""
    net = DeviceMLP.init(seed)
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
    """Trained surrogate loaded on the GPU."""

    def __init__(self, path=MODEL_PATH):
        z = np.load(path, allow_pickle=False)
        n_layers = int(z["n_layers"])
        self.net = DeviceMLP([z[f"W{i}"] for i in range(n_layers)], [z[f"b{i}"] for i in range(n_layers)])
        self.k = int(z["cmean"].shape[0])
        self.cscale = DeviceArray.from_numpy(z["cstd"], np.float32)
        self.cshift = DeviceArray.from_numpy(z["cmean"], np.float32)
        self.Pt = DeviceArray.from_numpy(np.ascontiguousarray(z["P"].T), np.float32)
        self.w0 = DeviceArray.from_numpy(z["w0"], np.float32)
        self.sq = DeviceArray.from_numpy(z["s_q"], np.float32)
        self.nq = int(z["s_q"].shape[0])
        self.info = {k: z[k] for k in z.files if k.startswith("meta_") or k in ("loss", "pod_ev")}

    def coefficients(self, P):
        """Normalised POD coefficients for parameter rows P (DeviceArray, n x k)."""
        X = self.net.buf("X", (len(P), 7))
        X.upload(normalise(P))
        return self.net.forward(X)

    def price(self, P, strike, otype, chunk=131072):
        """Prices of European calls (8) / puts (9) for in-box rows; NaN elsewhere.
        Returns (prices, mask of rows the surrogate priced)."""
        P = np.asarray(P, np.float64)
        n = len(P)
        otype = np.asarray(otype, np.int32)
        ok = in_box(P) & np.isin(otype, (8, 9))
        out = np.full(n, np.nan)
        idx = np.nonzero(ok)[0]
        for a in range(0, len(idx), chunk):
            rows = idx[a:a + chunk]
            m = len(rows)
            Cn = self.coefficients(P[rows])
            buf = self.net.buf
            K, ty = buf("K", (m,), np.float64), buf("type", (m,), np.int32)
            r, T, pr = buf("r", (m,), np.float64), buf("T", (m,), np.float64), buf("price", (m,), np.float64)
            K.upload(np.asarray(strike, np.float64)[rows])
            ty.upload(otype[rows])
            r.upload(P[rows, 5])
            T.upload(P[rows, 6])
            self.net._ok(self.net.lib.fpe_surrogate_price(Cn.c, m, self.k, self.cscale.c, self.cshift.c, self.Pt.c,
                                                          self.w0.c, self.sq.c, self.nq, K.c, ty.c, r.c, T.c, pr.c,
                                                          self.net.stream))
            out[rows] = pr.numpy()
        return out, ok

    def price_with_fallback(self, P, strike, otype):
        """Surrogate where it applies, the plain solve elsewhere; returns prices,
        the surrogate mask and the number of fallback solves."""
        out, ok = self.price(P, strike, otype)
        rest = np.nonzero(~ok)[0]
        if len(rest):
            out[rest], _ = solver_prices(np.asarray(P)[rest], np.asarray(strike)[rest], np.asarray(otype)[rest])
        return out, ok, len(rest)


def audit(sur, P, strike, otype, frac=0.01, seed=1):
    """Re-solves a random fraction of the surrogate-priced rows with the plain
    solver. Returns (rows, surrogate prices, solver prices, solver status)."""
    sp, ok = sur.price(P, strike, otype)
    rows = np.nonzero(ok)[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    pick = np.sort(rng.choice(rows, size=max(1, int(round(frac * len(rows)))), replace=False))
    ref, st = solver_prices(np.asarray(P)[pick], np.asarray(strike)[pick], np.asarray(otype)[pick])
    return pick, sp[pick], ref, st


def fit_pod(Q, marg, k=64):
    """POD of q(T) snapshot rows Q in the pricing inner product (float64 host
    step). Returns mu (n_dof), Phi (n_dof x k), P = marg Phi (n_qs x k, orthonormal
    columns), w0 = marg mu and the singular values."""
    mu = Q.mean(axis=0)
    X = Q - mu
    Wc = X @ marg.T
    U, sv, Vt = np.linalg.svd(Wc, full_matrices=False)
    k = min(k, len(sv))
    Phi = X.T @ (U[:, :k] / sv[:k])
    return mu, Phi, Vt[:k].T.copy(), marg @ mu, sv


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


def build_model(data_files, pod_file, out=MODEL_PATH, epochs=200, batch=1024, lr=2e-3, n_val=4096, log=print,
                seed=0):
    """Trains the MLP on generated data files (P, C arrays) against the POD in
    pod_file and saves the complete model."""
    pod = dict(np.load(pod_file, allow_pickle=False))
    Ps, Cs = [], []
    for f in data_files:
        z = np.load(f, allow_pickle=False)
        keep = z["ok"]
        Ps.append(z["P"][keep])
        Cs.append(z["C"][keep].astype(np.float64))
    P, Cf = np.concatenate(Ps), np.concatenate(Cs)
    cmean, cstd = Cf.mean(axis=0), Cf.std(axis=0) + 1e-30
    Y = (Cf - cmean) / cstd
    wcol = cstd ** 2 / np.mean(cstd ** 2)
    Pv, Yv = P[:n_val], Y[:n_val]
    Ptr, Ytr = P[n_val:], Y[n_val:]
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    Kv = rng.uniform(45.0, 75.0, n_val)
    tv = np.where(rng.uniform(0.0, 1.0, n_val) < 0.5, 8, 9)
    Wv = (Yv * cstd + cmean) @ pod["P"].T + pod["w0"]
    pay = np.maximum(np.where(tv[:, None] == 8, pod["s_q"][None] - Kv[:, None], Kv[:, None] - pod["s_q"][None]), 0)
    pv = np.exp(-Pv[:, 5] * Pv[:, 6]) * np.sum(pay * Wv, axis=1)
    Xv = DeviceArray.from_numpy(normalise(Pv))

    def val(net):
        Yp = net.forward(Xv).numpy().astype(np.float64)
        Wp = (Yp * cstd + cmean) @ pod["P"].T + pod["w0"]
        e = np.abs(np.exp(-Pv[:, 5] * Pv[:, 6]) * np.sum(pay * Wp, axis=1) - pv)
        return float(np.max(e))

    log(f"training on {len(Ptr)} samples, validating on {n_val}")
    net, hist = train(Ptr, Ytr, epochs=epochs, batch=batch, lr=lr, wcol=wcol, seed=seed, log=log, val=val)
    Ws, bs = net.numpy()
    arrays = {f"W{i}": w for i, w in enumerate(Ws)}
    arrays.update({f"b{i}": b for i, b in enumerate(bs)})
    arrays.update(n_layers=np.array(len(Ws)), cmean=cmean, cstd=cstd, box_lo=BOX_LO, box_hi=BOX_HI,
                  loss=np.array(hist), meta_samples=np.array(len(Ptr)), meta_val=np.array(n_val),
                  meta_val_max_abs_dprice=np.array(val(net)))
    for key in ("mu", "Phi", "P", "w0", "s_q", "pod_sv", "pod_snapshots"):
        arrays[key] = pod[key]
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
    g.add_argument("--pod", default=None, help="existing POD file; otherwise fit one and write --pod-out")
    g.add_argument("--pod-out", default=None)
    g.add_argument("--out", required=True)
    t = sub.add_parser("train", help="train the MLP on generated data and write the model")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--pod", required=True)
    t.add_argument("--epochs", type=int, default=200)
    t.add_argument("--batch", type=int, default=1024)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--out", default=MODEL_PATH)
    a = ap.parse_args()
    say = lambda m: print(m, flush=True)
    if a.cmd == "generate":
        pod = dict(np.load(a.pod, allow_pickle=False)) if a.pod else None
        P, Cf, ok, pod2 = generate(a.samples, a.seed, pod=pod, log=say)
        np.savez(a.out, P=P, C=Cf, ok=ok, seed=np.array(a.seed))
        if pod is None:
            np.savez(a.pod_out or a.out.replace(".npz", "_pod.npz"), **pod2)
    else:
        build_model(a.data, a.pod, out=a.out, epochs=a.epochs, batch=a.batch, lr=a.lr, log=say)
