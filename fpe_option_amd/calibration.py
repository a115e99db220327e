"""Heston calibration by Levenberg-Marquardt over batched GPU pricing.

The reference calibrator (src/engines/calibrator/calibrator.mojo:79-159,
src/numerics/optim/lm.mojo:20-124) fits x = (kappa, theta, sigma, rho, V0) to
market prices with LM: lambda starts at 1e-3, is multiplied by 10 on a
rejected step and by 0.1 on an accepted one (floor 1e-10), the step solves
(J^T J + lambda I) delta = -J^T r, and iteration stops when |J^T r|_inf < tol
or after max_iter. The Jacobian is central differences with
eps_j = 1e-6 (1 + |x_j|) (calibrator.mojo:55-77), and parameters are clamped
on the way into the model (calibrator.mojo:14-30). The reference evaluates
each residual vector and each of the 2n perturbed points one solve at a time
(objective.mojo:75-104).

Here every calibration problem of a batch, its current point and all its 2n
perturbed points, are priced in one fpe_batch_price launch per LM iteration
(plus one launch for the trial points): B problems x (2n + 1) points x m
quotes solves in a single batch. The model price is this engine's discounted
European price (calls, or puts by option_type) on the configured
discretisation; the reference objective's own quadrature (undiscounted,
trapezoid weights over the PDF grid, objective.mojo:26-48) is not reproduced.
"""
import numpy as np

from . import api

PARAM_NAMES = ("kappa", "theta", "sigma", "rho", "V0")


def clamp(X):
    """The reference's parameter clamps (calibrator.mojo:14-30), row-wise."""
    X = np.array(X, dtype=np.float64, copy=True)
    X[..., 0] = np.maximum(X[..., 0], 1e-4)
    X[..., 1] = np.maximum(X[..., 1], 1e-5)
    X[..., 2] = np.maximum(X[..., 2], 1e-4)
    X[..., 3] = np.clip(X[..., 3], -0.999, 0.999)
    X[..., 4] = np.maximum(X[..., 4], 1e-6)
    return X


def model_prices(X, strikes, expiries, base, option_type=8):
    """Prices for parameter rows X (..., 5) at each problem's quotes.
    X: (B, P, 5) points per problem; strikes/expiries: (B, m). Returns (B, P, m)."""
    X = clamp(X)
    B, P, _ = X.shape
    m = strikes.shape[1]
    cols = {name: np.repeat(X[:, :, j].reshape(B * P), m) for j, name in enumerate(PARAM_NAMES)}
    cols["strike"] = np.tile(strikes[:, None, :], (1, P, 1)).reshape(-1)
    cols["T"] = np.tile(expiries[:, None, :], (1, P, 1)).reshape(-1)
    cols["option_type"] = option_type
    cols["barrier"] = 0.0
    cols.update(base)
    prices, status = api.price_batch(**cols)
    if np.any(status != 0):
        bad = int(np.sum(status != 0))
        raise RuntimeError(f"{bad} calibration solves failed (status {np.unique(status[status != 0])})")
    return prices.reshape(B, P, m)


def levenberg_marquardt_batch(residuals, X0, max_iter=20, tol=1e-6, lambda_init=1e-3, lambda_up=10.0,
                              lambda_down=0.1, jac_eps=1e-6):
    """LM on B independent problems at once, each following lm.mojo:30-124.

    residuals(idx, pts) maps points pts (b, P, n) of the problems idx (b,) to
    residual vectors (b, P, m). Each iteration calls it once with the current
    points of the active problems and their 2n central-difference perturbations
    (calibrator.mojo:55-77), and once with their trial points. Returns the
    points, the iterations each problem ran and whether it met the gradient
    tolerance."""
    X = np.array(X0, dtype=np.float64, copy=True)
    B, n = X.shape
    lam = np.full(B, lambda_init)
    active = np.ones(B, bool)
    iters = np.zeros(B, int)
    for _ in range(max_iter):
        idx = np.nonzero(active)[0]
        if len(idx) == 0:
            break
        Xa = X[idx]
        eps = jac_eps * (1.0 + np.abs(Xa))
        pts = np.repeat(Xa[:, None, :], 2 * n + 1, axis=1)
        for j in range(n):
            pts[:, 1 + 2 * j, j] += eps[:, j]
            pts[:, 2 + 2 * j, j] -= eps[:, j]
        R = residuals(idx, pts)
        r = R[:, 0]
        J = np.stack([(R[:, 1 + 2 * j] - R[:, 2 + 2 * j]) / (2.0 * eps[:, j:j + 1]) for j in range(n)], axis=2)
        g = np.einsum("bmi,bm->bi", J, r)  # J^T r
        iters[idx] += 1
        done = np.max(np.abs(g), axis=1) < tol
        active[idx[done]] = False
        keep = ~done
        if not keep.any():
            break
        idx, J, r, g, Xa = idx[keep], J[keep], r[keep], g[keep], Xa[keep]
        A = np.einsum("bmi,bmj->bij", J, J) + lam[idx, None, None] * np.eye(n)[None]
        delta = np.linalg.solve(A, -g[..., None])[..., 0]
        Xn = Xa + delta
        rn = residuals(idx, Xn[:, None, :])[:, 0]
        better = np.sum(rn * rn, axis=1) < np.sum(r * r, axis=1)
        X[idx[better]] = Xn[better]
        lam[idx[better]] = np.maximum(lam[idx[better]] * lambda_down, 1e-10)
        lam[idx[~better]] = lam[idx[~better]] * lambda_up
    return X, iters, ~active


def calibrate_batch(market_prices, strikes, expiries, init_params, base=None, max_iter=20, tol=1e-6,
                    option_type=8):
    """Calibrate B independent problems.

    market_prices, strikes, expiries: (B, m) arrays; init_params: (B, 5) rows of
    (kappa, theta, sigma, rho, V0); base: the remaining pricing keywords shared by
    every solve (S0, r, s_min, s_max, n_s, n_v, disc_mode, ...). Returns a dict with
    the fitted parameter rows (clamped), the final residuals, LM iterations and
    convergence flags."""
    market = np.atleast_2d(np.asarray(market_prices, np.float64))
    K = np.atleast_2d(np.asarray(strikes, np.float64))
    T = np.atleast_2d(np.asarray(expiries, np.float64))
    X0 = np.atleast_2d(np.asarray(init_params, np.float64))
    base = dict(base or {})

    def residuals(idx, pts):
        return model_prices(pts, K[idx], T[idx], base, option_type) - market[idx][:, None, :]

    X, iters, conv = levenberg_marquardt_batch(residuals, X0, max_iter=max_iter, tol=tol)
    final = residuals(np.arange(len(K)), X[:, None, :])[:, 0]
    return {"params": clamp(X), "residuals": final, "iterations": iters, "converged": conv}


def calibrate(market_prices, strikes, expiries, init_params, base=None, max_iter=20, tol=1e-6, option_type=8):
    """Single calibration (calibrator.mojo:86-120): returns the fitted
    (kappa, theta, sigma, rho, V0) and the fit details."""
    out = calibrate_batch([market_prices], [strikes], [expiries], [init_params], base, max_iter, tol, option_type)
    return {k: v[0] for k, v in out.items()}
