"""FP64 numpy reference of the surrogate kernels, with the kernels' BF16
rounding points (test infrastructure only)."""
import numpy as np


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _r(x, fp32):
    return np.asarray(x, np.float64) if fp32 else bf16(x)


def linear_forward(X, W, b, act, fp32=False):
    Z = _r(X, fp32) @ _r(W, fp32).T + b
    return Z, (silu(Z) if act else Z)


def linear_backward(dZ, X, W, Zp, fp32=False):
    dW = _r(dZ, fp32).T @ _r(X, fp32)
    db = dZ.sum(axis=0)
    dX = _r(dZ, fp32) @ _r(W, fp32)
    if Zp is not None:
        dX = dX * dsilu(Zp)
    return dW, db, dX


def surrogate_price(Cn, cscale, cshift, P, w0, sq, strike, otype, r, T):
    c = Cn.astype(np.float64) * cscale + cshift
    w = c @ P.T + w0
    pay = np.where(otype[:, None] == 8, sq[None, :] - strike[:, None], strike[:, None] - sq[None, :])
    return np.exp(-r * T) * np.sum(np.maximum(pay, 0.0) * w, axis=1)
