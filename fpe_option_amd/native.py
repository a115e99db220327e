"""Drop-in for the reference's native extension module `fpe_engine._fpe_native`.

The reference package's Python layer (python/fpe_engine/__init__.py:31-32,
python/fpe_engine/pricer.py:9) imports `price` and `Compute` from a Mojo
extension (src/bindings/_fpe_native.mojo:34-53, src/bindings/_py_pipeline.mojo:
16-117) whose methods return plain dicts and lists. This module offers the same
two names with the same call shapes over libfpe_engine.so, so the reference's
own `__init__.py` and `pricer.py` run unchanged when `_fpe_native` resolves here.
"""
import numpy as np

from . import api


def _csr_dict(m) -> dict:
    # the dict form that pricer._csr_from_mojo reads (pricer.py:49-53)
    return {"data": m.data, "indices": m.indices, "indptr": m.indptr, "nrows": m.shape[0], "ncols": m.shape[1]}


def price(params: dict) -> dict:
    """py_price (_fpe_native.mojo:12-32): kwargs dict in, lists of prices and Greeks out."""
    kw = dict(params)
    r = api.price(**kw)
    return {"prices": r.prices, "deltas": r.deltas, "gammas": r.gammas, "vegas": r.vegas}


class Compute:
    """PyComputePipeline (_py_pipeline.mojo:16-117): raw-result methods."""

    def __init__(self, **kwargs):
        self._c = api.Compute(**kwargs)

    def knots(self) -> dict:
        k = self._c.knots
        return {"s": k.s, "v": k.v}

    def grid_points(self) -> dict:
        g = self._c.grid_points
        return {"s": g.s, "v": g.v, "s_weights": g.s_weights, "v_weights": g.v_weights}

    def basis_1d(self) -> dict:
        b = self._c.basis_1d
        return {"Bs": _csr_dict(b.Bs), "dBs": _csr_dict(b.dBs), "Bv": _csr_dict(b.Bv), "dBv": _csr_dict(b.dBv)}

    def basis_2d(self) -> dict:
        return _csr_dict(self._c.basis_2d)

    def initial_condition(self) -> np.ndarray:
        return self._c.initial_condition

    def solve(self) -> list:
        return self._c.solve

    def pdf(self) -> np.ndarray:
        return self._c.pdf

    def payoff_price(self, K) -> np.ndarray:
        if not isinstance(K, (list, float, int)):
            raise TypeError("K must be float, int, or list of floats")
        return self._c.payoff_price(K)

    def greeks(self, K) -> dict:
        if not isinstance(K, (list, float, int)):
            raise TypeError("K must be float, int, or list of floats")
        g = self._c.greeks(K)
        return {"delta": g.delta, "gamma": g.gamma, "vega": g.vega}
