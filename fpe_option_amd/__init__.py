"""MI355X Heston Fokker-Planck option pricing engine.

Drop-in for the reference `fpe_engine` Python package: `price`, `Compute`,
result dataclasses, `FpeParams`, `is_available`, plus the batched extension
`price_batch`. The numerics run in hand-written HIP kernels (libfpe_engine.so).
"""
import logging

from .api import (  # noqa: F401
    Basis1DResult, Compute, FpeParams, GreeksResult, GridPointsResult, KnotsResult, PriceResult, _OPTION_TYPES,
    is_available, price, price_batch)

__version__ = "0.1.0"
_logger = logging.getLogger("fpe_engine")
from .calibration import calibrate, calibrate_batch  # noqa: E402,F401
