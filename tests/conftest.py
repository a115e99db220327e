import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


# Options whose GPU and oracle solver counters differed, so that the Newton
# bound and not the 1e-8 bar applied (name, GPU record, oracle record, error).
FALLBACKS = []


def record_fallback(name, gpu_info, oracle_stats, rel_err):
    keys = ("n_accepted", "n_rejected", "n_lu", "n_newton", "n_err_solves")
    FALLBACKS.append((name, {k: gpu_info[k] for k in keys} | {"inner": gpu_info.get("n_inner_iters")},
                      {k: oracle_stats[k] for k in keys} | {"gmres": oracle_stats.get("n_gmres_iters")}, rel_err))


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if not FALLBACKS:
        return
    terminalreporter.section("solver-trajectory differences (Newton bound applied)")
    for name, g, o, e in FALLBACKS:
        terminalreporter.write_line(f"{name}: rel err {e:.2e}; gpu {g}; oracle {o}")
    terminalreporter.write_line(f"{len(FALLBACKS)} option(s) off the oracle's trajectory")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.build()
    return oracle_lib


@pytest.fixture(scope="session")
def engine():
    """The product library; GPU tests fail loudly if it is not built."""
    from fpe_option_amd import _lib
    L = _lib.lib()
    if L.fpe_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return _lib
