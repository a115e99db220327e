"""The per-call latency measurement (no GPU): every timed call has the reference's
published time beside it, and the case is the reference's README benchmark."""
from fpe_option_amd import latency


def test_every_call_has_a_reference_time():
    assert set(latency._calls()) == set(latency.REFERENCE_S)
    assert all(t > 0 for t in latency.REFERENCE_S.values())


def test_case_is_the_reference_benchmark():
    c = latency.CASE
    assert (c["n_s"], c["n_v"], c["num_insert"], c["s_max"], c["S0"], c["T"]) == (38, 38, 251, 150.0, 60.0, 0.6)
    assert latency.STRIKES == [65.0 + 5 * i for i in range(8)]
