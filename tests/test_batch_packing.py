"""Host-side packing of batch options (no GPU): the FpeBatchOption array that
price_batch hands to the C ABI, written column by column through a numpy view,
holds exactly the bytes that setting every field of every struct would, and
malformed columns are rejected before anything reaches the library."""
import numpy as np
import pytest

from fpe_option_amd import _lib as L
from fpe_option_amd import api, workloads


def _packed_field_by_field(columns, n):
    arr = L.batch_options(n)
    for k, v in columns.items():
        if k == "option_type":
            v = [api._option_type_int(x) for x in np.atleast_1d(v)] if not np.isscalar(v) else api._option_type_int(v)
        a = np.broadcast_to(np.asarray(v), (n,))
        for i in range(n):
            setattr(arr[i], k, a[i].item())
    return arr


def test_columns_pack_to_the_same_bytes():
    cols = workloads.batch_columns(257, seed=7)
    assert bytes(api._batch_array(cols)) == bytes(_packed_field_by_field(cols, 257))


def test_scalars_strings_and_defaults():
    cols = dict(kappa=[1.0, 2.0, 3.0], option_type="down_and_out_call", barrier=50.0, n_s=np.int64(40))
    arr = api._batch_array(cols)
    assert bytes(arr) == bytes(_packed_field_by_field(cols, 3))
    d = L.FpeBatchOption()
    L.lib().fpe_batch_option_default(d)
    # fields no column names keep the library's defaults
    assert [arr[i].rtol for i in range(3)] == [d.rtol] * 3
    assert [arr[i].option_type for i in range(3)] == [2] * 3


@pytest.mark.parametrize("bad, err", [
    (dict(n_s=40.5), TypeError), (dict(kappa="x"), TypeError), (dict(option_type=12), ValueError),
    (dict(option_type=[8, -1]), ValueError), (dict(option_type="nope"), ValueError), (dict(volatility=0.2), KeyError),
])
def test_malformed_columns_are_rejected(bad, err):
    with pytest.raises(err):
        api._batch_array(bad)
