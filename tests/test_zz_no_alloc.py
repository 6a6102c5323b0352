"""ksv_sample_noalloc, the allocation-free variant of ksv_sample, through the C ABI.

Collected last (it shares one fixed scratch array per GPU with every call).
"""
import numpy as np
import pytest

from ket_amd import workloads
from ket_amd._native import RangeError, State, check, lib, ptr

import ketsv_oracle as O
from helpers import build, dense_ref

pytestmark = pytest.mark.gpu


def noalloc(st, r):
    """ksv_sample_noalloc called through the C ABI."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty(len(r), dtype=np.uint64)
    check(lib().ksv_sample_noalloc(st.handle, ptr(r), len(r), ptr(out)))
    return out


def test_sample_noalloc_equals_sample_twice():
    """Two calls in a row, the second on a smaller state with fewer draws: the same
    indices as ksv_sample and as the oracle (the fixed scratch keeps the first call's
    sums, which the second call must not read)."""
    for n, shots, seed in ((16, 50000, 3), (9, 700, 4)):
        calls = workloads.random_layer_calls(n, 3, seed=seed)
        ref = dense_ref(n, calls, prune=True)
        st = build(n, calls).simulate()
        r = O.prng(seed, shots)
        got = noalloc(st, r)
        assert np.array_equal(got, st.sample(r))
        assert np.array_equal(got, O.sample(n, ref, r))
        st.free()


def test_sample_noalloc_limits():
    st = State(21, "f64")
    with pytest.raises(RangeError, match="at most 20"):
        noalloc(st, np.array([0.5]))
    st.free()
    st = State(4, "f64")
    with pytest.raises(RangeError, match="at most 65536"):
        noalloc(st, np.full(65537, 0.5))
    assert noalloc(st, np.array([0.3, 0.9])).tolist() == [0, 0]
    st.free()
