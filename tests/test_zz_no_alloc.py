"""ncsv_sample_noalloc (no counterpart in the reference): sampling over the
library's fixed per-GPU scratch array, for callers that cannot allocate.
Called through the C ABI and compared with ncsv_sample bit for bit, on two
calls in a row (the second smaller, so stale scratch contents from the first
must not leak into it), in both precisions, and its size limit is an error."""
import ctypes

import numpy as np
import pytest

import ncsim_amd as H
from ncsim_amd import _lib as L

pytestmark = pytest.mark.gpu


def sample_noalloc(state, us):
    us = np.ascontiguousarray(us, dtype=np.float64)
    out = np.zeros(max(1, len(us)), dtype=np.uint64)
    H._check(L.lib().ncsv_sample_noalloc(state._h, len(us), us.ctypes.data_as(L.c_dp),
                                         out.ctypes.data_as(L.c_u64p)))
    return out[:len(us)]


@pytest.mark.parametrize("precision", [H.C128, H.C64])
def test_noalloc_sampling_equals_the_allocating_entry_point(precision):
    rng = np.random.default_rng(77)
    first = (21, 300_000)
    second = (15, 20_000)
    for n, shots in (first, second):
        amp = (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)
        amp[rng.integers(0, 1 << n, 1 << (n - 4))] = 0.0
        d = H.HipStateVector.from_amplitudes(n, amp, precision)
        us = rng.random(shots)
        want = d.sample_with_uniforms(us)
        got = sample_noalloc(d, us)
        assert np.array_equal(want, got), (n, shots)
        assert np.array_equal(sample_noalloc(d, us[:7]), want[:7])
    # zero shots: ZeroState is still checked, nothing written
    z = H.HipStateVector.zero_state(4, precision)
    assert len(sample_noalloc(z, [])) == 0


def test_noalloc_sampling_rejects_more_than_its_scratch():
    d = H.HipStateVector.zero_state(20)
    with pytest.raises(H.NcsvError) as e:
        sample_noalloc(d, np.zeros(2_100_000))
    assert e.value.kind == "Invalid"
    # the handle keeps working after the refusal
    assert np.array_equal(sample_noalloc(d, [0.5, 0.25]), np.zeros(2, dtype=np.uint64))
