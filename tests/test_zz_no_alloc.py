"""naer_ref_poisson_epoch_noalloc (the reference's POISSON epoch staged
through a static device array, no allocation) against naer_ref_poisson_epoch
(buffers allocated per call) through the C ABI: bit-identical outputs on two
calls in a row, the second smaller (the array keeps the first call's
contents), and a clear error beyond the array's capacity."""
import ctypes

import numpy as np
import pytest

from naer_amd import _abi
from naer_amd.network import Context, NaerError

pytestmark = pytest.mark.gpu

SEED32 = 0x15AC7EBD
MAX_N = (1 << 20) // 11


def epoch(fn, n, n_pois, t0, rate=20.0):
    times = np.full((n, 10), 0xDEADBEEF, np.uint32)  # every entry is written by the call
    count = np.full(n, 0xDEADBEEF, np.uint32)
    _abi.check(fn(Context.get(0).handle, n, n_pois, SEED32, ctypes.c_float(rate), t0, t0 + 1500,
                  times.ctypes.data, count.ctypes.data))
    return times, count


def test_no_alloc_equals_allocating_entry_twice():
    L = _abi.lib()
    for n, n_pois, t0, rate in ((MAX_N, 120_000, 1500, 20.0), (2_000, 2_000, 4_500_000, 900.0)):
        t_a, c_a = epoch(L.naer_ref_poisson_epoch, n, n_pois, t0, rate)
        t_b, c_b = epoch(L.naer_ref_poisson_epoch_noalloc, n, n_pois, t0, rate)
        assert c_a.max() <= 10 and c_a.sum() > 0
        np.testing.assert_array_equal(c_b, c_a)
        np.testing.assert_array_equal(t_b, t_a)


def test_no_alloc_rejects_sizes_beyond_its_scratch():
    with pytest.raises(NaerError, match="static scratch"):
        epoch(_abi.lib().naer_ref_poisson_epoch_noalloc, MAX_N + 1, MAX_N + 1, 0)
