"""The allocation-free entry point clw_device_primes_noalloc (static device
scratch instead of hipMalloc) against clw_device_primes, the entry point it
varies, bit for bit: two calls in a row on the same scratch, the second on a
smaller range (its count pass must not see the first call's counts), a cap
below the count, and the width limit."""
import numpy as np
import pytest

import clwilson_amd as cw

pytestmark = pytest.mark.gpu


def test_noalloc_primes_equal_device_primes(dev):
    for lo, hi in [(10**12, 10**12 + 10**7), (2, 10**7 + 2), (999_000, 1_200_000)]:
        got, n = dev.primes_noalloc(lo, hi)
        ref = dev.primes(lo, hi)
        assert n == len(ref) and np.array_equal(got, ref), (lo, hi)
    got, n = dev.primes_noalloc(10**9, 10**9 + 100_000, cap=100)
    ref = dev.primes(10**9, 10**9 + 100_000)
    assert n == len(ref) and np.array_equal(got, ref[:100])


def test_noalloc_rejects_ranges_beyond_its_scratch(dev):
    with pytest.raises(cw.ClwError, match="scratch limit"):
        dev.primes_noalloc(10**12, 10**12 + 10**7 + 1)
    got, n = dev.primes_noalloc(10**12, 10**12)  # empty
    assert n == 0
