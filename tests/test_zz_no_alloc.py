"""GPU parity of xmpIntegersPowmNoAllocAsync, the allocation-free variant of
xmpIntegersPowmAsync (its workspace is a fixed device array inside the
library): through the C ABI, two calls in a row -- the second on a smaller
input -- must give exactly what xmpIntegersPowm gives on the same inputs, and
GMP agrees.  The array is shared and keeps the previous call's contents, so
the second, smaller call also shows that nothing stale leaks into a result."""
import ctypes

import numpy as np
import pytest

import xmp_amd
from helpers import O, assert_rows_equal, gmp, inputs

pytestmark = pytest.mark.gpu


def _noalloc(h, out, b, e, m, count):
    rc = xmp_amd.lib().xmpIntegersPowmNoAllocAsync(h.h, out.x, b.x, e.x, m.x, count)
    if rc != 0:
        raise xmp_amd.XmpError(rc, "xmpIntegersPowmNoAllocAsync")
    from xmp_amd import hip
    hip.device_sync()


def _case(h, seed, n, bits):
    limbs = bits // 32
    m = inputs.odd_moduli(seed, n, limbs)
    b = inputs.below(seed + 1, m)
    e = inputs.uniform(seed + 2, n, limbs)
    return m, b, e, h.from_array(m), h.from_array(b), h.from_array(e)


def test_noalloc_matches_powm_twice(fresh_handle):
    h = fresh_handle
    # first call: 16,384 x 2048-bit (window table ~310 MB at the tuned window:
    # the call narrows its window to fit the 256 MiB array); second call:
    # 1,000 x 1024-bit, smaller, on a different layout
    for seed, n, bits in ((7100, 16384, 2048), (7200, 1000, 1024)):
        m, b, e, M, B, E = _case(h, seed, n, bits)
        ref, got = h.integers(bits, n), h.integers(bits, n)
        h.powm(ref, B, E, M)
        _noalloc(h, got, B, E, M, n)
        want = ref.get()
        assert_rows_equal(got.get(), want, "no-alloc powm %d x %d-bit vs xmpIntegersPowm" % (n, bits))
        k = min(n, 512)
        assert_rows_equal(want[:k], gmp(O.GO_POWM, k, bits // 32, b[:k], e[:k], m[:k]),
                          "xmpIntegersPowm %d-bit vs GMP" % bits)
        for x in (M, B, E, ref, got):
            x.close()


def test_noalloc_in_place_indexed_and_limits(fresh_handle):
    """In place (out aliases the base) with a gathered modulus; count 0; a
    call whose smallest plan cannot fit the array is refused."""
    h = fresh_handle
    n, bits = 3000, 2048
    m, b, e, M, B, E = _case(h, 7300, n, bits)
    ref = h.integers(bits, n)
    h.powm(ref, B, E, M)
    _noalloc(h, B, B, E, M, n)
    assert_rows_equal(B.get(), ref.get(), "in-place no-alloc powm")
    assert xmp_amd.lib().xmpIntegersPowmNoAllocAsync(h.h, ref.x, B.x, E.x, M.x, 0) == xmp_amd.SUCCESS
    # 65,536 resident 32768-bit moduli need 256 MiB of R^2 mod m and 512 MiB
    # of division workspace: refused before anything is launched
    big_m = h.integers(32768, 65536)
    big_o = h.integers(32768, 65536)
    big_e = h.integers(64, 65536)
    rc = xmp_amd.lib().xmpIntegersPowmNoAllocAsync(h.h, big_o.x, big_o.x, big_e.x, big_m.x, 65536)
    assert rc == xmp_amd.INCREASE_SCRATCH_LIMIT
    for x in (M, B, E, ref, big_m, big_o, big_e):
        x.close()
