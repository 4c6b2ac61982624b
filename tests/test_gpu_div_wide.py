"""GPU: division by divisors beyond the register kernels (more than 128 limbs),
one wave per division with the divisor and the dividend in LDS, against GMP
(mpz_fdiv_qr, the reference test suite's oracle, unit_tests/test.cu:2784-2813).

Covers the first such size (129 limbs), a chunk per lane that ends exactly at
the window's top limb and one that does not, dividends shorter than, as long
as and several times the divisor, unnormalised divisors (leading zero limbs
and bits, down to b = 1), division by zero, all-ones operands (the largest
estimates), and limbs drawn from {0, 1, 2^31 - 1, 2^31, 2^32 - 2, 2^32 - 1},
whose quotient estimates are one too large often enough (about one digit in
a thousand) to run the add-back on every call.  The power-of-two dividend
(R^2 mod m for long moduli) is covered through the Montgomery contexts.
"""
import numpy as np
import pytest

from helpers import O, assert_rows_equal, gmp, inputs

pytestmark = pytest.mark.gpu


def _check(handle, a, b, la, lb, what):
    n = a.shape[0]
    A, B = handle.from_array(a), handle.from_array(b)
    Q, R = handle.integers(32 * la, n), handle.integers(32 * lb, n)
    handle.divmod(Q, R, A, B)
    wq, wr = gmp(O.GO_DIV, n, la, a, b, out2_limbs=lb)
    assert_rows_equal(Q.get(), wq, what + " q")
    assert_rows_equal(R.get(), wr, what + " r")
    Q2 = handle.integers(32 * la, n)
    handle.div(Q2, A, B)
    assert_rows_equal(Q2.get(), wq, what + " div")


@pytest.mark.parametrize("la,lb", [(258, 129), (512, 256), (1024, 512), (1251, 384), (191, 191), (900, 191),
                                   (200, 300)])
def test_div_wide_random(handle, la, lb):
    n = 130
    a = inputs.uniform(1000 + la, n, la)
    b = inputs.uniform(2000 + lb, n, lb)
    b[:, -1] |= np.uint32(0x80000000)
    _check(handle, a, b, la, lb, "%d/%d limbs" % (la, lb))


def test_div_wide_structured_limbs(handle):
    """Limbs at the carry and borrow extremes; the add-back runs."""
    n, la, lb = 300, 260, 130
    rng = np.random.default_rng(11)
    vals = np.array([0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF], np.uint32)
    a = vals[rng.integers(0, len(vals), (n, la))]
    b = vals[rng.integers(0, len(vals), (n, lb))]
    b[:, -1] |= np.uint32(0x80000000)
    a[0, :] = 0xFFFFFFFF
    b[0, :] = 0xFFFFFFFF
    a[1, :] = 0xFFFFFFFF
    b[1, :] = 0
    b[1, -1] = 0x80000000
    _check(handle, a, b, la, lb, "structured")


def test_div_wide_unnormalised_and_zero(handle):
    n, la, lb = 200, 400, 200
    a = inputs.uniform(31, n, la)
    b = inputs.uniform(32, n, lb)
    rng = np.random.default_rng(3)
    for i in range(n):
        top = int(rng.integers(1, lb + 1))
        b[i, top:] = 0
        b[i, top - 1] >>= int(rng.integers(0, 32))
        if not b[i].any():
            b[i, 0] = 1
    a[: n // 8, lb:] = 0   # some a < b
    b[-3:] = 0
    b[-3:, 0] = 1          # b = 1
    A, B = handle.from_array(a), handle.from_array(b)
    Q, R = handle.integers(32 * la, n), handle.integers(32 * lb, n)
    handle.divmod(Q, R, A, B)
    wq, wr = gmp(O.GO_DIV, n, la, a, b, out2_limbs=lb)
    assert_rows_equal(Q.get(), wq, "unnormalised q")
    assert_rows_equal(R.get(), wr, "unnormalised r")
    # division by zero: q = 0, r = a
    Z = handle.from_array(np.zeros((n, lb), np.uint32))
    handle.divmod(Q, R, A, Z)
    assert not Q.get().any()
    assert np.array_equal(R.get(), a[:, :lb])
