"""GPU: products beyond the row kernels (more than 512 limbs), which split by
Karatsuba into half-size products of the whole batch, against GMP's mpz_mul
(the reference test suite's oracle, unit_tests/test.cu:1616,1780).

The cases walk the split's branches: the smallest Karatsuba size (513 limbs,
h = 257), a multiplier no longer than h (a alone is split), unbalanced
operands whose sub-products split again, two levels (2048 limbs), truncated
and zero-padded outputs, all-ones operands (every term and the signed carry
of the recombination at their extremes), the one term that reaches past the
product (ca cb B^2 at limb 3h = len(a) + len(b)) with the product's end on a
chunk boundary of the recombination, squares, index arrays with a
broadcast multiplier, in place, the schoolbook fallback when the scratch
limit leaves no room for the workspace, and the bench's shape (4,096 pairs of
32,768 bits, whose half products take the 16-lane row groups).
"""
import numpy as np
import pytest

from helpers import O, assert_rows_equal, gmp, inputs

pytestmark = pytest.mark.gpu


def _pair(seed, n, la, lb):
    return inputs.uniform(seed, n, la), inputs.uniform(seed + 1, n, lb)


@pytest.mark.parametrize("bits_a,bits_b,bits_p", [
    (16416, 16416, 32832),     # 513 limbs: the first size that splits
    (16416, 8192, 24608),      # b.len <= h: a alone is split
    (40032, 30016, 70048),     # unbalanced; a1*b1 splits a alone
    (65536, 65536, 131072),    # two levels
    (32768, 32768, 40000),     # truncated product
    (20000, 20000, 40064),     # output wider than the product
    (20000, 96, 20096),        # short multiplier
])
def test_karatsuba_products(handle, bits_a, bits_b, bits_p):
    n = 70
    a, b = _pair(900 + bits_a // 32, n, bits_a // 32, bits_b // 32)
    A, B = handle.from_array(a), handle.from_array(b)
    P = handle.integers(bits_p, n)
    handle.mul(P, A, B)
    assert_rows_equal(P.get(), gmp(O.GO_MUL, n, bits_p // 32, a, b), "%dx%d->%d" % (bits_a, bits_b, bits_p))
    # the operands swapped: the longer one is split
    handle.mul(P, B, A)
    assert_rows_equal(P.get(), gmp(O.GO_MUL, n, bits_p // 32, a, b), "%dx%d->%d swapped" % (bits_a, bits_b, bits_p))


@pytest.mark.parametrize("bits", [16416, 32768, 65536])
def test_karatsuba_all_ones_and_squares(handle, bits):
    n, limbs = 67, bits // 32
    ones = np.full((n, limbs), 0xFFFFFFFF, np.uint32)
    A = handle.from_array(ones)
    P = handle.integers(2 * bits, n)
    handle.mul(P, A, A)   # a == b: the square path
    want = O.from_int(((1 << bits) - 1) ** 2, 2 * limbs)
    assert all(np.array_equal(r, want) for r in P.get()), "(2^%d - 1)^2" % bits
    # (2^k - 1) * (2^k - 2): a product, not a square, at the carry bound
    twos = ones.copy()
    twos[:, 0] = 0xFFFFFFFE
    B = handle.from_array(twos)
    handle.mul(P, A, B)
    want = O.from_int(((1 << bits) - 1) * ((1 << bits) - 2), 2 * limbs)
    assert all(np.array_equal(r, want) for r in P.get()), "(2^%d - 1)(2^%d - 2)" % (bits, bits)
    a = inputs.uniform(77 + bits, n, limbs)
    A = handle.from_array(a)
    handle.mul(P, A, A)
    assert_rows_equal(P.get(), gmp(O.GO_MUL, n, 2 * limbs, a, a), "random squares, %d bits" % bits)


def test_karatsuba_carry_term_past_the_product(handle):
    """1087 x 545 limbs: h = 544, len(a) + len(b) = 3h = 1632 (51 chunks of
    32 limbs), so the product of the half sums' carry limbs sits at limb
    1632, one past the product, where the carries from below cancel it; the
    output runs 32 limbs further.  All-ones operands set both carry limbs."""
    n, la, lb, lp = 66, 1087, 545, 1664
    ones_a = np.full((n, la), 0xFFFFFFFF, np.uint32)
    ones_b = np.full((n, lb), 0xFFFFFFFF, np.uint32)
    P = handle.integers(32 * lp, n)
    handle.mul(P, handle.from_array(ones_a), handle.from_array(ones_b))
    want = O.from_int(((1 << (32 * la)) - 1) * ((1 << (32 * lb)) - 1), lp)
    assert all(np.array_equal(r, want) for r in P.get()), "(2^34784 - 1)(2^17440 - 1)"
    a, b = _pair(611, n, la, lb)
    handle.mul(P, handle.from_array(a), handle.from_array(b))
    assert_rows_equal(P.get(), gmp(O.GO_MUL, n, lp, a, b), "1087 x 545 limbs -> 1664")


def test_karatsuba_indices_broadcast_and_in_place(handle):
    bits, n = 20480, 130
    limbs = bits // 32
    a = inputs.uniform(301, n, limbs)
    b = inputs.uniform(302, 3, limbs)
    rng = np.random.default_rng(7)
    ia, ip = rng.permutation(n).astype(np.uint32), rng.permutation(n).astype(np.uint32)
    A, B = handle.from_array(a), handle.from_array(b)
    P = handle.integers(2 * bits, n)
    handle.set_indices(0, ip)
    handle.set_indices(1, ia)
    try:
        handle.mul(P, A, B)
    finally:
        handle.clear_indices()
    rows = gmp(O.GO_MUL, n, 2 * limbs, a[ia], b[np.arange(n) % 3])
    want = np.zeros_like(rows)
    want[ip] = rows
    assert_rows_equal(P.get(), want, "indexed Karatsuba product")
    # in place, truncated to the operand's precision
    B2 = handle.from_array(inputs.uniform(303, n, limbs))
    b2 = B2.get()
    handle.mul(A, A, B2)
    assert_rows_equal(A.get(), gmp(O.GO_MUL, n, limbs, a, b2), "in-place Karatsuba product")


def test_karatsuba_scratch_limit_falls_back(fresh_handle):
    """A scratch limit below the workspace keeps the schoolbook kernel: the
    same products."""
    h = fresh_handle
    bits, n = 16416, 40
    a, b = _pair(401, n, bits // 32, bits // 32)
    A, B = h.from_array(a), h.from_array(b)
    P = h.integers(2 * bits, n)
    h.set_scratch_limit(4096)
    try:
        h.mul(P, A, B)
    finally:
        h.set_scratch_limit(0)
    assert_rows_equal(P.get(), gmp(O.GO_MUL, n, 2 * bits // 32, a, b), "schoolbook fallback")


def test_karatsuba_bench_shape(handle):
    """4,096 products and squares of 32,768 bits, the bench's shape: the
    half products then take the 16-lane row groups (below 4,096 products they
    take 32-lane ones), against GMP on a sample including the last rows."""
    n, limbs = 4096, 1024
    a, b = _pair(5151, n, limbs, limbs)
    a[0] = 0xFFFFFFFF
    b[0] = 0xFFFFFFFF
    A, B = handle.from_array(a), handle.from_array(b)
    P = handle.integers(2 * 32 * limbs, n)
    rows = np.unique(np.concatenate([np.linspace(0, n - 1, 200).astype(int), np.arange(n - 70, n)]))
    handle.mul(P, A, B)
    assert_rows_equal(P.get()[rows], gmp(O.GO_MUL, rows.size, 2 * limbs, a[rows], b[rows]), "products")
    handle.mul(P, A, A)
    assert_rows_equal(P.get()[rows], gmp(O.GO_MUL, rows.size, 2 * limbs, a[rows], a[rows]), "squares")
