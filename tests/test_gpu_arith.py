"""GPU parity: add, sub, mul, div, mod, divmod against GMP and the C oracle.

Mirrors the reference's genericTwoInOneOutTest / SubTest / genericTwoInTwoOutTest
(unit_tests/test.cu:1674-2376): seeded inputs, the device op through the
C ABI, GMP on the host, and the low `limbs` words of every instance compared
(test.cu:1223-1228).  Bit-exact.
"""
import numpy as np
import pytest

import xmp_amd
from helpers import O, assert_rows_equal, gmp, inputs

pytestmark = pytest.mark.gpu


def _pair(seed, count, la, lb):
    return inputs.uniform(seed, count, la), inputs.uniform(seed + 7, count, lb)


@pytest.mark.parametrize("bits_a,bits_b,bits_s", [(32, 32, 64), (64, 64, 64), (256, 128, 288), (2048, 2048, 2048),
                                                  (2048, 2048, 2080), (2048, 1024, 4096), (4096, 4096, 4096)])
def test_add(handle, bits_a, bits_b, bits_s):
    n = 1000
    a, b = _pair(11, n, bits_a // 32, bits_b // 32)
    A, B = handle.from_array(a), handle.from_array(b)
    S = handle.integers(bits_s, n)
    handle.add(S, A, B)
    assert_rows_equal(S.get(), gmp(O.GO_ADD, n, bits_s // 32, a, b), "add")


@pytest.mark.parametrize("bits_a,bits_b,bits_d", [(64, 64, 64), (2048, 2048, 2048), (2048, 1024, 2048),
                                                  (1024, 2048, 2048), (512, 512, 1024)])
def test_sub_wraps(handle, bits_a, bits_b, bits_d):
    n = 1000
    a, b = _pair(12, n, bits_a // 32, bits_b // 32)
    A, B = handle.from_array(a), handle.from_array(b)
    D = handle.integers(bits_d, n)
    handle.sub(D, A, B)
    assert_rows_equal(D.get(), gmp(O.GO_SUB, n, bits_d // 32, a, b), "sub")


def test_add_sub_roundtrip_full_size(handle):
    """Size-independent property at the benchmark shape: (a+b)-b == a (mod 2^P)."""
    n = 1 << 20
    a, b = inputs.ordered_pair(21, n, 64)
    A, B = handle.from_array(a), handle.from_array(b)
    S, D = handle.integers(2048, n), handle.integers(2048, n)
    handle.add(S, A, B)
    handle.sub(D, S, B)
    assert np.array_equal(D.get(), a)
    handle.sub(D, A, B)
    sample = np.arange(0, n, 997)
    assert_rows_equal(D.get()[sample], gmp(O.GO_SUB, sample.size, 64, a[sample], b[sample]), "sub sample")


def test_add_carry_extremes(handle):
    n = 64
    ones = np.full((n, 64), 0xFFFFFFFF, np.uint32)
    one = np.zeros((n, 64), np.uint32)
    one[:, 0] = 1
    A, B = handle.from_array(ones), handle.from_array(one)
    S = handle.integers(2080, n)
    handle.add(S, A, B)
    want = np.zeros((n, 65), np.uint32)
    want[:, 64] = 1
    assert np.array_equal(S.get(), want)
    D = handle.integers(2048, n)
    handle.sub(D, B, A)   # 1 - (2^2048-1) wraps to 2
    want = np.zeros((n, 64), np.uint32)
    want[:, 0] = 2
    assert np.array_equal(D.get(), want)


def test_add_broadcast_and_indices(handle):
    n, m = 512, 17
    a = inputs.uniform(31, n, 8)
    b = inputs.uniform(32, m, 8)
    A, B = handle.from_array(a), handle.from_array(b)
    S = handle.integers(256, n)
    handle.add(S, A, B)   # b read as i % m
    want = gmp(O.GO_ADD, n, 8, a, b)
    assert_rows_equal(S.get(), want, "broadcast")
    rng = np.random.default_rng(3)
    ia = rng.permutation(n).astype(np.uint32)
    ib = rng.integers(0, 1000, 100).astype(np.uint32)
    io = rng.permutation(n).astype(np.uint32)
    handle.set_indices(0, io)
    handle.set_indices(1, ia)
    handle.set_indices(2, ib)
    try:
        handle.add(S, A, B)
    finally:
        handle.clear_indices()   # export would scatter through slot 0 too (reference xmp.cu:586-682)
    got = S.get()
    want = np.zeros_like(got)
    for i in range(n):
        x = O.to_int(a[ia[i % n] % n]) + O.to_int(b[ib[i % 100] % m])
        want[io[i]] = O.from_int(x % (1 << 256), 8)
    assert_rows_equal(got, want, "indexed add")


def test_add_in_place(handle):
    n = 300
    a, b = _pair(41, n, 16, 16)
    A, B = handle.from_array(a), handle.from_array(b)
    handle.add(A, A, B)
    assert_rows_equal(A.get(), gmp(O.GO_ADD, n, 16, a, b), "in-place add")


@pytest.mark.parametrize("bits_a,bits_b,bits_p", [(32, 32, 64), (128, 128, 256), (512, 512, 1024),
                                                  (1024, 1024, 2048), (2048, 2048, 4096), (2048, 2048, 2048),
                                                  (2048, 1024, 3072), (96, 2048, 2144), (4096, 4096, 8192),
                                                  (1536, 1536, 3072), (1536, 512, 1600), (2048, 2048, 4160),
                                                  (2048, 1024, 2080), (4096, 2048, 4096), (8192, 8192, 16384),
                                                  (8192, 4096, 8192), (12288, 64, 12352), (16384, 16384, 32768),
                                                  (16384, 16384, 16384), (32, 2048, 2080),
                                                  (32768, 32768, 65536), (32768, 16384, 32768),
                                                  (1024, 544, 1568), (2048, 544, 2592), (1600, 1600, 3200),
                                                  (1056, 1056, 2112), (992, 992, 1984)])
def test_mul(handle, bits_a, bits_b, bits_p):
    n = 500
    a, b = _pair(51, n, bits_a // 32, bits_b // 32)
    A, B = handle.from_array(a), handle.from_array(b)
    P = handle.integers(bits_p, n)
    handle.mul(P, A, B)
    assert_rows_equal(P.get(), gmp(O.GO_MUL, n, bits_p // 32, a, b), "mul")


@pytest.mark.parametrize("bits", [1536, 2048, 4096, 8192, 16384])
def test_mul_all_ones_and_indices(handle, bits):
    """Operands of all ones (every accumulator at its bound, the longest carry
    chains) and shuffled index arrays on slots 0..2 with a broadcast multiplier."""
    n, limbs = 130, bits // 32
    ones = np.full((n, limbs), 0xFFFFFFFF, np.uint32)
    P = handle.integers(2 * bits, n)
    A = handle.from_array(ones)
    handle.mul(P, A, A)
    want = O.from_int(((1 << bits) - 1) ** 2, 2 * limbs)
    assert all(np.array_equal(r, want) for r in P.get())
    a = inputs.uniform(53 + bits, n, limbs)
    b = inputs.uniform(54 + bits, 3, limbs)
    rng = np.random.default_rng(bits)
    ia, ip = rng.permutation(n).astype(np.uint32), rng.permutation(n).astype(np.uint32)
    A, B = handle.from_array(a), handle.from_array(b)
    handle.set_indices(0, ip)
    handle.set_indices(1, ia)
    try:
        handle.mul(P, A, B)
    finally:
        handle.clear_indices()
    rows = gmp(O.GO_MUL, n, 2 * limbs, a[ia], b[np.arange(n) % 3])
    want = np.zeros_like(rows)
    want[ip] = rows
    assert_rows_equal(P.get(), want, "indexed mul")


@pytest.mark.parametrize("bits_a,bits_p", [(544, 1088), (992, 1984), (1024, 2048), (1024, 1024), (1024, 2080),
                                           (1568, 3136), (1600, 3200), (2048, 4096), (2048, 2048), (2048, 4160),
                                           (4096, 8192)])
def test_sqr(handle, bits_a, bits_p):
    """mul(a, a) takes the square kernels (each product of two different
    digits once, doubled); truncated and zero-extended outputs."""
    n = 300
    a = inputs.uniform(91 + bits_a, n, bits_a // 32)
    A = handle.from_array(a)
    P = handle.integers(bits_p, n)
    handle.mul(P, A, A)
    assert_rows_equal(P.get(), gmp(O.GO_MUL, n, bits_p // 32, a, a), "square %d -> %d bits" % (bits_a, bits_p))


def test_mul_full_size_sample(handle):
    """1,048,576 x 2048-bit products (the benchmark shape): a sample of rows
    against GMP, and a * 1 == a on every row."""
    n = 1 << 20
    a, b = inputs.ordered_pair(57, n, 64)
    A, B = handle.from_array(a), handle.from_array(b)
    P = handle.integers(4096, n)
    handle.mul(P, A, B)
    sample = np.arange(0, n, 1009)
    assert_rows_equal(P.get()[sample], gmp(O.GO_MUL, sample.size, 128, a[sample], b[sample]), "mul sample")
    # the same operands squared: a sample against GMP, and (a^2 == a*a) where b = a
    handle.mul(P, A, A)
    assert_rows_equal(P.get()[sample], gmp(O.GO_MUL, sample.size, 128, a[sample], a[sample]), "square sample")
    one = np.zeros((1, 64), np.uint32)
    one[0, 0] = 1
    handle.mul(P, A, handle.from_array(one))
    got = P.get()
    assert np.array_equal(got[:, :64], a) and not got[:, 64:].any()


@pytest.mark.parametrize("bits_a,bits_b,bits_p,square", [(1024, 1024, 2048, False), (2048, 1024, 3072, False),
                                                         (2048, 2048, 2048, False), (2048, 2048, 4160, False),
                                                         (1024, 1024, 2048, True), (2048, 2048, 2048, True),
                                                         (2048, 2048, 4096, True)])
def test_mul_large_batches(handle, bits_a, bits_b, bits_p, square):
    """400,000 products on the product-scanning kernels: every 997th row and
    the last 64 (the last, partial tile) against GMP, with truncated and
    zero-extended outputs, squares, and a shuffled output index array with a
    gathered first operand."""
    n = 400000
    a = inputs.uniform(61 + bits_a, n, bits_a // 32)
    b = a if square else inputs.uniform(62 + bits_b, n, bits_b // 32)
    A = handle.from_array(a)
    B = A if square else handle.from_array(b)
    P = handle.integers(bits_p, n)
    handle.mul(P, A, B)
    rows = np.concatenate([np.arange(0, n, 997), np.arange(n - 64, n)])
    got = P.get()
    assert_rows_equal(got[rows], gmp(O.GO_MUL, rows.size, bits_p // 32, a[rows], b[rows]), "mul, 400,000 products")
    if square:
        return
    rng = np.random.default_rng(bits_p)
    ia, ip = rng.permutation(n).astype(np.uint32), rng.permutation(n).astype(np.uint32)
    handle.set_indices(0, ip)
    handle.set_indices(1, ia)
    try:
        handle.mul(P, A, B)
    finally:
        handle.clear_indices()
    got = P.get()
    want = gmp(O.GO_MUL, rows.size, bits_p // 32, a[ia[rows]], b[rows])
    assert_rows_equal(got[ip[rows]], want, "indexed mul, 400,000 products")


@pytest.mark.parametrize("bits,count", [(8192, 4095), (8192, 4096), (16384, 4095), (16384, 4096), (6144, 2048),
                                        (12288, 2048)])
def test_mul_group_width_by_batch(handle, bits, count):
    """Above 4096 bits the lane-group width follows the batch (32 lanes below
    4,096 products, else 8 or 16): each side of each switch, and operands
    that fill only part of the group's digits, against GMP on a sample
    (all-ones rows first)."""
    limbs = bits // 32
    a = inputs.uniform(3100 + count, count, limbs)
    b = inputs.uniform(3200 + count, count, limbs)
    a[:3] = 0xFFFFFFFF
    b[:2] = 0xFFFFFFFF
    A, B = handle.from_array(a), handle.from_array(b)
    P = handle.integers(2 * bits, count)
    handle.mul(P, A, B)
    rows = np.unique(np.concatenate([np.arange(4), np.linspace(0, count - 1, 300).astype(int)]))
    assert_rows_equal(P.get()[rows], gmp(O.GO_MUL, rows.size, 2 * limbs, a[rows], b[rows]),
                      "%d bits x %d" % (bits, count))
    handle.mul(P, A, A)
    assert_rows_equal(P.get()[rows], gmp(O.GO_MUL, rows.size, 2 * limbs, a[rows], a[rows]),
                      "%d-bit squares x %d" % (bits, count))


@pytest.mark.parametrize("bits", [1024, 2048])
def test_mul_cols_big_batch_ragged_end(handle, bits):
    """Batches of 2^18 products and more (squares, and products up to 1024
    bits) run in 256-lane workgroups in XCD order: 262,244 of them end in a
    partial wave and a workgroup whose last waves are past the batch;
    products and squares against GMP on a sample that includes the last
    rows."""
    n, limbs = (1 << 18) + 100, bits // 32
    a = inputs.uniform(4100 + bits, n, limbs)
    b = inputs.uniform(4200 + bits, n, limbs)
    A, B = handle.from_array(a), handle.from_array(b)
    P = handle.integers(2 * bits, n)
    rows = np.unique(np.concatenate([np.linspace(0, n - 1, 400).astype(int), np.arange(n - 130, n)]))
    handle.mul(P, A, B)
    assert_rows_equal(P.get()[rows], gmp(O.GO_MUL, rows.size, 2 * limbs, a[rows], b[rows]), "products")
    handle.mul(P, A, A)
    assert_rows_equal(P.get()[rows], gmp(O.GO_MUL, rows.size, 2 * limbs, a[rows], a[rows]), "squares")


@pytest.mark.parametrize("bits,count,bits_p", [
    (96, 130, 192), (256, 130, 512), (512, 130, 1024), (512, 130, 704),     # 1x5, 1x10, 1x19 (truncated)
    (1536, 130, 3072), (3072, 130, 6144), (4096, 130, 4096),                # 2x28, 4x37 (truncated)
    (6144, 130, 12288), (8192, 4096, 16384), (8192, 130, 16416),            # 32x10, 8x38, 32x10 (padded)
    (12288, 130, 24576), (16384, 4096, 32768), (16384, 130, 20000)])        # 32x19, 16x38, 32x19 (truncated)
def test_sqr_lane_groups(handle, bits, count, bits_p):
    """Squares on the row kernels: the one-lane and 32-lane layouts take
    each product of two different digits once (doubled) or once in each
    order, by a tournament on the digits' places, and the other lane groups
    run full rows.  Every layout, truncated and padded outputs, all-ones rows
    (the column sums at their bound), against GMP."""
    limbs = bits // 32
    a = inputs.uniform(5300 + bits + count, count, limbs)
    a[:2] = 0xFFFFFFFF
    a[2, :] = 0
    a[2, -1] = 0xFFFFFFFF
    A = handle.from_array(a)
    P = handle.integers(bits_p, count)
    handle.mul(P, A, A)
    rows = np.unique(np.concatenate([np.arange(4), np.linspace(0, count - 1, 120).astype(int)]))
    assert_rows_equal(P.get()[rows], gmp(O.GO_MUL, rows.size, bits_p // 32, a[rows], a[rows]),
                      "%d-bit squares x %d -> %d bits" % (bits, count, bits_p))


@pytest.mark.parametrize("bits", [256, 4096, 16384])
def test_sqr_lane_groups_indexed(handle, bits):
    """xmpIntegersSqrAsync with an input gather (slot 1) and an output
    permutation (slot 0) on the row kernels' squares."""
    n, limbs = 200, bits // 32
    a = inputs.uniform(5400 + bits, 50, limbs)
    rng = np.random.default_rng(bits)
    ia = rng.integers(0, 50, n).astype(np.uint32)
    ip = rng.permutation(n).astype(np.uint32)
    A = handle.from_array(a)
    P = handle.integers(2 * bits, n)
    handle.set_indices(0, ip)
    handle.set_indices(1, ia)
    try:
        assert xmp_amd.lib().xmpIntegersSqrAsync(handle.h, P.x, A.x, n) == 0
    finally:
        handle.clear_indices()
    rows = gmp(O.GO_MUL, n, 2 * limbs, a[ia], a[ia])
    want = np.zeros_like(rows)
    want[ip] = rows
    assert_rows_equal(P.get(), want, "indexed %d-bit squares" % bits)


def test_mul_closed_form_and_square(handle):
    """(2^k - 1)^2 = 2^2k - 2^(k+1) + 1, and mul(a, a) takes the square path."""
    n = 64
    ones = np.full((n, 64), 0xFFFFFFFF, np.uint32)
    A = handle.from_array(ones)
    P = handle.integers(4096, n)
    handle.mul(P, A, A)
    k = 2048
    want = O.from_int((1 << (2 * k)) - (1 << (k + 1)) + 1, 128)
    assert all(np.array_equal(r, want) for r in P.get())
    a = inputs.uniform(52, n, 64)
    A = handle.from_array(a)
    handle.mul(P, A, A)
    assert_rows_equal(P.get(), gmp(O.GO_MUL, n, 128, a, a), "square")


@pytest.mark.parametrize("bits_a,bits_b", [(64, 32), (128, 64), (512, 256), (1024, 1024), (4096, 2048),
                                           (2048, 96), (8192, 4096), (3072, 3072), (8192, 3104),
                                           (16384, 8192)])
def test_div_mod(handle, bits_a, bits_b):
    n = 700
    a = inputs.uniform(61, n, bits_a // 32)
    b = inputs.uniform(62, n, bits_b // 32)
    b[:, -1] |= np.uint32(0x80000000) if bits_b >= 64 else np.uint32(0)
    b[b.sum(axis=1) == 0, 0] = 1
    A, B = handle.from_array(a), handle.from_array(b)
    Q, R = handle.integers(bits_a, n), handle.integers(bits_b, n)
    handle.div(Q, A, B)
    handle.mod(R, A, B)
    wq, wr = gmp(O.GO_DIV, n, bits_a // 32, a, b, out2_limbs=bits_b // 32)
    assert_rows_equal(Q.get(), wq, "div")
    assert_rows_equal(R.get(), wr, "mod")
    Q2, R2 = handle.integers(bits_a, n), handle.integers(bits_b, n)
    handle.divmod(Q2, R2, A, B)
    assert_rows_equal(Q2.get(), wq, "divmod q")
    assert_rows_equal(R2.get(), wr, "divmod r")


@pytest.mark.parametrize("nb", [64, 128, 192])
def test_div_short_and_unnormalised_divisors(handle, nb):
    """Divisors with leading zero limbs / bits, a < b, and b = 1, for the register
    kernels (64 and 128 limbs) and the workspace kernel (192)."""
    n = 800 if nb <= 64 else 256
    a = inputs.uniform(71, n, 2 * nb)
    b = inputs.uniform(72, n, nb)
    rng = np.random.default_rng(9)
    for i in range(n):
        top = int(rng.integers(1, nb + 1))
        b[i, top:] = 0
        b[i, top - 1] >>= int(rng.integers(0, 32))
        if not b[i].any():
            b[i, 0] = 1
    a[: n // 8, nb:] = 0   # some a < b
    b[n - 5:] = 0
    b[n - 5:, 0] = 1
    A, B = handle.from_array(a), handle.from_array(b)
    Q, R = handle.integers(64 * nb, n), handle.integers(32 * nb, n)
    handle.divmod(Q, R, A, B)
    wq, wr = gmp(O.GO_DIV, n, 2 * nb, a, b, out2_limbs=nb)
    assert_rows_equal(Q.get(), wq, "div q")
    assert_rows_equal(R.get(), wr, "div r")


def test_div_by_zero_defined(handle):
    n = 64
    a = inputs.uniform(81, n, 8)
    A = handle.from_array(a)
    B = handle.from_array(np.zeros((n, 4), np.uint32))
    Q, R = handle.integers(256, n), handle.integers(128, n)
    handle.divmod(Q, R, A, B)
    assert not Q.get().any()
    assert np.array_equal(R.get(), a[:, :4])


def test_div_invariant_full_size(handle):
    """Config shape: 65,536 x (4096 / 2048) -- q*b + r == a and r < b on every instance."""
    n = 65536
    a = inputs.uniform(91, n, 128)
    b = inputs.uniform(92, n, 64)
    b[:, -1] |= np.uint32(0x80000000)
    A, B = handle.from_array(a), handle.from_array(b)
    Q, R = handle.integers(4096, n), handle.integers(2048, n)
    handle.divmod(Q, R, A, B)
    q, r = Q.get(), R.get()
    QB = handle.integers(4096, n)
    handle.mul(QB, Q, B)
    S = handle.integers(4096, n)
    Rw = handle.from_array(np.concatenate([r, np.zeros((n, 64), np.uint32)], axis=1))
    handle.add(S, QB, Rw)
    assert np.array_equal(S.get(), a)
    sample = np.arange(0, n, 251)
    wq, wr = gmp(O.GO_DIV, sample.size, 128, a[sample], b[sample], out2_limbs=64)
    assert_rows_equal(q[sample], wq, "q sample")
    assert_rows_equal(r[sample], wr, "r sample")


def _addback_cases(nb, n, seed):
    """Dividends and divisors whose quotient estimate from the top limbs is one
    too large (Knuth D's add-back): V's limbs below its top two are all ones,
    U = (q + 1) * (V's top two limbs, zeros below), so U // V == q."""
    rng = np.random.default_rng(seed)
    B = 1 << 32
    a = np.zeros((n, nb + 1), np.uint32)
    b = np.zeros((n, nb), np.uint32)
    for i in range(n):
        vt = int(rng.integers(0, B)) | 0x80000000
        vt2 = int(rng.integers(0, B))
        top = vt * B ** (nb - 1) + vt2 * B ** (nb - 2)
        v = top + B ** (nb - 2) - 1
        q = int(rng.integers(1, B - 1))
        u = (q + 1) * top
        assert u // v == q
        a[i] = O.from_int(u, nb + 1)
        b[i] = O.from_int(v, nb)
    return a, b


@pytest.mark.parametrize("nb", [32, 64])
def test_div_addback_and_both_lane_layouts(handle, nb):
    """The add-back path on every instance, through the two-lane kernel
    (fewer than 131,072 divisions) and the one-lane kernel (more), and random
    divisions at both counts against GMP."""
    n = 256
    a, b = _addback_cases(nb, n, 1000 + nb)
    for reps in (1, 600):   # 256 and 153,600 divisions
        aa, bb = np.tile(a, (reps, 1)), np.tile(b, (reps, 1))
        A, B = handle.from_array(aa), handle.from_array(bb)
        Q, R = handle.integers(32 * (nb + 1), n * reps), handle.integers(32 * nb, n * reps)
        handle.divmod(Q, R, A, B)
        wq, wr = gmp(O.GO_DIV, n, nb + 1, a, b, out2_limbs=nb)
        assert_rows_equal(Q.get()[-n:], wq, "add-back q x%d" % reps)
        assert_rows_equal(R.get()[-n:], wr, "add-back r x%d" % reps)
    for cnt in (3000, 140000):
        a2 = inputs.uniform(1100 + nb, cnt, 2 * nb)
        b2 = inputs.uniform(1200 + nb, cnt, nb)
        b2[:, -1] |= np.uint32(0x80000000)
        A, B = handle.from_array(a2), handle.from_array(b2)
        Q, R = handle.integers(64 * nb, cnt), handle.integers(32 * nb, cnt)
        handle.divmod(Q, R, A, B)
        rows = np.unique(np.linspace(0, cnt - 1, 1000).astype(int))
        wq, wr = gmp(O.GO_DIV, rows.size, 2 * nb, a2[rows], b2[rows], out2_limbs=nb)
        assert_rows_equal(Q.get()[rows], wq, "q, %d divisions" % cnt)
        assert_rows_equal(R.get()[rows], wr, "r, %d divisions" % cnt)


def _later_digit_cases(nb, n, seed):
    """Dividends whose later quotient digits sit at the estimate's edges:
    (1) a window (q0 + 1) T after an exact digit, so that digit's estimate is
    one too large (add-back); (2) a next window whose top three limbs are one
    below a multiple of V's top two, so the borrows from the limbs below
    decide the estimate; (3) divisors of nearly all ones with remainders
    within a few units of V (the window's top limbs all ones).  Returns a
    (nb + 3 limbs) and b (nb limbs)."""
    rng = np.random.default_rng(seed)
    B = 1 << 32

    def r(bits):
        return int.from_bytes(rng.bytes((bits + 7) // 8), "little") % (1 << bits) if bits else 0

    a = np.zeros((n, nb + 3), np.uint32)
    b = np.zeros((n, nb), np.uint32)
    for i in range(n):
        kind = i % 3
        if kind == 0:
            vt, vt2 = r(32) | 0x80000000, r(32)
            t = vt * B ** (nb - 1) + vt2 * B ** (nb - 2)
            v = t + B ** (nb - 2) - 1
            w2 = (int(rng.integers(1, B - 1)) + 1) * t
            r1, u0 = divmod(w2, B)
        elif kind == 1:
            while True:
                v = r(32 * nb) | (1 << (32 * nb - 1))
                v2 = v >> (32 * (nb - 2))
                r1 = (int(rng.integers(2, B - 1)) * v2 - 1) * B ** (nb - 3) + r(32 * (nb - 3))
                if r1 < v:
                    break
            u0 = r(32)
        else:
            v = B ** nb - 1 - r(int(rng.choice([0, 1, 8, 32])))
            r1 = v - 1 - r(int(rng.choice([0, 1, 2, 40])))
            u0 = r(32)
        q1 = int(rng.integers(1, B))
        u = ((q1 * v + r1) * B + u0) * B + r(32)
        assert u < B ** (nb + 3)
        a[i] = O.from_int(u, nb + 3)
        b[i] = O.from_int(v, nb)
    return a, b


@pytest.mark.parametrize("nb", [32, 64, 96, 128])
def test_div_later_digits_two_lanes(handle, nb):
    """Every two-lane divisor size (16-, 32-, 48- and 64-limb halves) on the
    constructed later-digit cases, against GMP."""
    n = 384
    a, b = _later_digit_cases(nb, n, 2000 + nb)
    A, B = handle.from_array(a), handle.from_array(b)
    Q, R = handle.integers(32 * (nb + 3), n), handle.integers(32 * nb, n)
    handle.divmod(Q, R, A, B)
    wq, wr = gmp(O.GO_DIV, n, nb + 3, a, b, out2_limbs=nb)
    assert_rows_equal(Q.get(), wq, "q, %d-limb divisors" % nb)
    assert_rows_equal(R.get(), wr, "r, %d-limb divisors" % nb)
