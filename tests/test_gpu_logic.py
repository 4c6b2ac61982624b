"""GPU parity: ior, and, xor, not, shf, popc and cmp against GMP and the C oracle.

Mirrors the reference's tests of src/logic.cu (unit_tests/test.cu):
genericTwoInOneOutTest over mpz_and/mpz_ior/mpz_xor (:2516-2595), ShfTest
(:1272-1416, positive shifts are mpz_mul_2exp, negative mpz_tdiv_q_2exp, with
shifted index arrays and an in-place pass), NotTest (:1418-1520, ~a and all
ones above a's length), CmpTest (:2001-2114, sgn of mpz_cmp) and PopcTest
(:2116-2230, mpz_popcount).  Bit-exact.
"""
import numpy as np
import pytest

import xmp_amd
from xmp_amd import hip
from helpers import O, assert_rows_equal, gmp, inputs

pytestmark = pytest.mark.gpu

TWO_IN = [(32, 32, 32), (64, 64, 64), (128, 128, 128), (256, 256, 256), (512, 512, 512),
          (64, 64, 32), (128, 128, 64), (256, 256, 128), (512, 512, 256), (160, 160, 160), (384, 96, 384),
          (2048, 2048, 2048), (4096, 2048, 1024)]
OPS = [("ior", O.GO_IOR), ("and_", O.GO_AND), ("xor", O.GO_XOR)]


@pytest.mark.parametrize("name,gop", OPS)
@pytest.mark.parametrize("bits_c,bits_a,bits_b", TWO_IN)
def test_two_in_one_out(handle, name, gop, bits_c, bits_a, bits_b):
    n = 777   # not a multiple of 4: the tail instances take the scalar path
    a = inputs.uniform(41 + bits_a, n, bits_a // 32)
    b = inputs.uniform(42 + bits_b, n, bits_b // 32)
    A, B = handle.from_array(a), handle.from_array(b)
    C = handle.integers(bits_c, n)
    getattr(handle, name)(C, A, B)
    assert_rows_equal(C.get(), gmp(gop, n, bits_c // 32, a, b), name)
    # operands swapped (the shorter one first)
    getattr(handle, name)(C, B, A)
    assert_rows_equal(C.get(), gmp(gop, n, bits_c // 32, b, a), name + " swapped")


@pytest.mark.parametrize("name,gop", OPS)
def test_two_in_indices_broadcast_in_place(handle, name, gop):
    """Shuffled index arrays on slots 0..2 (test.cu TwoInOneOut with indices),
    a broadcast input, and the in-place form with an output index array."""
    n, na, nb = 512, 512, 37
    a = inputs.uniform(51, na, 16)
    b = inputs.uniform(52, nb, 16)
    rng = np.random.default_rng(9)
    ia, ib, ic = rng.permutation(na).astype(np.uint32), rng.permutation(nb).astype(np.uint32), \
        rng.permutation(n).astype(np.uint32)
    A, B = handle.from_array(a), handle.from_array(b)
    C = handle.integers(512, n)
    handle.set_indices(0, ic)
    handle.set_indices(1, ia)
    handle.set_indices(2, ib)
    try:
        getattr(handle, name)(C, A, B)
    finally:
        handle.clear_indices()
    rows = gmp(gop, n, 16, a[ia[np.arange(n) % na] % na], b[ib[np.arange(n) % nb] % nb])
    want = np.zeros_like(rows)
    want[ic] = rows
    assert_rows_equal(C.get(), want, name + " indexed")
    # in place: a = a op b, only the instances named by slot 0 written
    sel = rng.permutation(na)[:300].astype(np.uint32)
    handle.set_indices(0, sel)
    try:
        getattr(handle, name)(A, A, B, 300)
    finally:
        handle.clear_indices()
    want = a.copy()
    want[sel] = gmp(gop, 300, 16, a[np.arange(300) % na], b[np.arange(300) % nb])
    assert_rows_equal(A.get(), want, name + " in place")


@pytest.mark.parametrize("bits_c,bits_a", [(32, 32), (64, 64), (512, 512), (64, 32), (512, 256), (2048, 2048),
                                           (2080, 2048)])
def test_not(handle, bits_c, bits_a):
    n = 1001
    a = inputs.uniform(61 + bits_a, n, bits_a // 32)
    A = handle.from_array(a)
    C = handle.integers(bits_c, n)
    handle.not_(C, A)
    want = np.full((n, bits_c // 32), 0xFFFFFFFF, np.uint32)
    want[:, :bits_a // 32] = ~a
    assert_rows_equal(C.get(), want, "not")
    for i in range(0, n, 97):
        assert np.array_equal(O.not_(a[i], bits_c // 32), want[i])
    handle.not_(A, A)   # in place
    assert_rows_equal(A.get(), ~a, "not in place")


@pytest.mark.parametrize("bits_c,bits_a", [(32, 32), (64, 64), (128, 128), (256, 256), (512, 512), (64, 32),
                                           (512, 256), (160, 160), (352, 352), (2048, 2048), (1024, 2048)])
def test_shf(handle, bits_c, bits_a):
    """Shifts drawn as in test.cu:1305 (rand % (2aP-2) - aP + 1), plus the
    whole-limb and out-of-range cases; shuffled indices on slots 0 and 1."""
    n = 640
    la, lc = bits_a // 32, bits_c // 32
    a = inputs.uniform(71 + bits_a, n, la)
    rng = np.random.default_rng(bits_a + bits_c)
    shifts = rng.integers(-bits_a + 1, bits_a, n).astype(np.int32)
    shifts[:10] = [0, 1, -1, 32, -32, 31, -31, bits_a, -bits_a, 3 * bits_c]
    A = handle.from_array(a)
    C = handle.integers(bits_c, n)
    handle.shf(C, A, shifts)
    want = gmp(O.GO_SHF, n, lc, a, shifts.view(np.uint32).reshape(-1, 1))
    assert_rows_equal(C.get(), want, "shf")
    for i in range(0, n, 53):
        assert np.array_equal(O.shf(a[i], int(shifts[i]), lc), want[i])
    # indices on slots 0 and 1, and in place (test.cu:1384-1404)
    ia, ic = rng.permutation(n).astype(np.uint32), rng.permutation(n).astype(np.uint32)
    handle.set_indices(0, ic)
    handle.set_indices(1, ia)
    try:
        handle.shf(C, A, shifts)
    finally:
        handle.clear_indices()
    rows = gmp(O.GO_SHF, n, lc, a[ia], shifts.view(np.uint32).reshape(-1, 1))
    want = np.zeros_like(rows)
    want[ic] = rows
    assert_rows_equal(C.get(), want, "shf indexed")
    if bits_c == bits_a:
        handle.set_indices(0, ic)
        handle.set_indices(1, ia)
        try:
            handle.shf(A, A, shifts)
        finally:
            handle.clear_indices()
        assert_rows_equal(A.get(), want, "shf in place")


def test_shf_broadcast_shift_and_device_pointer(handle):
    """shift_count < count repeats the shifts (shift[i % shift_count]); the
    shift array may also live in device memory; shift_count 0 is rejected."""
    n = 300
    a = inputs.uniform(81, n, 8)
    A = handle.from_array(a)
    C = handle.integers(256, n)
    shifts = np.array([5, -77, 64, -1000, 2 ** 31 - 1, -2 ** 31], np.int32)
    handle.shf(C, A, shifts)
    rep = shifts[np.arange(n) % shifts.size]
    want = np.stack([O.shf(a[i], int(rep[i]), 8) for i in range(n)])
    assert_rows_equal(C.get(), want, "shf repeated shifts")
    d = hip.malloc(shifts.nbytes)
    try:
        hip.memcpy_htod(d, shifts)
        C2 = handle.integers(256, n)
        handle.shf_device(C2, A, d, shifts.size)
        assert_rows_equal(C2.get(), want, "shf device shifts")
    finally:
        hip.free(d)
    with pytest.raises(xmp_amd.XmpError) as err:
        handle.shf(C, A, np.zeros(0, np.int32))
    assert err.value.code == xmp_amd.INVALID_PARAMETER


@pytest.mark.parametrize("bits_a,bits_b", [(32, 32), (64, 64), (512, 512), (512, 256), (256, 512), (2048, 2048),
                                           (4096, 1024)])
def test_cmp(handle, bits_a, bits_b):
    n = 1500
    a = inputs.uniform(91 + bits_a, n, bits_a // 32)
    b = inputs.uniform(92 + bits_b, n, bits_b // 32)
    k = min(bits_a, bits_b) // 32
    # rows 0..49 equal values (the longer operand's extra limbs zero), rows 50..99 differ in limb 0 only
    for rows, flip in ((slice(0, 50), False), (slice(50, 100), True)):
        a[rows] = 0
        a[rows, :k] = b[rows, :k]
        b[rows, k:] = 0
        if flip:
            a[rows, 0] ^= 1
    A, B = handle.from_array(a), handle.from_array(b)
    got = handle.cmp(A, B, n)
    want = gmp(O.GO_CMP, n, 1, a, b).view(np.int32)[:, 0]
    assert np.array_equal(got, want)
    assert (got[:50] == 0).all()
    for i in range(0, n, 101):
        assert O.cmp(a[i], b[i]) == got[i]


def test_cmp_indices_and_broadcast(handle):
    n, na, nb = 700, 700, 3
    a = inputs.uniform(95, na, 8)
    b = inputs.uniform(96, nb, 8)
    b[0] = a[5]
    rng = np.random.default_rng(4)
    ia = rng.permutation(na).astype(np.uint32)
    A, B = handle.from_array(a), handle.from_array(b)
    handle.set_indices(0, ia)
    try:
        got = handle.cmp(A, B, n)
    finally:
        handle.clear_indices()
    want = gmp(O.GO_CMP, n, 1, a[ia], b[np.arange(n) % nb]).view(np.int32)[:, 0]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bits", [32, 64, 512, 2048, 8192])
def test_popc(handle, bits):
    n = 1200
    a = inputs.uniform(97 + bits, n, bits // 32)
    a[0] = 0
    a[1] = 0xFFFFFFFF
    A = handle.from_array(a)
    got = handle.popc(A)
    want = gmp(O.GO_POPC, n, 1, a)[:, 0]
    assert np.array_equal(got, want)
    assert got[0] == 0 and got[1] == bits
    rng = np.random.default_rng(bits)
    ia = rng.permutation(n).astype(np.uint32)
    handle.set_indices(0, ia)
    try:
        got = handle.popc(A)
    finally:
        handle.clear_indices()
    assert np.array_equal(got, want[ia])


def test_bitwise_full_size_properties(handle):
    """At 1,048,576 x 2048 bits: (a ^ b) ^ b == a, a & ~a == 0, a | ~a == all ones,
    (a << s) >> s == a when nothing leaves the top, popc(a) + popc(~a) == 2048."""
    n = 1 << 20
    a, b = inputs.ordered_pair(101, n, 64)
    A, B = handle.from_array(a), handle.from_array(b)
    X, Y = handle.integers(2048, n), handle.integers(2048, n)
    handle.xor(X, A, B)
    handle.xor(Y, X, B)
    assert np.array_equal(Y.get(), a)
    handle.not_(X, A)
    handle.and_(Y, A, X)
    assert not Y.get().any()
    handle.ior(Y, A, X)
    assert (Y.get() == 0xFFFFFFFF).all()
    assert (handle.popc(A) + handle.popc(X) == 2048).all()
    W = handle.integers(4096, n)
    shifts = (np.arange(n) % 2048).astype(np.int32)
    handle.shf(W, A, shifts)
    handle.shf(Y, W, -shifts)
    assert np.array_equal(Y.get(), a)


def test_bitwise_errors(handle):
    a = inputs.uniform(111, 8, 4)
    A = handle.from_array(a)
    C = handle.integers(64, 8)
    with pytest.raises(xmp_amd.XmpError) as err:
        handle.ior(C, A, A)   # output narrower than the inputs
    assert err.value.code == xmp_amd.INVALID_PRECISION
    with pytest.raises(xmp_amd.XmpError) as err:
        handle.not_(C, A)
    assert err.value.code == xmp_amd.INVALID_PRECISION
    C = handle.integers(128, 4)
    with pytest.raises(xmp_amd.XmpError) as err:
        handle.xor(C, A, A, 8)   # output holds fewer instances than count
    assert err.value.code == xmp_amd.INVALID_COUNT
    with pytest.raises(xmp_amd.XmpError) as err:
        handle.popc(A, 9)        # popc checks the input count (logic.cu:563-564)
    assert err.value.code == xmp_amd.INVALID_COUNT
