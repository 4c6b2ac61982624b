"""GPU parity: xmpIntegersPowm against GMP mpz_powm and the C oracle.

Mirrors the reference's PowmTest (unit_tests/test.cu:1098-1270) and its
instantiations over sizes and algorithms (test.cu:2815-2910): seeded odd
moduli, bases not reduced below m, the device op through the C ABI, GMP on the
host, every limb compared.  Bit-exact.
"""
import numpy as np
import pytest

import xmp_amd
from helpers import O, assert_rows_equal, gmp, inputs

pytestmark = pytest.mark.gpu


def _powm_case(handle, seed, n, bits, exp_bits=None, reduce_base=True, count_mod=None):
    limbs = bits // 32
    el = (exp_bits or bits) // 32
    m = inputs.odd_moduli(seed, count_mod or n, limbs)
    b = inputs.below(seed + 1, m) if reduce_base else inputs.uniform(seed + 1, n, limbs)
    if b.shape[0] != n:
        b = inputs.uniform(seed + 1, n, limbs)
    e = inputs.uniform(seed + 2, n, el)
    M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
    Out = handle.integers(bits, n)
    handle.powm(Out, B, E, M)
    return Out.get(), b, e, m


# the reference's instantiation sizes, its odd ones included (test.cu:2393-2910)
@pytest.mark.parametrize("bits", [64, 96, 128, 160, 256, 288, 320, 384, 416, 512, 544, 768, 1024, 1056, 1536, 2048,
                                  2080, 3072, 4096, 6144, 8192])
def test_powm_sizes(handle, bits):
    n = 128 if bits <= 2048 else 64
    got, b, e, m = _powm_case(handle, 100 + bits, n, bits)
    want = gmp(O.GO_POWM, n, bits // 32, b, e, m)
    assert_rows_equal(got, want, "powm %d" % bits)


@pytest.mark.parametrize("bits,exp_bits,n", [(12288, 128, 6), (16384, 256, 4), (16384, 16384, 2), (32768, 64, 3)])
def test_powm_above_8192_bits(handle, bits, exp_bits, n):
    """Moduli above 8192 bits (the reference's digit-serial fallback range,
    powm_implementations.h:154-286): half-wave and whole-wave layouts."""
    got, b, e, m = _powm_case(handle, 900 + bits + exp_bits, n, bits, exp_bits=exp_bits)
    assert_rows_equal(got, gmp(O.GO_POWM, n, bits // 32, b, e, m), "powm %d-bit, %d-bit exponent" % (bits, exp_bits))


def test_powm_above_8192_bits_special_values(handle):
    """e = 0, b = 0, b = m - 1 and a small modulus in a 16384-bit container."""
    limbs, n = 512, 4
    m = inputs.odd_moduli(951, n, limbs)
    b = inputs.uniform(952, n, limbs)
    e = inputs.uniform(953, n, 4)
    e[0] = 0
    b[1] = 0
    b[2] = m[2]
    b[2, 0] -= 1
    m[3] = 0
    m[3, 0] = 0xFFFFFFFB   # 2^32 - 5, prime
    M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
    Out = handle.integers(limbs * 32, n)
    handle.powm(Out, B, E, M)
    assert_rows_equal(Out.get(), gmp(O.GO_POWM, n, limbs, b, e, m), "special values above 8192 bits")


@pytest.mark.parametrize("bits,exp_bits,n", [(34816, 96, 2), (40960, 1024, 2), (65536, 64, 2), (131072, 32, 1),
                                             (262144, 32, 1), (393184, 32, 1)])
def test_powm_wider_than_lane_layouts(handle, bits, exp_bits, n):
    """Moduli above 33,278 bits, beyond every lane layout: one workgroup per
    instance (256 threads with 8, 16 and 32 digits each, then 512 threads with
    32), R^2 mod m from a short division and Montgomery squarings; against GMP."""
    got, b, e, m = _powm_case(handle, 3100 + bits + exp_bits, n, bits, exp_bits=exp_bits, reduce_base=False)
    assert_rows_equal(got, gmp(O.GO_POWM, n, bits // 32, b, e, m), "powm %d-bit, %d-bit exponent" % (bits, exp_bits))


@pytest.mark.parametrize("bits,exp_bits,n", [(393216, 32, 4), (409600, 64, 1), (524288, 32, 1)])
def test_powm_memory_tier(handle, bits, exp_bits, n):
    """Above 393,214 bits the digits live in global memory (1024 threads, C
    radix-2^20 digits each: C = 20, 21 and 26 here); against GMP, with the
    first instance's base at m - 1 and an exponent of all ones, and at
    393,216 bits also e = 0 and an even modulus (defined as 0)."""
    limbs = bits // 32
    m = inputs.odd_moduli(3300 + bits, n, limbs)
    b = inputs.below(3301 + bits, m)
    e = inputs.uniform(3302 + bits, n, exp_bits // 32)
    b[0] = m[0]
    b[0, 0] -= 1
    e[0] = 0xFFFFFFFF
    if n > 2:
        e[2] = 0
        m[3, 0] &= 0xFFFFFFFE
    M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
    Out = handle.integers(bits, n)
    handle.powm(Out, B, E, M)
    want = gmp(O.GO_POWM, n, limbs, b, e, m)
    if n > 2:
        want[3] = 0
    assert_rows_equal(Out.get(), want, "powm %d-bit in global memory" % bits)


def test_powm_beyond_the_widest_tier(handle):
    """Above 83,886,078 bits (4,096 radix-2^20 digits per thread of 1024) the
    lazy sums would overflow: xmpIntegersPowm returns xmpErrorUnsupported
    before launching anything."""
    arrs = [handle.integers(83886080, 1) for _ in range(3)] + [handle.integers(32, 1)]
    m, b, o, e = arrs
    rc = xmp_amd.lib().xmpIntegersPowm(handle.h, o.x, b.x, e.x, m.x, 1)
    assert rc == xmp_amd.UNSUPPORTED
    for x in arrs:
        x.close()


def test_powm_wider_than_lane_layouts_special_values(handle):
    """e = 0, b = 0, b = m - 1 with e = 2, a small modulus and an even modulus
    (defined as 0) in a 34,816-bit container."""
    limbs, n = 1088, 5
    m = inputs.odd_moduli(3201, n, limbs)
    b = inputs.uniform(3202, n, limbs)
    e = inputs.uniform(3203, n, 2)
    e[0] = 0
    b[1] = 0
    b[2] = m[2]
    b[2, 0] -= 1
    e[2] = 0
    e[2, 0] = 2
    m[3] = 0
    m[3, 0] = 0xFFFFFFFB   # 2^32 - 5, prime
    m[4, 0] &= 0xFFFFFFFE  # even
    M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
    Out = handle.integers(limbs * 32, n)
    handle.powm(Out, B, E, M)
    got = Out.get()
    want = gmp(O.GO_POWM, n, limbs, b, e, m)
    want[4] = 0
    assert_rows_equal(got, want, "special values above 33,278 bits")


@pytest.mark.parametrize("bits", [1024, 2048])
def test_powm_matches_c_oracle(handle, bits):
    n = 16
    got, b, e, m = _powm_case(handle, 7 + bits, n, bits)
    want = O.powm_batch(b, e, m, n)
    assert_rows_equal(got, want, "powm vs restated reference algorithm")


@pytest.mark.parametrize("alg", [xmp_amd.ALG_REGMP, xmp_amd.ALG_DISTRIBUTEDMP])
@pytest.mark.parametrize("bits", [512, 1024, 2048])
def test_powm_forced_layouts(handle, alg, bits):
    n = 96
    handle.set_algorithm(alg)
    try:
        got, b, e, m = _powm_case(handle, 300 + bits, n, bits)
    except xmp_amd.XmpError as err:
        assert err.code == xmp_amd.UNSUPPORTED
        return
    finally:
        handle.set_algorithm(xmp_amd.ALG_DEFAULT)
    assert_rows_equal(got, gmp(O.GO_POWM, n, bits // 32, b, e, m), "forced layout")


@pytest.mark.parametrize("bits", [512, 1024, 1056, 2080, 6144])
def test_powm_digitmp_policy(handle, bits):
    """The reference's DigitMP instantiations (test.cu:2890-2910): the policy
    takes every layout the modulus fits, and the results are GMP's."""
    n = 64 if bits <= 2080 else 16
    handle.set_algorithm(xmp_amd.ALG_DIGITMP)
    try:
        got, b, e, m = _powm_case(handle, 500 + bits, n, bits)
    finally:
        handle.set_algorithm(xmp_amd.ALG_DEFAULT)
    assert_rows_equal(got, gmp(O.GO_POWM, n, bits // 32, b, e, m), "DigitMP policy %d" % bits)


def test_cost_model_avoids_spilling_layout(handle):
    """Without a layout table the cost model picks the layout: for 39,514
    3072-bit instances it must not pick 2x55, which spills (over a kilobyte
    per lane) and ran 4x slower than the model's choice."""
    import os
    from xmp_amd.ext import Montgomery, tuner_load
    with open(os.path.join(os.path.dirname(xmp_amd.LIB_PATH), "tune_gfx950.json")) as f:
        table = f.read()
    tuner_load('{"device": "gfx950", "entries": []}')
    try:
        M = handle.from_array(inputs.odd_moduli(78, 64, 96))
        ctx = Montgomery(handle, M, 3072, 39514)
        lanes, digits, radix = ctx.layout()[:3]
        ctx.close()
        M.close()
    finally:
        tuner_load(table)
    assert (lanes, digits) != (2, 55)
    assert lanes * digits * radix >= 3072 + 2


def test_powm_unreduced_base_and_short_exponent(handle):
    n = 200
    got, b, e, m = _powm_case(handle, 401, n, 2048, exp_bits=32, reduce_base=False)
    assert_rows_equal(got, gmp(O.GO_POWM, n, 64, b, e, m), "base >= m, 32-bit exponent")
    got, b, e, m = _powm_case(handle, 402, n, 1024, exp_bits=2048, reduce_base=False)
    assert_rows_equal(got, gmp(O.GO_POWM, n, 32, b, e, m), "exponent longer than modulus")


def test_powm_special_values(handle):
    """e = 0 -> 1 (0 for m = 1); b = 0 -> 0; e = 1 -> b mod m; b = m - 1; m = 1; small m in wide precision."""
    limbs = 32
    n = 12
    m = inputs.odd_moduli(501, n, limbs)
    b = inputs.uniform(502, n, limbs)
    e = inputs.uniform(503, n, limbs)
    e[0] = 0
    b[1] = 0
    e[2] = 0
    e[2, 0] = 1
    b[3] = m[3]
    b[3, 0] -= 1
    m[4] = 0
    m[4, 0] = 1               # m = 1
    e[4] = 0                  # 0^0 mod 1 = 0
    m[5] = 0
    m[5, 0] = 0x10001         # small modulus in a 1024-bit container
    m[6, 20:] = 0
    m[6, 0] |= 1              # modulus with leading zero limbs
    b[7] = m[7]               # b == m
    e[8] = 0xFFFFFFFF         # all-ones exponent
    M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
    Out = handle.integers(1024, n)
    handle.powm(Out, B, E, M)
    assert_rows_equal(Out.get(), gmp(O.GO_POWM, n, limbs, b, e, m), "special values")


def test_powm_fermat_and_crt_sample_key(handle):
    """Toy RSA key of the reference samples (samples/crt_rsa/rsa.cc:167-178):
    n = 17460671 = 4931 * 3541, e = 65537, d = 16156673; m^(e*d) == m, and a
    prime p gives a^(p-1) == 1 (Fermat)."""
    msgs = np.array([[12345], [2], [17460670], [99999]], np.uint32)
    N = handle.from_array(np.full((4, 1), 17460671, np.uint32))
    E = handle.from_array(np.full((4, 1), 65537, np.uint32))
    D = handle.from_array(np.full((4, 1), 16156673, np.uint32))
    M = handle.from_array(msgs)
    C, P = handle.integers(32, 4), handle.integers(32, 4)
    handle.powm(C, M, E, N)
    handle.powm(P, C, D, N)
    assert np.array_equal(P.get(), msgs)
    p = (1 << 127) - 1   # Mersenne prime
    a = inputs.uniform(601, 32, 4)
    pa = np.tile(O.from_int(p, 4), (32, 1))
    pm1 = np.tile(O.from_int(p - 1, 4), (32, 1))
    Out = handle.integers(128, 32)
    handle.powm(Out, handle.from_array(a), handle.from_array(pm1), handle.from_array(pa))
    got = Out.get()
    for i in range(32):
        want = 0 if O.to_int(a[i]) % p == 0 else 1
        assert O.to_int(got[i]) == want


@pytest.mark.parametrize("nb,ne,nm", [(1, 1, 1), (None, 1, 1), (None, 7, 7), (1, 7, None), (None, None, 3),
                                      (2, None, None)])
def test_powm_broadcast_patterns(handle, nb, ne, nm):
    """The reference's instantiation patterns (test.cu:2393-2910): inputs with
    fewer instances than the output are read at i % count ((N,N,1,1),
    (N,N,M,M), (N,1,M,N), ...), and longer inputs are read from the start."""
    n, bits = 300, 1024
    limbs = bits // 32
    cb, ce, cm = (c if c is not None else 2 * n for c in (nb, ne, nm))
    m = inputs.odd_moduli(1300 + cm, cm, limbs)
    b = inputs.uniform(1301 + cb, cb, limbs)
    e = inputs.uniform(1302 + ce, ce, limbs)
    M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
    Out = handle.integers(bits, n)
    handle.powm(Out, B, E, M, n)
    i = np.arange(n)
    want = gmp(O.GO_POWM, n, limbs, b[i % cb], e[i % ce], m[i % cm])
    assert_rows_equal(Out.get(), want, "powm with base/exp/mod counts %d/%d/%d" % (cb, ce, cm))


@pytest.mark.parametrize("bits,n", [(2048, 64), (4096, 64), (6144, 32), (8192, 32), (16384, 8), (32768, 4)])
def test_montgomery_r2_canonical(handle, bits, n):
    """The resident R^2 mod m of a context is canonical (< m) and exact: one
    division up to 4096-bit moduli, Montgomery squarings of 2^t R mod m above.
    Checked against Python's pow(2, 2 rbits, m) on every modulus."""
    from xmp_amd.ext import Montgomery
    m = inputs.odd_moduli(4100 + bits, n, bits // 32)
    m[0] = 0
    m[0, 0] = 1 + 2          # a tiny modulus in a wide container
    m[1, -1] = 0x80000000    # top limb with only its top bit
    m[2, :] = 0xFFFFFFFF     # 2^bits - 1
    M = handle.from_array(m)
    ctx = Montgomery(handle, M, bits, n)
    try:
        lanes, digits, radix, _ = ctx.layout()
        # R spans the lanes the modulus needs: the fewest whole lanes of
        # digits*radix bits that reach bits + 2
        need = -(-(bits + 2) // radix)
        rbits = digits * radix * min(lanes, -(-need // digits))
        got = ctx.r2()
    finally:
        ctx.close()
    want = np.stack([O.from_int(pow(2, 2 * rbits, O.to_int(row)), bits // 32) for row in m])
    assert_rows_equal(got, want, "R^2 mod m, %d-bit moduli, R = 2^%d" % (bits, rbits))


def test_montgomery_context_indices_and_in_place(fresh_handle):
    """xmpIntegersPowmMontgomeryAsync with policy slots (0 scatter, 1/2 base and
    exponent gather, 3 the modulus inside the context) and in place (out ==
    base), against GMP; the same index semantics as xmpIntegersPowm."""
    from xmp_amd import hip
    from xmp_amd.ext import Montgomery
    h = fresh_handle
    n, nm, limbs = 256, 16, 64
    m = inputs.odd_moduli(1701, nm, limbs)
    b = inputs.uniform(1702, n, limbs)
    e = inputs.uniform(1703, 8, limbs)
    rng = np.random.default_rng(17)
    im = rng.integers(0, 10 * nm, 48).astype(np.uint32)
    ie = rng.integers(0, 8, 24).astype(np.uint32)
    io = rng.permutation(n).astype(np.uint32)
    M, B, E = h.from_array(m), h.from_array(b), h.from_array(e)
    Out = h.integers(2048, n)
    ctx = Montgomery(h, M, 2048, n)
    try:
        h.set_indices(0, io)
        h.set_indices(2, ie)
        h.set_indices(3, im)
        try:
            ctx.powm_async(Out, B, E, n)
            hip.device_sync()
        finally:
            h.clear_indices()
        rows = gmp(O.GO_POWM, n, limbs, b, e[ie[np.arange(n) % 24] % 8], m[im[np.arange(n) % 48] % nm])
        want = np.zeros_like(rows)
        want[io] = rows
        assert_rows_equal(Out.get(), want, "indexed powm through a Montgomery context")
        # in place: B <- B^E mod M[i % nm]
        ctx.powm_async(B, B, E, n)
        hip.device_sync()
        want = gmp(O.GO_POWM, n, limbs, b, e[np.arange(n) % 8], m[np.arange(n) % nm])
        assert_rows_equal(B.get(), want, "in-place powm through a Montgomery context")
    finally:
        ctx.close()
        for x in (M, B, E, Out):
            x.close()


def test_powm_indices_and_resident_moduli(handle):
    """Policy index gather (slots 1..3) and scatter (slot 0), as the reference
    powm kernels read them (powm_implementations.h:417-418, copy-out :480-516)."""
    n, nm = 256, 16
    m = inputs.odd_moduli(701, nm, 64)
    b = inputs.uniform(702, n, 64)
    e = inputs.uniform(703, 8, 64)
    rng = np.random.default_rng(5)
    im = rng.integers(0, 10 * nm, 64).astype(np.uint32)
    ie = rng.integers(0, 8, 32).astype(np.uint32)
    io = rng.permutation(n).astype(np.uint32)
    M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
    Out = handle.integers(2048, n)
    handle.set_indices(0, io)
    handle.set_indices(2, ie)
    handle.set_indices(3, im)
    try:
        handle.powm(Out, B, E, M)
    finally:
        handle.clear_indices()   # export would scatter through slot 0 too (reference xmp.cu:586-682)
    got = Out.get()
    bb = b
    ee = e[ie[np.arange(n) % 32] % 8]
    mm = m[im[np.arange(n) % 64] % nm]
    want_rows = gmp(O.GO_POWM, n, 64, bb, ee, mm)
    want = np.zeros_like(want_rows)
    want[io] = want_rows
    assert_rows_equal(got, want, "indexed powm")


def test_powm_in_place(handle):
    n = 100
    m = inputs.odd_moduli(801, n, 32)
    b = inputs.below(802, m)
    e = inputs.uniform(803, n, 32)
    M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
    handle.powm(B, B, E, M)
    assert_rows_equal(B.get(), gmp(O.GO_POWM, n, 32, b, e, m), "in-place powm")


def test_powm_config1_full(handle):
    """4096 x 1024-bit (the reference's CPU-runnable configuration) against GMP."""
    n = 4096
    got, b, e, m = _powm_case(handle, 1001, n, 1024)
    assert_rows_equal(got, gmp(O.GO_POWM, n, 32, b, e, m), "config 1")


def test_powm_2048_full_size_every_row(handle):
    """The headline shape, 65,536 x 2048-bit, every row against GMP (about
    7 s of GMP on 16 host threads); the bench checks only a sample."""
    n = 65536
    got, b, e, m = _powm_case(handle, 2048, n, 2048)
    assert_rows_equal(got, gmp(O.GO_POWM, n, 64, b, e, m), "65,536 x 2048-bit powm")
