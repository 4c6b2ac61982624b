"""CPU: pin the oracle before trusting it.

Every oracle function is checked against (a) the committed golden fixtures
(computed with Python's int and the reference's known-answer cases), (b) GMP
mpz -- the reference test suite's own oracle -- on seeded inputs, and (c)
closed forms.  The last tests plant plausible mistakes (a dropped carry, a
swapped operand, a wrong window offset) in copies of the oracle's formulas and
check that the golden fixtures catch them.
"""
import json
import os
import random

import numpy as np
import pytest

from helpers import O, gmp, inputs

GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))


def H(s):
    return int(s, 16)


def L(v, n):
    return O.from_int(v, n)


# ---------------------------------------------------------------- golden
def test_golden_add_sub_mul():
    for c in GOLDEN["addsubmul"]:
        a, b = L(H(c["a"]), c["an"]), L(H(c["b"]), c["bn"])
        assert O.to_int(O.add(a, b, c["on"])) == H(c["add"])
        assert O.to_int(O.sub(a, b, c["on"])) == H(c["sub"])
        assert O.to_int(O.mul(a, b, c["on"])) == H(c["mul"])


def test_golden_divmod():
    for c in GOLDEN["divmod"]:
        a, b = L(H(c["a"]), c["an"]), L(H(c["b"]), c["bn"])
        q, r, rc = O.divmod_(a, b, c["an"], c["bn"])
        assert rc == 0
        assert O.to_int(q) == H(c["q"]) and O.to_int(r) == H(c["r"])


def test_golden_powm():
    for c in GOLDEN["powm"]:
        out, rc = O.powm(L(H(c["b"]), c["n"]), L(H(c["e"]), c["en"]), L(H(c["m"]), c["n"]))
        assert rc == 0
        assert O.to_int(out) == H(c["out"]), c
        assert O.window_bits(32 * c["en"]) == c["window_bits"]


def test_golden_np0():
    for c in GOLDEN["np0"]:
        assert O.np0(H(c["m0"])) == H(c["np0"])
        assert (O.np0(H(c["m0"])) * H(c["m0"])) % (1 << 32) == (1 << 32) - 1   # np0*m0 == -1 mod 2^32


def test_golden_rsa_sample_key():
    s = GOLDEN["rsa_sample"]
    n = np.array([s["n"]], np.uint32)
    for msg, ct in zip(s["messages"], s["ciphertexts"]):
        c, _ = O.powm(np.array([msg], np.uint32), np.array([s["e"]], np.uint32), n)
        assert int(c[0]) == ct
        back, _ = O.powm(c, np.array([s["d"]], np.uint32), n)
        assert int(back[0]) == msg
        mp, _ = O.powm(c, np.array([s["dp"]], np.uint32), np.array([s["p"]], np.uint32))
        mq, _ = O.powm(c, np.array([s["dq"]], np.uint32), np.array([s["q"]], np.uint32))
        assert (int(mp[0]) * s["cp"] + int(mq[0]) * s["cq"]) % s["n"] == msg   # CRT (samples/crt_rsa/rsa.cc:205-219)


def test_golden_import_transform():
    for c in GOLDEN["import"]:
        data = np.frombuffer(bytes.fromhex(c["data"]), np.uint8)
        out = O.transform(data, c["count"], c["words"], c["order"], c["endian"], c["nails"], c["size"])
        got = np.frombuffer(out.tobytes(), np.uint32).reshape(c["count"], -1)
        assert got.tolist() == c["limbs"], c


def test_order_known_answer():
    """1+3 = 4 and 2+4 = 6 with most-significant-word-first order (unit_tests/test.cu:683-695)."""
    k = GOLDEN["order_add"]
    a = O.transform(np.array(k["a_msw_first"], np.uint32).view(np.uint8), 2, 2, 1, -1, 0, 4)
    b = O.transform(np.array(k["b_msw_first"], np.uint32).view(np.uint8), 2, 2, 1, -1, 0, 4)
    a = np.frombuffer(a.tobytes(), np.uint32).reshape(2, 2)
    b = np.frombuffer(b.tobytes(), np.uint32).reshape(2, 2)
    s = np.array([O.add(a[i], b[i], 2) for i in range(2)])
    back = O.transform(s.view(np.uint8).reshape(-1), 2, 2, 1, -1, 0, 4)
    assert np.frombuffer(back.tobytes(), np.uint32).tolist() == k["sum_msw_first"]


# ---------------------------------------------------------------- GMP and Python int
def test_oracle_vs_gmp_and_int():
    if not O.gmp_available():
        pytest.skip("libgmp.so.10 not present")
    r = random.Random(7)
    n = 40
    for limbs in (1, 3, 8, 32, 64):
        a = inputs.uniform(100 + limbs, n, limbs)
        b = inputs.uniform(200 + limbs, n, limbs)
        b[:, -1] |= np.uint32(1)
        for i in range(n):
            assert np.array_equal(O.add(a[i], b[i], limbs + 1), gmp(O.GO_ADD, 1, limbs + 1, a[i:i + 1], b[i:i + 1])[0])
            assert np.array_equal(O.sub(a[i], b[i], limbs), gmp(O.GO_SUB, 1, limbs, a[i:i + 1], b[i:i + 1])[0])
            assert np.array_equal(O.mul(a[i], b[i], 2 * limbs), gmp(O.GO_MUL, 1, 2 * limbs, a[i:i + 1], b[i:i + 1])[0])
        wide = inputs.uniform(300 + limbs, n, 2 * limbs)
        wq, wr = gmp(O.GO_DIV, n, 2 * limbs, wide, b, out2_limbs=limbs)
        for i in range(n):
            q, rr, _ = O.divmod_(wide[i], b[i], 2 * limbs, limbs)
            assert np.array_equal(q, wq[i]) and np.array_equal(rr, wr[i])
        m = inputs.odd_moduli(400 + limbs, 8, limbs)
        e = inputs.uniform(500 + limbs, 8, limbs)
        want = gmp(O.GO_POWM, 8, limbs, a[:8], e, m)
        got = O.powm_batch(a[:8], e, m, 8)
        assert np.array_equal(got, want)
        for i in range(8):
            assert O.to_int(got[i]) == pow(O.to_int(a[i]), O.to_int(e[i]), O.to_int(m[i]))
    # the GMP wrapper itself against Python int
    x = r.getrandbits(2048)
    y = r.getrandbits(1024) | 1
    g = gmp(O.GO_MUL, 1, 96, O.from_int(x, 64).reshape(1, -1), O.from_int(y, 32).reshape(1, -1))[0]
    assert O.to_int(g) == x * y


def test_closed_forms():
    p = (1 << 521) - 1                     # Mersenne prime M521
    nl = 17
    for a in (2, 3, 123456789):
        out, _ = O.powm(L(a, nl), L(p - 1, nl), L(p, nl))   # Fermat
        assert O.to_int(out) == 1
    k = 2048
    ones = L((1 << k) - 1, 64)
    assert O.to_int(O.mul(ones, ones, 128)) == (1 << (2 * k)) - (1 << (k + 1)) + 1
    r = random.Random(3)
    for _ in range(50):
        a, b = r.getrandbits(4096), r.getrandbits(r.randrange(1, 2048)) | 1
        q, rr, _ = O.divmod_(L(a, 128), L(b, 64), 128, 64)
        assert O.to_int(q) * b + O.to_int(rr) == a and O.to_int(rr) < b
    q, rr, rc = O.divmod_(L(12345, 2), L(0, 2), 2, 2)   # division by zero: defined q = 0, r = a
    assert rc == -1 and O.to_int(q) == 0 and O.to_int(rr) == 12345
    out, rc = O.powm(L(3, 2), L(5, 2), L(10, 2))        # even modulus is outside the contract
    assert rc == -1


# ---------------------------------------------------------------- planted mistakes
def _golden_catches(fn):
    """True when some golden add/sub/mul case disagrees with fn (a mutated oracle)."""
    for c in GOLDEN["addsubmul"]:
        a, b = H(c["a"]), H(c["b"])
        if fn(a, b, c) != H(c[fn.field]):
            return True
    return False


def test_planted_mistakes_are_caught():
    def add_no_carry(a, b, c):
        n = c["on"]
        return sum(((((a >> (32 * i)) & 0xFFFFFFFF) + ((b >> (32 * i)) & 0xFFFFFFFF)) & 0xFFFFFFFF) << (32 * i)
                   for i in range(n))
    add_no_carry.field = "add"

    def sub_swapped(a, b, c):
        return (b - a) % (1 << (32 * c["on"]))
    sub_swapped.field = "sub"

    def mul_untruncated_wrong_width(a, b, c):
        return (a * b) % (1 << (32 * (c["on"] - 1))) if c["on"] > 1 else (a * b + 1) % (1 << 32)
    mul_untruncated_wrong_width.field = "mul"

    for f in (add_no_carry, sub_swapped, mul_untruncated_wrong_width):
        assert _golden_catches(f), f.__name__

    # powm with the top window taken from the wrong offset (a plausible indexing slip)
    def powm_wrong_window(b, e, m, en):
        bits = 32 * en
        w = O.window_bits(bits)
        acc = 1
        pos = bits - w   # ignores bits % w
        while pos >= 0:
            acc = pow(acc, 1 << w, m) * pow(b, (e >> pos) & ((1 << w) - 1), m) % m
            pos -= w
        return acc
    caught = any(powm_wrong_window(H(c["b"]), H(c["e"]), H(c["m"]), c["en"]) != H(c["out"])
                 for c in GOLDEN["powm"] if H(c["m"]) > 1)
    assert caught


# ---------------------------------------------------------------- bitwise ops
def test_golden_bitwise():
    for c in GOLDEN["bitwise"]:
        a, b, on = L(H(c["a"]), c["an"]), L(H(c["b"]), c["bn"]), c["on"]
        assert O.to_int(O.ior(a, b, on)) == H(c["ior"])
        assert O.to_int(O.and_(a, b, on)) == H(c["and"])
        assert O.to_int(O.xor(a, b, on)) == H(c["xor"])
        assert O.to_int(O.not_(a, on)) == H(c["not"])
        for sh, v in c["shf"]:
            assert O.to_int(O.shf(a, sh, on)) == H(v), (c, sh)
        assert O.popc(a) == c["popc"]
        assert O.cmp(a, b) == c["cmp"] and O.cmp(b, a) == -c["cmp"]


def test_bitwise_oracle_vs_gmp_and_int():
    if not O.gmp_available():
        pytest.skip("libgmp.so.10 not present")
    r = random.Random(11)
    n = 48
    for an, bn, on in ((1, 1, 1), (3, 2, 4), (8, 8, 8), (16, 5, 16), (64, 64, 64)):
        a = inputs.uniform(600 + an, n, an)
        b = inputs.uniform(700 + bn, n, bn)
        a[0, :] = b[0, :min(an, bn)].tolist() + [0] * max(0, an - bn)   # an equal pair
        for op, fn in ((O.GO_IOR, O.ior), (O.GO_AND, O.and_), (O.GO_XOR, O.xor)):
            want = gmp(op, n, on, a, b)
            for i in range(n):
                assert np.array_equal(fn(a[i], b[i], on), want[i]), (op, i)
        shifts = np.array([r.randrange(-40 * an, 40 * an) for _ in range(n)], np.int32)
        shifts[:4] = [0, 32, -32, 32 * an]
        want = gmp(O.GO_SHF, n, on, a, shifts.view(np.uint32).reshape(-1, 1))
        for i in range(n):
            got = O.shf(a[i], int(shifts[i]), on)
            assert np.array_equal(got, want[i]), int(shifts[i])
            x, sh = O.to_int(a[i]), int(shifts[i])
            assert O.to_int(got) == (((x << sh) % (1 << (32 * on))) if sh >= 0 else x >> -sh)
        pc = gmp(O.GO_POPC, n, 1, a)
        cm = gmp(O.GO_CMP, n, 1, a, b).view(np.int32)
        for i in range(n):
            assert O.popc(a[i]) == int(pc[i, 0]) == bin(O.to_int(a[i])).count("1")
            assert O.cmp(a[i], b[i]) == int(cm[i, 0])
            x, y = O.to_int(a[i]), O.to_int(b[i])
            assert O.cmp(a[i], b[i]) == (x > y) - (x < y)


def test_bitwise_closed_forms():
    r = random.Random(5)
    for _ in range(40):
        n = r.randrange(1, 9)
        x, y = r.getrandbits(32 * n), r.getrandbits(32 * n)
        a, b = L(x, n), L(y, n)
        # De Morgan and self-inverse identities
        assert np.array_equal(O.not_(O.and_(a, b, n), n), O.ior(O.not_(a, n), O.not_(b, n), n))
        assert O.to_int(O.xor(a, a, n)) == 0 and np.array_equal(O.not_(O.not_(a, n), n), a)
        # a left shift undone by a right shift when nothing falls off the top
        s = r.randrange(0, 32 * n)
        wide = O.shf(a, s, 2 * n)
        assert np.array_equal(O.shf(wide, -s, n), a)
        # popcount(a) + popcount(~a) = width
        assert O.popc(a) + O.popc(O.not_(a, n)) == 32 * n
    assert O.cmp(L(5, 3), L(5, 1)) == 0 and O.cmp(L(1 << 64, 3), L((1 << 64) - 1, 2)) == 1


def test_bitwise_planted_mistakes_are_caught():
    """Plausible slips in a copy of the bitwise formulas disagree with the fixtures."""
    cases = GOLDEN["bitwise"]

    def not_zero_fill(a, an, on):          # limbs above a's length filled with 0, not 1
        return ((1 << (32 * an)) - 1) ^ a

    def shf_sign_flipped(a, sh, on):        # positive shift taken as a right shift
        return (a >> sh) if sh >= 0 else ((a << -sh) % (1 << (32 * on)))

    def shf_limb_off_by_one(a, sh, on):     # whole-limb shifts moved one limb too far
        t = sh + 32 if sh != 0 and sh % 32 == 0 else sh
        return ((a << t) % (1 << (32 * on))) if t >= 0 else (a >> -t)

    assert any(not_zero_fill(H(c["a"]), c["an"], c["on"]) != H(c["not"]) for c in cases)
    for bad in (shf_sign_flipped, shf_limb_off_by_one):
        assert any(bad(H(c["a"]), sh, c["on"]) != H(v) for c in cases for sh, v in c["shf"]), bad.__name__
    assert any(bin(H(c["a"]) & 0xFFFFFFFF).count("1") != c["popc"] for c in cases)   # low limb only
    assert any(-c["cmp"] != c["cmp"] for c in cases)                                   # sign convention


def test_input_stream_is_portable():
    """The seeded generator is a fixed bit stream (PCG64 raw output)."""
    x = inputs.uniform(12345, 2, 4)
    assert x.tolist() == GOLDEN_STREAM


GOLDEN_STREAM = None


def setup_module(module):
    global GOLDEN_STREAM
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pcg64_stream.json")
    GOLDEN_STREAM = json.load(open(path))
    module.GOLDEN_STREAM = GOLDEN_STREAM
