"""Write tests/golden/golden.json: small input/expected-output vectors.

The reference commits no golden vectors (its tests compute expected values
with GMP at run time, unit_tests/test.cu).  These fixtures are computed here
with Python's built-in int -- an oracle independent of both GMP and the C
restatement -- on seeded inputs, plus the reference's own known-answer
cases:
  * limb pattern i*limbs+j round trip (unit_tests/test.cu:209-289);
  * word order: 1+3 = 4 and 2+4 = 6 with most-significant-first words
    (test.cu:683-695);
  * the toy RSA key of samples/crt_rsa/rsa.cc:167-178 and
    samples/simple_rsa/rsa.cc:126-130;
  * import/export repacking for every order, endian and word size
    (test.cu:296-579), computed from the word-transform definition of
    src/xmp.cu:461-489;
  * bitwise ops as the reference tests define them (test.cu:1366-1372
    shifts, :1480-1488 not, :2091 cmp, :2191 popcount, :2516-2595
    and/ior/xor).

Run: python tests/golden/make_golden.py   (deterministic; rewrites the file)
"""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))


def limbs(v, n):
    return [(v >> (32 * i)) & 0xFFFFFFFF for i in range(n)]


def hexs(v):
    return "%x" % v


def rnd(r, bits):
    return r.getrandbits(bits)


def window_bits(exp_bits):
    return 4 if exp_bits < 400 else 5 if exp_bits < 800 else 6 if exp_bits < 1600 else 7 if exp_bits <= 4096 else 8


def main():
    r = random.Random(20261202)
    g = {"about": "inputs and expected outputs as hex strings; limb counts n/an/bn/... are 32-bit limbs"}

    # add / sub / mul with zero-extension and truncation to the output width
    cases = []
    for an, bn, on in [(1, 1, 2), (2, 2, 2), (4, 3, 5), (8, 8, 8), (8, 4, 16), (16, 16, 17), (3, 7, 7), (64, 64, 64),
                       (64, 64, 65)]:
        for k in range(4):
            a, b = rnd(r, 32 * an), rnd(r, 32 * bn)
            if k == 0:
                a, b = (1 << (32 * an)) - 1, 1
            mod = 1 << (32 * on)
            cases.append({"an": an, "bn": bn, "on": on, "a": hexs(a), "b": hexs(b), "add": hexs((a + b) % mod),
                          "sub": hexs((a - b) % mod), "mul": hexs((a * b) % mod)})
    g["addsubmul"] = cases

    # division: q = floor(a/b), r = a mod b
    cases = []
    for an, bn in [(2, 1), (4, 2), (8, 3), (16, 8), (128, 64), (5, 5), (3, 4), (64, 64)]:
        for k in range(4):
            a, b = rnd(r, 32 * an), rnd(r, 32 * bn)
            if k == 1:
                b |= 1 << (32 * bn - 1)
            if k == 2:
                b >>= r.randrange(1, 32 * bn)
            if b == 0:
                b = 1
            cases.append({"an": an, "bn": bn, "a": hexs(a), "b": hexs(b), "q": hexs(a // b), "r": hexs(a % b)})
    g["divmod"] = cases

    # powm: odd moduli, bases not reduced (as the reference test, test.cu:1144-1149)
    cases = []
    for n, en in [(1, 1), (2, 2), (4, 4), (4, 1), (8, 8), (16, 16), (32, 32), (32, 1), (64, 64), (64, 2)]:
        for k in range(3):
            m = rnd(r, 32 * n) | 1
            if k == 0:
                m |= 1 << (32 * n - 1)
            b = rnd(r, 32 * n)
            e = rnd(r, 32 * en)
            cases.append({"n": n, "en": en, "b": hexs(b), "e": hexs(e), "m": hexs(m), "out": hexs(pow(b, e, m)),
                          "window_bits": window_bits(32 * en)})
    # closed forms
    cases.append({"n": 4, "en": 4, "b": hexs(12345), "e": "0", "m": hexs((1 << 127) - 1), "out": "1", "window_bits": 4})
    cases.append({"n": 4, "en": 4, "b": "0", "e": hexs(77), "m": hexs((1 << 127) - 1), "out": "0", "window_bits": 4})
    cases.append({"n": 4, "en": 4, "b": hexs(98765), "e": "1", "m": "1", "out": "0", "window_bits": 4})
    cases.append({"n": 1, "en": 1, "b": hexs(12345), "e": hexs(65537), "m": hexs(17460671),
                  "out": hexs(pow(12345, 65537, 17460671)), "window_bits": 4})
    g["powm"] = cases

    # np0 = -m0^-1 mod 2^32
    g["np0"] = [{"m0": hexs(m0), "np0": hexs((-pow(m0, -1, 1 << 32)) % (1 << 32))}
                for m0 in [1, 3, 0xFFFFFFFF, 0x80000001] + [rnd(r, 32) | 1 for _ in range(8)]]

    # toy RSA key from the reference samples
    p, q, n, e, d = 4931, 3541, 17460671, 65537, 16156673
    dp, dq = d % (p - 1), d % (q - 1)
    cp, cq = 15212136, 2248536
    g["rsa_sample"] = {"p": p, "q": q, "n": n, "e": e, "d": d, "dp": dp, "dq": dq, "cp": cp, "cq": cq,
                       "messages": [2, 12345, 99999, 17460670],
                       "ciphertexts": [pow(x, e, n) for x in [2, 12345, 99999, 17460670]]}

    # word-order known answer: 1 + 3 = 4, 2 + 4 = 6 with order = +1 (most significant word first)
    g["order_add"] = {"words": 2, "a_msw_first": [0, 1, 0, 2], "b_msw_first": [0, 3, 0, 4],
                      "sum_msw_first": [0, 4, 0, 6]}

    # import transform (reference src/xmp.cu:461-489): compact words -> native limbs
    cases = []
    for size in (1, 2, 4, 8):
        for order in (-1, 1):
            for endian in (-1, 1):
                for nails in (0, 3):
                    count, nbytes = 3, 16
                    words = nbytes // size
                    data = bytes(r.getrandbits(8) for _ in range(count * nbytes))
                    out = []
                    mask = ((1 << (8 * size)) - 1) >> nails
                    for i in range(count):
                        v = 0
                        for j in range(words):
                            pos = j if order == -1 else words - 1 - j
                            wb = data[(i * words + pos) * size:(i * words + pos + 1) * size]
                            w = int.from_bytes(wb, "little" if endian == -1 else "big") & mask
                            v |= w << (8 * size * j)
                        out.append(limbs(v, nbytes // 4))
                    cases.append({"size": size, "order": order, "endian": endian, "nails": nails, "count": count,
                                  "words": words, "data": data.hex(), "limbs": out})
    g["import"] = cases

    # bitwise ops (src/logic.cu): inputs zero-extended to the output width,
    # not fills with ones, shift positive = left (truncated), negative = right;
    # popcount and sign comparison
    r2 = random.Random(20261203)
    cases = []
    for an, bn, on in [(1, 1, 1), (2, 1, 2), (4, 4, 4), (3, 5, 6), (8, 2, 8), (16, 16, 16)]:
        for k in range(3):
            a, b = rnd(r2, 32 * an), rnd(r2, 32 * bn)
            if k == 0:
                a = (1 << (32 * an)) - 1
            mod = 1 << (32 * on)
            shifts = sorted({0, 1, -1, 31, -31, 32, -32, 33, -33, 32 * an, -32 * an, 32 * on + 5, -(32 * an) - 5,
                             r2.randrange(-64 * an, 64 * an)})
            cases.append({"an": an, "bn": bn, "on": on, "a": hexs(a), "b": hexs(b),
                          "ior": hexs(a | b), "and": hexs(a & b), "xor": hexs(a ^ b),
                          "not": hexs((mod - 1) ^ a),
                          "shf": [[sh, hexs(((a << sh) % mod) if sh >= 0 else (a >> -sh))] for sh in shifts],
                          "popc": bin(a).count("1"), "cmp": (a > b) - (a < b)})
    # equal values of different lengths compare equal; a leading-zero longer operand compares by value
    cases.append({"an": 4, "bn": 2, "on": 4, "a": hexs(0x1234), "b": hexs(0x1234), "ior": hexs(0x1234),
                  "and": hexs(0x1234), "xor": "0", "not": hexs(((1 << 128) - 1) ^ 0x1234),
                  "shf": [[0, hexs(0x1234)]], "popc": bin(0x1234).count("1"), "cmp": 0})
    g["bitwise"] = cases

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=0, sort_keys=True)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
