"""Python access to the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (built by ``oracle/Makefile``) and exposes:

* the C restatement of XMP's arithmetic (``xo_*``: add, sub, mul, divmod,
  powm, transform) operating on compact little-endian uint32 limb arrays;
* GMP mpz through ``go_batch`` -- the reference test suite's own oracle
  (unit_tests/test.cu:1214 and friends) and the CPU baseline of the
  reference perf harness (perf_tests/main.cu:153-229).

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module.  Arrays are numpy ``uint32`` of shape ``(count, limbs)``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

(GO_ADD, GO_SUB, GO_MUL, GO_DIV, GO_MOD, GO_POWM, GO_POWM_SEC,
 GO_IOR, GO_AND, GO_XOR, GO_SHF, GO_POPC, GO_CMP) = range(13)

_u32p = ctypes.POINTER(ctypes.c_uint32)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        for name in ("xo_add", "xo_sub", "xo_mul", "xo_ior", "xo_and", "xo_xor"):
            getattr(L, name).argtypes = [_u32p, ctypes.c_int, _u32p, ctypes.c_int, _u32p, ctypes.c_int]
            getattr(L, name).restype = None
        L.xo_not.argtypes = [_u32p, ctypes.c_int, _u32p, ctypes.c_int]
        L.xo_not.restype = None
        L.xo_shf.argtypes = [_u32p, ctypes.c_int, _u32p, ctypes.c_int, ctypes.c_int32]
        L.xo_shf.restype = None
        L.xo_popc.argtypes = [_u32p, ctypes.c_int]
        L.xo_popc.restype = ctypes.c_uint32
        L.xo_cmp.argtypes = [_u32p, ctypes.c_int, _u32p, ctypes.c_int]
        L.xo_cmp.restype = ctypes.c_int32
        L.xo_divmod.argtypes = [_u32p, ctypes.c_int, _u32p, ctypes.c_int, _u32p, ctypes.c_int, _u32p, ctypes.c_int]
        L.xo_divmod.restype = ctypes.c_int
        L.xo_np0.argtypes = [ctypes.c_uint32]
        L.xo_np0.restype = ctypes.c_uint32
        L.xo_window_bits.argtypes = [ctypes.c_uint32]
        L.xo_window_bits.restype = ctypes.c_int
        L.xo_powm.argtypes = [_u32p, _u32p, _u32p, ctypes.c_int, _u32p, ctypes.c_int]
        L.xo_powm.restype = ctypes.c_int
        L.xo_transform.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32,
                                   _u32p, ctypes.c_uint32, _u32p, ctypes.c_uint32]
        L.xo_transform.restype = None
        L.xo_powm_batch.argtypes = [_u32p, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u32p, ctypes.c_uint32,
                                    ctypes.c_int, _u32p, ctypes.c_uint32, ctypes.c_int]
        L.xo_powm_batch.restype = None
        L.go_load.restype = ctypes.c_int
        L.go_version.restype = ctypes.c_char_p
        L.go_batch.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                               _u32p, ctypes.c_int, _u32p, ctypes.c_int,
                               _u32p, ctypes.c_int, ctypes.c_uint32,
                               _u32p, ctypes.c_int, ctypes.c_uint32,
                               _u32p, ctypes.c_int, ctypes.c_uint32]
        L.go_batch.restype = ctypes.c_double
        _LIB = L
    return _LIB


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- C restatement
def add(a, b, out_limbs):
    a, b = _c(a), _c(b)
    out = np.zeros(out_limbs, np.uint32)
    lib().xo_add(_p(out), out_limbs, _p(a), a.size, _p(b), b.size)
    return out


def sub(a, b, out_limbs):
    a, b = _c(a), _c(b)
    out = np.zeros(out_limbs, np.uint32)
    lib().xo_sub(_p(out), out_limbs, _p(a), a.size, _p(b), b.size)
    return out


def mul(a, b, out_limbs):
    a, b = _c(a), _c(b)
    out = np.zeros(out_limbs, np.uint32)
    lib().xo_mul(_p(out), out_limbs, _p(a), a.size, _p(b), b.size)
    return out


def _bitwise(name, a, b, out_limbs):
    a, b = _c(a), _c(b)
    out = np.zeros(out_limbs, np.uint32)
    getattr(lib(), name)(_p(out), out_limbs, _p(a), a.size, _p(b), b.size)
    return out


def ior(a, b, out_limbs):
    return _bitwise("xo_ior", a, b, out_limbs)


def and_(a, b, out_limbs):
    return _bitwise("xo_and", a, b, out_limbs)


def xor(a, b, out_limbs):
    return _bitwise("xo_xor", a, b, out_limbs)


def not_(a, out_limbs):
    a = _c(a)
    out = np.zeros(out_limbs, np.uint32)
    lib().xo_not(_p(out), out_limbs, _p(a), a.size)
    return out


def shf(a, shift, out_limbs):
    a = _c(a)
    out = np.zeros(out_limbs, np.uint32)
    lib().xo_shf(_p(out), out_limbs, _p(a), a.size, int(shift))
    return out


def popc(a):
    a = _c(a)
    return int(lib().xo_popc(_p(a), a.size))


def cmp(a, b):
    a, b = _c(a), _c(b)
    return int(lib().xo_cmp(_p(a), a.size, _p(b), b.size))


def divmod_(a, b, q_limbs, r_limbs):
    a, b = _c(a), _c(b)
    q = np.zeros(max(q_limbs, 1), np.uint32)
    r = np.zeros(max(r_limbs, 1), np.uint32)
    rc = lib().xo_divmod(_p(q), q_limbs, _p(r), r_limbs, _p(a), a.size, _p(b), b.size)
    return q[:q_limbs], r[:r_limbs], rc


def np0(m0):
    return int(lib().xo_np0(int(m0)))


def window_bits(exp_bits):
    return int(lib().xo_window_bits(int(exp_bits)))


def powm(base, exp, mod):
    base, exp, mod = _c(base), _c(exp), _c(mod)
    assert base.size == mod.size
    out = np.zeros(mod.size, np.uint32)
    rc = lib().xo_powm(_p(out), _p(base), _p(exp), exp.size, _p(mod), mod.size)
    return out, rc


def powm_batch(base, exp, mod, count):
    """Compact arrays (k, n); instance i reads row i % k of each input."""
    base, exp, mod = _c(base), _c(exp), _c(mod)
    n = mod.shape[1]
    out = np.zeros((count, n), np.uint32)
    lib().xo_powm_batch(_p(out), count, _p(base), base.shape[0], _p(exp), exp.shape[0], exp.shape[1],
                        _p(mod), mod.shape[0], n)
    return out


def transform(inp, count, words, order, endian, nails, size, out_indices=None, in_indices=None, out_rows=None):
    """Restated xmpTransform over a byte buffer of count*words words of `size` bytes."""
    inp = np.ascontiguousarray(inp)
    rows = out_rows if out_rows is not None else count
    out = np.zeros(rows * words * size, np.uint8)
    oi = _c(out_indices) if out_indices is not None else None
    ii = _c(in_indices) if in_indices is not None else None
    lib().xo_transform(out.ctypes.data, inp.ctypes.data, count, words, order, endian, nails, size,
                       _p(oi), 0 if oi is None else oi.size, _p(ii), 0 if ii is None else ii.size)
    return out


# ---------------------------------------------------------------- GMP
def gmp_available():
    return lib().go_load() == 0


def gmp_version():
    lib().go_load()
    return lib().go_version().decode()


def gmp_batch(op, count, out_limbs, a, b=None, c=None, out2_limbs=0, threads=1):
    """Run a GMP op over `count` instances (input row i % rows).  Returns
    (out, out2, seconds).  op: GO_ADD/SUB/MUL/DIV/MOD/POWM/POWM_SEC and the
    bitwise GO_IOR/AND/XOR/SHF/POPC/CMP; for powm a=base, b=exp, c=mod; for
    GO_SHF b holds one int32 shift per row (viewed as uint32); GO_POPC and
    GO_CMP write one word per instance (cmp as int32 -1/0/1).  Results are
    truncated to out_limbs limbs."""
    a = _c(a)
    b = _c(b) if b is not None else None
    c = _c(c) if c is not None else None
    out = np.zeros((count, out_limbs), np.uint32)
    out2 = np.zeros((count, out2_limbs), np.uint32) if out2_limbs else None
    secs = lib().go_batch(op, count, threads, _p(out), out_limbs, _p(out2), out2_limbs,
                          _p(a), a.shape[1], a.shape[0],
                          _p(b), 0 if b is None else b.shape[1], 0 if b is None else b.shape[0],
                          _p(c), 0 if c is None else c.shape[1], 0 if c is None else c.shape[0])
    if secs < 0:
        raise RuntimeError("GMP (libgmp.so.10) could not be loaded")
    return out, out2, secs


# ---------------------------------------------------------------- helpers
def to_int(limbs):
    v = 0
    for x in reversed(np.asarray(limbs, dtype=np.uint64).tolist()):
        v = (v << 32) | int(x)
    return v


def from_int(v, n):
    out = np.zeros(n, np.uint32)
    for i in range(n):
        out[i] = v & 0xFFFFFFFF
        v >>= 32
    return out


def serving_expected(keys, kinds, key_idx, bases, exps, threads=8):
    """Expected results of served requests: kind 0 is base^exp mod n (GMP);
    kind 1 is the CRT pair of samples/crt_rsa/rsa.cc:205-219,
    (c^dp mod p)*cp + (c^dq mod q)*cq mod n, with c reduced mod p and q first
    (Python int)."""
    n32 = keys["n"].shape[1]
    out = np.zeros((len(kinds), n32), np.uint32)
    a = np.nonzero(kinds == 0)[0]
    if a.size:
        out[a], _, _ = gmp_batch(GO_POWM, a.size, n32, bases[a], exps[a], keys["n"][key_idx[a]], threads=threads)
    for i in np.nonzero(kinds == 1)[0]:
        k = key_idx[i]
        p, q, n = to_int(keys["p"][k]), to_int(keys["q"][k]), to_int(keys["n"][k])
        c = to_int(bases[i])
        mp = pow(c % p, to_int(keys["dp"][k]), p)
        mq = pow(c % q, to_int(keys["dq"][k]), q)
        out[i] = from_int((mp * to_int(keys["cp"][k]) + mq * to_int(keys["cq"][k])) % n, n32)
    return out
