"""GPU: the drop-in boundary -- handle, integers, transfer, Set, error codes.

Mirrors the reference's handle, null/error, query, transfer and
set/order/index tests (unit_tests/test.cu:42-173, 296-850) through the C ABI.
Import/export results are checked against the oracle's restatement of
xmpTransform (src/xmp.cu:461-489) for every word order, endianness, word size
and nails width; error codes follow the reference's validation order
(xmp.cu:492-682, 734-813, powm.cu:250-258).
"""
import ctypes

import numpy as np
import pytest

import xmp_amd
from helpers import O, assert_rows_equal, inputs
from xmp_amd import hip

pytestmark = pytest.mark.gpu

L = xmp_amd.lib
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32


def _ptr(arr):
    return _vp(arr.ctypes.data)


def _import(h, x, buf, words, order, size, endian, nails, count):
    return L().xmpIntegersImport(h.h, x.x, words, order, size, endian, nails, _ptr(buf), count)


def _export(h, buf, x, order, size, endian, nails, count):
    w = _u32()
    rc = L().xmpIntegersExport(h.h, _ptr(buf), ctypes.byref(w), order, size, endian, nails, x.x, count)
    return rc, w.value


def _native(h, x, count):
    """x's values as (count, limbs) uint32 through the native export."""
    out = np.zeros((count, x.limbs), np.uint32)
    rc, _ = _export(h, out, x, -1, 4, 0, 0, count)
    assert rc == xmp_amd.SUCCESS
    return out


# ---------------------------------------------------------------- handle
def test_handle_device_stream_and_memory_functions(fresh_handle):
    h = fresh_handle
    dev = ctypes.c_int32(-1)
    assert L().xmpHandleGetDevice(h.h, ctypes.byref(dev)) == 0 and dev.value == 0
    s = hip.stream_create()
    got = _vp()
    assert L().xmpHandleSetStream(h.h, _vp(s)) == 0
    assert L().xmpHandleGetStream(h.h, ctypes.byref(got)) == 0 and got.value == s
    # declared but never defined by the reference (xmp.h:135-138): aliases here
    assert L().xmpHandleSetCudaStream(h.h, None) == 0
    assert L().xmpHandleGetCudaStream(h.h, ctypes.byref(got)) == 0 and not got.value
    fns = [_vp() for _ in range(4)]
    assert L().xmpHandleGetMemoryFunctions(h.h, *[ctypes.byref(f) for f in fns]) == 0
    assert all(f.value for f in fns)


_ALLOC = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_size_t)
_FREE = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


def test_custom_memory_functions_round_trip():
    """xmpHandleCreateWithMemoryFunctions (reference xmp.cu:43-64): every
    allocation goes through the caller's functions and is returned to them."""
    libc = ctypes.CDLL(None)
    libc.malloc.restype, libc.malloc.argtypes = ctypes.c_void_p, [ctypes.c_size_t]
    libc.free.argtypes = [ctypes.c_void_p]
    live = {"host": set(), "dev": set()}
    calls = {"ha": 0, "da": 0}

    def ha(n):
        p = libc.malloc(max(n, 1))
        live["host"].add(p)
        calls["ha"] += 1
        return p

    def hf(p):
        live["host"].discard(p)
        libc.free(p)

    def da(n):
        p = hip.malloc(max(n, 1))
        live["dev"].add(p)
        calls["da"] += 1
        return p

    def df(p):
        if p in live["dev"]:
            live["dev"].discard(p)
            hip.free(p)

    cbs = (_ALLOC(ha), _FREE(hf), _ALLOC(da), _FREE(df))
    hh = _vp()
    assert L().xmpHandleCreateWithMemoryFunctions(ctypes.byref(hh), *[ctypes.cast(c, _vp) for c in cbs]) == 0
    h = xmp_amd.Handle.__new__(xmp_amd.Handle)
    h.h, h._policy = hh, None
    n = 300
    a, b = inputs.ordered_pair(31, n, 16)
    A, B = h.from_array(a), h.from_array(b)
    S = h.integers(512, n)
    h.add(S, A, B)
    want, _, _ = O.gmp_batch(O.GO_ADD, n, 16, a, b)
    assert_rows_equal(S.get(), want, "add with custom allocators")
    for x in (A, B, S):
        x.close()
    h.close()
    assert calls["ha"] > 0 and calls["da"] > 0
    assert not live["host"] and not live["dev"], "allocations not returned: %r" % live


# ---------------------------------------------------------------- integers
def test_null_out_parameters_and_queries(fresh_handle):
    """reference test.cu:103-173."""
    h = fresh_handle
    assert L().xmpHandleCreate(None) != 0
    assert L().xmpHandleCreateWithMemoryFunctions(None, None, None, None, None) != 0
    assert L().xmpHandleGetStream(h.h, None) != 0
    assert L().xmpHandleGetDevice(h.h, None) != 0
    assert L().xmpIntegersCreate(h.h, None, 64, 100) != 0
    for p, n in ((64, 1000), (1024, 100), (8192, 1)):
        x = h.integers(p, n)
        pp, nn = _u32(), _u32()
        assert L().xmpIntegersGetPrecision(h.h, x.x, ctypes.byref(pp)) == 0 and pp.value == p
        assert L().xmpIntegersGetCount(h.h, x.x, ctypes.byref(nn)) == 0 and nn.value == n
        assert L().xmpIntegersGetPrecision(h.h, x.x, None) != 0
        assert L().xmpIntegersGetCount(h.h, x.x, None) != 0
        x.close()
    x = h.integers(64, 100)
    assert L().xmpIntegersImport(h.h, x.x, 1, 1, 1, 1, 0, None, 1) != 0
    assert L().xmpIntegersExport(h.h, None, None, 1, 1, 1, 0, x.x, 1) != 0
    x.close()


# ---------------------------------------------------------------- transfer
@pytest.mark.parametrize("size", [1, 2, 4, 8])
def test_import_export_every_format(handle, size):
    """Import with every order/endian/nails, checked as native limbs against
    the oracle's xmpTransform; export of the same values back into the
    format, checked the same way (reference transfer tests test.cu:296-579)."""
    h = handle
    n, bits = 97, 512
    words = bits // (8 * size)
    rng = np.random.default_rng(size)
    raw = rng.integers(0, 256, n * words * size, dtype=np.uint8)
    X = h.integers(bits, n)
    for order in (-1, 1):
        for endian in (-1, 0, 1):
            for nails in sorted({0, 1, 3, 8 * size - 1}):
                assert _import(h, X, raw, words, order, size, endian, nails, n) == 0
                e = -1 if endian == 0 else endian   # native is little-endian here
                want = O.transform(raw, n, words, order, e, nails, size).view(np.uint32).reshape(n, bits // 32)
                assert_rows_equal(_native(h, X, n), want,
                                  "import order %d endian %d size %d nails %d" % (order, endian, size, nails))
                out = np.zeros(n * words * size, np.uint8)
                rc, w = _export(h, out, X, order, size, endian, 0, n)
                assert rc == 0 and w == words
                back = O.transform(want.view(np.uint8).reshape(-1), n, words, order, e, 0, size)
                assert np.array_equal(out, back), "export order %d endian %d size %d" % (order, endian, size)
    X.close()


def test_import_export_device_buffers(handle):
    """User buffers in device memory take the same transforms (reference
    classifies the pointer, xmp.cu:522-531,619-628)."""
    h = handle
    n, bits, size = 64, 256, 2
    words = bits // (8 * size)
    raw = np.random.default_rng(5).integers(0, 256, n * words * size, dtype=np.uint8)
    dbuf = hip.malloc(raw.nbytes)
    try:
        hip.memcpy_htod(dbuf, raw)
        X = h.integers(bits, n)
        assert L().xmpIntegersImport(h.h, X.x, words, 1, size, 1, 0, _vp(dbuf), n) == 0
        want = O.transform(raw, n, words, 1, 1, 0, size).view(np.uint32).reshape(n, bits // 32)
        assert_rows_equal(_native(h, X, n), want, "import from device memory")
        w = _u32()
        assert L().xmpIntegersExport(h.h, _vp(dbuf), ctypes.byref(w), -1, 4, -1, 0, X.x, n) == 0
        back = np.zeros(n * bits // 32, np.uint32)
        hip.hip().hipMemcpy(_ptr(back), _vp(dbuf), back.nbytes, 2)   # device to host
        assert np.array_equal(back.reshape(n, -1), want)
        X.close()
    finally:
        hip.free(dbuf)


def test_order_known_answer(handle):
    """reference xmpOrderTest (test.cu:672-700): 1+3 = 4 and 2+4 = 6 through
    most-significant-first import/export of one-word integers."""
    h = handle
    a, b, c = h.integers(32, 2), h.integers(32, 2), h.integers(32, 2)
    assert _import(h, a, np.array([1, 2], np.uint32), 1, 1, 4, 0, 0, 2) == 0
    assert _import(h, b, np.array([3, 4], np.uint32), 1, 1, 4, 0, 0, 2) == 0
    h.add(c, a, b)
    out = np.zeros(2, np.uint32)
    rc, w = _export(h, out, c, 1, 4, 0, 0, 2)
    assert rc == 0 and w == 1 and out.tolist() == [4, 6]
    for x in (a, b, c):
        x.close()


def test_import_export_set_indices(fresh_handle):
    """reference xmpIndexTests (test.cu:703-783): slot 0 gathers on import and
    scatters on export; Set reads through slot 1 and writes through slot 0."""
    h = fresh_handle
    n = 10
    hv = np.arange(n, dtype=np.uint32)
    idx = np.random.default_rng(3).permutation(n).astype(np.uint32)
    A, B = h.integers(32, n), h.integers(32, n)
    h.set_indices(0, idx)
    assert _import(h, A, hv, 1, -1, 4, 0, 0, n) == 0          # A[i] = h[idx[i]]
    h.clear_indices()
    assert _native(h, A, n)[:, 0].tolist() == hv[idx].tolist()
    res = np.zeros(n, np.uint32)
    h.set_indices(0, idx)
    assert _export(h, res, A, -1, 4, 0, 0, n)[0] == 0          # res[idx[i]] = A[i] = h[idx[i]]
    assert res[idx].tolist() == hv[idx].tolist()
    h.clear_indices()
    assert _import(h, A, hv, 1, -1, 4, 0, 0, n) == 0
    h.set_indices(0, idx)
    assert _export(h, res, A, -1, 4, 0, 0, n)[0] == 0          # res[idx[i]] = h[i]
    assert res[idx].tolist() == hv.tolist()
    h.clear_indices()
    h.set_indices(1, idx)
    assert L().xmpIntegersSet(h.h, B.x, A.x, n) == 0           # B[i] = A[idx[i]]
    h.clear_indices()
    assert _native(h, B, n)[:, 0].tolist() == hv[idx].tolist()
    # gather with a short, wrapping index array and a scatter permutation
    C = h.integers(32, 2 * n)
    h.set_indices(1, np.array([7, 12, 3], np.uint32))           # read modulo A's count
    h.set_indices(0, np.arange(2 * n, dtype=np.uint32)[::-1].copy())
    assert L().xmpIntegersSet(h.h, C.x, B.x, n) == 0
    h.clear_indices()
    got = _native(h, C, 2 * n)[:, 0]
    bv = hv[idx]
    want = [bv[[7, 12 % n, 3][k % 3]] for k in range(n)]
    assert got[::-1][:n].tolist() == want
    for x in (A, B, C):
        x.close()


def test_set_plain_and_in_place(handle):
    h = handle
    n = 500
    a = inputs.uniform(41, n, 32)
    A, B = h.from_array(a), h.integers(1024, n)
    assert L().xmpIntegersSet(h.h, B.x, A.x, n) == 0
    assert_rows_equal(B.get(), a, "set")
    perm = np.random.default_rng(8).permutation(n).astype(np.uint32)
    h.set_indices(1, perm)
    try:
        assert L().xmpIntegersSet(h.h, A.x, A.x, n) == 0       # in place, through scratch
    finally:
        h.clear_indices()
    assert_rows_equal(A.get(), a[perm], "in-place permuting set")
    A.close()
    B.close()


# ---------------------------------------------------------------- errors
def test_error_codes(handle):
    """The reference's error tests (test.cu:785-850) with the codes its
    validation returns: import/export parameters (xmp.cu:492-682), Set counts
    and precisions (xmp.cu:734-752), powm counts and precisions (powm.cu:250-258)."""
    h = handle
    E = xmp_amd
    a = h.integers(64, 100)
    buf = np.zeros(400, np.uint32)
    assert _import(h, a, buf, 2, 1, 4, 1, 0, 200) == E.INVALID_PARAMETER     # count > a.count
    assert _import(h, a, buf, 2, 1, 4, 1, 0, 0) == E.INVALID_PARAMETER       # count 0
    assert _import(h, a, buf, 2, 0, 4, 1, 0, 10) == E.INVALID_PARAMETER      # order 0
    assert _import(h, a, buf, 2, 1, 3, 1, 0, 10) == E.INVALID_PARAMETER      # size 3
    assert _import(h, a, buf, 3, 1, 4, 1, 0, 10) == E.INVALID_PRECISION      # 3*4*8 != 64
    assert _export(h, buf, a, 1, 4, 1, 0, 200)[0] == E.INVALID_PARAMETER
    b = h.integers(64, 200)
    assert L().xmpIntegersSet(h.h, a.x, b.x, 200) == E.INVALID_COUNT
    c128 = h.integers(128, 100)
    assert L().xmpIntegersSet(h.h, c128.x, a.x, 10) == E.INVALID_PRECISION
    c, bb, e, m = (h.integers(64, 100) for _ in range(4))
    assert L().xmpIntegersPowmAsync(h.h, c.x, bb.x, e.x, m.x, 200) == E.INVALID_COUNT
    assert L().xmpIntegersPowmAsync(h.h, c.x, bb.x, e.x, m.x, 0) == E.SUCCESS
    assert L().xmpIntegersPowmAsync(h.h, c128.x, bb.x, e.x, m.x, 100) == E.INVALID_PRECISION
    assert L().xmpIntegersAddAsync(h.h, c.x, bb.x, e.x, 0) == E.SUCCESS
    assert L().xmpIntegersAddAsync(h.h, c.x, bb.x, b.x, 150) == E.INVALID_COUNT
    hip.device_sync()
    for x in (a, b, c128, c, bb, e, m):
        x.close()


def test_policy_parameters_and_scratch_limit(fresh_handle):
    """xmpExecutionPolicySet/GetParameter (reference xmp.cu:283-314), and the
    scratch limit as the reference's failure handling uses it
    (powm_operators.h:308-328): a tight limit shrinks the launch (smaller
    window, chunks of instances) without changing any result, and a limit
    below the smallest plan gives xmpErrorIncreaseScratchLimit."""
    h = fresh_handle
    p = h.policy()
    v = ctypes.c_uint64(99)
    assert L().xmpExecutionPolicyGetParameter(h.h, p, xmp_amd.PARAM_ALGORITHM, ctypes.byref(v)) == 0
    assert v.value == xmp_amd.ALG_DEFAULT
    assert L().xmpExecutionPolicySetParameter(h.h, p, xmp_amd.PARAM_ALGORITHM, xmp_amd.ALG_DISTRIBUTEDMP) == 0
    assert L().xmpExecutionPolicyGetParameter(h.h, p, xmp_amd.PARAM_ALGORITHM, ctypes.byref(v)) == 0
    assert v.value == xmp_amd.ALG_DISTRIBUTEDMP
    assert L().xmpExecutionPolicySetParameter(h.h, p, xmp_amd.PARAM_ALGORITHM, 7) == xmp_amd.INVALID_PARAMETER
    assert L().xmpExecutionPolicySetParameter(h.h, p, 5, 0) == xmp_amd.INVALID_PARAMETER
    h.set_algorithm(xmp_amd.ALG_DEFAULT)
    h.set_scratch_limit(1 << 20)
    assert L().xmpExecutionPolicyGetParameter(h.h, p, xmp_amd.PARAM_SCRATCH_LIMIT, ctypes.byref(v)) == 0
    assert v.value == 1 << 20

    # 3,000 x 2048-bit: the full window table (~57 MB) does not fit 10 MB
    n, limbs = 3000, 64
    m = inputs.odd_moduli(61, n, limbs)
    b = inputs.uniform(62, n, limbs)
    e = inputs.uniform(63, n, limbs)
    M, B, E = h.from_array(m), h.from_array(b), h.from_array(e)
    Out = h.integers(2048, n)
    want = O.gmp_batch(O.GO_POWM, n, limbs, b, e, m, threads=8)[0]
    h.set_scratch_limit(10 << 20)
    h.powm(Out, B, E, M)
    assert_rows_equal(Out.get(), want, "powm under a 10 MB scratch limit")
    h.set_scratch_limit(1000)
    with pytest.raises(xmp_amd.XmpError) as err:
        h.powm(Out, B, E, M)
    assert err.value.code == xmp_amd.INCREASE_SCRATCH_LIMIT
    h.set_scratch_limit(0)   # the limit covers every op's scratch, imports included
    for x in (M, B, E, Out):
        x.close()

    # 140,000 x 512-bit under 24 MB: the instances run in chunks
    n, limbs = 140000, 16
    m = inputs.odd_moduli(64, n, limbs)
    b = inputs.uniform(65, n, limbs)
    e = inputs.uniform(66, n, limbs)
    M, B, E = h.from_array(m), h.from_array(b), h.from_array(e)
    Out = h.integers(512, n)
    h.set_scratch_limit(24 << 20)   # above the 9 MB of R^2 mod m, below one plain table
    h.powm(Out, B, E, M)
    h.set_scratch_limit(0)
    want = O.gmp_batch(O.GO_POWM, n, limbs, b, e, m, threads=16)[0]
    assert_rows_equal(Out.get(), want, "chunked powm under a 24 MB scratch limit")
    for x in (M, B, E, Out):
        x.close()


def test_cxx_caller_crt_decrypt():
    """A C++ program built against include/xmp.h and linked to the library
    (tests/cxx/crt_decrypt.cpp, compiled by __graft_entry__.build()): RSA
    encryption and CRT decryption of 4,096 messages round-trip exactly."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "_build", "crt_decrypt")
    assert os.path.exists(exe), "tests/cxx/_build/crt_decrypt is not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "crt ok" in r.stdout, r.stdout + r.stderr
