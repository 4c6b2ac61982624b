"""xmp_amd -- batched big-integer arithmetic for the AMD MI355X (gfx950).

The product is the C-ABI library ``xmp_amd/libxmp_amd.so`` whose entry points
are declared in ``include/xmp.h`` (the reference XMP API, src/include/xmp.h)
and ``include/xmp_ext.h`` (batch, serving and tuning extensions).  This
module is a thin ctypes mirror of that API for Python callers, tests and the
benchmark: it never computes anything itself, and it raises if the library is
missing.

Integers are passed as numpy ``uint32`` arrays of shape ``(count, limbs)``
(32-bit little-endian limbs, least significant first), the reference's
native import format (xmpIntegersImport with order=-1, size=4, endian=0).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxmp_amd.so")

# xmpError_t (reference src/include/xmp.h:40-54)
ERRORS = [
    "xmpErrorSuccess", "xmpErrorInvalidDispatchTable", "xmpErrorInvalidParameter", "xmpErrorInvalidMalloc",
    "xmpErrorInvalidCudaMalloc", "xmpErrorInvalidCount", "xmpErrorInvalidDevice", "xmpErrorInvalidFormat",
    "xmpErrorInvalidPrecision", "xmpErrorIncreaseScratchLimit", "xmpErrorUnsupported", "xmpErrorUnspecified",
    "xmpErrorCuda",
]
(SUCCESS, INVALID_DISPATCH, INVALID_PARAMETER, INVALID_MALLOC, INVALID_CUDA_MALLOC, INVALID_COUNT,
 INVALID_DEVICE, INVALID_FORMAT, INVALID_PRECISION, INCREASE_SCRATCH_LIMIT, UNSUPPORTED, UNSPECIFIED,
 CUDA_ERROR) = range(13)

# xmpExecutionPolicyParam_t / xmpAlgorithm_t (xmp.h:57-67)
PARAM_ALGORITHM, PARAM_SCRATCH_LIMIT = 0, 1
ALG_DEFAULT, ALG_REGMP, ALG_DIGITMP, ALG_DISTRIBUTEDMP = range(4)


class XmpError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        name = ERRORS[code] if 0 <= code < len(ERRORS) else "xmpErrorUnknown"
        super().__init__("%s failed: %s (%d)" % (where, name, code))


_lib = None
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int32
_sz = ctypes.c_size_t


def _declare(L):
    E = ctypes.c_int
    sig = {
        "xmpHandleCreate": [ctypes.POINTER(_vp)],
        "xmpHandleCreateWithMemoryFunctions": [ctypes.POINTER(_vp), _vp, _vp, _vp, _vp],
        "xmpHandleDestroy": [_vp],
        "xmpHandleSetStream": [_vp, _vp],
        "xmpHandleGetStream": [_vp, ctypes.POINTER(_vp)],
        "xmpHandleSetCudaStream": [_vp, _vp],
        "xmpHandleGetCudaStream": [_vp, ctypes.POINTER(_vp)],
        "xmpHandleSetExecutionPolicy": [_vp, _vp],
        "xmpHandleGetMemoryFunctions": [_vp, _vp, _vp, _vp, _vp],
        "xmpHandleGetDevice": [_vp, ctypes.POINTER(_i32)],
        "xmpExecutionPolicyCreate": [_vp, ctypes.POINTER(_vp)],
        "xmpExecutionPolicyDestroy": [_vp, _vp],
        "xmpExecutionPolicySetIndices": [_vp, _vp, _u32, _vp, _u32],
        "xmpExecutionPolicySetIndicesAsync": [_vp, _vp, _u32, _vp, _u32],
        # the value union is 8 bytes and travels like a size_t
        "xmpExecutionPolicySetParameter": [_vp, _vp, ctypes.c_int, ctypes.c_uint64],
        "xmpExecutionPolicyGetParameter": [_vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)],
        "xmpIntegersCreate": [_vp, ctypes.POINTER(_vp), _u32, _u32],
        "xmpIntegersDestroy": [_vp, _vp],
        "xmpIntegersGetPrecision": [_vp, _vp, ctypes.POINTER(_u32)],
        "xmpIntegersGetCount": [_vp, _vp, ctypes.POINTER(_u32)],
        "xmpIntegersImport": [_vp, _vp, _u32, _i32, _sz, _i32, _u32, _vp, _u32],
        "xmpIntegersImportAsync": [_vp, _vp, _u32, _i32, _sz, _i32, _u32, _vp, _u32],
        "xmpIntegersExport": [_vp, _vp, ctypes.POINTER(_u32), _i32, _sz, _i32, _u32, _vp, _u32],
        "xmpIntegersExportAsync": [_vp, _vp, ctypes.POINTER(_u32), _i32, _sz, _i32, _u32, _vp, _u32],
        "xmpIntegersSet": [_vp, _vp, _vp, _u32],
        "xmpIntegersSetAsync": [_vp, _vp, _vp, _u32],
        "xmpIntegersSqrAsync": [_vp, _vp, _vp, _u32],
        "xmpIntegersDivMod": [_vp, _vp, _vp, _vp, _vp, _u32],
        "xmpIntegersDivModAsync": [_vp, _vp, _vp, _vp, _vp, _u32],
        "xmpIntegersPowm": [_vp, _vp, _vp, _vp, _vp, _u32],
        "xmpIntegersPowmAsync": [_vp, _vp, _vp, _vp, _vp, _u32],
        "xmpIntegersNot": [_vp, _vp, _vp, _u32],
        "xmpIntegersNotAsync": [_vp, _vp, _vp, _u32],
        "xmpIntegersShf": [_vp, _vp, _vp, _vp, _u32, _u32],
        "xmpIntegersShfAsync": [_vp, _vp, _vp, _vp, _u32, _u32],
        "xmpIntegersPopc": [_vp, _vp, _vp, _u32],
        "xmpIntegersPopcAsync": [_vp, _vp, _vp, _u32],
        "xmpIntegersCmp": [_vp, _vp, _vp, _vp, _u32],
        "xmpIntegersCmpAsync": [_vp, _vp, _vp, _vp, _u32],
    }
    for n in ("Add", "Sub", "Mul", "Div", "Mod", "Ior", "And", "Xor"):
        sig["xmpIntegers" + n] = [_vp, _vp, _vp, _vp, _u32]
        sig["xmpIntegers" + n + "Async"] = [_vp, _vp, _vp, _vp, _u32]
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = E
    return L


def lib():
    """Load the HIP library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("xmp_amd/libxmp_amd.so is not built: run __graft_entry__.build() "
                               "or `python -m xmp_amd.build`")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _lib = _declare(L)
        try:
            from . import ext as _ext
            _ext.declare(L)
        except ImportError:
            pass
    return _lib


def check(code, where):
    if code != SUCCESS:
        raise XmpError(code, where)


class Handle:
    """xmpHandle_t on the current HIP device (reference xmp.cu:39-102)."""

    def __init__(self):
        self.h = _vp()
        check(lib().xmpHandleCreate(ctypes.byref(self.h)), "xmpHandleCreate")
        self._policy = None

    def close(self):
        if self._policy is not None:
            lib().xmpExecutionPolicyDestroy(self.h, self._policy)
            self._policy = None
        if self.h:
            check(lib().xmpHandleDestroy(self.h), "xmpHandleDestroy")
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_stream(self, stream_ptr):
        check(lib().xmpHandleSetStream(self.h, _vp(stream_ptr)), "xmpHandleSetStream")

    # ---- execution policy
    def policy(self):
        if self._policy is None:
            p = _vp()
            check(lib().xmpExecutionPolicyCreate(self.h, ctypes.byref(p)), "xmpExecutionPolicyCreate")
            self._policy = p
            check(lib().xmpHandleSetExecutionPolicy(self.h, p), "xmpHandleSetExecutionPolicy")
        return self._policy

    def set_indices(self, which, indices):
        p = self.policy()
        if indices is None:
            check(lib().xmpExecutionPolicySetIndices(self.h, p, which, None, 0), "SetIndices")
            return
        arr = np.ascontiguousarray(indices, dtype=np.uint32)
        check(lib().xmpExecutionPolicySetIndices(self.h, p, which, arr.ctypes.data, arr.size), "SetIndices")

    def clear_indices(self):
        for w in range(4):
            self.set_indices(w, None)

    def set_algorithm(self, alg):
        check(lib().xmpExecutionPolicySetParameter(self.h, self.policy(), PARAM_ALGORITHM, alg), "SetParameter")

    def set_scratch_limit(self, nbytes):
        check(lib().xmpExecutionPolicySetParameter(self.h, self.policy(), PARAM_SCRATCH_LIMIT, nbytes),
              "SetParameter")

    # ---- integers
    def integers(self, precision, count):
        return Integers(self, precision, count)

    def from_array(self, arr, precision=None):
        arr = np.ascontiguousarray(arr, dtype=np.uint32)
        if arr.ndim == 1:
            arr = arr.reshape(1, -1)
        x = Integers(self, precision or 32 * arr.shape[1], arr.shape[0])
        x.load(arr)
        return x

    # ---- math (reference xmp.h:180-199)
    def _bin(self, name, out, a, b, count):
        check(getattr(lib(), name)(self.h, out.x, a.x, b.x, count), name)

    def add(self, out, a, b, count=None):
        self._bin("xmpIntegersAdd", out, a, b, count if count is not None else out.count)

    def sub(self, out, a, b, count=None):
        self._bin("xmpIntegersSub", out, a, b, count if count is not None else out.count)

    def mul(self, out, a, b, count=None):
        self._bin("xmpIntegersMul", out, a, b, count if count is not None else out.count)

    def div(self, out, a, b, count=None):
        self._bin("xmpIntegersDiv", out, a, b, count if count is not None else out.count)

    def mod(self, out, a, b, count=None):
        self._bin("xmpIntegersMod", out, a, b, count if count is not None else out.count)

    def divmod(self, q, r, a, b, count=None):
        c = count if count is not None else q.count
        check(lib().xmpIntegersDivMod(self.h, q.x, r.x, a.x, b.x, c), "xmpIntegersDivMod")

    def powm(self, out, base, exp, mod, count=None, asynchronous=False):
        c = count if count is not None else out.count
        f = lib().xmpIntegersPowmAsync if asynchronous else lib().xmpIntegersPowm
        check(f(self.h, out.x, base.x, exp.x, mod.x, c), "xmpIntegersPowm")

    # ---- bitwise (reference xmp.h:206-226)
    def ior(self, out, a, b, count=None):
        self._bin("xmpIntegersIor", out, a, b, count if count is not None else out.count)

    def and_(self, out, a, b, count=None):
        self._bin("xmpIntegersAnd", out, a, b, count if count is not None else out.count)

    def xor(self, out, a, b, count=None):
        self._bin("xmpIntegersXor", out, a, b, count if count is not None else out.count)

    def not_(self, out, a, count=None):
        c = count if count is not None else out.count
        check(lib().xmpIntegersNot(self.h, out.x, a.x, c), "xmpIntegersNot")

    def shf(self, out, a, shifts, count=None):
        """shifts: int32 array (host numpy) or a device pointer (int); instance i uses shifts[i % len]."""
        c = count if count is not None else out.count
        if isinstance(shifts, int):
            raise TypeError("pass a device pointer with shf_device")
        sh = np.ascontiguousarray(shifts, dtype=np.int32)
        check(lib().xmpIntegersShf(self.h, out.x, a.x, sh.ctypes.data, sh.size, c), "xmpIntegersShf")

    def shf_device(self, out, a, shift_ptr, shift_count, count=None):
        c = count if count is not None else out.count
        check(lib().xmpIntegersShf(self.h, out.x, a.x, shift_ptr, shift_count, c), "xmpIntegersShf")

    def popc(self, a, count=None):
        c = count if count is not None else a.count
        res = np.zeros(c, np.uint32)
        check(lib().xmpIntegersPopc(self.h, res.ctypes.data, a.x, c), "xmpIntegersPopc")
        return res

    def cmp(self, a, b, count=None):
        c = count if count is not None else max(a.count, b.count)
        res = np.zeros(c, np.int32)
        check(lib().xmpIntegersCmp(self.h, res.ctypes.data, a.x, b.x, c), "xmpIntegersCmp")
        return res


class Integers:
    """xmpIntegers_t: `count` integers of `precision` bits on the device."""

    def __init__(self, handle, precision, count):
        self.handle = handle
        self.precision = precision
        self.count = count
        self.limbs = precision // 32
        self.x = _vp()
        check(lib().xmpIntegersCreate(handle.h, ctypes.byref(self.x), precision, count), "xmpIntegersCreate")

    def close(self):
        if self.x:
            lib().xmpIntegersDestroy(self.handle.h, self.x)
            self.x = _vp()

    def __del__(self):
        try:
            if self.x and self.handle.h:
                self.close()
        except Exception:
            pass

    def load(self, arr, count=None):
        arr = np.ascontiguousarray(arr, dtype=np.uint32)
        c = count if count is not None else arr.shape[0]
        check(lib().xmpIntegersImport(self.handle.h, self.x, self.limbs, -1, 4, 0, 0, arr.ctypes.data, c),
              "xmpIntegersImport")

    def get(self, count=None):
        c = count if count is not None else self.count
        out = np.zeros((c, self.limbs), np.uint32)
        words = _u32()
        check(lib().xmpIntegersExport(self.handle.h, out.ctypes.data, ctypes.byref(words), -1, 4, 0, 0, self.x, c),
              "xmpIntegersExport")
        return out
