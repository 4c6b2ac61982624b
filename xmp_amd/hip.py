"""Minimal HIP runtime access (streams, events, sync) through libamdhip64.

Used by the benchmark to time kernels with HIP events on the exact stream the
library launches on.  Streams are created at the default priority.
"""
import ctypes
import ctypes.util

_hip = None
_vp = ctypes.c_void_p


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        H = _hip
        H.hipStreamCreate.argtypes = [ctypes.POINTER(_vp)]
        H.hipStreamDestroy.argtypes = [_vp]
        H.hipStreamSynchronize.argtypes = [_vp]
        H.hipEventCreate.argtypes = [ctypes.POINTER(_vp)]
        H.hipEventDestroy.argtypes = [_vp]
        H.hipEventRecord.argtypes = [_vp, _vp]
        H.hipEventSynchronize.argtypes = [_vp]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
        H.hipDeviceSynchronize.argtypes = []
        H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        H.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
        H.hipFree.argtypes = [_vp]
        H.hipMemset.argtypes = [_vp, ctypes.c_int, ctypes.c_size_t]
        H.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with hipError %d" % (what, rc))


def device_count():
    n = ctypes.c_int(0)
    if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def stream_create():
    s = _vp()
    _ok(hip().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
    return s.value


def stream_sync(s):
    _ok(hip().hipStreamSynchronize(_vp(s)), "hipStreamSynchronize")


def device_sync():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Event:
    def __init__(self):
        e = _vp()
        _ok(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.e = e

    def record(self, stream):
        _ok(hip().hipEventRecord(self.e, _vp(stream)), "hipEventRecord")

    def sync(self):
        _ok(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.e, later.e), "hipEventElapsedTime")
        return ms.value


def malloc(nbytes, fill=None):
    p = _vp()
    _ok(hip().hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
    if fill is not None:
        _ok(hip().hipMemset(p, fill, nbytes), "hipMemset")
    return p.value


def free(ptr):
    _ok(hip().hipFree(_vp(ptr)), "hipFree")


def memcpy_htod(ptr, arr):
    """Synchronous copy of a contiguous numpy array to device memory at ptr."""
    import numpy as np
    a = np.ascontiguousarray(arr)
    _ok(hip().hipMemcpy(_vp(ptr), _vp(a.ctypes.data), a.nbytes, 1), "hipMemcpy")   # 1 = hipMemcpyHostToDevice
