"""ctypes declarations and thin wrappers for include/xmp_ext.h."""
import ctypes

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32


def declare(L):
    E = ctypes.c_int
    sig = {
        "xmpMontgomeryCreate": [_vp, ctypes.POINTER(_vp), _vp, _u32, _u32],
        "xmpMontgomeryDestroy": [_vp, _vp],
        "xmpMontgomeryGetLayout": [_vp, _vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32), ctypes.POINTER(_u32),
                                   ctypes.POINTER(_u32)],
        "xmpIntegersPowmMontgomeryAsync": [_vp, _vp, _vp, _vp, _vp, _u32],
        "xmpBatchCreate": [_vp, ctypes.POINTER(_vp), _u32, _vp, _vp, _vp, _vp, _vp],
        "xmpBatchDestroy": [_vp, _vp],
        "xmpBatchPrepare": [_vp, _vp],
        "xmpBatchLaunchAsync": [_vp, _vp],
        "xmpBatchLaunchPlainAsync": [_vp, _vp],
        "xmpTunerRun": [_vp, _u32, _u32, _u32, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_u32)],
        "xmpTunerLoad": [ctypes.c_char_p],
    }
    for name, args in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = E


class Montgomery:
    """Resident moduli with precomputed R^2 mod m (xmpMontgomeryCreate)."""

    def __init__(self, handle, mod, exp_precision, count_hint):
        from . import check, lib
        self.handle, self.mod = handle, mod
        self.ctx = _vp()
        check(lib().xmpMontgomeryCreate(handle.h, ctypes.byref(self.ctx), mod.x, exp_precision, count_hint),
              "xmpMontgomeryCreate")

    def layout(self):
        from . import check, lib
        v = [_u32() for _ in range(4)]
        check(lib().xmpMontgomeryGetLayout(self.handle.h, self.ctx, *[ctypes.byref(x) for x in v]), "GetLayout")
        return tuple(x.value for x in v)

    def powm_async(self, out, base, exp, count):
        from . import check, lib
        check(lib().xmpIntegersPowmMontgomeryAsync(self.handle.h, out.x, base.x, exp.x, self.ctx, count),
              "xmpIntegersPowmMontgomeryAsync")

    def close(self):
        from . import lib
        if self.ctx:
            lib().xmpMontgomeryDestroy(self.handle.h, self.ctx)
            self.ctx = _vp()


class Batch:
    """Mixed-size persistent batch (xmpBatchCreate)."""

    def __init__(self, handle, outs, bases, exps, mods, counts):
        from . import check, lib
        n = len(outs)
        arr = lambda xs: (_vp * n)(*[x.x for x in xs])  # noqa: E731
        self.handle = handle
        self._keep = (outs, bases, exps, mods)
        self.b = _vp()
        cnt = (_u32 * n)(*counts)
        check(lib().xmpBatchCreate(handle.h, ctypes.byref(self.b), n, arr(outs), arr(bases), arr(exps), arr(mods),
                                   cnt), "xmpBatchCreate")

    def prepare(self):
        from . import check, lib
        check(lib().xmpBatchPrepare(self.handle.h, self.b), "xmpBatchPrepare")

    def launch_async(self):
        from . import check, lib
        check(lib().xmpBatchLaunchAsync(self.handle.h, self.b), "xmpBatchLaunchAsync")

    def launch_plain_async(self):
        from . import check, lib
        check(lib().xmpBatchLaunchPlainAsync(self.handle.h, self.b), "xmpBatchLaunchPlainAsync")

    def close(self):
        from . import lib
        if self.b:
            lib().xmpBatchDestroy(self.handle.h, self.b)
            self.b = _vp()


def tune(handle, precision, exp_precision, count):
    """Run the on-device tuner; returns the best entry as JSON text."""
    from . import check, lib
    buf = ctypes.create_string_buffer(4096)
    n = _u32()
    check(lib().xmpTunerRun(handle.h, precision, exp_precision, count, buf, 4096, ctypes.byref(n)), "xmpTunerRun")
    return buf.value.decode(), n.value


def tuner_load(json_text):
    from . import check, lib
    check(lib().xmpTunerLoad(json_text.encode()), "xmpTunerLoad")
