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
        "xmpMontgomeryGetR2": [_vp, _vp, _vp],
        "xmpIntegersPowmMontgomeryAsync": [_vp, _vp, _vp, _vp, _vp, _u32],
        "xmpBatchCreate": [_vp, ctypes.POINTER(_vp), _u32, _vp, _vp, _vp, _vp, _vp],
        "xmpBatchDestroy": [_vp, _vp],
        "xmpBatchPrepare": [_vp, _vp],
        "xmpBatchLaunchAsync": [_vp, _vp],
        "xmpBatchLaunchPlainAsync": [_vp, _vp],
        "xmpTunerRun": [_vp, _u32, _u32, _u32, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_u32)],
        "xmpTunerLoad": [ctypes.c_char_p],
        "xmpIntegersPowmNoAllocAsync": [_vp, _vp, _vp, _vp, _vp, _u32],
        "xmpProbeMadRate": [_vp, _u32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
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

    def r2(self):
        """The resident R^2 mod m as a (count, limbs) array."""
        from . import check, lib
        out = self.handle.integers(self.mod.precision, self.mod.count)
        try:
            check(lib().xmpMontgomeryGetR2(self.handle.h, self.ctx, out.x), "xmpMontgomeryGetR2")
            return out.get()
        finally:
            out.close()

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
    buf = ctypes.create_string_buffer(1 << 18)
    n = _u32()
    check(lib().xmpTunerRun(handle.h, precision, exp_precision, count, buf, len(buf), ctypes.byref(n)), "xmpTunerRun")
    return buf.value.decode(), n.value


def probe_mad_rate(handle, waves_per_simd):
    """Measured v_mad_u64_u32 multiply-adds per clock per CU (xmpProbeMadRate)."""
    from . import check, lib
    r, t = ctypes.c_double(), ctypes.c_double()
    check(lib().xmpProbeMadRate(handle.h, waves_per_simd, ctypes.byref(r), ctypes.byref(t)), "xmpProbeMadRate")
    return r.value


def tuner_load(json_text):
    from . import check, lib
    check(lib().xmpTunerLoad(json_text.encode()), "xmpTunerLoad")


class ServerConfig(ctypes.Structure):
    _fields_ = [("streams", _u32), ("max_batch", _u32), ("hbm_budget", ctypes.c_size_t), ("reserved_cus", _u32),
                ("layout_scale", _u32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("slot_capacity", _u32), ("max_inflight", _u32),
                ("peak_inflight", _u32), ("peak_batch", _u32), ("bytes_per_request", ctypes.c_size_t),
                ("hbm_budget", ctypes.c_size_t)]


def declare_server(L):
    E = ctypes.c_int
    u32p = ctypes.POINTER(_u32)
    sig = {
        "xmpServerCreate": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(ServerConfig), _u32, _u32, _vp, _vp, _vp, _vp,
                            _vp, _vp, _vp],
        "xmpServerDestroy": [_vp],
        "xmpServerSubmit": [_vp, _u32, _u32, _vp, _vp, ctypes.c_uint64],
        "xmpServerPoll": [_vp, _u32, _vp, _vp, _vp, u32p],
        "xmpServerDrain": [_vp],
        "xmpServerGetStats": [_vp, ctypes.POINTER(ServerStats)],
        "xmpServerRunPoisson": [_vp, ctypes.c_double, _u32, ctypes.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp,
                                ctypes.POINTER(ctypes.c_double)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = E


def make_keys(nkeys, bits, seed):
    """RSA-style key material with random odd coprime halves p, q (bits/2
    each, top bit set), n = p*q, random half-size exponents dp, dq and the
    CRT coefficients cp = q*(q^-1 mod p), cq = p*(p^-1 mod q)."""
    import math
    import numpy as np
    from . import inputs
    h32, n32 = bits // 64, bits // 32
    raw = inputs.uniform(seed, 8 * nkeys, h32)
    keys = {k: np.zeros((nkeys, w), np.uint32) for k, w in
            (("n", n32), ("p", h32), ("q", h32), ("dp", h32), ("dq", h32), ("cp", n32), ("cq", n32))}
    to_int = lambda row: int.from_bytes(row.astype("<u4").tobytes(), "little")  # noqa: E731
    to_row = lambda v, w: np.frombuffer(v.to_bytes(4 * w, "little"), "<u4")  # noqa: E731
    r = 0
    for k in range(nkeys):
        while True:
            p = to_int(raw[r]) | 1 | (1 << (bits // 2 - 1))
            q = to_int(raw[r + 1]) | 1 | (1 << (bits // 2 - 1))
            r += 2
            if p != q and math.gcd(p, q) == 1:
                break
        n = p * q
        keys["n"][k] = to_row(n, n32)
        keys["p"][k] = to_row(p, h32)
        keys["q"][k] = to_row(q, h32)
        keys["dp"][k] = raw[r]
        keys["dq"][k] = raw[r + 1]
        r += 2
        keys["cp"][k] = to_row(q * pow(q, -1, p), n32)
        keys["cq"][k] = to_row(p * pow(p, -1, q), n32)
    return keys


class Server:
    """Continuous-batching powm server over resident keys (xmpServerCreate)."""

    def __init__(self, handle, keys, bits, streams=4, max_batch=1024, hbm_budget=0, reserved_cus=0, layout_scale=1):
        import numpy as np
        from . import check, lib
        L = lib()
        declare_server(L)
        self.handle, self.bits, self.n32 = handle, bits, bits // 32
        self.keys = {k: np.ascontiguousarray(v, np.uint32) for k, v in keys.items()}
        cfg = ServerConfig(streams, max_batch, hbm_budget, reserved_cus, layout_scale)
        self.s = _vp()
        kk = self.keys
        check(L.xmpServerCreate(handle.h, ctypes.byref(self.s), ctypes.byref(cfg), bits, kk["n"].shape[0],
                                *[kk[x].ctypes.data for x in ("n", "p", "q", "dp", "dq", "cp", "cq")]),
              "xmpServerCreate")

    def submit(self, kind, key, base, exp, tag):
        import numpy as np
        from . import check, lib
        b = np.ascontiguousarray(base, np.uint32)
        e = np.ascontiguousarray(exp, np.uint32)
        check(lib().xmpServerSubmit(self.s, kind, key, b.ctypes.data, e.ctypes.data, tag), "xmpServerSubmit")

    def drain(self):
        from . import check, lib
        check(lib().xmpServerDrain(self.s), "xmpServerDrain")

    def poll(self, max_n=4096):
        import numpy as np
        from . import check, lib
        tags = np.zeros(max_n, np.uint64)
        res = np.zeros((max_n, self.n32), np.uint32)
        lat = np.zeros(max_n, np.float64)
        n = _u32()
        check(lib().xmpServerPoll(self.s, max_n, tags.ctypes.data, res.ctypes.data, lat.ctypes.data,
                                  ctypes.byref(n)), "xmpServerPoll")
        k = n.value
        return tags[:k], res[:k], lat[:k]

    def stats(self):
        """Admission counters (xmpServerGetStats) as a dict."""
        from . import check, lib
        st = ServerStats()
        check(lib().xmpServerGetStats(self.s, ctypes.byref(st)), "xmpServerGetStats")
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def run_poisson(self, rate, kinds, keys, bases, exps, seed=1):
        import numpy as np
        from . import check, lib
        n = len(kinds)
        kinds = np.ascontiguousarray(kinds, np.uint32)
        keys = np.ascontiguousarray(keys, np.uint32)
        bases = np.ascontiguousarray(bases, np.uint32)
        exps = np.ascontiguousarray(exps, np.uint32)
        res = np.zeros((n, self.n32), np.uint32)
        lat = np.zeros(n, np.float64)
        wall = ctypes.c_double()
        check(lib().xmpServerRunPoisson(self.s, float(rate), n, seed, kinds.ctypes.data, keys.ctypes.data,
                                        bases.ctypes.data, exps.ctypes.data, res.ctypes.data, lat.ctypes.data,
                                        ctypes.byref(wall)), "xmpServerRunPoisson")
        return res, lat, wall.value

    def close(self):
        from . import lib
        if self.s:
            lib().xmpServerDestroy(self.s)
            self.s = _vp()

