"""GPU: mixed-size persistent batch (xmpBatch*) against GMP and the plain path.

Precision classes 512..8192 bits run in one persistent, graph-captured
launch.  Each replay is checked against GMP, with the inputs changed by eager
imports between replays (a graph must not replay stale state), and the
batched results must equal the plain per-class launches bit for bit.
"""
import numpy as np
import pytest

from helpers import O, assert_rows_equal, gmp, inputs
from xmp_amd.ext import Batch

pytestmark = pytest.mark.gpu

CLASSES = [(512, 150), (1024, 200), (2048, 120), (3072, 40), (4096, 40), (8192, 12)]


def _class_inputs(seed, bits, n):
    limbs = bits // 32
    m = inputs.odd_moduli(seed, n, limbs)
    b = inputs.below(seed + 1, m)
    e = inputs.uniform(seed + 2, n, limbs)
    return m, b, e


@pytest.mark.parametrize("classes", [CLASSES, [(2048, 300)], [(8192, 20)]], ids=["mixed", "one-2048", "one-8192"])
def test_batch_replays(handle, classes):
    """Mixed classes run in the persistent kernel; a one-class batch runs
    that class's own kernel in the captured graph."""
    CLASSES = classes
    arrays, host = [], []
    for k, (bits, n) in enumerate(CLASSES):
        m, b, e = _class_inputs(1000 + k, bits, n)
        host.append([m, b, e])
        arrays.append((handle.integers(bits, n), handle.from_array(b), handle.from_array(e), handle.from_array(m)))
    outs, bases, exps, mods = zip(*arrays)
    batch = Batch(handle, list(outs), list(bases), list(exps), list(mods), [n for _, n in CLASSES])
    try:
        for replay in range(3):
            if replay:
                # eager work between replays: new bases and exponents (same moduli, constants stay valid)
                for k, (bits, n) in enumerate(CLASSES):
                    _, b, e = _class_inputs(5000 + 17 * replay + k, bits, n)
                    b = inputs.below(5000 + 17 * replay + k, host[k][0])
                    host[k][1], host[k][2] = b, e
                    bases[k].load(b)
                    exps[k].load(e)
            batch.launch_async()
            for k, (bits, n) in enumerate(CLASSES):
                m, b, e = host[k]
                got = outs[k].get()
                want = gmp(O.GO_POWM, n, bits // 32, b, e, m)
                assert_rows_equal(got, want, "batch replay %d class %d-bit" % (replay, bits))
        # plain per-class launches give the same bits
        batched = [o.get() for o in outs]
        batch.launch_plain_async()
        for k in range(len(CLASSES)):
            assert np.array_equal(outs[k].get(), batched[k])
        # new moduli need the constants recomputed
        for k, (bits, n) in enumerate(CLASSES):
            m, b, e = _class_inputs(9000 + k, bits, n)
            host[k] = [m, b, e]
            mods[k].load(m)
            bases[k].load(b)
            exps[k].load(e)
        batch.prepare()
        batch.launch_async()
        for k, (bits, n) in enumerate(CLASSES):
            m, b, e = host[k]
            assert_rows_equal(outs[k].get(), gmp(O.GO_POWM, n, bits // 32, b, e, m), "after prepare %d" % bits)
    finally:
        batch.close()


def test_batch_wide_classes(handle):
    """Classes above 8192 bits run in the persistent kernel on the half-wave
    (32 x 20) and whole-wave (64 x 20) layouts, beside a 2048-bit class;
    512-bit exponents keep the GMP check short."""
    spec = [(2048, 2048, 64), (16384, 512, 8), (32768, 512, 4)]
    host, arrays = [], []
    for k, (bits, ebits, n) in enumerate(spec):
        m, b, e = inputs.powm_class(2100 + k, n, bits, ebits)
        host.append((m, b, e))
        arrays.append((handle.integers(bits, n), handle.from_array(b), handle.from_array(e), handle.from_array(m)))
    outs, bases, exps, mods = zip(*arrays)
    batch = Batch(handle, list(outs), list(bases), list(exps), list(mods), [n for _, _, n in spec])
    try:
        for replay in range(2):
            batch.launch_async()
            for k, (bits, ebits, n) in enumerate(spec):
                m, b, e = host[k]
                assert_rows_equal(outs[k].get(), gmp(O.GO_POWM, n, bits // 32, b, e, m),
                                  "batch replay %d, %d-bit class" % (replay, bits))
        batched = [o.get() for o in outs]
        batch.launch_plain_async()
        for k in range(len(spec)):
            assert np.array_equal(outs[k].get(), batched[k])
    finally:
        batch.close()
        for t in arrays:
            for x in t:
                x.close()


def test_batch_size_limit(handle):
    """Above 33,278 bits no lane layout holds an instance: a batch class of
    65,536 bits is refused with xmpErrorUnsupported (single calls take such
    moduli on the workgroup layout instead)."""
    import ctypes

    import xmp_amd
    arrs = [handle.integers(65536, 2) for _ in range(4)]
    b = ctypes.c_void_p()
    one = lambda x: (ctypes.c_void_p * 1)(x.x)  # noqa: E731
    rc = xmp_amd.lib().xmpBatchCreate(handle.h, ctypes.byref(b), 1, one(arrs[0]), one(arrs[1]), one(arrs[2]),
                                      one(arrs[3]), (ctypes.c_uint32 * 1)(2))
    assert rc == xmp_amd.UNSUPPORTED
    for x in arrs:
        x.close()


def test_batch_full_mixed_draw(handle):
    """The full mixed draw the benchmark times: 262,144 instances over
    512..8192 bits (seed 777).  The persistent launch equals the plain
    per-class launches on every row, and GMP agrees on 256 rows of every
    class spread over the class (first and last rows included)."""
    counts = inputs.mixed_counts(262144, 777)
    host, arrays = [], []
    for k, (bits, n) in enumerate(zip(inputs.MIXED_SIZES, counts)):
        m, b, e = inputs.powm_class(777 + 10 * k, n, bits)
        host.append((m, b, e))
        arrays.append((handle.integers(bits, n), handle.from_array(b), handle.from_array(e), handle.from_array(m)))
    outs, bases, exps, mods = zip(*arrays)
    batch = Batch(handle, list(outs), list(bases), list(exps), list(mods), counts)
    try:
        batch.launch_async()
        batched = [o.get() for o in outs]
        batch.launch_plain_async()
        for k, bits in enumerate(inputs.MIXED_SIZES):
            plain = outs[k].get()
            assert_rows_equal(batched[k], plain, "batched vs plain, %d-bit class of %d" % (bits, counts[k]))
            m, b, e = host[k]
            rows = np.unique(np.linspace(0, counts[k] - 1, num=256).astype(int))
            assert_rows_equal(batched[k][rows], gmp(O.GO_POWM, rows.size, bits // 32, b[rows], e[rows], m[rows]),
                              "batched vs GMP, %d-bit class" % bits)
    finally:
        batch.close()
        for t in arrays:
            for x in t:
                x.close()
