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
