"""GPU: the autotuner (xmpTunerRun / xmpTunerLoad), the project's
replacement for the reference's offline tuner (src/tune/tune.cu:70-275).

A small run times every fitting layout, checks each one's results against
the library's own xmpIntegersPowm (the reference tuner's gold-standard check,
tune.cu:93-97,131-140), and writes JSON that parses and loads; the runtime
then picks the fastest measured layout, whose results are bit-exact against
GMP.  The shipped table is restored afterwards."""
import json
import os

import pytest

import xmp_amd
from helpers import O, assert_rows_equal, gmp, inputs
from xmp_amd import ext

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits,count", [(512, 256), (1024, 512), (2048, 256)])
def test_tuner_run_load_and_choice(handle, bits, count):
    text, n = ext.tune(handle, bits, bits, count)
    rows = json.loads("[" + text + "]")
    assert n == len(rows) and n >= 2
    for r in rows:
        assert r["precision"] == bits and r["count"] == count and r["powm_per_s"] > 0
        assert r["lanes"] * r["digits"] * r["radix"] >= bits + 2
    best = max(rows, key=lambda r: r["powm_per_s"])
    shipped = os.path.join(os.path.dirname(xmp_amd.LIB_PATH), "tune_gfx950.json")
    try:
        table = json.dumps({"device": "gfx950", "entries": rows})
        assert xmp_amd.lib().xmpTunerLoad(table.encode()) == 0
        m, b, e = inputs.powm_class(3300 + bits, count, bits)
        M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
        ctx = ext.Montgomery(handle, M, bits, count)
        try:
            lanes, digits, radix, window = ctx.layout()
            assert (lanes, digits, radix, window) == (best["lanes"], best["digits"], best["radix"], best["window"])
            out = handle.integers(bits, count)
            ctx.powm_async(out, B, E, count)
            assert_rows_equal(out.get(), gmp(O.GO_POWM, count, bits // 32, b, e, m),
                              "powm on the tuned layout %dx%d" % (lanes, digits))
            out.close()
        finally:
            ctx.close()
            for x in (M, B, E):
                x.close()
    finally:
        with open(shipped) as f:
            assert xmp_amd.lib().xmpTunerLoad(f.read().encode()) == 0


def _entry(bits, count, lanes, digits, ms, window=6):
    return {"precision": bits, "exp_bits": bits, "count": count, "lanes": lanes, "digits": digits, "radix": 28,
            "window": window, "ms": ms, "powm_per_s": count / (ms * 1e-3)}


@pytest.mark.parametrize("count,want", [(1000, (16, 3, 6)), (4096, (16, 3, 6)), (4097, (4, 10, 5)),
                                        (8192, (4, 10, 5)), (16384, (4, 10, 5)), (100000, (4, 10, 5))])
def test_choice_between_measured_batch_sizes(handle, count, want):
    """A batch between two measured sizes is priced at the larger one: 16x3 is
    faster at 4,096 (2.4 against 2.8 ms) and 4x10 at 16,384 (2.84 against
    6.0 ms), so every batch of 4,097 to 16,384 goes to 4x10 with the window
    measured at 16,384; below the smallest measured size the smallest decides,
    above the largest the largest."""
    rows = [_entry(1024, 4096, 16, 3, 2.4, 6), _entry(1024, 4096, 4, 10, 2.8, 6),
            _entry(1024, 16384, 16, 3, 6.0, 6), _entry(1024, 16384, 4, 10, 2.84, 5)]
    shipped = os.path.join(os.path.dirname(xmp_amd.LIB_PATH), "tune_gfx950.json")
    M = handle.from_array(inputs.odd_moduli(3400, 64, 32))
    try:
        assert xmp_amd.lib().xmpTunerLoad(json.dumps({"device": "gfx950", "entries": rows}).encode()) == 0
        ctx = ext.Montgomery(handle, M, 1024, count)
        try:
            lanes, digits, radix, window = ctx.layout()
        finally:
            ctx.close()
        assert (lanes, digits, window) == want
    finally:
        M.close()
        with open(shipped) as f:
            assert xmp_amd.lib().xmpTunerLoad(f.read().encode()) == 0
