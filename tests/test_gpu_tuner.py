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
