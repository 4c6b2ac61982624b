"""GPU: every compiled Montgomery lane layout, forced through a one-entry
layout table, against GMP.

The runtime normally runs only the layouts its measured table picks, so a
layout that no table entry selects would go untested.  Here each compiled
layout (lanes x digits x radix) runs at the widest standard size it holds and
at a narrower one, where the steps stop at the lanes that hold digits, in the
plain call, a Montgomery context (resident R^2 mod m, by squarings above 128
limbs) and the persistent batch.  Exponents are kept short (the exponent
schedule is tested elsewhere); the moduli, bases and products are full size.
"""
import json
import os

import pytest

import xmp_amd
from helpers import O, assert_rows_equal, gmp, inputs
from xmp_amd import ext

pytestmark = pytest.mark.gpu

# (lanes, digits, radix): (widest size, narrower size) in bits
LAYOUTS = {
    (1, 5, 28): (128, 96), (1, 10, 28): (256, 160), (1, 12, 28): (320, 288), (1, 20, 28): (512, 416),
    (1, 28, 28): (768, 544), (1, 37, 28): (1024, 800), (2, 19, 28): (1056, 544), (2, 28, 28): (1536, 1056),
    (2, 37, 28): (2048, 1088), (4, 19, 28): (2080, 1056), (4, 28, 28): (3072, 2080), (2, 55, 28): (3072, 2048),
    (4, 38, 27): (4096, 3072), (8, 19, 27): (4096, 2080), (8, 29, 27): (6144, 4096), (8, 38, 27): (8192, 6144),
    (16, 19, 27): (8192, 4096), (4, 10, 28): (1088, 544), (8, 5, 28): (1088, 512), (8, 10, 28): (2208, 1088),
    (16, 5, 28): (2208, 1024), (16, 3, 28): (1312, 512), (4, 37, 28): (4096, 3072), (8, 37, 28): (8256, 6144),
    (32, 20, 26): (16384, 12288), (64, 20, 26): (32768, 20480),
}
EXP_BITS = 96


def _load(entries):
    assert xmp_amd.lib().xmpTunerLoad(json.dumps({"device": "gfx950", "entries": entries}).encode()) == 0


def _restore():
    with open(os.path.join(os.path.dirname(xmp_amd.LIB_PATH), "tune_gfx950.json")) as f:
        assert xmp_amd.lib().xmpTunerLoad(f.read().encode()) == 0


def _entry(layout, bits, count):
    lanes, digits, radix = layout
    return {"precision": bits, "exp_bits": EXP_BITS, "count": count, "lanes": lanes, "digits": digits,
            "radix": radix, "window": 3, "ms": 1.0, "powm_per_s": 1.0}


def _case(seed, n, bits):
    m, b, e = inputs.powm_class(seed, n, bits, exp_bits=EXP_BITS)
    m[0, :] = 0xFFFFFFFF          # 2^bits - 1 (odd, every digit at its maximum)
    b[0, :] = 0xFFFFFFFF
    b[0, 0] = 0xFFFFFFFE          # m - 1
    e[0, :] = 0xFFFFFFFF
    return m, b, e


@pytest.mark.parametrize("layout", sorted(LAYOUTS))
def test_every_layout(handle, layout):
    lanes = layout[0]
    n = max(64 // lanes, 1) * 2 + 1   # two groups of waves and one instance more
    try:
        for bits in LAYOUTS[layout]:
            m, b, e = _case(700 + bits + lanes, n, bits)
            want = gmp(O.GO_POWM, n, bits // 32, b, e, m)
            _load([_entry(layout, bits, n)])
            M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
            out = handle.integers(bits, n)
            handle.powm(out, B, E, M)
            assert_rows_equal(out.get(), want, "%dx%d radix %d, %d bits" % (layout + (bits,)))
            ctx = ext.Montgomery(handle, M, EXP_BITS, n)
            try:
                assert tuple(ctx.layout()[:3]) == layout
                out2 = handle.integers(bits, n)
                ctx.powm_async(out2, B, E, n)
                assert_rows_equal(out2.get(), want, "context on %dx%d, %d bits" % (layout[:2] + (bits,)))
            finally:
                ctx.close()
    finally:
        _restore()


# Layouts with a quotient-free kernel (the reduction by m * (-m^-1 mod 2^D),
# taken when the active lanes hold D bits more than R >= 4m needs): the
# largest size that takes it and the largest that does not.
QUOTIENT_FREE = {
    (1, 20, 28): (512, 544), (2, 19, 28): (1024, 1056), (2, 28, 28): (1536, 768), (4, 19, 28): (2080, 2112),
    (4, 28, 28): (3104, 2336), (8, 29, 27): (6208, 6240), (32, 20, 26): (16608, 16096),
    (64, 20, 26): (33248, 32736), (8, 37, 28): (8256, 7232), (4, 10, 28): (1088, 832), (8, 5, 28): (1088, 960),
    (8, 10, 28): (2208, 1952), (16, 5, 28): (2208, 2080), (16, 3, 28): (1312, 1248),
}


def _spare_digit(layout, bits):
    """Whether the lanes holding a bits-wide modulus have D bits to spare."""
    lanes, digits, radix = layout
    nd = -(-(bits + 2) // radix)
    active = min(lanes, -(-nd // digits))
    return digits * radix * active >= bits + 2 + radix


@pytest.mark.parametrize("layout", sorted(QUOTIENT_FREE))
def test_quotient_free_boundary(handle, layout):
    """Each such layout on both sides of the boundary, against GMP."""
    lanes = layout[0]
    n = max(64 // lanes, 1) * 2 + 1
    on, off = QUOTIENT_FREE[layout]
    assert _spare_digit(layout, on) and not _spare_digit(layout, off)
    try:
        for bits in (on, off):
            m, b, e = _case(500 + bits + lanes, n, bits)
            want = gmp(O.GO_POWM, n, bits // 32, b, e, m)
            _load([_entry(layout, bits, n)])
            M, B, E = handle.from_array(m), handle.from_array(b), handle.from_array(e)
            out = handle.integers(bits, n)
            handle.powm(out, B, E, M)
            assert_rows_equal(out.get(), want, "%dx%d radix %d, %d bits" % (layout + (bits,)))
    finally:
        _restore()


@pytest.mark.parametrize("layouts", [((4, 37, 28), (8, 38, 27)), ((4, 38, 27), (16, 19, 27))])
def test_batch_on_wide_layouts(handle, layouts):
    """The persistent batch with a 4096-bit and an 8192-bit class: the 4x37
    radix-2^28 group with carry passes beside 8x38, and the radix-2^27
    layouts (8x37 is not a batch layout)."""
    sizes, counts = (4096, 8192), (21, 9)
    try:
        _load([_entry(lay, bits, c) for lay, bits, c in zip(layouts, sizes, counts)])
        outs, bases, exps, mods, wants = [], [], [], [], []
        for bits, c in zip(sizes, counts):
            m, b, e = _case(900 + bits, c, bits)
            wants.append(gmp(O.GO_POWM, c, bits // 32, b, e, m))
            mods.append(handle.from_array(m))
            bases.append(handle.from_array(b))
            exps.append(handle.from_array(e))
            outs.append(handle.integers(bits, c))
        batch = ext.Batch(handle, outs, bases, exps, mods, list(counts))
        try:
            batch.launch_async()
            for out, want, bits in zip(outs, wants, sizes):
                assert_rows_equal(out.get(), want, "batched %d-bit class" % bits)
        finally:
            batch.close()
    finally:
        _restore()
