"""Autotuner driver: measure every compiled Montgomery layout on this device.

Replaces the reference's offline tuner (src/tune/tune.cu:160-275), which swept
blocks/SM per algorithm on one broadcast modulus and compiled the resulting
latency tables into the library (src/tune/tune_maxwell.h, src/powm.cu:117-224).
Here, for each precision class and batch size, xmpTunerRun times every lane
layout (lanes x digits x radix) and the window widths around the modelled best
on random per-instance moduli; the table is written as JSON next to the
library (xmp_amd/tune_gfx950.json), where the runtime loads it by default
($XMP_TUNE_FILE overrides).  The choice for a call is the fastest layout at
the smallest measured batch size that holds its batch.

Usage (on the GPU):  python -m xmp_amd.tune [--out PATH] [--quick]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import xmp_amd  # noqa: E402
from xmp_amd import ext  # noqa: E402

# (precision, exponent precision, batch sizes)
PLAN = [
    (128, 128, [4096, 65536, 262144]),
    (256, 256, [4096, 65536]),
    (512, 512, [4096, 65536, 262144]),
    (768, 768, [16384]),
    (1024, 1024, [1024, 2048, 4096, 6144, 8192, 12288, 16384, 32768, 65536]),
    (1536, 1536, [16384]),
    (2048, 2048, [1024, 2048, 4096, 6144, 8192, 12288, 16384, 32768, 65536]),
    (3072, 3072, [4096, 16384]),
    (4096, 4096, [4096, 16384]),
    (6144, 6144, [4096]),
    (8192, 8192, [4096, 16384]),
    (16384, 16384, [1024]),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tune_gfx950.json"))
    ap.add_argument("--quick", action="store_true", help="one batch size per precision")
    args = ap.parse_args()
    entries = []
    with xmp_amd.Handle() as h:
        for bits, ebits, counts in PLAN:
            for count in (counts[-1:] if args.quick else counts):
                t = time.time()
                text, n = ext.tune(h, bits, ebits, count)
                rows = json.loads("[" + text + "]")
                entries += rows
                best = max(rows, key=lambda r: r["powm_per_s"])
                print("%5d-bit x %6d: %2d measured in %5.1f s, best %dx%d radix %d window %d: %.0f powm/s" % (
                    bits, count, n, time.time() - t, best["lanes"], best["digits"], best["radix"], best["window"],
                    best["powm_per_s"]), flush=True)
    # keep the best window per (size, batch, layout): the runtime chooses among layouts
    best = {}
    for r in entries:
        key = (r["precision"], r["exp_bits"], r["count"], r["lanes"], r["digits"], r["radix"])
        if key not in best or r["powm_per_s"] > best[key]["powm_per_s"]:
            best[key] = r
    entries = sorted(best.values(), key=lambda r: (r["precision"], r["count"], -r["powm_per_s"]))
    table = {"device": "gfx950",
             "note": "measured by xmp_amd/tune.py (xmpTunerRun) on one MI355X; best window per layout and batch size",
             "entries": entries}
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", args.out, len(entries), "entries")


if __name__ == "__main__":
    main()
