"""Shared helpers for the parity tests (the oracle is the checker)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from xmp_amd import inputs  # noqa: E402

THREADS = min(16, os.cpu_count() or 1)


def gmp(op, count, out_limbs, a, b=None, c=None, out2_limbs=0):
    out, out2, _ = O.gmp_batch(op, count, out_limbs, a, b, c, out2_limbs=out2_limbs, threads=THREADS)
    return out if out2 is None else (out, out2)


def first_mismatch(got, want):
    bad = np.nonzero((got != want).any(axis=1))[0]
    return None if bad.size == 0 else int(bad[0])


def assert_rows_equal(got, want, what):
    i = first_mismatch(got, want)
    assert i is None, "%s: %d/%d rows differ, first at %d: got %s want %s" % (
        what, int((got != want).any(axis=1).sum()), got.shape[0], i, got[i][:6], want[i][:6])
