"""Seeded, portable synthetic inputs (numpy PCG64 raw stream).

``numpy.random.PCG64(seed).random_raw`` is a fixed bit stream on every
platform and numpy version, so the same seed yields the same integers on the
build container, the GPU box and in the golden fixtures.  Arrays are uint32
``(count, limbs)``, little-endian limbs.
"""
import numpy as np


def raw_u32(seed, n):
    bg = np.random.PCG64(seed)
    words = bg.random_raw((n + 1) // 2).astype(np.uint64)
    out = np.empty(2 * words.size, np.uint32)
    out[0::2] = (words & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    out[1::2] = (words >> np.uint64(32)).astype(np.uint32)
    return out[:n]


def uniform(seed, count, limbs):
    """Uniform random limbs."""
    return raw_u32(seed, count * limbs).reshape(count, limbs)


def odd_moduli(seed, count, limbs, top_bit=True):
    """Uniform limbs with bit 0 (and the top bit) set."""
    m = uniform(seed, count, limbs).copy()
    m[:, 0] |= np.uint32(1)
    if top_bit:
        m[:, -1] |= np.uint32(0x80000000)
    return m


def below(seed, mods):
    """Uniform in [0, m): a uniform draw of the same limb count with its top
    limb reduced below the modulus' top limb (rejection-free, slightly
    non-uniform in the top limb; exact arithmetic is unaffected)."""
    count, limbs = mods.shape
    x = uniform(seed, count, limbs).copy()
    top = mods[:, -1].astype(np.uint64)
    x[:, -1] = (x[:, -1].astype(np.uint64) % np.maximum(top, 1)).astype(np.uint32)
    return x


def ordered_pair(seed, count, limbs):
    """(a, b) with a >= b elementwise (swap where needed)."""
    a = uniform(seed, count, limbs)
    b = uniform(seed + 1, count, limbs)
    # compare from the top limb
    gt = np.zeros(count, bool)
    decided = np.zeros(count, bool)
    for j in range(limbs - 1, -1, -1):
        d = (a[:, j] != b[:, j]) & ~decided
        gt[d] = a[d, j] > b[d, j]
        decided |= d
    swap = ~gt & decided
    a2, b2 = a.copy(), b.copy()
    a2[swap], b2[swap] = b[swap], a[swap]
    return a2, b2


# The mixed-size draw: modulus bit-lengths with their weights (exponents as
# long as the modulus).
MIXED_SIZES = (512, 1024, 2048, 3072, 4096, 8192)
MIXED_WEIGHTS = (0.10, 0.20, 0.35, 0.15, 0.15, 0.05)


def mixed_counts(total, seed):
    """Class sizes of a mixed batch: `total` draws over MIXED_SIZES with MIXED_WEIGHTS."""
    rng = np.random.Generator(np.random.PCG64(seed))
    draws = rng.choice(len(MIXED_SIZES), size=total, p=MIXED_WEIGHTS)
    return [int((draws == k).sum()) for k in range(len(MIXED_SIZES))]


def powm_class(seed, n, bits, exp_bits=None):
    """One precision class of powm inputs: odd moduli (top bit set), bases
    below them, uniform exponents of exp_bits (default: bits)."""
    limbs = bits // 32
    m = odd_moduli(seed, n, limbs)
    b = below(seed + 1, m)
    e = uniform(seed + 2, n, (exp_bits or bits) // 32)
    return m, b, e
