"""Seeded synthetic polynomial coefficients (counter-based splitmix64).

value_i = lo + floor(mix64(seed + (i+1) * 0x9E3779B97F4A7C15) * (hi - lo) / 2^64)

Values are uniform in [lo, hi) (exact power-of-two spans are unbiased; other
spans carry a bias below (hi-lo)/2^64).  The same definition is used by the
test oracle, so a product measured here can be re-derived there.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED_BASE = 0x6770756666740000


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, lo, hi, count):
    """int64 array of `count` values uniform in [lo, hi)."""
    span = np.uint64(hi - lo)
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        r = mix64(np.uint64(seed) + idx * GOLDEN)
    # high 64 bits of r * span, via 32-bit halves
    r_hi, r_lo = r >> np.uint64(32), r & np.uint64(0xFFFFFFFF)
    s_hi, s_lo = span >> np.uint64(32), span & np.uint64(0xFFFFFFFF)
    with np.errstate(over="ignore"):
        ll = r_lo * s_lo
        lh = r_lo * s_hi
        hl = r_hi * s_lo
        hh = r_hi * s_hi
        mid = (ll >> np.uint64(32)) + (lh & np.uint64(0xFFFFFFFF)) + (hl & np.uint64(0xFFFFFFFF))
        high = hh + (lh >> np.uint64(32)) + (hl >> np.uint64(32)) + (mid >> np.uint64(32))
    return (np.int64(lo) + high.astype(np.int64)).astype(np.int64)
