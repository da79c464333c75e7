"""Order-sensitive 64-bit hash of a process list (used by fixtures and tests).

h = sum_p key_p * M^(n-1-p)  (mod 2^64), key_p = (i2 << 42) ^ (i3 << 21) ^ i4,
so swapping two entries or changing one changes the hash.
"""
import numpy as np

M = np.uint64(0x9E3779B97F4A7C15)


def list_hash(i2, i3, i4) -> int:
    i2, i3, i4 = (np.asarray(x, dtype=np.uint64) for x in (i2, i3, i4))
    n = i2.size
    if n == 0:
        return 0
    key = (i2 << np.uint64(42)) ^ (i3 << np.uint64(21)) ^ i4
    pw = np.empty(n, dtype=np.uint64)
    pw[0] = 1
    if n > 1:
        with np.errstate(over="ignore"):
            pw[1:] = M
            pw = np.cumprod(pw, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum(key * pw[::-1], dtype=np.uint64))


def segment_hashes(N, lists):
    """Hashes of each mode's slice of the concatenated lists (counts N)."""
    acc = np.concatenate([[0], np.cumsum(N)])
    return np.array([list_hash(*(x[acc[m]:acc[m + 1]] for x in lists)) for m in range(len(N))], dtype=np.uint64)
