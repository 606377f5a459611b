"""Deterministic synthetic inputs shared by the tests, the smoke check and the bench.

Generator: SplitMix64 as a counter-based stream, draw i of stream `seed` is
mix(seed + (i+1)*0x9E3779B97F4A7C15); uniform residues in [0, p) by rejection
(draws at or above 2^64 - (2^64 mod p) are discarded), so the bytes do not
depend on numpy's own generators.  Seeds follow 0x4C454C4100000000 + 16*cfg +
matrix_index for the five workloads.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED_BASE = 0x4C454C4100000000
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def seed_for(cfg, matrix_index):
    return (SEED_BASE + 16 * cfg + matrix_index) & 0xFFFFFFFFFFFFFFFF


def _mix(z):
    z = z ^ (z >> np.uint64(30))
    z = z * _M1
    z = z ^ (z >> np.uint64(27))
    z = z * _M2
    return z ^ (z >> np.uint64(31))


def raw_draws(seed, start, count):
    """Draws start .. start+count-1 of stream `seed` as uint64."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        return _mix(np.uint64(seed) + idx * GOLDEN)


def uniform_mod(seed, count, p, chunk=1 << 22):
    """`count` residues uniform in [0, p) as uint64 (rejection sampling)."""
    p = int(p)
    limit = (1 << 64) - ((1 << 64) % p)
    out = np.empty(count, dtype=np.uint64)
    pos = 0
    for start in range(0, count, chunk):
        n = min(chunk, count - start)
        out[start:start + n] = raw_draws(seed, start, n)
    extra = count
    if limit < (1 << 64):
        lim = np.uint64(limit)
        bad = np.nonzero(out >= lim)[0]
        for i in bad:  # probability p/2^64 per draw: essentially never taken
            while True:
                z = raw_draws(seed, extra, 1)[0]
                extra += 1
                if z < lim:
                    out[i] = z
                    break
    del pos
    return out % np.uint64(p)


def zp_matrix(seed, rows, cols, p, dtype=np.uint32):
    """rows x cols matrix of uniform residues in [0, p)."""
    return uniform_mod(seed, rows * cols, p).astype(dtype).reshape(rows, cols)


def gf2_matrix(seed, rows, cols):
    """rows x cols Bernoulli(1/2) bit matrix, packed LSB-first into uint64 words
    (row stride ceil(cols/64) words, unused tail bits zero)."""
    words = (cols + 63) // 64
    W = raw_draws(seed, 0, rows * words).reshape(rows, words)
    r = cols % 64
    if r:
        W[:, -1] &= np.uint64((1 << r) - 1)
    return W
