"""Deterministic synthetic inputs shared by the tests, the smoke check and the bench.

Generator: SplitMix64 as a counter-based stream, draw i of stream `seed` is
mix(seed + (i+1)*0x9E3779B97F4A7C15); uniform residues in [0, p) by rejection
(draws at or above 2^64 - (2^64 mod p) are discarded).  Every generator here
draws from such streams (the structured ones through `Stream`, which hands
out consecutive draws), so the bytes do not depend on numpy's own generators
or its version.  Seeds follow 0x4C454C4100000000 + 16*w + matrix_index for
workload w.
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
    """Draws start .. start+count-1 of stream `seed` as uint64 (the mix of
    _mix, computed in place)."""
    with np.errstate(over="ignore"):
        z = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z *= GOLDEN
        z += np.uint64(seed)
        t = np.empty_like(z)
        np.right_shift(z, np.uint64(30), out=t)
        z ^= t
        z *= _M1
        np.right_shift(z, np.uint64(27), out=t)
        z ^= t
        z *= _M2
        np.right_shift(z, np.uint64(31), out=t)
        z ^= t
        return z


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


class Stream:
    """Consecutive draws of the SplitMix64 stream `seed`, shaped the way the
    generators below need them.  A draw is split into two 32-bit halves (low
    half first) where 32 bits suffice: integers in [lo, hi) for hi - lo <=
    2^32 by the multiply-shift method with rejection (x * span >> 32 for a
    half x, rejected when the low 32 bits of x * span fall below 2^32 mod
    span, so every value is equally likely); masks compare a half / 2^32 with
    the density.  Floats in [0, 1) from the top 53 bits of a draw."""

    def __init__(self, seed):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def draws(self, count):
        out = raw_draws(self.seed, self.pos, count)
        self.pos += count
        return out

    def halves(self, count):
        """`count` 32-bit halves as uint64 (ceil(count / 2) draws)."""
        d = self.draws((count + 1) // 2)
        h = np.empty(2 * d.size, dtype=np.uint64)
        np.bitwise_and(d, np.uint64(0xFFFFFFFF), out=h[0::2])
        np.right_shift(d, np.uint64(32), out=h[1::2])
        return h[:count]

    def random(self, shape):
        n = int(np.prod(shape))
        return ((self.draws(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))).reshape(shape)

    def mask(self, shape, density):
        """Boolean array, each entry true with probability `density` (to 2^-32)."""
        n = int(np.prod(shape))
        return (self.halves(n) < np.uint64(min(int(density * 4294967296.0), 1 << 32))).reshape(shape)

    def uniform(self, a, b):
        return a + (b - a) * float(self.random(1)[0])

    def integers(self, lo, hi, size):
        """Uniform integers in [lo, hi) (hi - lo <= 2^32) as int64, shape `size`."""
        span = int(hi) - int(lo)
        if not 0 < span <= (1 << 32):
            raise ValueError("integers: range must hold 1 .. 2^32 values")
        shape = (size,) if np.isscalar(size) else tuple(size)
        n = int(np.prod(shape))
        thr = np.uint64((1 << 32) % span)
        sp = np.uint64(span)
        m = self.halves(n) * sp
        bad = np.nonzero((m & np.uint64(0xFFFFFFFF)) < thr)[0]
        for i in bad:   # probability below span / 2^32 per value
            while True:
                x = self.halves(1)[0] * sp
                if (x & np.uint64(0xFFFFFFFF)) >= thr:
                    m[i] = x
                    break
        return (m >> np.uint64(32)).astype(np.int64).reshape(shape) + int(lo)


def zp_matrix(seed, rows, cols, p, dtype=np.uint32):
    """rows x cols matrix of uniform residues in [0, p)."""
    return uniform_mod(seed, rows * cols, p).astype(dtype).reshape(rows, cols)


def fl_matrix(seed, p, na=15000, nc=5000, nb=15000, band=600, nnz_a=150, density=0.02,
              chunk=500):
    """Faugere-Lachartre-shaped X = [A B; C D] in CSR (row_ptr, col, val uint32):
    A na x na upper triangular, A[i][i] = 1 plus nnz_a-1 distinct columns drawn
    from (i, min(i+band, na)); B, C, D entries nonzero with probability
    `density`, values uniform in [1, p).  Each C|D row whose first nonzero is in
    column h is placed right after A's row h (F4 row order); an all-zero C part
    is redrawn so the row cannot become a spurious pivot row.  A's extra
    columns are the nnz_a-1 candidates of smallest random key (a uniform
    subset).  All draws come from one SplitMix64 stream."""
    rng = Stream(seed)
    n = na + nb
    a_rows = []
    for start in range(0, na, chunk):
        stop = min(na, start + chunk)
        Bmask = rng.mask((stop - start, nb), density)
        # candidate offsets 1 .. band-1 right of the diagonal, those past na excluded
        keys = rng.random((stop - start, max(band - 1, 0)))
        for i in range(start, stop):
            hi = min(i + band, na)
            avail = hi - (i + 1)
            k = min(nnz_a - 1, avail)
            off = (np.argsort(keys[i - start, :avail], kind="stable")[:k] + i + 1 if k > 0
                   else np.zeros(0, np.int64))
            bcols = np.nonzero(Bmask[i - start])[0] + na
            cols = np.concatenate(([i], np.sort(off), bcols)).astype(np.uint32)
            vals = rng.integers(1, p, cols.size).astype(np.uint32)
            vals[0] = 1
            a_rows.append((cols, vals))
    c_rows = []
    for start in range(0, nc, chunk):
        stop = min(nc, start + chunk)
        Cmask = rng.mask((stop - start, na), density)
        Dmask = rng.mask((stop - start, nb), density)
        for t in range(stop - start):
            while not Cmask[t].any():
                Cmask[t] = rng.mask(na, density)
            cols = np.concatenate((np.nonzero(Cmask[t])[0], np.nonzero(Dmask[t])[0] + na))
            vals = rng.integers(1, p, cols.size).astype(np.uint32)
            c_rows.append((cols.astype(np.uint32), vals))
    # F4 order: C rows after the A row of their head, stable in generation order
    heads = np.array([r[0][0] for r in c_rows], dtype=np.int64)
    order = np.argsort(heads, kind="stable")
    rows = []
    ci = 0
    for i in range(na):
        rows.append(a_rows[i])
        while ci < nc and heads[order[ci]] == i:
            rows.append(c_rows[order[ci]])
            ci += 1
    lens = np.array([r[0].size for r in rows], dtype=np.uint64)
    row_ptr = np.zeros(len(rows) + 1, dtype=np.uint64)
    row_ptr[1:] = np.cumsum(lens)
    col = np.concatenate([r[0] for r in rows])
    val = np.concatenate([r[1] for r in rows])
    return (row_ptr, col, val), (len(rows), n)


def csr_to_dense(csr, shape, dtype=np.uint32):
    row_ptr, col, val = csr
    X = np.zeros(shape, dtype=dtype)
    rows = np.repeat(np.arange(shape[0]), np.diff(row_ptr.astype(np.int64)))
    X[rows, col.astype(np.int64)] = val
    return X


def gf2_matrix(seed, rows, cols):
    """rows x cols Bernoulli(1/2) bit matrix, packed LSB-first into uint64 words
    (row stride ceil(cols/64) words, unused tail bits zero)."""
    words = (cols + 63) // 64
    W = raw_draws(seed, 0, rows * words).reshape(rows, words)
    r = cols % 64
    if r:
        W[:, -1] &= np.uint64((1 << r) - 1)
    return W


def batch_instances(seed, count=4096, p=65521, lo=32, hi=1024, dtype=np.uint16, sparse_lo=0.01,
                    sparse_hi=0.20):
    """Uneven batch of independent Z/p instances (as dicts).

    Instance i has m, n, k uniform in [lo, hi].  i % 4 in {0, 1}: gemm
    C <- A*B + C (A m x k, B k x n, C m x n uniform); otherwise the reduced
    echelon form of an m x n matrix A.  Odd i are sparse: entries of A (and B)
    nonzero with probability d uniform in [sparse_lo, sparse_hi], values
    uniform in [1, p); even i are dense uniform residues (d = 1).  All draws
    come from one SplitMix64 stream."""
    rng = Stream(seed)
    out = []
    for i in range(count):
        m, n, k = (int(v) for v in rng.integers(lo, hi + 1, 3))
        sparse = i % 2 == 1
        d = rng.uniform(sparse_lo, sparse_hi) if sparse else 1.0

        def mat(r, c):
            if not sparse:
                return rng.integers(0, p, (r, c)).astype(dtype)
            mask = rng.mask((r, c), d)
            out = np.zeros((r, c), dtype=dtype)
            out[mask] = rng.integers(1, p, int(mask.sum())).astype(dtype)
            return out

        if i % 4 in (0, 1):
            out.append(dict(op=1, m=m, n=n, k=k, density=d, A=mat(m, k), B=mat(k, n),
                            C=rng.integers(0, p, (m, n)).astype(dtype)))
        else:
            out.append(dict(op=2, m=m, n=n, k=min(m, n), density=d, A=mat(m, n)))
    return out


def gemm_density_instances(seed, count, p=65521, lo=64, hi=1024, dmin=0.0005, dmax=0.2, dtype=np.uint16):
    """Gemm instances (as dicts, the form of batch_instances) whose A has a
    density log-uniform in [dmin, dmax]; B and C dense uniform residues,
    m, n, k uniform in [lo, hi].  All draws come from one SplitMix64 stream."""
    rng = Stream(seed)
    out = []
    for _ in range(count):
        m, n, k = (int(v) for v in rng.integers(lo, hi + 1, 3))
        d = float(np.exp(rng.uniform(np.log(dmin), np.log(dmax))))
        mask = rng.mask((m, k), d)
        A = np.zeros((m, k), dtype=dtype)
        A[mask] = rng.integers(1, p, int(mask.sum())).astype(dtype)
        out.append(dict(op=1, m=m, n=n, k=k, density=d, A=A,
                        B=rng.integers(0, p, (k, n)).astype(dtype),
                        C=rng.integers(0, p, (m, n)).astype(dtype)))
    return out
