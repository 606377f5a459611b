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


def fl_matrix(seed, p, na=15000, nc=5000, nb=15000, band=600, nnz_a=150, density=0.02,
              chunk=500):
    """Faugere-Lachartre-shaped X = [A B; C D] in CSR (row_ptr, col, val uint32):
    A na x na upper triangular, A[i][i] = 1 plus nnz_a-1 distinct columns drawn
    from (i, min(i+band, na)); B, C, D entries nonzero with probability
    `density`, values uniform in [1, p).  Each C|D row whose first nonzero is in
    column h is placed right after A's row h (F4 row order); an all-zero C part
    is redrawn so the row cannot become a spurious pivot row."""
    rng = np.random.default_rng(seed)
    n = na + nb
    a_rows = []
    for start in range(0, na, chunk):
        stop = min(na, start + chunk)
        Bmask = rng.random((stop - start, nb)) < density
        for i in range(start, stop):
            hi = min(i + band, na)
            avail = hi - (i + 1)
            k = min(nnz_a - 1, avail)
            off = rng.choice(avail, size=k, replace=False) + i + 1 if k > 0 else np.zeros(0, np.int64)
            bcols = np.nonzero(Bmask[i - start])[0] + na
            cols = np.concatenate(([i], np.sort(off), bcols)).astype(np.uint32)
            vals = rng.integers(1, p, size=cols.size, dtype=np.uint64).astype(np.uint32)
            vals[0] = 1
            a_rows.append((cols, vals))
    c_rows = []
    for start in range(0, nc, chunk):
        stop = min(nc, start + chunk)
        Cmask = rng.random((stop - start, na)) < density
        Dmask = rng.random((stop - start, nb)) < density
        for t in range(stop - start):
            while not Cmask[t].any():
                Cmask[t] = rng.random(na) < density
            cols = np.concatenate((np.nonzero(Cmask[t])[0], np.nonzero(Dmask[t])[0] + na))
            vals = rng.integers(1, p, size=cols.size, dtype=np.uint64).astype(np.uint32)
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
    uniform in [1, p); even i are dense uniform residues (d = 1)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m, n, k = (int(v) for v in rng.integers(lo, hi + 1, size=3))
        sparse = i % 2 == 1
        d = float(rng.uniform(sparse_lo, sparse_hi)) if sparse else 1.0

        def mat(r, c):
            if not sparse:
                return rng.integers(0, p, size=(r, c), dtype=np.int64).astype(dtype)
            mask = rng.random((r, c)) < d
            vals = rng.integers(1, p, size=(r, c), dtype=np.int64)
            return np.where(mask, vals, 0).astype(dtype)

        if i % 4 in (0, 1):
            out.append(dict(op=1, m=m, n=n, k=k, density=d, A=mat(m, k), B=mat(k, n),
                            C=rng.integers(0, p, size=(m, n), dtype=np.int64).astype(dtype)))
        else:
            out.append(dict(op=2, m=m, n=n, k=min(m, n), density=d, A=mat(m, n)))
    return out


def gemm_density_instances(seed, count, p=65521, lo=64, hi=1024, dmin=0.0005, dmax=0.2, dtype=np.uint16):
    """Gemm instances (as dicts, the form of batch_instances) whose A has a
    density log-uniform in [dmin, dmax]; B and C dense uniform residues,
    m, n, k uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        m, n, k = (int(v) for v in rng.integers(lo, hi + 1, size=3))
        d = float(np.exp(rng.uniform(np.log(dmin), np.log(dmax))))
        A = np.where(rng.random((m, k)) < d, rng.integers(1, p, size=(m, k), dtype=np.int64), 0).astype(dtype)
        out.append(dict(op=1, m=m, n=n, k=k, density=d, A=A,
                        B=rng.integers(0, p, size=(k, n), dtype=np.int64).astype(dtype),
                        C=rng.integers(0, p, size=(m, n), dtype=np.int64).astype(dtype)))
    return out
