"""GPU parity of mlz_zp_fl_echelonize against the oracle's restatement of
FaugereLachartre::echelonize (bit-exact RREF, rank and det, whose D factor
comes from the heads of D's non-reduced echelon form),
the reference test's own assertion that FL equals Elimination's reduced form
(test-faugere-lachartre.C:227-243), WrongMatrixForm, and the 20000 x 30000
workload through size-independent checks."""
import ctypes

import numpy as np
import pytest

import lela_amd as la
import oracle_lib as orc
from lela_amd import synth
from test_oracle import f4_matrix

pytestmark = pytest.mark.gpu


def x_scan_det(X, p):
    """Product of the heads of the rows scanned by setup_splicer (:61-81,119)."""
    d = 1 % p
    for row in X:
        nz = np.nonzero(row)[0]
        if nz.size == 0:
            break
        d = d * int(row[nz[0]]) % p
    return d


def run_fl(ctx, p, storage, X, algo=la.ALGO_AUTO):
    F = la.Modular(p, storage)
    ctx.F = F
    R = np.zeros(X.shape, dtype=F.dtype)
    FL = la.FaugereLachartre(ctx)
    rank, det = FL.echelonize(R, F.from_canonical(X), algo=algo)
    return F.to_canonical(R), rank, det, FL.pivots


@pytest.mark.parametrize("p", [101, 65521, 2147483647])
@pytest.mark.parametrize("m,n,seed", [(24, 32, 1), (96, 128, 3), (30, 30, 4), (200, 260, 5), (300, 200, 6)])
def test_fl_matches_oracle(gpu_ctx, p, m, n, seed):
    X = f4_matrix(seed, m, n, p)
    got, rank, det, piv = run_fl(gpu_ctx, p, "uint32", X)
    rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X)
    assert rc == 0
    assert rank == wrank and np.array_equal(got, want)
    # the reference test's check: FL result == Elimination's reduced form
    rc2, E, erank, epiv, _ = orc.zp_echelonize(p, X, reduced=True)
    assert np.array_equal(got, E) and list(piv) == list(epiv)
    assert det == wdet


@pytest.mark.parametrize("storage,p", [("uint16", 65521), ("double", 67108859)])
def test_fl_storages(gpu_ctx, storage, p):
    X = f4_matrix(7, 150, 220, p)
    got, rank, det, _ = run_fl(gpu_ctx, p, storage, X)
    rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X)
    assert rank == wrank and np.array_equal(got, want) and det == wdet


@pytest.mark.parametrize("algo", [la.ALGO_PLAIN, la.ALGO_INT8])
def test_fl_small_f4_shape(gpu_ctx, algo):
    """The 20000 x 30000 F4-shaped generator at 1/20 scale, against the oracle."""
    p = 65521
    csr, shape = synth.fl_matrix(11, p, na=750, nc=250, nb=750, band=60, nnz_a=15)
    X = synth.csr_to_dense(csr, shape)
    got, rank, det, _ = run_fl(gpu_ctx, p, "uint32", X, algo=algo)
    rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X)
    assert rc == 0 and rank == wrank == 1000
    assert np.array_equal(got, want)
    assert det == wdet


@pytest.mark.parametrize("algo", [la.ALGO_AUTO, la.ALGO_PLAIN])
def test_fl_multiblock_banded_pivot_block(gpu_ctx, algo):
    """1000 pivot rows (two blocks of the pivot-block solve), banded A."""
    p = 65521
    csr, shape = synth.fl_matrix(13, p, na=1000, nc=300, nb=1000, band=90, nnz_a=20)
    X = synth.csr_to_dense(csr, shape)
    got, rank, det, _ = run_fl(gpu_ctx, p, "uint32", X, algo=algo)
    rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X)
    assert rc == 0 and rank == wrank
    assert np.array_equal(got, want) and det == wdet


def test_fl_multiblock_dense_pivot_block(gpu_ctx):
    """A dense upper-triangular pivot block spanning several solve blocks."""
    p = 65521
    X = f4_matrix(14, 1100, 1300, p)
    got, rank, det, _ = run_fl(gpu_ctx, p, "uint32", X)
    rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X)
    assert rc == 0 and rank == wrank and rank > 600
    assert np.array_equal(got, want) and det == wdet


@pytest.mark.parametrize("which", ["first", "middle"])
def test_fl_zero_columns_in_d(gpu_ctx, which):
    """Columns that are zero in B and D: D's pivot columns start after them
    ("first") or skip one ("middle"), both forms of the back-substitution's
    pivot-column handling."""
    p = 65521
    X = f4_matrix(21, 200, 260, p)
    heads = {int(np.nonzero(row)[0][0]) for row in X if row.any()}
    free = [j for j in range(X.shape[1]) if j not in heads]
    cols = free[:3] if which == "first" else free[len(free) // 3: len(free) // 3 + 2]
    X[:, cols] = 0
    got, rank, det, piv = run_fl(gpu_ctx, p, "uint32", X)
    rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X)
    assert rc == 0 and rank == wrank
    assert np.array_equal(got, want) and det == wdet


def test_fl_wrong_matrix_form(gpu_ctx):
    p = 101
    X = np.zeros((4, 6), dtype=np.uint32)
    X[0, 1] = 1
    X[1, 2] = 1
    X[2, 0] = 1
    X[3, 4] = 1
    with pytest.raises(la.WrongMatrixForm) as ei:
        run_fl(gpu_ctx, p, "uint32", X)
    assert ei.value.row == 2


def test_fl_zero_row_ends_scan(gpu_ctx):
    # rows after the first zero row join the C|D part (faugere-lachartre.tcc:64-70, :137-142)
    p = 101
    X = f4_matrix(12, 40, 50, p)
    X[10] = 0
    X[25, :] = 0
    X[25, 3] = 5                   # a small head after the zero row: not WrongMatrixForm
    got, rank, det, _ = run_fl(gpu_ctx, p, "uint32", X)
    rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X)
    assert rc == 0 and rank == wrank and np.array_equal(got, want) and det == wdet


def csr_matvec_mod(csr, shape, v, p):
    """(X @ v) mod p for X in CSR with entries < p < 2^32 and v < p (exact)."""
    row_ptr, col, val = csr
    prod = (val.astype(np.uint64) * v[col.astype(np.int64)].astype(np.uint64)) % np.uint64(p)
    rows = np.repeat(np.arange(shape[0]), np.diff(row_ptr.astype(np.int64)))
    out = np.zeros(shape[0], dtype=np.uint64)
    np.add.at(out, rows, prod)          # at most n terms below 2^32 per row
    return out % np.uint64(p)


def check_rref_all_rows(R, piv, rank, csr, shape, p, seed, chunk=1000):
    """Every row of R: the standard RREF structure (rows past the rank zero,
    row t's first nonzero a 1 in column piv[t], the pivot columns the identity
    on the first rank rows), and X = X[:, piv] R at a random vector over every
    row of X: X v == X[:, piv] (R v) mod p."""
    m, n = shape
    piv = piv.astype(np.int64)
    assert rank == piv.size and (rank == 0 or np.all(np.diff(piv) > 0))
    assert not R[rank:].any()
    rng = np.random.default_rng(seed)
    v = rng.integers(0, p, size=n, dtype=np.uint64)
    Rv = np.zeros(rank, dtype=np.uint64)
    P = np.uint64(p)
    for s0 in range(0, rank, chunk):
        blk = R[s0:min(rank, s0 + chunk)]
        nzfirst = (blk != 0).argmax(axis=1)
        assert np.array_equal(nzfirst, piv[s0:s0 + blk.shape[0]]), f"row heads near row {s0}"
        sub = blk[:, piv]
        assert np.count_nonzero(sub) == blk.shape[0], f"pivot columns near row {s0}"
        assert np.all(sub[np.arange(blk.shape[0]), np.arange(s0, s0 + blk.shape[0])] == 1)
        Rv[s0:s0 + blk.shape[0]] = (blk.astype(np.uint64) @ v) % P   # n terms < 2^32: below 2^64
    lhs = csr_matvec_mod(csr, shape, v, p)
    # X[:, piv] @ Rv: the entries of X in pivot columns, weighted by Rv at their position
    pos = np.full(n, -1, dtype=np.int64)
    pos[piv] = np.arange(rank)
    row_ptr, col, val = csr
    w = np.zeros(n, dtype=np.uint64)
    w[piv] = Rv
    sel = pos[col.astype(np.int64)] >= 0
    row_ptr2 = np.zeros(m + 1, dtype=np.uint64)
    rows = np.repeat(np.arange(m), np.diff(row_ptr.astype(np.int64)))
    np.add.at(row_ptr2, rows[sel] + 1, 1)
    csr2 = (np.cumsum(row_ptr2, dtype=np.uint64), col[sel], val[sel])
    rhs = csr_matvec_mod(csr2, shape, w, p)
    assert np.array_equal(lhs, rhs), "a row of X is not X[:, piv] R"


def test_fl_20000x30000_full(gpu_ctx):
    """20000 x 30000, p = 65521: rank 20000, the RREF structure of every row
    of R, and every row of X equals X[:, piv] R (checked at a random vector)."""
    p = 65521
    csr, shape = synth.fl_matrix(synth.seed_for(4, 0), p)
    F = la.Modular(p, "uint32")
    gpu_ctx.F = F
    FL = la.FaugereLachartre(gpu_ctx)
    R = np.zeros(shape, dtype=np.uint32)
    print("fl 20000 x 30000", flush=True)
    rank, det = FL.echelonize(R, csr, shape=shape)
    assert rank == 20000
    print("all-row check", flush=True)
    check_rref_all_rows(R, FL.pivots, rank, csr, shape, p, seed=11)

def test_fl_banded_inverse_solve(gpu_ctx):
    """A pivot block of 2100 rows (two full 1024-row blocks inverted up
    front, a 52-row tail on the 64-row chain) with a 600-wide band: the whole
    reduction bit-exact against the oracle."""
    p = 65521
    csr, shape = synth.fl_matrix(synth.seed_for(4, 5), p, na=2100, nc=100, nb=200, band=600, nnz_a=150)
    X = synth.csr_to_dense(csr, shape)
    got, rank, det, piv = run_fl(gpu_ctx, p, "uint32", X)
    rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X)
    assert rc == 0 and rank == wrank and np.array_equal(got, want) and det == wdet


@pytest.mark.parametrize("storage,p", [("uint32", 65521), ("uint16", 65521), ("double", 67108859),
                                       ("uint32", 4294967291)])
def test_fl_compact_output(gpu_ctx, storage, p):
    """mlz_zp_fl_echelonize_compact: the pivots, det and the rows of R on the
    non-pivot columns equal the dense reduced form's (the oracle's), for
    shapes with and without a D part and with zero columns."""
    F = la.Modular(p, storage)
    gpu_ctx.F = F
    for seed, kw in enumerate([dict(na=150, nc=60, nb=120, band=40, nnz_a=8, density=0.05),
                               dict(na=300, nc=0, nb=80, band=30, nnz_a=5, density=0.1),
                               dict(na=64, nc=200, nb=300, band=20, nnz_a=4, density=0.03)]):
        csr, shape = synth.fl_matrix(40 + seed, p, **kw)
        X = synth.csr_to_dense(csr, shape)
        rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X.astype(np.uint32))
        assert rc == 0
        FL = la.FaugereLachartre(gpu_ctx)
        rank, det, piv, NP = FL.echelonize_compact((csr[0], csr[1], F.from_canonical(csr[2])), shape=shape)
        free = np.setdiff1d(np.arange(shape[1]), piv)
        assert rank == wrank and det == wdet
        assert np.array_equal(want[np.arange(rank)[:, None], piv[None, :]], np.eye(rank, dtype=np.uint32))
        assert np.array_equal(F.to_canonical(NP), want[:rank][:, free]), (seed, storage)


def test_fl_compact_capacity(gpu_ctx):
    p = 65521
    F = la.Modular(p, "uint32")
    gpu_ctx.F = F
    csr, shape = synth.fl_matrix(7, p, na=100, nc=30, nb=60, band=20, nnz_a=5, density=0.05)
    m, n = shape
    NP = np.zeros(10, dtype=np.uint32)
    rank = ctypes.c_size_t(0)
    rc = gpu_ctx.lib.mlz_zp_fl_echelonize_compact(
        gpu_ctx.handle, F.elem, p, m, n, np.ascontiguousarray(csr[0]).ctypes.data,
        np.ascontiguousarray(csr[1]).ctypes.data, np.ascontiguousarray(csr[2], dtype=np.uint32).ctypes.data,
        NP.ctypes.data, NP.size, ctypes.byref(rank), None, None, None, 0)
    assert rc == 1 and rank.value * (n - rank.value) > NP.size and "compact block" in gpu_ctx.last_error()
