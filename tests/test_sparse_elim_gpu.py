"""GPU parity of mlz_zp_echelonize_sparse (Elimination with LELA's
SparsePartialPivotStrategy, the rule its sparse, sparse 0-1 and hybrid rows
select) against the oracle's restatement, bit for bit: the output matrix,
rank, determinant and transposition lists, for the three modes (U, the
compute_L transform, pluq), every storage, moduli from 2 to 2^32-5, entry and
word counts, shapes from empty to 700 x 900, ties, zero rows and columns, and
the composite-modulus error."""
import numpy as np
import pytest

import lela_amd as la
import oracle_lib as orc

pytestmark = pytest.mark.gpu

MODES = {"U": 0, "transform": 1, "pluq": 2}


def _sparse(seed, m, n, dens, p):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((m, n)) < dens, rng.integers(1, p, size=(m, n)), 0).astype(np.uint32)


def _inputs(p):
    out = [("hand", np.array([[1, 2, 3, 4, 5], [0, 1, 1, 0, 0], [7 % p, 0, 0, 0, 1], [0, 1, 0, 1, 0]],
                             dtype=np.uint32) % p),
           ("20x30 at 20%", _sparse(1, 20, 30, 0.2, p)),
           ("120x90 at 5%", _sparse(2, 120, 90, 0.05, p)),
           ("90x300 at 3%", _sparse(3, 90, 300, 0.03, p)),
           ("300x260 at 1%", _sparse(4, 300, 260, 0.01, p)),
           ("dense 64x64", _sparse(5, 64, 64, 1.0, p)),
           ("700x900 at 0.5%", _sparse(6, 700, 900, 0.005, p))]
    Z = _sparse(7, 150, 200, 0.04, p)
    Z[:, [0, 1, 63, 64, 130]] = 0
    Z[[3, 70]] = 0
    out.append(("zero rows and columns", Z))
    # many rows share a first column and a count: ties go to the first
    T = np.zeros((80, 120), dtype=np.uint32)
    T[:, 0] = 1
    T[np.arange(80), 1 + np.arange(80) % 7] = 2 % p or 1
    out.append(("ties", T % p))
    return out


def run(ctx, p, storage, A, mode, words):
    F = la.Modular(p, storage)
    ctx.F = F
    X = F.from_canonical(A)
    E = la.Elimination(ctx)
    PS = la.SparsePartialPivotStrategy(words)
    if mode == "pluq":
        P, Q, rank, det = E.pluq(X, PS=PS)
        return F.to_canonical(X), rank, det, P, Q
    P, rank, det = E.echelonize(X, compute_L=(mode == "transform"), PS=PS)
    return F.to_canonical(X), rank, det, P, None


def lela_lists(perm):
    return [(i, int(j)) for i, j in enumerate(perm) if j != i]


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("p,storage", [(65521, "uint16"), (3, "uint16"), (2147483647, "uint32"),
                                       (4294967291, "uint32"), (67108859, "double"), (101, "double")])
def test_sparse_elim_matches_oracle(gpu_ctx, p, storage, mode):
    for name, A in _inputs(p):
        got, rank, det, P, Q = run(gpu_ctx, p, storage, A, mode, False)
        rc, want, wrank, wdet, pr, pc = orc.zp_elim_sparse(p, A, MODES[mode])
        case = (name, p, storage, mode)
        assert rc == 0, case
        assert rank == wrank and det == wdet, case
        assert np.array_equal(got, want), case
        if mode == "pluq":
            assert P == lela_lists(pr)[::-1] and Q == lela_lists(pc)[::-1], case
        else:
            assert P == lela_lists(pr), case


@pytest.mark.parametrize("mode", ["U", "transform"])
def test_sparse_elim_word_counts_gf2(gpu_ctx, mode):
    """GF(2) hybrid rows as Z/2 entries: the count is of nonzero 64-column words."""
    for seed, (m, n, dens) in enumerate([(40, 200, 0.03), (150, 700, 0.004), (64, 130, 0.1), (300, 257, 0.01)]):
        A = _sparse(50 + seed, m, n, dens, 2)
        for words in (True, False):
            got, rank, det, P, _ = run(gpu_ctx, 2, "uint16", A, mode, words)
            rc, want, wrank, wdet, pr, _ = orc.zp_elim_sparse(2, A, MODES[mode], words)
            case = (m, n, dens, words, mode)
            assert rc == 0 and rank == wrank and np.array_equal(got, want) and P == lela_lists(pr), case


def test_sparse_elim_edges(gpu_ctx):
    F = la.Modular(101, "uint32")
    gpu_ctx.F = F
    E = la.Elimination(gpu_ctx)
    PS = la.SparsePartialPivotStrategy()
    for shape in ((0, 5), (5, 0), (0, 0)):
        X = np.zeros(shape, dtype=np.uint32)
        assert E.echelonize(X, PS=PS) == ([], 0, 1)
    X = np.zeros((6, 9), dtype=np.uint32)
    assert E.echelonize(X, compute_L=False, PS=PS) == ([], 0, 1) and not X.any()
    X = np.array([[0, 0, 5]], dtype=np.uint32)
    assert E.echelonize(X, compute_L=False, PS=PS)[1:] == (1, 5)
    # composite modulus: a pivot without an inverse
    G = la.Modular(12, "uint32")
    gpu_ctx.F = G
    X = np.array([[2, 1], [1, 1]], dtype=np.uint32)
    with pytest.raises(la.LELAError):
        la.Elimination(gpu_ctx).echelonize(X, PS=PS)
    rc, *_ = orc.zp_elim_sparse(12, np.array([[2, 1], [1, 1]], dtype=np.uint32), 1)
    assert rc == orc.ORC_E_NOT_INVERTIBLE
