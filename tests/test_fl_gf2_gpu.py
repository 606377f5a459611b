"""GPU parity of mlz_gf2_fl_echelonize, FaugereLachartre<GF2> on hybrid rows
(lela/algorithms/faugere-lachartre.tcc:735-917, lela/vector/hybrid.h:30-33),
against the equality LELA's own test asserts: the result is the reduced
row-echelon form of X, with the same rank (tests/test-faugere-lachartre.C:
227-243, run over GF(2) at :278-285).  The reduced form is the oracle's
(orc_gf2_echelonize), bit for bit, with the pivot list; WrongMatrixForm at a
decreasing head, a zero row ending the scan, explicit zero words in the
rows, and a larger matrix through the all-row check."""
import numpy as np
import pytest

import bruteforce as bf
import lela_amd as la
import oracle_lib as orc
from test_gf2_gpu import check_gf2_rref_all_rows

pytestmark = pytest.mark.gpu


def f4_matrix_gf2(seed, m, n):
    """F4-shaped GF(2) matrix (bits) in the style of createRandomF4Matrix
    (test-faugere-lachartre.C:103-175): each row starts at a column that stays,
    advances by one or jumps; the row is random from that column, its first
    bit set."""
    rng = np.random.default_rng(seed)
    X = np.zeros((m, n), dtype=np.uint8)
    col = 0
    for i in range(m):
        v = int(rng.integers(0, 7))
        if 1 <= v <= 5:
            col += 1
        elif v == 6:
            hi = n - (m - i + 1)
            if hi > col:
                col = int(rng.integers(col, hi))
        if col >= n:
            break
        X[i, col:] = rng.integers(0, 2, size=n - col)
        X[i, col] = 1
    return X


def run(ctx, bits, hybrid=None):
    ctx.F = la.GF2()
    m, n = bits.shape
    R = la.Dense01Matrix(np.zeros((m, (n + 63) // 64), dtype=np.uint64), n)
    FL = la.FaugereLachartre(ctx)
    if hybrid is None:
        rank, det = FL.echelonize(R, la.Dense01Matrix(bf.gf2_pack(bits), n))
    else:
        rank, det = FL.echelonize(R, hybrid, shape=(m, n))
    return R.words, rank, det, FL.pivots


@pytest.mark.parametrize("m,n,seed", [(96, 128, 1), (24, 32, 2), (30, 30, 3), (300, 200, 4), (200, 330, 5),
                                      (500, 700, 6), (640, 1000, 7)])
def test_gf2_fl_matches_rref(gpu_ctx, m, n, seed):
    X = f4_matrix_gf2(seed, m, n)
    got, rank, det, piv = run(gpu_ctx, X)
    want, wrank, wpiv = orc.gf2_echelonize(m, n, bf.gf2_pack(X), reduced=True)
    assert rank == wrank and list(piv) == list(wpiv)
    assert np.array_equal(got, want) and det == 1


def test_gf2_fl_wrong_matrix_form(gpu_ctx):
    X = np.zeros((4, 70), dtype=np.uint8)
    X[0, 1] = X[1, 66] = X[2, 0] = X[3, 68] = 1
    with pytest.raises(la.WrongMatrixForm) as ei:
        run(gpu_ctx, X)
    assert ei.value.row == 2


def test_gf2_fl_zero_row_and_zero_words(gpu_ctx):
    # the first zero row ends the scan; rows after it join C|D (:64-70,
    # :137-142), a small head among them is no error; explicit zero words
    m, n = 120, 200
    X = f4_matrix_gf2(8, m, n)
    X[40] = 0
    X[70] = 0
    X[70, 3] = 1
    W = bf.gf2_pack(X)
    rows, blks, words = [0], [], []
    for i in range(m):
        for b in range(W.shape[1]):
            if W[i, b] or (i + b) % 3 == 0:   # a zero word now and then
                blks.append(b)
                words.append(W[i, b])
        rows.append(len(blks))
    hybrid = (np.array(rows, dtype=np.uint64), np.array(blks, dtype=np.uint32), np.array(words, dtype=np.uint64))
    got, rank, _, piv = run(gpu_ctx, X, hybrid)
    want, wrank, wpiv = orc.gf2_echelonize(m, n, W, reduced=True)
    assert rank == wrank and list(piv) == list(wpiv) and np.array_equal(got, want)


def test_gf2_fl_4000x6000(gpu_ctx):
    """4000 x 6000: the RREF structure of every row and X = X[:, piv] R at
    random vectors over every row."""
    m, n = 4000, 6000
    X = f4_matrix_gf2(9, m, n)
    print("gf2 fl 4000x6000", flush=True)
    got, rank, _, piv = run(gpu_ctx, X)
    check_gf2_rref_all_rows(bf.gf2_pack(X), got, piv, rank, n, seed=5)
