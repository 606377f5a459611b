"""Pins the CPU oracle (oracle/liboracle.so) before anything is compared to it.

LELA ships no golden vectors for this path (its tests draw /dev/urandom-seeded
inputs), and its compiled form cannot be built here, so each oracle function
is pinned by checks that a plausible mistake would fail:
  * gemm   -- exact big-integer products (bruteforce.matmul_mod), including a
              modulus near 2^32 where every 64-bit partial sum wraps and
              LELA's two_64 correction (level1-modular.tcc:150-151) matters;
  * trsm   -- A*X == a*B exactly; LELA's reported index rule for a
              non-invertible diagonal (level3-generic.tcc:441-491);
  * RREF   -- an independent Gauss-Jordan with a different pivot choice plus
              structural RREF checks; det against the field determinant up to
              the sign LELA drops (it never negates for transpositions);
  * GF(2)  -- unpacked integer matmul mod 2, independent bit RREF;
  * FL     -- equality with Elimination's reduced form and rank, which is the
              assertion of LELA's own test (test-faugere-lachartre.C:227-243),
              plus WrongMatrixForm on a decreasing head.
"""
import numpy as np
import pytest

import bruteforce as bf
import oracle_lib as orc
from lela_amd import synth

PRIMES = [2, 3, 101, 251, 65521, 2147483647, 4294967291]


def rnd(seed, rows, cols, p):
    return synth.zp_matrix(seed, rows, cols, p)


@pytest.mark.parametrize("p", PRIMES + [12, 65536])
@pytest.mark.parametrize("shape", [(7, 5, 9), (16, 16, 33), (1, 1, 1), (3, 4, 0)])
def test_gemm_u32_bruteforce(p, shape):
    m, n, k = shape
    A, B, C = rnd(1, m, k, p), rnd(2, k, n, p), rnd(3, m, n, p)
    for a, b in [(1, 1), (0, 1), (1, 0), (p - 1, 5 % p), (7 % p, p - 1)]:
        got = orc.zp_gemm(p, a, A, B, b, C)
        want = bf.gemm_mod(p, a, A, B, b, C)
        assert np.array_equal(got.astype(object), want), (p, shape, a, b)


def test_gemm_u32_wrap_correction():
    # (p-1)^2 ~ 2^64: the uint64 running sum wraps at almost every step.
    p = 4294967291
    A = np.full((2, 64), p - 1, dtype=np.uint32)
    B = np.full((64, 2), p - 2, dtype=np.uint32)
    C = np.zeros((2, 2), dtype=np.uint32)
    got = orc.zp_gemm(p, 1, A, B, 0, C)
    assert int(got[0, 0]) == (64 * (p - 1) * (p - 2)) % p


@pytest.mark.parametrize("p", [2, 251, 65521, 65535])
def test_gemm_u16_bruteforce(p):
    A, B, C = rnd(4, 9, 70, p), rnd(5, 70, 6, p), rnd(6, 9, 6, p)
    got = orc.zp_gemm(p, 3 % p, A, B, 2 % p, C, storage="uint16")
    want = bf.gemm_mod(p, 3 % p, A, B, 2 % p, C)
    assert np.array_equal(got.astype(object), want)


def _tri(seed, n, p, upper, unit, zero_diag=()):
    A = rnd(seed, n, n, p)
    A = np.triu(A) if upper else np.tril(A)
    d = np.arange(n)
    A[d, d] = np.where(A[d, d] == 0, 1, A[d, d])
    if unit:
        A[d, d] = rnd(seed + 1, 1, n, p)[0]      # ignored by a unit-diagonal solve
    for i in zero_diag:
        A[i, i] = 0
    return A


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("p", [3, 101, 65521, 2147483647])
@pytest.mark.parametrize("n,nrhs", [(1, 3), (2, 2), (13, 5), (40, 17)])
def test_trsm_solves(upper, unit, p, n, nrhs):
    A = _tri(10 + n, n, p, upper, unit)
    B = rnd(20 + n, n, nrhs, p)
    a = 7 % p or 1
    rc, _, X = orc.zp_trsm(p, upper, unit, a, A, B)
    assert rc == orc.ORC_OK
    Aeff = A.astype(object)
    if unit:
        Aeff[np.arange(n), np.arange(n)] = 1
    lhs = bf.matmul_mod(Aeff, X, p)
    assert np.array_equal(lhs, (a * B.astype(object)) % p)


def test_trsm_a_zero_clears_B():
    p = 101
    A = np.zeros((5, 5), dtype=np.uint32)      # singular, but a = 0 returns first
    B = rnd(30, 5, 4, p)
    rc, _, X = orc.zp_trsm(p, False, False, 0, A, B)
    assert rc == orc.ORC_OK and not X.any()


@pytest.mark.parametrize("upper,expect", [(False, 3), (True, 11)])
def test_trsm_bad_diag_index(upper, expect):
    p = 65521
    A = _tri(40, 16, p, upper, False, zero_diag=(3, 7, 11))
    rc, bad, _ = orc.zp_trsm(p, upper, False, 1, A, rnd(41, 16, 3, p))
    assert rc == orc.ORC_E_NOT_INVERTIBLE and bad == expect


def test_trsm_composite_modulus_not_invertible():
    p = 12
    A = np.array([[5, 0], [3, 4]], dtype=np.uint32)  # 4 has no inverse mod 12
    rc, bad, _ = orc.zp_trsm(p, False, False, 1, A, np.ones((2, 1), dtype=np.uint32))
    assert rc == orc.ORC_E_NOT_INVERTIBLE and bad == 1


def _lowrank(seed, m, n, r, p):
    X, Y = rnd(seed, m, r, p), rnd(seed + 1, r, n, p)
    return np.array(bf.matmul_mod(X, Y, p), dtype=np.uint32)


@pytest.mark.parametrize("p", [2, 3, 101, 65521, 2147483647])
@pytest.mark.parametrize("m,n,r", [(8, 8, 8), (12, 9, 5), (6, 20, 6), (20, 6, 3), (10, 10, 0)])
def test_echelonize_reduced(p, m, n, r):
    A = _lowrank(50 + m + n, m, n, r, p)
    rc, R, rank, piv, det = orc.zp_echelonize(p, A, reduced=True)
    assert rc == orc.ORC_OK
    Rb, rb, pb = bf.rref_mod(A, p)
    assert np.array_equal(R.astype(object), Rb) and rank == rb and list(piv) == pb
    assert bf.is_rref(R, p) == list(piv)


@pytest.mark.parametrize("p", [3, 101, 65521])
def test_echelonize_det_and_nonreduced(p):
    A = rnd(60, 9, 9, p)
    rc, E, rank, piv, det = orc.zp_echelonize(p, A, reduced=False)
    assert rc == orc.ORC_OK
    d = bf.det_mod(A, p)
    if rank == A.shape[0]:
        assert det in (d, (-d) % p)
    else:
        assert d == 0       # LELA's det is then the product of the pivots found
    # the non-reduced form is upper echelon with the same RREF as A
    for i in range(rank):
        assert all(int(x) == 0 for x in E[i, :piv[i]]) and E[i, piv[i]] != 0
    assert not E[rank:].any()
    assert np.array_equal(bf.rref_mod(E, p)[0], bf.rref_mod(A, p)[0])
    _, R2, rank2, _, det2 = orc.zp_echelonize(p, A, reduced=True)
    assert det2 == det and rank2 == rank


@pytest.mark.parametrize("m,n,k", [(5, 70, 9), (64, 64, 64), (3, 130, 129), (1, 1, 1)])
@pytest.mark.parametrize("a,b", [(1, 1), (1, 0), (0, 1)])
def test_gf2_gemm(m, n, k, a, b):
    A = synth.gf2_matrix(70, m, k)
    B = synth.gf2_matrix(71, k, n)
    C = synth.gf2_matrix(72, m, n)
    got = orc.gf2_gemm(m, n, k, a, A, B, b, C)
    want = (a * bf.gf2_matmul_bits(bf.gf2_unpack(A, k), bf.gf2_unpack(B, n))
            + b * bf.gf2_unpack(C, n)) & 1
    assert np.array_equal(bf.gf2_unpack(got, n), want)


def gf2_triangular(seed, n, upper, diag_zeros=()):
    """n x n GF(2) triangular bits (word rows) with ones on the diagonal
    except at diag_zeros."""
    rng = np.random.default_rng(seed)
    T = rng.integers(0, 2, size=(n, n)).astype(np.uint8)
    T = np.triu(T) if upper else np.tril(T)
    T[np.arange(n), np.arange(n)] = 1
    for i in diag_zeros:
        T[i, i] = 0
    return T, bf.gf2_pack(T)


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (2, 70), (63, 64), (64, 1), (65, 70), (127, 64), (128, 130), (129, 3),
                                    (200, 70)])
def test_gf2_trsm_solves(upper, n, nrhs):
    # A X = B over GF(2) checked with numpy's product of the unpacked bits
    T, A = gf2_triangular(300 + n, n, upper)
    B = synth.gf2_matrix(301 + n, n, nrhs)
    rc, bad, X = orc.gf2_trsm(upper, False, n, nrhs, 1, A, B)
    assert rc == orc.ORC_OK
    assert np.array_equal(bf.gf2_matmul_bits(T, bf.gf2_unpack(X, nrhs)), bf.gf2_unpack(B, nrhs))
    # bits past nrhs keep their values
    tail = nrhs % 64
    if tail:
        m = np.uint64((1 << 64) - (1 << tail))
        assert np.array_equal(X[:, -1] & m, B[:, -1] & m)


@pytest.mark.parametrize("upper", [False, True])
def test_gf2_trsm_diagonal(upper):
    # A zero diagonal: the recursion meets lower blocks top-down (the smallest
    # such index is reported) and upper ones bottom-up (the largest); with
    # unit_diag the diagonal is taken as ones.  a = 0 clears B.
    n, nrhs = 200, 70
    T, A = gf2_triangular(310, n, upper, diag_zeros=(5, 70, 130))
    B = synth.gf2_matrix(311, n, nrhs)
    rc, bad, X = orc.gf2_trsm(upper, False, n, nrhs, 1, A, B)
    assert rc == orc.ORC_E_NOT_INVERTIBLE and bad == (130 if upper else 5)
    assert np.array_equal(X, B)
    rc, bad, X = orc.gf2_trsm(upper, True, n, nrhs, 1, A, B)
    T1 = T.copy()
    T1[np.arange(n), np.arange(n)] = 1
    assert rc == orc.ORC_OK
    assert np.array_equal(bf.gf2_matmul_bits(T1, bf.gf2_unpack(X, nrhs)), bf.gf2_unpack(B, nrhs))
    rc, bad, X = orc.gf2_trsm(upper, False, n, nrhs, 0, A, B)
    assert rc == orc.ORC_OK and not bf.gf2_unpack(X, nrhs).any()
    m = np.uint64((1 << 64) - (1 << (nrhs % 64)))
    assert np.array_equal(X[:, -1] & m, B[:, -1] & m)


@pytest.mark.parametrize("m,n", [(10, 70), (70, 10), (64, 64), (100, 130)])
def test_gf2_echelonize(m, n):
    A = synth.gf2_matrix(80 + m, m, n)
    R, rank, piv = orc.gf2_echelonize(m, n, A, reduced=True)
    Rb, rb, pb = bf.gf2_rref_bits(bf.gf2_unpack(A, n))
    assert rank == rb and list(piv) == pb
    assert np.array_equal(bf.gf2_unpack(R, n), Rb)


def f4_matrix(seed, m, n, p, rng_moves=None):
    """F4-shaped matrix in the style of createRandomF4Matrix
    (test-faugere-lachartre.C:142-175): each row starts at a column that
    stays, advances by one, or jumps; the row is random from that column."""
    rng = np.random.default_rng(seed)
    X = np.zeros((m, n), dtype=np.uint32)
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
        row = rng.integers(0, p, size=n - col).astype(np.uint32)
        row[0] = int(rng.integers(1, p))
        X[i, col:] = row
    return X


@pytest.mark.parametrize("p", [101, 65521])
@pytest.mark.parametrize("m,n,seed", [(24, 32, 1), (48, 64, 2), (96, 128, 3), (30, 30, 4)])
def test_fl_equals_elimination(p, m, n, seed):
    X = f4_matrix(seed, m, n, p)
    rc, R, rank, det, _ = orc.fl_echelonize(p, X)
    assert rc == orc.ORC_OK
    rc2, R2, rank2, _, _ = orc.zp_echelonize(p, X, reduced=True)
    assert np.array_equal(R, R2) and rank == rank2


def test_fl_det_when_all_rows_are_pivots():
    # No C/D rows: det is the product of the row heads (faugere-lachartre.tcc:119).
    p = 65521
    X = np.triu(rnd(90, 6, 9, p))
    X[np.arange(6), np.arange(6)] = [3, 5, 7, 11, 13, 17]
    rc, R, rank, det, _ = orc.fl_echelonize(p, X)
    assert rc == orc.ORC_OK and rank == 6 and det == (3 * 5 * 7 * 11 * 13 * 17) % p


@pytest.mark.parametrize("p", [3, 101])
@pytest.mark.parametrize("m,n,seed", [(16, 20, 1), (24, 30, 2), (20, 20, 3), (30, 24, 5)])
def test_fl_det_from_first_principles(p, m, n, seed):
    # The D factor (:833) depends on the pivoting of D's non-reduced echelon
    # form; bruteforce.fl_det builds D - C A^-1 B by Gauss-Jordan and takes
    # the heads from the independent characterisation of LELA's pivots.
    X = f4_matrix(seed, m, n, p)
    rc, R, rank, det, _ = orc.fl_echelonize(p, X)
    assert rc == orc.ORC_OK
    want, rd = bf.fl_det(X, p)
    assert det == want


def test_fl_wrong_matrix_form():
    p = 101
    X = np.zeros((4, 6), dtype=np.uint32)
    X[0, 1] = 1
    X[1, 2] = 1
    X[2, 0] = 1        # head decreases at row 2
    X[3, 4] = 1
    rc, _, _, _, bad = orc.fl_echelonize(p, X)
    assert rc == orc.ORC_E_WRONG_FORM and bad == 2


def test_inverse():
    assert orc.inv(65521, 2) == 32761
    assert orc.inv(12, 4) is None
    for p in (3, 101, 65521, 4294967291):
        for y in (1, 2, p - 1):
            assert (orc.inv(p, y) * y) % p == 1


def test_full_size_checks_accept_rref_and_reject_planted_errors():
    """The all-row checks that the full-size GPU tests apply (Z/p FL and the
    GF(2) echelon form) accept the oracle's RREF and reject an R with one
    entry changed."""
    from lela_amd import dense_to_csr
    from test_fl_gpu import check_rref_all_rows
    from test_gf2_gpu import check_gf2_rref_all_rows
    p = 65521
    X = f4_matrix(5, 200, 260, p)
    rc, R, rank, piv, _ = orc.zp_echelonize(p, X, reduced=True)
    csr = dense_to_csr(X)
    check_rref_all_rows(R.copy(), piv, rank, csr, X.shape, p, seed=3)
    free = [j for j in range(X.shape[1]) if j not in set(piv.tolist())]
    bad = R.copy()
    bad[7, free[5]] = (int(bad[7, free[5]]) + 1) % p
    with pytest.raises(AssertionError):
        check_rref_all_rows(bad, piv, rank, csr, X.shape, p, seed=3)
    n = 300
    Xg, Yg = synth.gf2_matrix(1, n, 200), synth.gf2_matrix(2, 200, n)
    M = orc.gf2_gemm(n, n, 200, 1, Xg, Yg, 0, np.zeros((n, (n + 63) // 64), dtype=np.uint64))
    Rg, rk, pv = orc.gf2_echelonize(n, n, M)
    check_gf2_rref_all_rows(M, Rg, pv, rk, n, seed=4)
    freeg = [j for j in range(n) if j not in set(pv.tolist())]
    badg = Rg.copy()
    badg[3, freeg[2] // 64] ^= np.uint64(1 << (freeg[2] % 64))
    with pytest.raises(AssertionError):
        check_gf2_rref_all_rows(M, badg, pv, rk, n, seed=4)


def _echelon_inputs():
    """Dense, rank-deficient, zero-column and sparse small cases."""
    out = []
    for seed, (m, n, r), p in [(1, (12, 15, 12), 101), (2, (20, 14, 9), 65521), (3, (9, 30, 6), 7),
                               (4, (16, 16, 16), 2147483647), (5, (25, 18, 11), 3)]:
        out.append((p, _lowrank(seed, m, n, r, p)))
    rng = np.random.default_rng(9)
    Z = rng.integers(0, 101, size=(14, 20)).astype(np.uint32)
    Z[:, [0, 3, 7]] = 0                          # zero columns, including the first
    Z[5] = 0
    out.append((101, Z))
    S = np.where(rng.random((18, 22)) < 0.15, rng.integers(1, 65521, size=(18, 22)), 0).astype(np.uint32)
    out.append((65521, S))                        # sparse: first nonzeros far down the rows
    return out


@pytest.mark.parametrize("case", range(7))
def test_elim_matches_independent_pivot_rule(case):
    """orc_zp_elim (Elimination::echelonize) against the independent
    characterisation of LELA's pivots and the closed form of U given them:
    the same transposition sequence, rank and pivot columns, U, det, zero
    rows past the rank."""
    p, A = _echelon_inputs()[case]
    rc, U, rank, det, perm = orc.zp_elim(p, A, compute_L=False)
    assert rc == orc.ORC_OK
    steps, order, cols = bf.lela_pivot_sequence(A, p)
    assert rank == len(steps) and list(perm) == steps
    want = bf.echelon_given_order(A, order, cols, p)
    assert np.array_equal(U.astype(object), want)
    d = 1
    for t, c in enumerate(cols):
        d = d * int(want[t, c]) % p
    assert det == d % p
    # the reduced form of U is the canonical RREF of A
    R, r2, piv = bf.rref_mod(U, p)
    R0, r0, piv0 = bf.rref_mod(A, p)
    assert r2 == r0 == rank and piv == piv0 == cols and np.array_equal(R, R0)


@pytest.mark.parametrize("case", range(7))
def test_elim_compute_L_is_the_transform(case):
    """compute_L: the strict lower part of the output is T with T P A = U (T
    unit lower), and the upper part equals the compute_L = false result
    (tests/test-gauss-jordan.C:96-99 compares the two forms the same way)."""
    p, A = _echelon_inputs()[case]
    rc, O, rank, det, perm = orc.zp_elim(p, A, compute_L=True)
    rc0, U0, rank0, det0, perm0 = orc.zp_elim(p, A, compute_L=False)
    assert rc == rc0 == orc.ORC_OK and rank == rank0 and det == det0 and list(perm) == list(perm0)
    m, n = A.shape
    O = O.astype(object)
    T = np.zeros((m, m), dtype=object)
    U = O.copy()
    for t in range(m):
        T[t, t] = 1
        for s_ in range(min(t, n)):
            T[t, s_] = O[t, s_]
            U[t, s_] = 0
    assert np.array_equal(U, U0.astype(object))
    order = list(range(m))
    for i, q in enumerate(perm):
        order[i], order[q] = order[q], order[i]
    PA = np.asarray(A).astype(object)[order]
    assert np.array_equal(T.dot(PA) % p, U)


@pytest.mark.parametrize("case", range(7))
def test_pluq_factors(case):
    """orc_zp_pluq: P A Q = L U with L unit lower (A's part left of the
    diagonal, first rank columns) and U upper (A's part from the diagonal on,
    first rank rows); its pivots are Elimination::echelonize's (the same rows,
    the same columns in the original numbering)."""
    p, A = _echelon_inputs()[case]
    rc, O, rank, det, pr, pc = orc.zp_pluq(p, A)
    assert rc == orc.ORC_OK
    m, n = A.shape
    rc2, _, rank2, det2, perm2 = orc.zp_elim(p, A)
    steps, order, cols = bf.lela_pivot_sequence(A, p)
    assert rank == rank2 and det == det2 and list(pr) == list(perm2)
    corder = list(range(n))
    for i, q in enumerate(pc):
        corder[i], corder[q] = corder[q], corder[i]
    assert corder[:rank] == cols
    rorder = list(range(m))
    for i, q in enumerate(pr):
        rorder[i], rorder[q] = rorder[q], rorder[i]
    O = O.astype(object)
    L = np.zeros((m, rank), dtype=object)
    U = np.zeros((rank, n), dtype=object)
    for t in range(m):
        for s_ in range(min(t, rank)):
            L[t, s_] = O[t, s_]
        if t < rank:
            L[t, t] = 1
            U[t, t:] = O[t, t:]
    PAQ = np.asarray(A).astype(object)[rorder][:, corder]
    assert np.array_equal(L.dot(U) % p, PAQ % p)


def _sparse_pivot_inputs():
    """Cases where the sparse rule and the dense one part ways: rows sharing a
    first column with different counts, ties, zero rows and columns, words."""
    out = []
    # row 0 and row 2 start in column 0; row 2 has fewer nonzeros; rows 1, 3 tie
    out.append((101, np.array([[1, 2, 3, 4, 5], [0, 1, 1, 0, 0], [7, 0, 0, 0, 1], [0, 1, 0, 1, 0]],
                              dtype=np.uint32), False))
    rng = np.random.default_rng(21)
    for seed, (m, n, dens, p) in enumerate([(14, 20, 0.2, 101), (20, 14, 0.3, 65521), (25, 30, 0.1, 7),
                                            (16, 16, 0.5, 2147483647), (30, 12, 0.25, 3), (12, 40, 0.15, 2)]):
        r = np.random.default_rng(100 + seed)
        S = np.where(r.random((m, n)) < dens, r.integers(1, p, size=(m, n)), 0).astype(np.uint32)
        if seed % 2 == 0:
            S[3] = 0
            S[:, 1] = 0
        out.append((p, S, False))
    # GF(2) rows over 200 columns: counts of 64-column words
    for seed in range(3):
        r = np.random.default_rng(200 + seed)
        B = (r.random((24, 200)) < 0.04).astype(np.uint32)
        B[r.integers(0, 24, 6), r.integers(0, 200, 6)] = 1
        out.append((2, B, True))
    return out


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("case", range(10))
def test_elim_sparse_follows_the_rule(case, mode):
    """orc_zp_elim_sparse: every pivot satisfies SparsePartialPivotStrategy's
    definition against the matrix as it stands at that step, and the result
    is the textbook elimination with those pivots (bruteforce.sparse_rule_replay);
    the forms then satisfy the factor identities of the dense rule's forms."""
    p, A, words = _sparse_pivot_inputs()[case]
    rc, O, rank, det, pr, pc = orc.zp_elim_sparse(p, A, mode, words)
    assert rc == orc.ORC_OK
    want, wdet = bf.sparse_rule_replay(A, p, mode, words, pr, pc)
    assert np.array_equal(O.astype(object), want) and det == wdet
    m, n = A.shape
    R0, r0, piv0 = bf.rref_mod(A, p)
    assert rank == r0
    if mode == 0:
        # an echelon form of A: same row space, so the same RREF
        R, r2, piv = bf.rref_mod(O, p)
        assert r2 == rank and piv == piv0 and np.array_equal(R, R0)
        for t in range(rank, m):
            assert not O[t].any()
    if mode == 1:
        rc0, U0, rank0, det0, pr0, _ = orc.zp_elim_sparse(p, A, 0, words)
        assert list(pr0) == list(pr) and det0 == det
        T = np.zeros((m, m), dtype=object)
        U = O.astype(object).copy()
        for t in range(m):
            T[t, t] = 1
            for s_ in range(min(t, n)):
                T[t, s_] = U[t, s_]
                U[t, s_] = 0
        assert np.array_equal(U, U0.astype(object))
        order = list(range(m))
        for i, q in enumerate(pr):
            order[i], order[q] = order[q], order[i]
        assert np.array_equal(T.dot(np.asarray(A).astype(object)[order]) % p, U)
    if mode == 2:
        rorder, corder = list(range(m)), list(range(n))
        for i, q in enumerate(pr):
            rorder[i], rorder[q] = rorder[q], rorder[i]
        for i, q in enumerate(pc):
            corder[i], corder[q] = corder[q], corder[i]
        Oo = O.astype(object)
        L = np.zeros((m, rank), dtype=object)
        U = np.zeros((rank, n), dtype=object)
        for t in range(m):
            for s_ in range(min(t, rank)):
                L[t, s_] = Oo[t, s_]
            if t < rank:
                L[t, t] = 1
                U[t, t:] = Oo[t, t:]
        assert np.array_equal(L.dot(U) % p, np.asarray(A).astype(object)[rorder][:, corder] % p)


def test_elim_sparse_hand_case():
    """Worked by hand: rows 0 and 2 start in column 0, row 2 has 2 nonzeros
    against 5, so the sparse rule takes row 2 (the dense rule takes row 0)."""
    p = 101
    A = np.array([[1, 2, 3, 4, 5], [0, 1, 1, 0, 0], [7, 0, 0, 0, 1], [0, 1, 0, 1, 0]], dtype=np.uint32)
    rc, U, rank, det, pr, _ = orc.zp_elim_sparse(p, A, 0)
    assert rc == 0 and pr[0] == 2
    # step 1 from column 1: rows 1 (1 1 0 0 -> two entries), 2 (the old row 0,
    # 2 3 4 5 - 1/7 -> five), 3 (1 0 1 0 -> two): 1 and 3 tie, row 1 is first
    assert pr[1] == 1
    rc_d, _, _, _, perm_d = orc.zp_elim(p, A)
    assert perm_d[0] == 0
    # the sparse-rule replay fails a sequence with the dense rule's first pivot
    with pytest.raises(AssertionError):
        bf.sparse_rule_replay(A, p, 0, False, [0] + list(pr[1:]), None)
