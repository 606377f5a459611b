"""GPU parity of the GF(2) kernels (mlz_gf2_gemm, mlz_gf2_rref) against the
oracle (bit-exact), plus the 16384^2 workloads through size-independent
checks: Freivalds for the product; rank, RREF structure and row-space
membership for the echelon form of a rank-12000 product."""
import numpy as np
import pytest

import bruteforce as bf
import lela_amd as la
import oracle_lib as orc
from lela_amd import synth

pytestmark = pytest.mark.gpu


def gpu_gemm(ctx, m, n, k, a, A, B, b, C, kernel=0):
    """kernel: 0 Four Russians tables, 1 the AND/XOR-popcount dot form."""
    ctx.F = la.GF2()
    Cw = C.copy()
    ctx.set_option("gf2.gemm.kernel", kernel)
    try:
        la.BLAS3.gemm(ctx, a, la.Dense01Matrix(A, k), la.Dense01Matrix(B, n), b, la.Dense01Matrix(Cw, n))
    finally:
        ctx.set_option("gf2.gemm.kernel", 0)
    return Cw


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 70, 9), (64, 64, 64), (130, 300, 129),
                                   (1100, 257, 200), (300, 1030, 700), (2049, 64, 63)])
@pytest.mark.parametrize("a,b", [(1, 1), (1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize("kernel", [0, 1])
def test_gf2_gemm_matches_oracle(gpu_ctx, m, n, k, a, b, kernel):
    A, B, C = synth.gf2_matrix(1, m, k), synth.gf2_matrix(2, k, n), synth.gf2_matrix(3, m, n)
    got = gpu_gemm(gpu_ctx, m, n, k, a, A, B, b, C, kernel)
    want = orc.gf2_gemm(m, n, k, a, A, B, b, C)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kernel", [0, 1])
def test_gf2_gemm_ignores_padding_bits(gpu_ctx, kernel):
    m, n, k = 40, 100, 70
    A, B, C = synth.gf2_matrix(4, m, k), synth.gf2_matrix(5, k, n), synth.gf2_matrix(6, m, n)
    want = orc.gf2_gemm(m, n, k, 1, A, B, 1, C)
    A2, B2 = A.copy(), B.copy()
    A2[:, -1] |= ~np.uint64((1 << (k % 64)) - 1)       # garbage beyond column k
    B2[:, -1] |= ~np.uint64((1 << (n % 64)) - 1)
    got = gpu_gemm(gpu_ctx, m, n, k, 1, A2, B2, 1, C, kernel)
    assert np.array_equal(bf.gf2_unpack(got, n), bf.gf2_unpack(want, n))
    assert np.array_equal(got[:, -1] >> np.uint64(n % 64), C[:, -1] >> np.uint64(n % 64))


def gpu_rref(ctx, m, n, A):
    ctx.F = la.GF2()
    R = la.Dense01Matrix(A.copy(), n)
    EF = la.EchelonForm(ctx)
    EF.echelonize(R)
    return R.words, EF.rank(R), EF.pivots


def lowrank_gf2(seed, m, n, r):
    X, Y = synth.gf2_matrix(seed, m, r), synth.gf2_matrix(seed + 1, r, n)
    Z = np.zeros((m, (n + 63) // 64), dtype=np.uint64)
    return orc.gf2_gemm(m, n, r, 1, X, Y, 0, Z)


def holes_gf2(seed, m, n):
    """A random matrix with columns j = 1, 3 (mod 5) cleared: every cleared
    column sits between two pivot columns."""
    A = synth.gf2_matrix(seed, m, n)
    keep = np.array([j % 5 not in (1, 3) for j in range(A.shape[1] * 64)], dtype=np.uint64)
    return A & (keep.reshape(-1, 64) << np.arange(64, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)


# "holes": zero columns between pivot columns, through the single-set path
# (m <= 512) and the tournament (m > 512, rank-deficient panels); a search
# that skips a zero column must not pick the previous column's pivot row again.
@pytest.mark.parametrize("m,n,kind", [(10, 70, "rand"), (70, 10, "rand"), (300, 300, "rand"),
                                      (500, 700, "low"), (1500, 400, "low"), (257, 1000, "sparse"),
                                      (64, 64, "zero"), (2100, 900, "low"), (847, 597, "low260"),
                                      (300, 200, "holes"), (1100, 600, "holes")])
def test_gf2_rref_matches_oracle(gpu_ctx, m, n, kind):
    if kind == "rand":
        A = synth.gf2_matrix(7, m, n)
    elif kind == "low":
        A = lowrank_gf2(8, m, n, min(m, n) // 3)
    elif kind == "low260":
        A = lowrank_gf2(20, m, n, 260)
    elif kind == "sparse":
        A = synth.gf2_matrix(9, m, n) & synth.gf2_matrix(10, m, n) & synth.gf2_matrix(11, m, n)
    elif kind == "holes":
        A = holes_gf2(12, m, n)
    else:
        A = np.zeros((m, (n + 63) // 64), dtype=np.uint64)
    got, rank, piv = gpu_rref(gpu_ctx, m, n, A)
    want, wrank, wpiv = orc.gf2_echelonize(m, n, A, reduced=True)
    assert rank == wrank and list(piv) == list(wpiv)
    assert np.array_equal(got, want)


def test_gf2_rref_zero_column_between_pivots(gpu_ctx):
    A = np.array([[1], [4]], dtype=np.uint64)          # rows e0 and e2 of GF(2)^3
    got, rank, piv = gpu_rref(gpu_ctx, 2, 3, A)
    assert rank == 2 and list(piv) == [0, 2]
    assert np.array_equal(got, A)


def _gf2_matvec(W, ncols, x_bits):
    """(W as bits) @ x mod 2 for a packed matrix W and a 0/1 vector x."""
    out = np.zeros(W.shape[0], dtype=np.int64)
    for start in range(0, W.shape[0], 1024):
        out[start:start + 1024] = bf.gf2_unpack(W[start:start + 1024], ncols).astype(np.int64) @ x_bits
    return out & 1


@pytest.mark.parametrize("kernel", [0, 1])
def test_gf2_gemm_full_size_16384(gpu_ctx, kernel):
    n = 16384
    A = synth.gf2_matrix(synth.seed_for(3, 0), n, n)
    B = synth.gf2_matrix(synth.seed_for(3, 1), n, n)
    C = np.zeros_like(A)
    got = gpu_gemm(gpu_ctx, n, n, n, 1, A, B, 0, C, kernel)
    rng = np.random.default_rng(12)
    for _ in range(2):
        x = rng.integers(0, 2, size=n).astype(np.int64)
        assert np.array_equal(_gf2_matvec(got, n, x), _gf2_matvec(A, n, _gf2_matvec(B, n, x)))


def gf2_matvec_packed(W, v_words):
    """(W @ v) mod 2 for packed rows W and a packed bit vector v: parity of the
    popcount of each row AND v."""
    return (np.bitwise_count(W & v_words[None, :]).sum(axis=1, dtype=np.int64) & 1).astype(np.uint8)


def check_gf2_rref_all_rows(M, R, piv, rank, n, seed, chunk=2048):
    """Every row of R in standard RREF (rows past the rank zero, row t's lowest
    set bit at piv[t], pivot columns the identity on the first rank rows), and
    every row of M equals M[:, piv] R: M v == M[:, piv] (R v) over GF(2) at
    random bit vectors v."""
    piv = piv.astype(np.int64)
    assert rank == piv.size and (rank == 0 or np.all(np.diff(piv) > 0))
    assert not R[rank:].any()
    for s0 in range(0, rank, chunk):
        bits = bf.gf2_unpack(R[s0:min(rank, s0 + chunk)], n)
        t = np.arange(s0, s0 + bits.shape[0])
        assert np.array_equal(bits.argmax(axis=1), piv[t]), f"row heads near row {s0}"
        sub = bits[:, piv]
        assert np.count_nonzero(sub) == bits.shape[0] and np.all(sub[np.arange(bits.shape[0]), t] == 1)
    words = (n + 63) // 64
    for it in range(2):
        v = synth.raw_draws(seed, it * words, words)
        if n % 64:
            v[-1] &= np.uint64((1 << (n % 64)) - 1)
        lhs = gf2_matvec_packed(M, v)
        Rv = gf2_matvec_packed(R[:rank], v)
        # M[:, piv] @ Rv: pack Rv at the pivot columns and take M's parity against it
        w_bits = np.zeros((1, n), dtype=np.uint8)
        w_bits[0, piv] = Rv
        w = bf.gf2_pack(w_bits)[0]
        rhs = gf2_matvec_packed(M, w)
        assert np.array_equal(lhs, rhs), "a row of M is not M[:, piv] R"


def test_gf2_rref_rank12000_16384(gpu_ctx):
    n, r = 16384, 12000
    X = synth.gf2_matrix(synth.seed_for(3, 2), n, r)
    Y = synth.gf2_matrix(synth.seed_for(3, 3), r, n)
    M = gpu_gemm(gpu_ctx, n, n, r, 1, X, Y, 0, np.zeros((n, n // 64), dtype=np.uint64))
    R, rank, piv = gpu_rref(gpu_ctx, n, n, M)
    assert rank == r
    check_gf2_rref_all_rows(M, R, piv, rank, n, seed=13)


def gpu_trsm(ctx, upper, unit, a, A, n, B, nrhs):
    ctx.F = la.GF2()
    Bm = la.Dense01Matrix(B.copy(), nrhs)
    la.BLAS3.trsm(ctx, a, la.Dense01Matrix(A, n), Bm, la.UpperTriangular if upper else la.LowerTriangular, unit)
    return Bm.words


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (2, 70), (63, 64), (64, 1), (65, 70), (127, 64), (128, 130), (129, 3),
                                    (200, 70), (700, 300), (1100, 2100)])
def test_gf2_trsm_matches_oracle(gpu_ctx, upper, n, nrhs):
    """GF(2) trsm (level3-generic.tcc:421-498) bit for bit against the oracle,
    padding bits of B kept, leaf and recursion sizes around 64 and 128."""
    from test_oracle import gf2_triangular
    _, A = gf2_triangular(400 + n, n, upper)
    B = synth.gf2_matrix(401 + n, n, nrhs)
    B[:, -1] |= np.uint64(0xF000000000000000) if nrhs % 64 and nrhs % 64 < 60 else np.uint64(0)
    rc, _, want = orc.gf2_trsm(upper, False, n, nrhs, 1, A, B)
    assert rc == orc.ORC_OK
    got = gpu_trsm(gpu_ctx, upper, False, 1, A, n, B, nrhs)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("upper", [False, True])
def test_gf2_trsm_diagonal_and_scalar(gpu_ctx, upper):
    """A zero on the diagonal: DiagonalEntryNotInvertible at the index LELA's
    recursion reports, B unchanged; unit_diag takes the diagonal as ones;
    a = 0 clears B (bits past nrhs kept)."""
    from test_oracle import gf2_triangular
    n, nrhs = 300, 70
    _, A = gf2_triangular(410, n, upper, diag_zeros=(5, 70, 130, 257))
    B = synth.gf2_matrix(411, n, nrhs)
    with pytest.raises(la.DiagonalEntryNotInvertible) as ei:
        gpu_trsm(gpu_ctx, upper, False, 1, A, n, B, nrhs)
    assert ei.value.idx == (257 if upper else 5)
    rc, _, want = orc.gf2_trsm(upper, True, n, nrhs, 1, A, B)
    assert rc == orc.ORC_OK
    assert np.array_equal(gpu_trsm(gpu_ctx, upper, True, 1, A, n, B, nrhs), want)
    rc, _, want = orc.gf2_trsm(upper, False, n, nrhs, 0, A, B)
    assert np.array_equal(gpu_trsm(gpu_ctx, upper, False, 0, A, n, B, nrhs), want)


@pytest.mark.parametrize("upper", [False, True])
def test_gf2_trsm_8192(gpu_ctx, upper):
    """8192 x 8192 triangular, 1024 right-hand sides: A X = B at random bit
    vectors over every row (A (X v) == B v)."""
    n, nrhs = 8192, 1024
    rng = np.random.default_rng(12 + upper)
    T = synth.gf2_matrix(420 + upper, n, n)
    bits = bf.gf2_unpack(T, n)
    bits = np.triu(bits) if upper else np.tril(bits)
    bits[np.arange(n), np.arange(n)] = 1
    A = bf.gf2_pack(bits)
    B = synth.gf2_matrix(422, n, nrhs)
    print("gf2 trsm 8192", upper, flush=True)
    X = gpu_trsm(gpu_ctx, upper, False, 1, A, n, B, nrhs)
    for _ in range(3):
        v = rng.integers(0, 2 ** 63, size=nrhs // 64, dtype=np.uint64)
        xv = gf2_matvec_packed(X, v)                  # X v, n bits
        lhs = (bits.astype(np.int32) @ xv.astype(np.int32)) & 1
        assert np.array_equal(lhs.astype(np.uint8), gf2_matvec_packed(B, v))
