"""GPU parity of mlz_zp_trsm against the oracle's restatement of LELA's
recursive trsm (bit-exact), error indices, and the 8192 x 1024 workload by a
Freivalds check of A*X == a*B."""
import numpy as np
import pytest

import lela_amd as la
import oracle_lib as orc
from lela_amd import synth

pytestmark = pytest.mark.gpu


def tri(seed, n, p, upper, unit=False, zero_diag=()):
    A = synth.zp_matrix(seed, n, n, p)
    A = np.triu(A) if upper else np.tril(A)
    d = np.arange(n)
    A[d, d] = np.where(A[d, d] == 0, 1, A[d, d])
    if unit:
        A[d, d] = synth.zp_matrix(seed + 1, 1, n, p)[0]   # must be ignored
    for i in zero_diag:
        A[i, i] = 0
    return A


def run_trsm(ctx, p, storage, upper, unit, a, A, B, algo=la.ALGO_AUTO):
    F = la.Modular(p, storage)
    ctx.F = F
    Bs = F.from_canonical(B)
    la.BLAS3.trsm(ctx, a, F.from_canonical(A), Bs,
                  la.UpperTriangular if upper else la.LowerTriangular, unit, algo=algo)
    return F.to_canonical(Bs)


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("storage,p", [("uint16", 65521), ("uint32", 65521), ("uint32", 2147483647),
                                       ("double", 67108859), ("uint16", 251)])
def test_trsm_matches_oracle(gpu_ctx, upper, unit, storage, p):
    print(f"trsm upper={upper} unit={unit} {storage} p={p}", flush=True)
    for n, nrhs, seed in [(1, 3, 1), (63, 5, 2), (64, 64, 3), (65, 70, 4), (200, 33, 5), (300, 129, 6)]:
        A = tri(seed, n, p, upper, unit)
        B = synth.zp_matrix(seed + 100, n, nrhs, p)
        a = 7 % p
        got = run_trsm(gpu_ctx, p, storage, upper, unit, a, A, B)
        rc, _, want = orc.zp_trsm(p, upper, unit, a, A, B)
        assert rc == 0 and np.array_equal(got, want), (n, nrhs)


@pytest.mark.parametrize("algo", [la.ALGO_PLAIN, la.ALGO_INT8, la.ALGO_FP64])
def test_trsm_each_gemm_family(gpu_ctx, algo):
    p = 65521
    A = tri(7, 257, p, False)
    B = synth.zp_matrix(8, 257, 300, p)
    got = run_trsm(gpu_ctx, p, "uint32", False, False, 3, A, B, algo=algo)
    assert np.array_equal(got, orc.zp_trsm(p, False, False, 3, A, B)[2])


@pytest.mark.parametrize("upper,expect", [(False, 3), (True, 150)])
def test_trsm_not_invertible_index(gpu_ctx, upper, expect):
    p = 65521
    A = tri(9, 160, p, upper, zero_diag=(3, 77, 150))
    B = synth.zp_matrix(10, 160, 4, p)
    rc, bad, _ = orc.zp_trsm(p, upper, False, 1, A, B)
    assert rc == orc.ORC_E_NOT_INVERTIBLE and bad == expect
    with pytest.raises(la.DiagonalEntryNotInvertible) as ei:
        run_trsm(gpu_ctx, p, "uint32", upper, False, 1, A, B)
    assert ei.value.idx == expect


def test_trsm_composite_modulus(gpu_ctx):
    p = 12
    A = np.array([[5, 0], [3, 4]], dtype=np.uint32)
    with pytest.raises(la.DiagonalEntryNotInvertible) as ei:
        run_trsm(gpu_ctx, p, "uint32", False, False, 1, A, np.ones((2, 1), dtype=np.uint32))
    assert ei.value.idx == 1


def test_trsm_a_zero(gpu_ctx):
    p = 101
    A = np.zeros((70, 70), dtype=np.uint32)
    B = synth.zp_matrix(11, 70, 5, p)
    got = run_trsm(gpu_ctx, p, "uint32", False, False, 0, A, B)
    assert not got.any()


def test_trsm_full_size_8192x1024(gpu_ctx):
    """Lower, unit diagonal, n = 8192, 1024 right-hand sides, p = 65521."""
    p, n, nrhs = 65521, 8192, 1024
    L = np.tril(synth.zp_matrix(synth.seed_for(2, 3), n, n, p, np.uint16), -1)
    L[np.arange(n), np.arange(n)] = 1
    B = synth.zp_matrix(synth.seed_for(2, 4), n, nrhs, p, np.uint16)
    F = la.Modular(p, "uint16")
    gpu_ctx.F = F
    X = B.copy()
    la.BLAS3.trsm(gpu_ctx, 1, L, X, la.LowerTriangular, True)
    rng = np.random.default_rng(3)
    P = np.uint64(p)
    for _ in range(2):
        y = rng.integers(0, p, size=nrhs, dtype=np.uint64)
        Xy = (X.astype(np.uint64) @ y) % P
        lhs = (L.astype(np.uint64) @ Xy) % P
        rhs = (B.astype(np.uint64) @ y) % P
        assert np.array_equal(lhs, rhs)


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("storage", ["uint16", "uint32", "double"])
def test_trsm_block_inverse_path(gpu_ctx, upper, unit, storage):
    """n = 2 * 1024 + 333, p <= 2^16: the two full 1024-row diagonal blocks
    are inverted up front (doubling from the 64-row inverses) and solved as
    gemms, the last partial block by the 64-row chain; bit-exact against
    the oracle."""
    p = 65521
    n, nrhs = 2381, 40
    print(f"trsm inverse path upper={upper} unit={unit} {storage}", flush=True)
    A = tri(31 + upper, n, p, upper, unit)
    B = synth.zp_matrix(33, n, nrhs, p)
    got = run_trsm(gpu_ctx, p, storage, upper, unit, 5, A, B)
    rc, _, want = orc.zp_trsm(p, upper, unit, 5, A, B)
    assert rc == 0 and np.array_equal(got, want)


@pytest.mark.parametrize("upper,unit,storage", [(False, True, "uint16"), (True, False, "uint16"),
                                                (False, False, "uint32"), (True, True, "double")])
def test_trsm_packed_rhs_path(gpu_ctx, upper, unit, storage):
    """n = 2 * 1024 and 128 right-hand sides: every block is an inverted
    leaf and the columns fill whole tiles, so the right-hand sides stay
    packed between the products (each product's epilogue packs the next
    one's operand); bit-exact against the oracle."""
    p = 65521
    n, nrhs = 2048, 128
    print(f"trsm packed rhs upper={upper} unit={unit} {storage}", flush=True)
    A = tri(35 + upper, n, p, upper, unit)
    B = synth.zp_matrix(37, n, nrhs, p)
    got = run_trsm(gpu_ctx, p, storage, upper, unit, 7, A, B)
    rc, _, want = orc.zp_trsm(p, upper, unit, 7, A, B)
    assert rc == 0 and np.array_equal(got, want)


def test_trsm_upper_nonunit_4096(gpu_ctx):
    """Upper, non-unit diagonal, n = 4096 (four inverted blocks), checked by
    Freivalds: A X == a B."""
    from test_gemm_gpu import matvec_mod
    p, n, nrhs = 65521, 4096, 300
    A = tri(41, n, p, True)
    B = synth.zp_matrix(42, n, nrhs, p)
    X = run_trsm(gpu_ctx, p, "uint32", True, False, 9, A, B)
    rng = np.random.default_rng(4)
    for _ in range(2):
        y = rng.integers(0, p, size=nrhs, dtype=np.uint64)
        lhs = matvec_mod(A, matvec_mod(X, y, p), p)
        rhs = (np.uint64(9) * matvec_mod(B, y, p)) % np.uint64(p)
        assert np.array_equal(lhs, rhs)
