"""GPU parity of mlz_zp_echelonize (LELA's non-reduced echelon forms with
transposition pivoting) against the oracle's restatement of
Elimination::echelonize and Elimination::pluq, bit for bit: the output
matrix, the rank, the determinant and the transposition lists, for all three
modes (U, the compute_L transform, pluq), every storage, moduli from 3 to
2^32-5, shapes that fit one 128-row window and shapes that need several,
rank-deficient, zero-column and sparse inputs (whose pivots lie beyond the
window, so the search over all rows runs), and the composite-modulus error.
At a larger size the factorisation is checked through T P A = U."""
import numpy as np
import pytest

import lela_amd as la
import oracle_lib as orc
from lela_amd import synth

pytestmark = pytest.mark.gpu


def _lowrank(seed, m, n, r, p):
    rng = np.random.default_rng(seed)
    L = rng.integers(0, p, size=(m, r)).astype(object)
    R = rng.integers(0, p, size=(r, n)).astype(object)
    return (L.dot(R) % p).astype(np.uint32)


def _inputs(p):
    rng = np.random.default_rng(p % 1000 + 7)
    out = [("dense 50x70", synth.zp_matrix(3, 50, 70, p)),
           ("tall 300x90", synth.zp_matrix(4, 300, 90, p)),
           ("wide 140x400", synth.zp_matrix(5, 140, 400, p)),
           ("rank 60 of 200x180", _lowrank(6, 200, 180, 60, p)),
           ("rank 130 of 260x260", _lowrank(7, 260, 260, 130, p))]
    Z = synth.zp_matrix(8, 180, 150, p)
    Z[:, [0, 1, 64, 65, 100]] = 0
    Z[3] = 0
    out.append(("zero columns", Z))
    S = np.where(rng.random((400, 300)) < 0.004, rng.integers(1, p, size=(400, 300)), 0).astype(np.uint32)
    out.append(("sparse 400x300", S))
    T = np.zeros((300, 200), dtype=np.uint32)
    T[250:, :] = synth.zp_matrix(9, 50, 200, p)      # the first nonzero rows far down
    out.append(("late rows", T))
    return out


def run(ctx, p, storage, A, mode):
    F = la.Modular(p, storage)
    ctx.F = F
    X = F.from_canonical(A)
    E = la.Elimination(ctx)
    if mode == "pluq":
        P, Q, rank, det = E.pluq(X)
        return F.to_canonical(X), rank, det, P, Q
    P, rank, det = E.echelonize(X, compute_L=(mode == "transform"))
    return F.to_canonical(X), rank, det, P, None


def lela_lists(perm):
    return [(i, int(j)) for i, j in enumerate(perm) if j != i]


@pytest.mark.parametrize("mode", ["u", "transform", "pluq"])
@pytest.mark.parametrize("storage,p", [("uint16", 65521), ("uint32", 65521), ("uint32", 101), ("uint16", 3),
                                       ("uint32", 2147483647), ("uint32", 4294967291), ("double", 67108859)])
def test_echelonize_matches_oracle(gpu_ctx, mode, storage, p):
    for name, A in _inputs(p):
        print(f"echelonize {mode} {storage} p={p} {name}", flush=True)
        got, rank, det, P, Q = run(gpu_ctx, p, storage, A, mode)
        if mode == "pluq":
            rc, want, wrank, wdet, pr, pc = orc.zp_pluq(p, A)
            assert rc == 0
            assert Q == lela_lists(pc)[::-1], name
            assert P == lela_lists(pr)[::-1], name
        else:
            rc, want, wrank, wdet, pr = orc.zp_elim(p, A, compute_L=(mode == "transform"))
            assert rc == 0
            assert P == lela_lists(pr), name
        assert rank == wrank and det == wdet, name
        assert np.array_equal(got, want), name


def test_echelon_form_non_reduced(gpu_ctx):
    """EchelonForm::echelonize (A, false) is Elimination's U (echelon-form.h:113-131)."""
    p = 65521
    A = _lowrank(11, 150, 170, 90, p)
    F = la.Modular(p, "uint32")
    gpu_ctx.F = F
    X = A.copy()
    EF = la.EchelonForm(gpu_ctx)
    EF.echelonize(X, reduced=False)
    rc, want, wrank, _, _ = orc.zp_elim(p, A, compute_L=False)
    assert EF.rank(X) == wrank and np.array_equal(X, want)


def test_echelonize_composite_modulus(gpu_ctx):
    """A pivot without an inverse mod 12 raises as LELA does (elimination.tcc:103-104)."""
    p = 12
    A = np.array([[4, 1, 0], [0, 2, 3], [1, 1, 1]], dtype=np.uint32)
    rc, *_ = orc.zp_elim(p, A)
    assert rc == orc.ORC_E_NOT_INVERTIBLE
    with pytest.raises(la.LELAError):
        run(gpu_ctx, p, "uint32", A, "u")


@pytest.mark.timeout(600)
def test_transform_property_2048(gpu_ctx):
    """2048 x 2500 rank 1900: T P A = U checked at random vectors, U in echelon
    form with the pivots' product as det."""
    from test_gemm_gpu import matvec_mod
    p = 65521
    rng = np.random.default_rng(12)
    # a product of 16-bit factors over k = 1900: every partial sum is below
    # 2^43, exact in float64
    Lf = rng.integers(0, p, size=(2048, 1900)).astype(np.float64)
    Rf = rng.integers(0, p, size=(1900, 2500)).astype(np.float64)
    A = np.mod(Lf @ Rf, p).astype(np.uint32)
    del Lf, Rf
    print("transform 2048 x 2500", flush=True)
    got, rank, det, P, _ = run(gpu_ctx, p, "uint32", A, "transform")
    assert rank == 1900
    m, n = A.shape
    order = list(range(m))
    for i, j in P:
        order[i], order[j] = order[j], order[i]
    PA = A[order]
    U = np.triu(got.astype(np.uint64), 0)
    T = np.tril(got.astype(np.uint64)[:, :m], -1)
    rng = np.random.default_rng(5)
    for _ in range(2):
        v = rng.integers(0, p, size=n, dtype=np.uint64)
        w = matvec_mod(PA, v, p)
        lhs = (w + matvec_mod(T, w, p)) % np.uint64(p)
        assert np.array_equal(lhs, matvec_mod(U, v, p))
    heads = [int(np.nonzero(U[t])[0][0]) for t in range(rank)]
    assert all(b > a for a, b in zip(heads, heads[1:]))
    assert not U[rank:].any()
    d = 1
    for t, h in enumerate(heads):
        d = d * int(U[t, h]) % p
    assert d == det
