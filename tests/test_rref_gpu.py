"""GPU parity of mlz_zp_rref against the oracle's Elimination::echelonize_reduced
(bit-exact RREF, rank and pivot columns): random, low-rank, sparse, tall,
wide, zero and ragged matrices, every storage and gemm family."""
import numpy as np
import pytest

import bruteforce as bf
import lela_amd as la
import oracle_lib as orc
from lela_amd import synth

pytestmark = pytest.mark.gpu


def lowrank(seed, m, n, r, p):
    X = synth.zp_matrix(seed, m, r, p).astype(np.uint64)
    Y = synth.zp_matrix(seed + 1, r, n, p).astype(np.uint64)
    out = np.zeros((m, n), dtype=np.uint64)
    for i in range(0, m, 256):
        out[i:i + 256] = (X[i:i + 256] @ Y) % np.uint64(p) if p < (1 << 20) else \
            np.array(bf.matmul_mod(X[i:i + 256], Y, p), dtype=np.uint64)
    return out.astype(np.uint32)


def sparse(seed, m, n, p, density):
    rng = np.random.default_rng(seed)
    A = synth.zp_matrix(seed, m, n, p)
    A[rng.random((m, n)) >= density] = 0
    return A


def run_rref(ctx, p, storage, A, algo=la.ALGO_AUTO):
    F = la.Modular(p, storage)
    ctx.F = F
    X = F.from_canonical(A)
    EF = la.EchelonForm(ctx)
    EF.echelonize(X, reduced=True, algo=algo)
    return F.to_canonical(X), EF.rank(X), EF.pivots


def check(ctx, p, storage, A, algo=la.ALGO_AUTO):
    got, rank, piv = run_rref(ctx, p, storage, A, algo)
    rc, want, wrank, wpiv, _ = orc.zp_echelonize(p, A, reduced=True)
    assert rc == 0
    assert rank == wrank and list(piv) == list(wpiv)
    assert np.array_equal(got, want)


CASES = [
    ("square", lambda p: synth.zp_matrix(1, 100, 100, p)),
    ("lowrank", lambda p: lowrank(2, 150, 170, 37, p)),
    ("tall", lambda p: synth.zp_matrix(3, 700, 90, p)),
    ("wide", lambda p: synth.zp_matrix(4, 60, 400, p)),
    ("sparse1pct", lambda p: sparse(5, 300, 260, p, 0.01)),
    ("sparse10pct", lambda p: sparse(6, 257, 333, p, 0.10)),
    ("zero", lambda p: np.zeros((70, 80), dtype=np.uint32)),
    ("one", lambda p: np.ones((1, 1), dtype=np.uint32)),
    ("ragged", lambda p: lowrank(7, 513, 129, 100, p)),
    # n >= 1024: 512-column blocks with a trailing gemm per block (the last
    # block partial)
    ("blocks_wide", lambda p: synth.zp_matrix(12, 120, 1100, p)),
    ("blocks_lowrank", lambda p: lowrank(13, 400, 1300, 300, p)),
    ("blocks_sparse", lambda p: sparse(14, 300, 1100, p, 0.02)),
    ("blocks_square", lambda p: synth.zp_matrix(15, 600, 1030, p)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("p", [2, 3, 65521, 1000003, 8388593, 2147483647])
def test_rref_matches_oracle(gpu_ctx, name, make, p):
    print(f"rref {name} p={p}", flush=True)
    check(gpu_ctx, p, "uint32", make(p))


@pytest.mark.parametrize("storage,p", [("uint16", 65521), ("double", 67108859), ("uint16", 251)])
def test_rref_storages(gpu_ctx, storage, p):
    check(gpu_ctx, p, storage, lowrank(8, 200, 230, 150, p))
    check(gpu_ctx, p, storage, lowrank(16, 260, 1100, 240, p))


@pytest.mark.parametrize("algo", [la.ALGO_PLAIN, la.ALGO_INT8, la.ALGO_FP64])
def test_rref_gemm_families(gpu_ctx, algo):
    check(gpu_ctx, 65521, "uint32", lowrank(9, 300, 280, 200, 65521), algo=algo)
    check(gpu_ctx, 65521, "uint32", lowrank(17, 350, 1200, 330, 65521), algo=algo)


def test_rref_many_pivot_chunks(gpu_ctx):
    # more than 256 rows per panel with pivots spread over several chunks and a
    # rank deficit inside panels
    p = 65521
    A = lowrank(10, 1500, 600, 420, p)
    A[::7] = 0
    check(gpu_ctx, p, "uint32", A)


def test_rref_structure_large(gpu_ctx):
    """2000 x 3000 rank 1500: RREF structure, rank and row-space checks
    (size-independent properties; the oracle is too slow at full size)."""
    p = 65521
    A = lowrank(11, 2000, 3000, 1500, p)
    got, rank, piv = run_rref(gpu_ctx, p, "uint32", A)
    assert rank == 1500
    R = got[:rank].astype(np.uint64)
    P = np.uint64(p)
    # pivot columns are an identity block, rows below rank are zero
    assert np.array_equal(R[:, piv], np.eye(rank, dtype=np.uint64))
    assert not got[rank:].any()
    assert all(R[i, :piv[i]].sum() == 0 for i in range(0, rank, 97))
    # every row of A is the combination A[i, piv] of the RREF rows
    sel = np.arange(0, 2000, 37)
    recon = (A[sel][:, piv].astype(np.uint64) @ R) % P
    assert np.array_equal(recon, A[sel].astype(np.uint64))


@pytest.mark.parametrize("threads", [256, 512, 1024])
@pytest.mark.parametrize("blocks", [1, 2, 4])
def test_rref_tuning_options(gpu_ctx, threads, blocks):
    """The pivot-search workgroup size and the update strip width change speed only."""
    p = 65521
    gpu_ctx.set_option("rref.search.threads", threads)
    gpu_ctx.set_option("rref.update.blocks", blocks)
    try:
        for A in (synth.zp_matrix(31, 300, 500, p), lowrank(32, 260, 300, 70, p), sparse(33, 200, 330, p, 0.02)):
            got, rank, piv = run_rref(gpu_ctx, p, "uint32", A)
            rc, want, wrank, wpiv, _ = orc.zp_echelonize(p, A, reduced=True)
            assert rank == wrank and np.array_equal(got, want) and list(piv) == list(wpiv)
    finally:
        gpu_ctx.set_option("rref.search.threads", 1024)
        gpu_ctx.set_option("rref.update.blocks", 1)


@pytest.mark.parametrize("lds", [0, 1, 2, 4])
def test_update_staging_options(gpu_ctx, lds):
    """The INT8 trailing update with its tiles in registers or staged in LDS
    (1, 2, 4 per wave; partial row blocks, ragged column tails, a row stride
    that is not a multiple of 4) changes speed only: the reduced form and the
    non-reduced LU form against the oracle."""
    p = 65521
    gpu_ctx.set_option("update.i8.lds", lds)
    try:
        for A in (synth.zp_matrix(41, 333, 520, p), lowrank(42, 300, 416, 90, p), synth.zp_matrix(43, 200, 203, p)):
            got, rank, piv = run_rref(gpu_ctx, p, "uint32", A)
            rc, want, wrank, wpiv, _ = orc.zp_echelonize(p, A, reduced=True)
            assert rank == wrank and np.array_equal(got, want) and list(piv) == list(wpiv)
            F = la.Modular(p, "uint32")
            gpu_ctx.F = F
            X = F.from_canonical(A)
            P, lrank, det = la.Elimination(gpu_ctx).echelonize(X, compute_L=False)
            rc, U, erank, edet, _ = orc.zp_elim(p, A, compute_L=False)
            assert lrank == erank and det == edet and np.array_equal(F.to_canonical(X), U)
    finally:
        gpu_ctx.set_option("update.i8.lds", -1)


def test_rref_composite_modulus(gpu_ctx):
    """p = 15: a pivot without inverse is reported (LELA's check,
    elimination.tcc:217-220); with invertible pivots the result is the
    oracle's."""
    with pytest.raises(la.LELAError):
        run_rref(gpu_ctx, 15, "uint32", np.array([[3, 1], [6, 2]], dtype=np.uint32))
    check(gpu_ctx, 15, "uint32", np.array([[2, 1, 7], [1, 1, 4]], dtype=np.uint32))
