"""Randomized GPU parity against the oracle (bit-exact): every entry point at
seeded random shapes, moduli, storages and kernel families, on inputs drawn
from four kinds (uniform, sparse, low rank, uniform with zero columns).
Zero columns and rank-deficient panels reach the branches that skip columns
and the tournament pivot search; ragged shapes reach every tile edge.

Entry points and the reference interfaces they replace:
  gemm  BLAS3::gemm (lela/blas/level3.h:81-83)
  trsm  BLAS3::trsm (lela/blas/level3.h:120-122)
  rref  EchelonForm::echelonize(reduced=true) (lela/solutions/echelon-form.h:103-140)
  fl    FaugereLachartre::echelonize (lela/algorithms/faugere-lachartre.tcc:735-917);
        det is the product of the scanned heads of X (:119), the factor the
        GPU call returns
  gf2   word-packed GF(2) gemm and echelon form
  batch batched gemm / echelon instances under both planning policies"""
import numpy as np
import pytest

import lela_amd as la
import oracle_lib as orc
from lela_amd import batch as lb
from lela_amd import synth

pytestmark = pytest.mark.gpu

ALG = {"auto": la.ALGO_AUTO, "plain": la.ALGO_PLAIN, "int8": la.ALGO_INT8, "fp64": la.ALGO_FP64}
P16 = [2, 3, 5, 251, 32749, 65521]
P32 = P16 + [65537, 1000003, 33554393, 67108859, 2147483647, 4294967291]


def pick_field(rng):
    storage = str(rng.choice(["uint16", "uint32", "double"]))
    if storage == "uint16":
        pool = P16
    elif storage == "double":
        pool = [p for p in P32 if p < 1 << 26]
    else:
        pool = P32
    p = int(rng.choice(pool))
    algs = ["auto", "plain"] + (["int8"] if p <= 1 << 16 else []) + (["fp64"] if p < 1 << 26 else [])
    return storage, p, str(rng.choice(algs))


def rand_mat(rng, m, n, p, kind):
    """kind 0 uniform, 1 sparse, 2 low rank (a product over Z), 3 zero columns."""
    if kind == 0:
        return rng.integers(0, p, size=(m, n)).astype(np.uint64)
    if kind == 1:
        d = float(rng.uniform(0.002, 0.3))
        return np.where(rng.random((m, n)) < d, rng.integers(1, p, size=(m, n)), 0).astype(np.uint64)
    if kind == 2:
        r = int(rng.integers(0, max(1, min(m, n)) + 1))
        if r == 0:
            return np.zeros((m, n), dtype=np.uint64)
        L = rng.integers(0, p, size=(m, r)).astype(object)
        R = rng.integers(0, p, size=(r, n)).astype(object)
        return (L.dot(R) % p).astype(np.uint64)
    X = rng.integers(0, p, size=(m, n)).astype(np.uint64)
    X[:, rng.random(n) < float(rng.uniform(0.1, 0.6))] = 0
    return X


def test_random_gemm(gpu_ctx):
    rng = np.random.default_rng(101)
    for it in range(16):
        storage, p, alg = pick_field(rng)
        F = la.Modular(p, storage)
        gpu_ctx.F = F
        m, n = (int(x) for x in rng.integers(1, 400, size=2))
        k = int(rng.integers(1, 1200))
        A, B, C = rand_mat(rng, m, k, p, it % 4), rand_mat(rng, k, n, p, (it + 1) % 4), rand_mat(rng, m, n, p, 0)
        a, b = int(rng.integers(0, p)), int(rng.integers(0, p))
        Cs = F.from_canonical(C)
        la.BLAS3.gemm(gpu_ctx, a, F.from_canonical(A), F.from_canonical(B), b, Cs, algo=ALG[alg])
        want = orc.zp_gemm(p, a, A.astype(np.uint32), B.astype(np.uint32), b, C.astype(np.uint32))
        assert np.array_equal(F.to_canonical(Cs), want), (storage, p, alg, m, n, k, a, b)


def test_random_trsm(gpu_ctx):
    rng = np.random.default_rng(102)
    for it in range(12):
        storage, p, alg = pick_field(rng)
        F = la.Modular(p, storage)
        gpu_ctx.F = F
        n, nrhs = int(rng.integers(1, 700)), int(rng.integers(1, 500))
        upper, unit = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
        A = rand_mat(rng, n, n, p, int(rng.integers(0, 2)))
        A = np.triu(A) if upper else np.tril(A)
        d = np.arange(n)
        A[d, d] = rng.integers(1, p, size=n) if p > 2 else 1
        B = rand_mat(rng, n, nrhs, p, int(rng.integers(0, 4)))
        a = int(rng.integers(0, p))
        Bs = F.from_canonical(B)
        la.BLAS3.trsm(gpu_ctx, a, F.from_canonical(A), Bs, la.UpperTriangular if upper else la.LowerTriangular,
                      unit, algo=ALG[alg])
        rc, _, want = orc.zp_trsm(p, upper, unit, a, A.astype(np.uint32), B.astype(np.uint32))
        assert rc == 0 and np.array_equal(F.to_canonical(Bs), want), (storage, p, alg, n, nrhs, upper, unit)


def test_random_rref(gpu_ctx):
    rng = np.random.default_rng(103)
    for it in range(16):
        storage, p, alg = pick_field(rng)
        F = la.Modular(p, storage)
        gpu_ctx.F = F
        m, n = (int(x) for x in rng.integers(1, 600, size=2))
        X = rand_mat(rng, m, n, p, it % 4)
        Xs = F.from_canonical(X)
        EF = la.EchelonForm(gpu_ctx)
        EF.echelonize(Xs, algo=ALG[alg])
        rc, R, rank, piv, _ = orc.zp_echelonize(p, X.astype(np.uint32), reduced=True)
        assert rc == 0 and EF.rank(Xs) == rank and list(EF.pivots) == list(piv), (storage, p, alg, m, n, it % 4)
        assert np.array_equal(F.to_canonical(Xs), R), (storage, p, alg, m, n, it % 4)


def test_random_fl(gpu_ctx):
    rng = np.random.default_rng(104)
    for it in range(8):
        p = int(rng.choice([3, 101, 65521, 2147483647]))
        storage = "uint16" if p < 1 << 16 and rng.random() < 0.5 else "uint32"
        F = la.Modular(p, storage)
        gpu_ctx.F = F
        na, nb, nc = int(rng.integers(20, 700)), int(rng.integers(1, 700)), int(rng.integers(0, 300))
        band = int(rng.integers(1, 100))
        nnz_a = int(rng.integers(1, min(band, 30) + 1))
        csr, shape = synth.fl_matrix(int(rng.integers(1, 1 << 30)), p, na=na, nc=nc, nb=nb, band=band, nnz_a=nnz_a,
                                     density=float(rng.uniform(0.005, 0.2)))
        X = synth.csr_to_dense(csr, shape)
        alg = str(rng.choice(["auto", "plain"] + (["int8"] if p <= 1 << 16 else [])))
        R = np.zeros(X.shape, dtype=F.dtype)
        FL = la.FaugereLachartre(gpu_ctx)
        rank, det = FL.echelonize(R, F.from_canonical(X.astype(np.uint64)), algo=ALG[alg])
        rc, want, wrank, wdet, _ = orc.fl_echelonize(p, X.astype(np.uint32))
        case = (storage, p, alg, na, nb, nc, band)
        assert rc == 0 and rank == wrank and np.array_equal(F.to_canonical(R), want), case
        assert det == wdet, case


def test_random_gf2(gpu_ctx):
    rng = np.random.default_rng(105)
    gpu_ctx.F = la.GF2()
    for it in range(8):
        m, n, k = (int(x) for x in rng.integers(1, 1500, size=3))
        A, B, C = (synth.gf2_matrix(int(rng.integers(1, 1 << 30)), r, c) for r, c in ((m, k), (k, n), (m, n)))
        a, b = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        Cw = C.copy()
        la.BLAS3.gemm(gpu_ctx, a, la.Dense01Matrix(A, k), la.Dense01Matrix(B, n), b, la.Dense01Matrix(Cw, n))
        assert np.array_equal(Cw, orc.gf2_gemm(m, n, k, a, A, B, b, C)), (m, n, k, a, b)
    for it in range(24):
        m, n = (int(x) for x in rng.integers(1, 1800, size=2))
        nw = (n + 63) // 64
        kind = it % 3
        if kind == 0:
            A = synth.gf2_matrix(int(rng.integers(1, 1 << 30)), m, n)
        else:
            r = int(rng.integers(1, max(2, min(m, n))))
            X = synth.gf2_matrix(int(rng.integers(1, 1 << 30)), m, r)
            Y = synth.gf2_matrix(int(rng.integers(1, 1 << 30)), r, n)
            A = orc.gf2_gemm(m, n, r, 1, X, Y, 0, np.zeros((m, nw), dtype=np.uint64))
        if kind == 2 or rng.random() < 0.5:
            keep = (rng.random(nw * 64) >= float(rng.uniform(0.05, 0.7))).astype(np.uint64)
            A = A & (keep.reshape(-1, 64) << np.arange(64, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
        D = la.Dense01Matrix(A.copy(), n)
        EF = la.EchelonForm(gpu_ctx)
        EF.echelonize(D)
        want, wrank, wpiv = orc.gf2_echelonize(m, n, A, reduced=True)
        assert EF.rank(D) == wrank and list(EF.pivots) == list(wpiv), (m, n, kind)
        assert np.array_equal(D.words, want), (m, n, kind)


def test_random_batch(gpu_ctx):
    rng = np.random.default_rng(106)
    for it in range(2):
        storage = str(rng.choice(["uint16", "uint32"]))
        p = int(rng.choice(P16))
        F = la.Modular(p, storage)
        gpu_ctx.F = F
        ins = []
        for i in range(50):
            m, n = (int(x) for x in rng.integers(1, 400, size=2))
            if i % 2:
                k = int(rng.integers(1, 400))
                ins.append(lb.Instance(lb.TASK_GEMM, F.from_canonical(rand_mat(rng, m, k, p, i % 4)),
                                       F.from_canonical(rand_mat(rng, k, n, p, 1)),
                                       F.from_canonical(rand_mat(rng, m, n, p, 0)),
                                       a=int(rng.integers(0, p)), b=int(rng.integers(0, p))))
            else:
                ins.append(lb.Instance(lb.TASK_RREF, F.from_canonical(rand_mat(rng, m, n, p, (i // 2) % 4))))
        ws = lb.Workload(gpu_ctx, F, ins)
        try:
            for pol in (lb.POLICY_TABLE, lb.POLICY_MODEL):
                ws.upload()
                ws.set_variants([-1] * ws.n)
                plan = lb.Plan(gpu_ctx, F, ws.tasks, pol)
                plan.launch()
                plan.finish()
                plan.close()
                raw = ws.download()
                for i, x in enumerate(ins):
                    case = (storage, p, pol, i, x.m, x.n)
                    if x.op == lb.TASK_GEMM:
                        want = orc.zp_gemm(p, x.a % p, F.to_canonical(x.A), F.to_canonical(x.B), x.b % p,
                                           F.to_canonical(x.C))
                        assert np.array_equal(F.to_canonical(ws.result(i, raw)), want), case
                    else:
                        rc, R, rank, piv, _ = orc.zp_echelonize(p, F.to_canonical(x.A), reduced=True)
                        gR, gp = ws.result(i, raw)
                        assert int(ws.tasks[i].rank) == rank and np.array_equal(gp, piv), case
                        assert np.array_equal(F.to_canonical(gR), R), case
        finally:
            ws.free()
