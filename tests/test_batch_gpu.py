"""Batches of independent instances on the GPU (mlz_batch_*), against the oracle.

Every kernel variant of every op is exact, so each variant's output must equal
the oracle's bit for bit: gemm against LELA's ZpModule product
(orc_zp_gemm_u32), echelon forms against LELA's Elimination
(orc_zp_echelonize, reduced): the same R, rank and pivot columns.  The full
4096-instance workload is checked through size-independent properties:
Freivalds' check of every product, the row-echelon shape and row-space
containment of every R, and agreement of two different variant assignments;
plus the oracle on a subset."""
import numpy as np
import pytest

import oracle_lib as orc
from lela_amd import Modular, LELAError, synth
from lela_amd import batch as lb

pytestmark = pytest.mark.gpu

P16 = 65521


def make_instances(F, specs, seed):
    """specs: (op, m, n, k, density)."""
    rng = np.random.default_rng(seed)
    p = F.p
    out = []
    for op, m, n, k, d in specs:
        def mat(r, c):
            if d >= 1.0:
                X = rng.integers(0, p, size=(r, c), dtype=np.int64)
            else:
                X = np.where(rng.random((r, c)) < d, rng.integers(1, p, size=(r, c), dtype=np.int64), 0)
            return F.from_canonical(X.astype(np.uint64))
        if op == lb.TASK_GEMM:
            C = F.from_canonical(rng.integers(0, p, size=(m, n), dtype=np.int64).astype(np.uint64))
            out.append(lb.Instance(op, mat(m, k), mat(k, n), C, a=int(rng.integers(0, p)),
                                   b=int(rng.integers(0, p)), density=min(d, 1.0)))
        else:
            out.append(lb.Instance(op, mat(m, n), density=min(d, 1.0)))
    return out


SPECS = [
    (1, 1, 1, 1, 1.0), (1, 1, 77, 3, 1.0), (1, 65, 1, 130, 1.0), (1, 64, 64, 64, 1.0),
    (1, 100, 37, 200, 0.05), (1, 129, 257, 65, 1.0), (1, 33, 300, 500, 0.0), (1, 200, 150, 97, 0.2),
    (2, 1, 1, 0, 1.0), (2, 1, 50, 0, 1.0), (2, 50, 1, 0, 1.0), (2, 70, 70, 0, 1.0),
    (2, 130, 60, 0, 1.0), (2, 60, 130, 0, 1.0), (2, 90, 90, 0, 0.02), (2, 40, 200, 0, 0.0),
    (2, 150, 150, 0, 0.1), (2, 33, 97, 0, 0.5),
]


def rank_deficient(F, m, n, r, seed):
    rng = np.random.default_rng(seed)
    L = rng.integers(0, F.p, size=(m, r), dtype=np.int64)
    R = rng.integers(0, F.p, size=(r, n), dtype=np.int64)
    X = (L.astype(object).dot(R.astype(object)) % F.p).astype(np.uint64)
    return lb.Instance(lb.TASK_RREF, F.from_canonical(X))


def check_against_oracle(F, ws, raw, ins_list):
    p = F.p
    for i, ins in enumerate(ins_list):
        if ins.op == lb.TASK_GEMM:
            want = orc.zp_gemm(p, ins.a % p, F.to_canonical(ins.A), F.to_canonical(ins.B), ins.b % p,
                               F.to_canonical(ins.C))
            got = F.to_canonical(ws.result(i, raw))
            assert np.array_equal(got, want), f"gemm instance {i} ({ins.m}x{ins.n}x{ins.k})"
        else:
            rc, R, rank, piv, _ = orc.zp_echelonize(p, F.to_canonical(ins.A), reduced=True)
            assert rc == orc.ORC_OK
            gotR, gotp = ws.result(i, raw)
            assert int(ws.tasks[i].rank) == rank, f"rref instance {i} rank"
            assert np.array_equal(gotp, piv), f"rref instance {i} pivots"
            assert np.array_equal(F.to_canonical(gotR), R), f"rref instance {i} ({ins.m}x{ins.n})"


@pytest.mark.parametrize("storage,p", [("uint16", P16), ("uint32", P16), ("uint32", 4294967291),
                                       ("double", 67108859), ("uint16", 3)])
def test_every_variant_matches_oracle(gpu_ctx, storage, p):
    F = Modular(p, storage)
    ins = make_instances(F, SPECS, seed=100 + p % 97) + [rank_deficient(F, 120, 90, 37, 5),
                                                        rank_deficient(F, 66, 140, 65, 6)]
    ws = lb.Workload(gpu_ctx, F, ins)
    try:
        for v in range(4):
            if not any(lb.variant_ok(x.op, v, p, x.k) for x in ins):
                continue
            ws.upload()
            ws.set_variants([v if lb.variant_ok(x.op, v, p, x.k) else 0 for x in ins])
            plan = lb.Plan(gpu_ctx, F, ws.tasks, lb.POLICY_FIXED)
            plan.launch()
            plan.finish()
            plan.close()
            print(f"variant {v}: storage {storage} p {p}", flush=True)
            check_against_oracle(F, ws, ws.download(), ins)
    finally:
        ws.free()


def test_graph_replays_match_eager(gpu_ctx):
    """A captured plan replayed several times, with other GPU work and host
    synchronisation in between, gives the eager result every time (gemm with
    b = 0 and echelon forms are idempotent under replay)."""
    F = Modular(P16, "uint16")
    ins = make_instances(F, SPECS, seed=7)
    for x in ins:
        x.b = 0
    ws = lb.Workload(gpu_ctx, F, ins)
    try:
        eager = lb.Plan(gpu_ctx, F, ws.tasks, lb.POLICY_TABLE)
        eager.launch()
        eager.finish()
        eager.close()
        want = ws.download()
        ranks = [int(t.rank) for t in ws.tasks]
        ws.upload()
        plan = lb.Plan(gpu_ctx, F, ws.tasks, lb.POLICY_TABLE, graph=True)
        for rep in range(4):
            plan.launch()
            plan.finish()
            got = ws.download()
            assert np.array_equal(got, want), f"replay {rep}"
            assert [int(t.rank) for t in ws.tasks] == ranks
            # unrelated eager work between replays
            from lela_amd import BLAS3
            gpu_ctx.F = F
            A = np.ones((64, 64), dtype=np.uint16)
            C = np.zeros((64, 64), dtype=np.uint16)
            BLAS3.gemm(gpu_ctx, 1, A, A, 0, C)
            assert np.all(C == 64)
        plan.close()
    finally:
        ws.free()


def test_policies_give_identical_results(gpu_ctx):
    F = Modular(P16, "uint16")
    ins = make_instances(F, SPECS, seed=8)
    ws = lb.Workload(gpu_ctx, F, ins)
    try:
        outs = []
        for policy, graph, serial in ((lb.POLICY_TABLE, False, False), (lb.POLICY_MODEL, True, False),
                                      (lb.POLICY_TABLE, False, True)):
            ws.upload()
            ws.set_variants([-1] * len(ins))
            plan = lb.Plan(gpu_ctx, F, ws.tasks, policy, graph=graph, serial=serial)
            plan.launch()
            plan.finish()
            plan.close()
            outs.append(ws.download())
        assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
        check_against_oracle(F, ws, outs[0], ins)
    finally:
        ws.free()


def test_composite_modulus_reports_non_invertible_pivot(gpu_ctx):
    F = Modular(15, "uint16")
    # rank one; the pivot block [3] has no inverse mod 15
    ins = [lb.Instance(lb.TASK_RREF, np.array([[3, 1], [6, 2]], dtype=np.uint16))]
    ws = lb.Workload(gpu_ctx, F, ins)
    try:
        plan = lb.Plan(gpu_ctx, F, ws.tasks, lb.POLICY_TABLE)
        plan.launch()
        with pytest.raises(LELAError):
            plan.finish()
        plan.close()
    finally:
        ws.free()


def test_timing_reports_positive_costs(gpu_ctx):
    F = Modular(P16, "uint16")
    ins = make_instances(F, SPECS, seed=9)
    ws = lb.Workload(gpu_ctx, F, ins)
    try:
        plan = lb.Plan(gpu_ctx, F, ws.tasks, lb.POLICY_TABLE, timing=True, serial=True)
        plan.launch()
        cost = plan.finish()
        plan.close()
        assert cost.shape == (len(ins),) and np.all(cost > 0) and np.all(np.isfinite(cost))
    finally:
        ws.free()


def freivalds_ok(F, ins, C_new, rng):
    p = F.p
    x = rng.integers(0, p, size=ins.n, dtype=np.int64)
    A = F.to_canonical(ins.A).astype(np.int64)
    B = F.to_canonical(ins.B).astype(np.int64)
    C0 = F.to_canonical(ins.C).astype(np.int64)
    C1 = F.to_canonical(C_new).astype(np.int64)
    Bx = (B @ x) % p
    lhs = (C1 @ x) % p
    rhs = (ins.a * ((A @ Bx) % p) + ins.b * ((C0 @ x) % p)) % p
    return np.array_equal(lhs, rhs)


def rref_ok(F, ins, R, piv, rank, rng):
    """R (canonical) in reduced echelon form with pivots piv, and X = X[:, piv] R
    (every row of X is in R's row space), checked at a random vector."""
    p = F.p
    R = R.astype(np.int64)
    if rank and not np.all(np.diff(piv.astype(np.int64)) > 0):
        return False
    if np.any(R[rank:]):
        return False
    for t, c in enumerate(piv):
        col = R[:rank, c]
        if col[t] != 1 or np.count_nonzero(col) != 1 or np.any(R[t, :c]):
            return False
    X = F.to_canonical(ins.A).astype(np.int64)
    v = rng.integers(0, p, size=ins.n, dtype=np.int64)
    lhs = (X @ v) % p
    rhs = (X[:, piv] @ ((R[:rank] @ v) % p)) % p if rank else np.zeros_like(lhs)
    return np.array_equal(lhs, rhs)


@pytest.mark.timeout(900)
def test_uneven_batch_full_size(gpu_ctx):
    """The 4096-instance workload (m, n, k uniform in [32, 1024], half sparse):
    Freivalds on every product, echelon shape and row-space containment of
    every R, two variant assignments bit-identical, oracle on a subset."""
    F = Modular(P16, "uint16")
    raw = synth.batch_instances(synth.seed_for(5, 0), count=4096, p=P16)
    ins = [lb.Instance(d["op"], d["A"], d.get("B"), d.get("C"), density=d["density"]) for d in raw]
    del raw
    print("workload generated", flush=True)
    ws = lb.Workload(gpu_ctx, F, ins)
    try:
        plan = lb.Plan(gpu_ctx, F, ws.tasks, lb.POLICY_TABLE, graph=True)
        plan.launch()
        plan.finish()
        plan.close()
        out1 = ws.download()
        ranks1 = [int(t.rank) for t in ws.tasks]
        print("first assignment done", flush=True)
        # second assignment: another applicable variant for every instance
        alt = []
        for x, t in zip(ins, ws.tasks):
            nv = lb.nvariants(x.op)
            alt.append((t.variant + 1) % nv)
        ws.upload()
        ws.set_variants(alt)
        plan = lb.Plan(gpu_ctx, F, ws.tasks, lb.POLICY_FIXED)
        plan.launch()
        plan.finish()
        plan.close()
        out2 = ws.download()
        assert [int(t.rank) for t in ws.tasks] == ranks1
        assert np.array_equal(out1, out2), "variant assignments disagree"
        print("second assignment agrees", flush=True)
        rng = np.random.default_rng(1)
        for i, x in enumerate(ins):
            if x.op == lb.TASK_GEMM:
                assert freivalds_ok(F, x, ws.result(i, out1), rng), f"gemm instance {i}"
            else:
                R, piv = ws.result(i, out1)
                assert rref_ok(F, x, F.to_canonical(R), piv, ranks1[i], rng), f"rref instance {i}"
        print("properties checked", flush=True)
        # oracle on the smallest instances of each kind (seconds of CPU work)
        sizes = sorted(range(len(ins)), key=lambda i: ins[i].m * ins[i].n * max(ins[i].k, min(ins[i].m, ins[i].n)))
        for i in sizes[:24]:
            x = ins[i]
            if x.op == lb.TASK_GEMM:
                want = orc.zp_gemm(P16, 1, x.A, x.B, 1, x.C)
                assert np.array_equal(ws.result(i, out1).astype(np.uint32), want), f"instance {i}"
            else:
                rc, R, rank, piv, _ = orc.zp_echelonize(P16, x.A.astype(np.uint32))
                gotR, gotp = ws.result(i, out1)
                assert rank == ranks1[i] and np.array_equal(gotp, piv)
                assert np.array_equal(gotR.astype(np.uint32), R), f"instance {i}"
    finally:
        ws.free()
