"""Benchmark: exact Z/p gemm, 8192 x 8192 x 8192, p = 65521, uint16 storage,
C <- A*B + C on one MI355X with split-int8 MFMA; inputs resident in HBM.

Prints one JSON line.  `value` is modular MACs/s of the int8 kernel family
over the timed steps; `roofline` covers the MFMA kernel alone, timed with HIP
events on the library's stream; `variants` holds the plain VALU kernel and the
FP64-carrier kernel on the same shape; `fl_echelon` is the Faugere-Lachartre
reduction of the 20000 x 30000 F4-shaped matrix (seconds, MFMA families vs
plain kernels); `rows` times the other level-3 workloads; `batched` runs
4096 uneven gemm/echelon-form instances in captured batches (cost model,
autotuner table, plain kernels, one call per instance); `cpu_baseline`
times the CPU restatement of LELA's ZpModule<uint16> gemm on a bounded block
of rows of the same workload, one core.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 8192
P = 65521
P_FP64 = 67108859
INT8_PEAK_TOPS = 5000.0      # MI355X dense INT8 MFMA, 5.0 POP/s
FP64_PEAK_TFLOPS = 78.6
# LDS: 256 CUs x 128 bytes per clock at 2.4 GHz; VALU: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12


def mfma_roofline(modular_macs, seconds, kernel=None):
    """INT8-MFMA roofline of a Z/p workload: 4 limb products (8 ops) per modular MAC."""
    achieved = 8.0 * modular_macs / seconds / 1e12
    out = {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
           "frac": achieved / INT8_PEAK_TOPS, "modular_MACs": modular_macs}
    if kernel:
        out["kernel"] = kernel
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip plain/FP64 gemm, FL and other rows")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--only", choices=["fl", "flmfma", "rows", "trsm", "gf2", "fp64", "batched", "cpu"],
                    default=None,
                    help="run one secondary leg alone and print its JSON (for profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run one device-to-device copy of BYTES (counter calibration)")
    return ap.parse_args()


def time_steps(ctx, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(seconds):
    """CPU restatement (oracle) of ZpModule<uint16> gemm on whole rows of C."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    from lela_amd import synth
    lib = orc.load()
    A = synth.zp_matrix(synth.seed_for(2, 0), 64, N, P, np.uint16)
    B = synth.zp_matrix(synth.seed_for(2, 1), N, N, P, np.uint16)
    C = synth.zp_matrix(synth.seed_for(2, 2), 64, N, P, np.uint16)

    def run(rows):
        t0 = time.perf_counter()
        lib.orc_zp_gemm_u16(P, rows, N, N, 1, A.ctypes.data, N, B.ctypes.data, N, 1,
                            C.ctypes.data, N)
        return time.perf_counter() - t0

    t1 = run(1)
    rows = max(1, min(64, int(seconds / max(t1, 1e-6))))
    t = run(rows)
    macs = rows * N * N
    return {"value": macs / t, "unit": "modular MACs/s", "cores": 1, "kind": "port",
            "sample": f"{rows} of {N} rows of C for the {N}^3 uint16 gemm "
                      f"({rows}*{N}*{N} MACs, {t:.1f} s), CPU restatement of "
                      "ZpModule<uint16> dot + generic gemm, -O2, one core"}


def host_cores():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"nproc": avail, "cpu_count": os.cpu_count()}


def cpu_baselines(seconds):
    """The CPU restatements (oracle) of LELA's paths for the trsm, the two
    GF(2) workloads, the Faugere-Lachartre reduction and the uneven batch,
    each on a bounded sample of the workload, one core, extrapolated as each
    entry's `sample` says."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    from lela_amd import synth
    lib = orc.load()
    out = {}
    # trsm: right-hand-side columns are independent -- solve a block of right-hand sides
    L = np.tril(synth.zp_matrix(synth.seed_for(2, 3), N, N, P), -1)
    L[np.arange(N), np.arange(N)] = 1
    cols = 32
    Bs = np.ascontiguousarray(synth.zp_matrix(synth.seed_for(2, 4), N, 1024, P)[:, :cols])
    import ctypes
    bad = ctypes.c_size_t(0)
    t0 = time.perf_counter()
    lib.orc_zp_trsm(P, 0, 1, N, cols, 1, L.ctypes.data, N, Bs.ctypes.data, cols, ctypes.byref(bad))
    t = time.perf_counter() - t0
    W = N * (N - 1) / 2 * cols
    out["trsm_8192x1024"] = {"value": W / t, "unit": "modular MACs/s", "cores": 1, "kind": "port",
                                "est_full_s": t * 1024 / cols,
                                "sample": f"{cols} of 1024 right-hand sides ({t:.1f} s), recursive trsm "
                                          "restatement, -O2"}
    del L, Bs
    # GF(2) gemm: rows are independent; the transposition of B is a fixed cost
    n = 16384
    A = synth.gf2_matrix(synth.seed_for(3, 0), n, n)
    B = synth.gf2_matrix(synth.seed_for(3, 1), n, n)
    C = np.zeros((129, n // 64), dtype=np.uint64)

    def gf2_rows(r):
        t0 = time.perf_counter()
        lib.orc_gf2_gemm(r, n, n, 1, A.ctypes.data, n // 64, B.ctypes.data, n // 64, 0, C.ctypes.data, n // 64)
        return time.perf_counter() - t0
    t1, t2 = gf2_rows(1), gf2_rows(129)
    per_row = max(1e-9, (t2 - t1) / 128)
    est = t1 - per_row + per_row * n
    out["gf2_gemm_16384"] = {"value": float(n) ** 3 / est, "unit": "bit-MACs/s", "cores": 1, "kind": "port",
                                "est_full_s": est,
                                "sample": f"rows 1 and 129 of {n} ({t1:.1f} s, {t2:.1f} s: transposition of B plus "
                                          f"{per_row * 1e3:.0f} ms per row), word AND/XOR-parity dots, -O2"}
    del A, B
    # GF(2) echelon form: the echelon form of a 4096^2 rank-3000 product, scaled by m*r*n to 16384^2 rank 12000
    m4, r4 = 4096, 3000
    X = synth.gf2_matrix(synth.seed_for(3, 2), m4, r4)
    Y = synth.gf2_matrix(synth.seed_for(3, 3), r4, m4)
    M = orc.gf2_gemm(m4, m4, r4, 1, X, Y, 0, np.zeros((m4, m4 // 64), dtype=np.uint64))
    rank = ctypes.c_size_t(0)
    piv = np.zeros(m4, dtype=np.uint32)
    t0 = time.perf_counter()
    lib.orc_gf2_echelonize(m4, m4, M.ctypes.data, m4 // 64, 1, ctypes.byref(rank), piv.ctypes.data)
    t = time.perf_counter() - t0
    scale = (16384 * 12000 * 16384) / (m4 * r4 * m4)
    out["gf2_rref_16384_rank12000"] = {"value": 16384 * 12000 * 16384 / (t * scale), "unit": "bit-ops/s",
                                          "cores": 1, "kind": "port", "est_full_s": t * scale,
                                          "sample": f"{m4}^2 rank {rank.value} product ({t:.1f} s), scaled by "
                                                    f"m*r*n ({scale:.0f}x), Gauss-Jordan restatement, -O2"}
    # Faugere-Lachartre phases -- the dense Schur product D - C A^-1 B on sampled rows of C,
    # the other phases at the same rate from their MAC counts
    rows = 2
    Cs = synth.zp_matrix(synth.seed_for(4, 1), rows, 15000, P)
    Bf = synth.zp_matrix(synth.seed_for(4, 2), 15000, 15000, P)
    Ds = synth.zp_matrix(synth.seed_for(4, 3), rows, 15000, P)
    t0 = time.perf_counter()
    lib.orc_zp_gemm_u32(P, rows, 15000, 15000, P - 1, Cs.ctypes.data, 15000, Bf.ctypes.data, 15000, 1,
                        Ds.ctypes.data, 15000)
    t = time.perf_counter() - t0
    rate = rows * 15000 * 15000 / t
    phases = {"trsm_A^-1B": 2.25e6 * 15000, "schur": 5000.0 * 15000 * 15000,
              "rref_D": 5000.0 * 5000 * 12500, "back_substitution": 15000.0 * 5000 * 10000}
    tot = sum(phases.values())
    out["fl_20000x30000"] = {"value": tot / (tot / rate), "unit": "modular MACs/s", "cores": 1, "kind": "port",
                               "est_full_s": tot / rate, "phase_MACs": phases,
                               "sample": f"{rows} of 5000 rows of the Schur product ({t:.1f} s, uint32 ZpModule "
                                         "dots); the other phases extrapolated at that rate from their MAC counts"}
    del Bf
    # uneven batch: the first 16 instances of the 4096-instance batch, one at a time
    raw = synth.batch_instances(synth.seed_for(5, 0), count=16, p=P)
    t0 = time.perf_counter()
    for d in raw:
        if d["op"] == 1:
            orc.zp_gemm(P, 1, d["A"].astype(np.uint32), d["B"].astype(np.uint32), 1, d["C"].astype(np.uint32))
        else:
            orc.zp_echelonize(P, d["A"].astype(np.uint32))
    t = time.perf_counter() - t0
    out["batch_4096"] = {"value": len(raw) / t, "unit": "instances/s", "cores": 1, "kind": "port",
                           "est_full_s": t * 4096 / len(raw),
                           "sample": f"instances 0..15 of the 4096 ({t:.1f} s; gemm via ZpModule dots, echelon "
                                     "forms via Elimination), one core"}
    return out


def bench_fl(la, synth, np):
    """Faugere-Lachartre on the 20000 x 30000 F4-shaped matrix."""
    ctx = la.Context(la.Modular(P, "uint32"), device=0)
    csr, shape = synth.fl_matrix(synth.seed_for(4, 0), P)
    drp = la.DeviceMatrix.from_host(ctx, csr[0].reshape(-1, 1))
    dc = la.DeviceMatrix.from_host(ctx, csr[1].reshape(-1, 1))
    dv = la.DeviceMatrix.from_host(ctx, csr[2].reshape(-1, 1))
    dR = la.DeviceMatrix(ctx, shape[0], shape[1], np.uint32)
    FL = la.FaugereLachartre(ctx)
    out = {"shape": list(shape), "nnz": int(csr[0][-1])}
    for algo, name in [(la.ALGO_AUTO, "mfma"), (la.ALGO_PLAIN, "plain")]:
        res = {}
        t = time_steps(ctx, lambda: res.update(r=FL.echelonize(dR, (drp, dc, dv), shape=shape,
                                                                 algo=algo)), 2, 1)
        out[f"{name}_s"] = t
        out["rank"] = int(res["r"][0])
    out["speedup"] = out["plain_s"] / out["mfma_s"]
    # the compact result (pivots + the rank x (n - rank) block on the
    # non-pivot columns, what LELA's sparse R holds) instead of a dense R
    dNP = la.DeviceMatrix(ctx, shape[0], shape[1] - min(shape), np.uint32)
    res = {}
    t = time_steps(ctx, lambda: res.update(r=FL.echelonize_compact((drp, dc, dv), shape=shape, NP=dNP)), 2, 1)
    out["compact_s"] = t
    out["compact_rank"] = int(res["r"][0])
    dNP.free()
    # algorithmic modular MACs: sparse A^-1 B, Schur product, RREF of D, back-substitution
    W = 2.25e6 * 15000 + 5000.0 * 15000 * 15000 + 5000.0 * 5000 * 12500 + 15000.0 * 5000 * 10000
    out["roofline"] = mfma_roofline(W, out["mfma_s"])
    for d in (drp, dc, dv, dR):
        d.free()
    ctx.close()
    return out


def bench_fl_mfma(la, synth, np):
    """The 20000 x 30000 Faugere-Lachartre reduction on the MFMA path only
    (dense R, det), three calls: the leg a kernel trace of the pipeline is
    taken on."""
    ctx = la.Context(la.Modular(P, "uint32"), device=0)
    csr, shape = synth.fl_matrix(synth.seed_for(4, 0), P)
    drp, dc, dv = (la.DeviceMatrix.from_host(ctx, a.reshape(-1, 1)) for a in csr)
    dR = la.DeviceMatrix(ctx, shape[0], shape[1], np.uint32)
    FL = la.FaugereLachartre(ctx)
    res = {}
    t = time_steps(ctx, lambda: res.update(r=FL.echelonize(dR, (drp, dc, dv), shape=shape)), 2, 1)
    for d in (drp, dc, dv, dR):
        d.free()
    ctx.close()
    return {"shape": list(shape), "mfma_s": t, "rank": int(res["r"][0]), "det": int(res["r"][1])}


def bench_batch(la, synth, np):
    """Batches of 4096 uneven instances (m, n, k in [32, 1024], half sparse,
    gemm and echelon forms mixed), resident in HBM, one captured graph per
    policy.  The autotuner table and the cost model are fitted on a separate
    1024-instance training batch (outside the timed region)."""
    from lela_amd import batch as lb
    F = la.Modular(P, "uint16")
    ctx = la.Context(F, device=0)

    def workload(seed, count):
        raw = synth.batch_instances(seed, count=count, p=P)
        ins = [lb.Instance(d["op"], d["A"], d.get("B"), d.get("C"), density=d["density"]) for d in raw]
        return lb.Workload(ctx, F, ins)

    train = workload(synth.seed_for(5, 1), 1024)
    _, losses = lb.fit(ctx, F, train.tasks, reps=3, steps=3000)
    train.free()
    ev = workload(synth.seed_for(5, 0), 4096)
    ops = np.array([t.op for t in ev.tasks])
    gemm_macs = float(sum(t.m * t.n * t.k for t in ev.tasks if t.op == lb.TASK_GEMM))
    out = {"instances": ev.n, "gemm": int((ops == 1).sum()), "rref": int((ops == 2).sum()),
           "gemm_modular_MACs": gemm_macs, "model_train_loss": {str(k): v for k, v in losses.items()}}

    def timed(policy, variants=None, reps=5):
        ev.set_variants(variants if variants is not None else [-1] * ev.n)
        plan = lb.Plan(ctx, F, ev.tasks, policy, graph=True)
        v = plan.variants()
        ts = []
        for r in range(reps + 1):
            ev.upload()
            ctx.synchronize()
            ctx.timer_start()
            plan.launch()
            ms = ctx.timer_stop()
            plan.finish()
            if r:
                ts.append(ms)
        plan.close()
        return float(np.median(ts)), v

    chosen = {}
    for name, policy, var in (("model", lb.POLICY_MODEL, None), ("table", lb.POLICY_TABLE, None),
                              ("plain", lb.POLICY_FIXED, [0] * ev.n)):
        ms, v = timed(policy, var)
        chosen[name] = v
        out[f"{name}_ms"] = ms
    out["instances_per_s"] = ev.n / (out["model_ms"] * 1e-3)
    out["model_over_plain"] = out["plain_ms"] / out["model_ms"]
    # one library call per instance, as a caller looping over LELA calls would
    import ctypes
    ev.upload()
    ctx.synchronize()
    ctx.set_ptr_mode(la.PTR_DEVICE)
    rank = ctypes.c_size_t()
    t0 = time.perf_counter()
    for t in ev.tasks:
        if t.op == lb.TASK_GEMM:
            rc = ctx.lib.mlz_zp_gemm(ctx.handle, F.elem, P, t.m, t.n, t.k, t.a, t.A, t.lda, t.B, t.ldb, t.b,
                                     t.C, t.ldc, 0)
        else:
            rc = ctx.lib.mlz_zp_rref(ctx.handle, F.elem, P, t.m, t.n, t.A, t.lda, ctypes.byref(rank), None, 0)
        ctx.check(rc, "per-instance call")
    ctx.synchronize()
    out["per_instance_calls_ms"] = (time.perf_counter() - t0) * 1e3
    out["batched_over_per_instance"] = out["per_instance_calls_ms"] / out["model_ms"]
    # measured cost (CU-microseconds) of every variant of every held-out instance,
    # on the original inputs (the calls above reduced the matrices in place)
    ev.upload()
    cost = lb.tune(ctx, F, ev.tasks, reps=2)
    idx = np.arange(ev.n)
    out["cu_us"] = {name: float(np.nansum(cost[idx, v])) for name, v in chosen.items()}
    out["cu_us"]["best"] = float(np.nansum(np.nanmin(cost, axis=1)))
    out["model_gain_over_table"] = out["cu_us"]["table"] / out["cu_us"]["model"]
    ev.free()
    out["mixed_density"] = bench_mixed_density(ctx, F, la, synth, np, lb)
    ctx.close()
    return out


def bench_mixed_density(ctx, F, la, synth, np, lb):
    """Where the choice matters: 2048 gemm instances (m, n, k in [64, 1024])
    whose A holds 0.05% to 20% nonzeros, log-uniform.  Below about 0.5% the
    sparse-A variant is the fastest, above it INT8 MFMA; the table keys on
    shape only, the model also sees the density.  Both are fitted on a
    separate 1024-instance batch; then each policy's plan runs as one captured
    graph (wall time) and the measured cost of its choices is summed."""
    def workload(seed, count):
        raw = synth.gemm_density_instances(seed, count, p=P)
        return lb.Workload(ctx, F, [lb.Instance(1, d["A"], d["B"], d["C"], density=d["density"]) for d in raw])

    train = workload(synth.seed_for(5, 3), 1024)
    lb.fit(ctx, F, train.tasks, reps=3, steps=3000)
    train.free()
    ev = workload(synth.seed_for(5, 4), 2048)
    out = {"instances": ev.n, "density": [0.0005, 0.2]}
    chosen = {}
    for name, policy in (("model", lb.POLICY_MODEL), ("table", lb.POLICY_TABLE)):
        ev.set_variants([-1] * ev.n)
        plan = lb.Plan(ctx, F, ev.tasks, policy, graph=True)
        chosen[name] = plan.variants()
        ts = []
        for r in range(6):
            ev.upload()
            ctx.synchronize()
            ctx.timer_start()
            plan.launch()
            ms = ctx.timer_stop()
            plan.finish()
            if r:
                ts.append(ms)
        plan.close()
        out[f"{name}_ms"] = float(np.median(ts))
    ev.upload()
    cost = lb.tune(ctx, F, ev.tasks, reps=2)
    idx = np.arange(ev.n)
    out["cu_us"] = {name: float(np.nansum(cost[idx, v])) for name, v in chosen.items()}
    out["cu_us"]["best"] = float(np.nansum(np.nanmin(cost, axis=1)))
    out["model_gain_over_table"] = out["cu_us"]["table"] / out["cu_us"]["model"]
    out["model_wall_gain_over_table"] = out["table_ms"] / out["model_ms"]
    out["sparse_a_chosen"] = {name: int(np.sum(np.asarray(v) == lb.GEMM_SPARSE)) for name, v in chosen.items()}
    ev.free()
    return out


def bench_fp64(la, synth, np):
    """The FP64-carrier gemm, 8192^3, p = 67108859, uint32 storage."""
    ctx = la.Context(la.Modular(P_FP64, "uint32"), device=0)
    W = float(N) ** 3
    dA2 = la.DeviceMatrix.from_host(ctx, synth.zp_matrix(synth.seed_for(2, 5), N, N, P_FP64))
    dB2 = la.DeviceMatrix.from_host(ctx, synth.zp_matrix(synth.seed_for(2, 6), N, N, P_FP64))
    dC2 = la.DeviceMatrix.from_host(ctx, synth.zp_matrix(synth.seed_for(2, 7), N, N, P_FP64))
    ctx.prof_reset()
    ctx.prof_enable(True)
    t = time_steps(ctx, lambda: la.BLAS3.gemm(ctx, 1, dA2, dB2, 1, dC2, algo=la.ALGO_FP64), 2, 1)
    ctx.prof_enable(False)
    cnt, kms = ctx.prof_query("gemm_f64_mfma")
    fp64_kernel_ms = kms / max(1, cnt)
    tf = 2 * 2 * W / (fp64_kernel_ms * 1e-3) / 1e12     # 2 limbs x (mul + add)
    for d in (dA2, dB2, dC2):
        d.free()
    ctx.close()
    return {"ms_per_step": t * 1e3, "value": W / t, "unit": "modular MACs/s",
            "roofline": {"bound": "mfma", "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tf / FP64_PEAK_TFLOPS, "kernel": "gemm_f64_kernel",
                         "kernel_avg_ms": fp64_kernel_ms, "op": "2 FP64 FMAs (4 flops) per modular MAC"}}


def bench_trsm(la, synth, np):
    """trsm lower, unit diagonal, 8192 x 8192 with 1024 right-hand sides."""
    out = {}
    ctx = la.Context(la.Modular(P, "uint16"), device=0)
    L = np.tril(synth.zp_matrix(synth.seed_for(2, 3), N, N, P, np.uint16), -1)
    L[np.arange(N), np.arange(N)] = 1
    dL = la.DeviceMatrix.from_host(ctx, L)
    dB = la.DeviceMatrix.from_host(ctx, synth.zp_matrix(synth.seed_for(2, 4), N, 1024, P, np.uint16))
    t = time_steps(ctx, lambda: la.BLAS3.trsm(ctx, 1, dL, dB, la.LowerTriangular, True), 5, 1)
    out["trsm_8192x1024_ms"] = t * 1e3
    out["trsm_modular_MACs_per_s"] = N * (N - 1) / 2 * 1024 / t
    out["trsm_roofline"] = mfma_roofline(N * (N - 1) / 2 * 1024, t)
    dL.free()
    dB.free()
    ctx.close()
    return out


def bench_gf2(la, synth, np):
    """GF(2) 16384^2 gemm and the echelon form of a rank-12000 product."""
    out = {}
    ctx = la.Context(la.GF2(), device=0)
    n = 16384
    dA = la.DeviceMatrix.from_host(ctx, synth.gf2_matrix(synth.seed_for(3, 0), n, n))
    dB = la.DeviceMatrix.from_host(ctx, synth.gf2_matrix(synth.seed_for(3, 1), n, n))
    dC = la.DeviceMatrix(ctx, n, n // 64, np.uint64)
    gf2 = lambda: la.BLAS3.gemm(ctx, 1, la.Dense01Matrix(dA, n), la.Dense01Matrix(dB, n), 0,
                                la.Dense01Matrix(dC, n))
    t = time_steps(ctx, gf2, 5, 1)
    out["gf2_gemm_16384_ms"] = t * 1e3
    out["gf2_bit_MACs_per_s"] = n ** 3 / t
    # Four Russians: 8 table reads of 8 bytes per (C row, 64-bit word of A, C word)
    lds = n * (n // 64) * (n // 64) * 64 / t / 1e12
    out["gf2_gemm_roofline"] = {"bound": "lds", "achieved": lds, "peak": LDS_PEAK_TBS, "unit": "TB/s",
                                "frac": lds / LDS_PEAK_TBS, "kernel": "gf2_gemm_m4rm_kernel",
                                "bytes": "table reads, 64 B per (C row, A word, C word)"}
    ctx.set_option("gf2.gemm.kernel", 1)
    t = time_steps(ctx, gf2, 5, 1)
    ctx.set_option("gf2.gemm.kernel", 0)
    out["gf2_gemm_dot_16384_ms"] = t * 1e3
    # dot form: one AND+XOR of 32-bit halves (2 lane ops) per 64-bit word of a dot product
    ops = 2.0 * n * n * (n // 64) / t / 1e12
    out["gf2_gemm_dot_roofline"] = {"bound": "valu", "achieved": ops, "peak": VALU_PEAK_TOPS, "unit": "Tops/s",
                                    "frac": ops / VALU_PEAK_TOPS, "kernel": "gf2_gemm_dot_kernel",
                                    "ops": "2 per (C entry, 64-bit word of k)"}
    r = 12000
    dX = la.DeviceMatrix.from_host(ctx, synth.gf2_matrix(synth.seed_for(3, 2), n, r))
    dY = la.DeviceMatrix.from_host(ctx, synth.gf2_matrix(synth.seed_for(3, 3), r, n))
    la.BLAS3.gemm(ctx, 1, la.Dense01Matrix(dX, r), la.Dense01Matrix(dY, n), 0, la.Dense01Matrix(dC, n))
    ctx.synchronize()
    M = dC.download()
    EF = la.EchelonForm(ctx)
    times = []
    for _ in range(2):
        dC.upload(M)
        ctx.synchronize()
        t0 = time.perf_counter()
        EF.echelonize(la.Dense01Matrix(dC, n))
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    out["gf2_rref_16384_rank12000_ms"] = min(times) * 1e3
    for d in (dA, dB, dC, dX, dY):
        d.free()
    ctx.close()
    return out


def bench_rows(la, synth, np):
    """trsm 8192 x 1024 and the GF(2) 16384^2 gemm and rank-12000 RREF."""
    out = bench_trsm(la, synth, np)
    out.update(bench_gf2(la, synth, np))
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} (WORLD_SIZE={world}) is not "
              "supported: every box here has one GPU and this path is not sharded.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import lela_amd as la
    from lela_amd import synth

    if args.only:
        if args.only == "cpu":
            print(json.dumps({"leg": "cpu", "cpu_baselines": cpu_baselines(args.cpu_seconds),
                              "host": host_cores()}), flush=True)
            return
        leg = {"fl": bench_fl, "flmfma": bench_fl_mfma, "rows": bench_rows, "trsm": bench_trsm, "gf2": bench_gf2,
               "fp64": bench_fp64, "batched": bench_batch}[args.only]
        print(json.dumps({"leg": args.only, args.only: leg(la, synth, np)}), flush=True)
        return

    F = la.Modular(P, "uint16")
    ctx = la.Context(F, device=0)
    A = synth.zp_matrix(synth.seed_for(2, 0), N, N, P, np.uint16)
    B = synth.zp_matrix(synth.seed_for(2, 1), N, N, P, np.uint16)
    C = synth.zp_matrix(synth.seed_for(2, 2), N, N, P, np.uint16)
    dA, dB, dC = (la.DeviceMatrix.from_host(ctx, X) for X in (A, B, C))
    del A, B

    def step_int8():
        la.BLAS3.gemm(ctx, 1, dA, dB, 1, dC, algo=la.ALGO_INT8)

    for _ in range(args.warmup):
        step_int8()
    ctx.synchronize()
    ctx.prof_reset()
    ctx.prof_enable(True)
    t0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        step_int8()
    ev_ms = ctx.timer_stop()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.prof_enable(False)
    launches, kernel_ms = ctx.prof_query("gemm_i8_mfma")
    _, pack_ms = ctx.prof_query("pack_i8")
    ms_per_step = wall * 1e3 / args.steps
    W = float(N) ** 3
    value = W / (ms_per_step * 1e-3)
    avg_kernel_ms = kernel_ms / max(1, launches)
    achieved = 8.0 * W / (avg_kernel_ms * 1e-3) / 1e12   # int8 ops: 4 limb MACs per modular MAC

    if args.copy_calibration:
        nb = int(args.copy_calibration)
        s = la.DeviceMatrix(ctx, 1, nb, np.uint8)
        d = la.DeviceMatrix(ctx, 1, nb, np.uint8)
        ctx.check(ctx.lib.mlz_memcpy_dtod(ctx.handle, d.ptr, s.ptr, nb), "memcpy_dtod")
        s.free()
        d.free()

    traffic, traffic_commit = None, None
    tfile = os.path.join(ROOT, "profiles", "gemm_i8_mfma_8192_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic, traffic_commit = tj.get("hbm_bytes_per_launch"), tj.get("measured_on_commit")

    out = {
        "metric": "Z/p gemm modular MACs/s @8192^3 p=65521 (split-int8 MFMA) + F-L echelon s",
        "value": value,
        "unit": "modular MACs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic (SplitMix64 uniform residues, seeds 0x4C454C4100000020..22; every secondary "
                "workload also drawn from SplitMix64 streams)",
        "config": {"workload": "Z/p gemm C <- A*B + C, m=n=k=8192, p=65521, Modular<uint16> "
                               "storage, inputs resident in HBM",
                   "m": N, "n": N, "k": N, "p": P, "storage": "uint16",
                   "algo": "split-int8 MFMA (2x2 8-bit limbs, int32 accumulate)"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS,
                     "unit": "TFLOP/s", "frac": achieved / INT8_PEAK_TOPS, "traffic": traffic,
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated, measured on commit "
                                       f"{traffic_commit} (profiles/gemm_i8_mfma_8192_traffic.json)",
                     "kernel": "gemm_i8_async_kernel", "kernel_avg_ms": avg_kernel_ms,
                     "ops_per_launch": 8.0 * W, "op": "int8 multiply-add counted as 2 ops"},
        "timing": {"event_ms_per_step": ev_ms / args.steps,
                   "pack_ms_per_step": pack_ms / args.steps},
    }

    if not args.no_variants:
        variants = {}
        t = time_steps(ctx, lambda: la.BLAS3.gemm(ctx, 1, dA, dB, 1, dC, algo=la.ALGO_PLAIN), 2, 1)
        variants["plain_u16_p65521"] = {"ms_per_step": t * 1e3, "value": W / t,
                                        "unit": "modular MACs/s"}
        out["int8_over_plain"] = value / variants["plain_u16_p65521"]["value"]
        for d in (dA, dB, dC):
            d.free()
        ctx.close()
        variants["fp64_u32_p67108859"] = bench_fp64(la, synth, np)
        out["variants"] = variants
        out["fl_echelon"] = bench_fl(la, synth, np)
        out["rows"] = bench_rows(la, synth, np)
        out["batched"] = bench_batch(la, synth, np)
    else:
        for d in (dA, dB, dC):
            d.free()
        ctx.close()

    out["host"] = host_cores()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        if not args.no_variants:
            out["cpu_baselines"] = cpu_baselines(args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
