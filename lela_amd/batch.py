"""Batches of independent Z/p instances (mlz_batch_*, mlz_tune, mlz_model_*).

LELA has no batch call: a caller loops over BLAS3::gemm
(lela/blas/level3.h:81-83) and EchelonForm::echelonize
(lela/solutions/echelon-form.h:103-154).  Here the instances of such a loop
are handed over at once and run in persistent, work-list driven launches
(optionally one captured hipGraph).  Each instance gets one of several exact
kernel variants, picked by a fixed choice, the autotuner's table, or the cost
model; results do not depend on the choice.

    ws = Workload(ctx, F, instances)          # matrices resident in one arena
    plan = Plan(ctx, F, ws.tasks, POLICY_MODEL, graph=True)
    plan.launch(); plan.finish()              # ranks in ws.tasks[i].rank
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (BATCH_GRAPH, BATCH_SERIAL, BATCH_TIMING, GEMM_I8L, GEMM_I8S, GEMM_PLAIN,
                   GEMM_SPARSE, MAX_VARIANTS, MODEL_PARAMS, POLICY_FIXED, POLICY_MODEL,
                   POLICY_TABLE, PTR_DEVICE, RREF_I8_W32, RREF_I8_W64, RREF_W32, RREF_W64, TASK_GEMM,
                   TASK_RREF, Task)

__all__ = [
    "Instance", "Workload", "Plan", "tune", "table_get", "table_set", "task_features", "CostModel",
    "TASK_GEMM", "TASK_RREF", "POLICY_FIXED", "POLICY_TABLE", "POLICY_MODEL", "GEMM_PLAIN",
    "GEMM_SPARSE", "GEMM_I8S", "GEMM_I8L", "RREF_W64", "RREF_W32", "RREF_I8_W64", "RREF_I8_W32",
    "MAX_VARIANTS",
    "MODEL_PARAMS", "nvariants", "variant_ok", "fit",
]

_ALIGN = 256


def nvariants(op):
    return 4


def variant_ok(op, v, p, k=1):
    if op == TASK_GEMM:
        if v in (GEMM_I8S, GEMM_I8L):
            return p <= 65536 and k <= 65504
        return 0 <= v < 4
    if v in (RREF_I8_W64, RREF_I8_W32):
        return p <= 65536
    return 0 <= v < 4


class Instance:
    """One instance: gemm (A m x k, B k x n, C m x n; C <- a*A*B + b*C) or an
    echelon form of A (m x n, in place).  Matrices are numpy arrays in the
    field's storage type."""

    def __init__(self, op, A, B=None, C=None, a=1, b=1, density=1.0, variant=-1):
        self.op = op
        self.A, self.B, self.C = A, B, C
        self.a, self.b = int(a), int(b)
        self.density = float(density)
        self.variant = int(variant)
        if op == TASK_GEMM:
            self.m, self.k = A.shape
            self.n = B.shape[1]
            assert B.shape[0] == self.k and C.shape == (self.m, self.n)
        elif op == TASK_RREF:
            self.m, self.n = A.shape
            self.k = 0
        else:
            raise ValueError(f"unknown op {op}")


class Workload:
    """The matrices of a list of instances, uploaded into one device arena.

    tasks is the ctypes array of struct mlz_task pointing into the arena;
    outputs are read back with result(i)."""

    def __init__(self, ctx, F, instances):
        self.ctx, self.F = ctx, F
        self.instances = list(instances)
        dt = np.dtype(F.dtype)
        offs = []
        off = 0

        def take(nbytes):
            nonlocal off
            o = off
            off = (off + max(int(nbytes), 1) + _ALIGN - 1) // _ALIGN * _ALIGN
            return o

        for ins in self.instances:
            if ins.op == TASK_GEMM:
                offs.append((take(ins.A.size * dt.itemsize), take(ins.B.size * dt.itemsize),
                             take(ins.C.size * dt.itemsize), None))
            else:
                offs.append((take(ins.A.size * dt.itemsize), None, None,
                             take(min(ins.m, ins.n) * 4)))
        self.nbytes = max(off, _ALIGN)
        self.offsets = offs
        host = np.zeros(self.nbytes, dtype=np.uint8)
        for ins, (oa, ob, oc, _) in zip(self.instances, offs):
            for X, o in ((ins.A, oa), (ins.B, ob), (ins.C, oc)):
                if o is not None:
                    raw = np.ascontiguousarray(X, dtype=dt).view(np.uint8).ravel()
                    host[o:o + raw.size] = raw
        ptr = ctypes.c_void_p()
        ctx.check(ctx.lib.mlz_dev_alloc(ctx.handle, self.nbytes, ctypes.byref(ptr)), "dev_alloc")
        self.base = ptr.value
        self._host = host
        self.upload()
        if ctx._ptr_mode != PTR_DEVICE:
            ctx.set_ptr_mode(PTR_DEVICE)
        self.tasks = (Task * len(self.instances))()
        p = F.p
        for i, (ins, (oa, ob, oc, opv)) in enumerate(zip(self.instances, offs)):
            t = self.tasks[i]
            t.op = ins.op
            t.variant = ins.variant
            t.m, t.n, t.k = ins.m, ins.n, ins.k
            t.density = ins.density
            t.A = self.base + oa
            if ins.op == TASK_GEMM:
                t.a, t.b = ins.a % p, ins.b % p
                t.lda, t.ldb, t.ldc = ins.k, ins.n, ins.n
                t.B = self.base + ob
                t.C = self.base + oc
            else:
                t.lda = ins.n
                t.pivots = self.base + opv
        self.n = len(self.instances)

    def upload(self):
        """Copies the original matrices (again) into the arena."""
        self.ctx.check(self.ctx.lib.mlz_memcpy_htod(self.ctx.handle, self.base, self._host.ctypes.data,
                                                    self.nbytes), "memcpy_htod")

    def download(self):
        out = np.empty(self.nbytes, dtype=np.uint8)
        self.ctx.check(self.ctx.lib.mlz_memcpy_dtoh(self.ctx.handle, out.ctypes.data, self.base,
                                                    self.nbytes), "memcpy_dtoh")
        return out

    def result(self, i, raw=None):
        """Output of instance i: C for gemm, (R, pivots[:rank]) for an echelon form."""
        raw = self.download() if raw is None else raw
        ins, (oa, ob, oc, opv) = self.instances[i], self.offsets[i]
        dt = np.dtype(self.F.dtype)
        if ins.op == TASK_GEMM:
            return raw[oc:oc + ins.C.size * dt.itemsize].view(dt).reshape(ins.C.shape).copy()
        R = raw[oa:oa + ins.A.size * dt.itemsize].view(dt).reshape(ins.A.shape).copy()
        r = int(self.tasks[i].rank)
        piv = raw[opv:opv + 4 * min(ins.m, ins.n)].view(np.uint32)[:r].copy()
        return R, piv

    def set_variants(self, variants):
        for t, v in zip(self.tasks, variants):
            t.variant = int(v)

    def free(self):
        if self.base:
            self.ctx.lib.mlz_dev_free(self.ctx.handle, ctypes.c_void_p(self.base))
            self.base = None


class Plan:
    """A planned batch (mlz_batch_plan): launch() queues one execution,
    finish() waits and returns per-task costs (microseconds of one CU) when
    timing is on."""

    def __init__(self, ctx, F, tasks, policy=POLICY_TABLE, graph=False, timing=False, serial=False):
        self.ctx, self.F, self.tasks = ctx, F, tasks
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_TIMING if timing else 0) | \
            (BATCH_SERIAL if serial else 0)
        self.timing = timing
        if ctx._ptr_mode != PTR_DEVICE:
            ctx.set_ptr_mode(PTR_DEVICE)
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.mlz_batch_plan(ctx.handle, F.elem, F.p, ctypes.addressof(tasks), len(tasks),
                                         policy, flags, ctypes.byref(h)), "batch_plan")
        self.handle = h

    def variants(self):
        return np.array([t.variant for t in self.tasks], dtype=np.int32)

    def launch(self):
        self.ctx.check(self.ctx.lib.mlz_batch_launch(self.ctx.handle, self.handle), "batch_launch")

    def finish(self):
        cost = np.zeros(len(self.tasks), dtype=np.float32) if self.timing else None
        bad = ctypes.c_size_t(0)
        rc = self.ctx.lib.mlz_batch_finish(self.ctx.handle, self.handle, ctypes.addressof(self.tasks),
                                           cost.ctypes.data if cost is not None else None,
                                           ctypes.byref(bad))
        if rc == _lib.MLZ_E_NOT_INVERTIBLE:
            from . import LELAError
            raise LELAError(f"task {bad.value}: a pivot has no inverse mod p")
        self.ctx.check(rc, "batch_finish")
        return cost

    def close(self):
        if self.handle:
            self.ctx.lib.mlz_batch_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune(ctx, F, tasks, reps=3):
    """Autotuner (mlz_tune): measured cost of every applicable variant of
    every task, shape [ntasks, MAX_VARIANTS] (NaN where a variant does not
    apply); also fills the context's table."""
    cost = np.full((len(tasks), MAX_VARIANTS), np.nan, dtype=np.float32)
    if ctx._ptr_mode != PTR_DEVICE:
        ctx.set_ptr_mode(PTR_DEVICE)
    ctx.check(ctx.lib.mlz_tune(ctx.handle, F.elem, F.p, ctypes.addressof(tasks), len(tasks), int(reps),
                               cost.ctypes.data), "tune")
    return cost


def table_get(ctx, op):
    t = np.zeros(16 * 16 * 16, dtype=np.int8)
    ctx.check(ctx.lib.mlz_tune_table_get(ctx.handle, op, t.ctypes.data), "tune_table_get")
    return t


def table_set(ctx, op, table):
    t = np.ascontiguousarray(table, dtype=np.int8)
    assert t.size == 16 * 16 * 16
    ctx.check(ctx.lib.mlz_tune_table_set(ctx.handle, op, t.ctypes.data), "tune_table_set")


def task_features(task, p):
    """The cost model's input for a task (mlz_task_features)."""
    x = np.zeros(5, dtype=np.float32)
    rc = _lib.load().mlz_task_features(ctypes.addressof(task), p, x.ctypes.data)
    if rc != 0:
        raise ValueError("bad task for features")
    return x


class CostModel:
    """The per-op MLP 5 -> 32 -> 32 -> 4 of the context (mlz_model_*)."""

    @staticmethod
    def train(ctx, op, x, y, mask, steps=2000, lr=3e-3, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        n = x.shape[0]
        assert x.shape == (n, 5) and y.shape == (n, MAX_VARIANTS) and mask.shape == y.shape
        loss = ctypes.c_float()
        ctx.check(ctx.lib.mlz_model_train(ctx.handle, op, x.ctypes.data, y.ctypes.data, mask.ctypes.data, n,
                                          int(steps), float(lr), int(seed), ctypes.byref(loss)), "model_train")
        return loss.value

    @staticmethod
    def predict(ctx, op, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], MAX_VARIANTS), dtype=np.float32)
        ctx.check(ctx.lib.mlz_model_predict(ctx.handle, op, x.ctypes.data, x.shape[0], y.ctypes.data),
                  "model_predict")
        return y

    @staticmethod
    def get(ctx, op):
        w = np.zeros(MODEL_PARAMS, dtype=np.float32)
        ctx.check(ctx.lib.mlz_model_get(ctx.handle, op, w.ctypes.data), "model_get")
        return w

    @staticmethod
    def set(ctx, op, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == MODEL_PARAMS
        ctx.check(ctx.lib.mlz_model_set(ctx.handle, op, w.ctypes.data), "model_set")

    @staticmethod
    def grad(ctx, w, x, y, mask):
        w = np.ascontiguousarray(w, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        g = np.zeros(MODEL_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        ctx.check(ctx.lib.mlz_model_grad(ctx.handle, w.ctypes.data, x.ctypes.data, y.ctypes.data,
                                         mask.ctypes.data, x.shape[0], g.ctypes.data, ctypes.byref(loss)),
                  "model_grad")
        return g, loss.value


def fit(ctx, F, tasks, reps=3, steps=3000, lr=3e-3):
    """Autotunes on `tasks` (mlz_tune: measured cost of every variant, which
    also fills the context's table), then trains each op's cost model on
    log2 of those costs.  Returns the cost matrix [ntasks, MAX_VARIANTS]
    and the final training loss per op."""
    cost = tune(ctx, F, tasks, reps=reps)
    x = np.stack([task_features(t, F.p) for t in tasks])
    ops = np.array([t.op for t in tasks])
    losses = {}
    for op in (TASK_GEMM, TASK_RREF):
        sel = ops == op
        if not sel.any():
            continue
        c = cost[sel]
        mask = ~np.isnan(c)
        y = np.log2(np.where(mask, np.maximum(c, 1e-3), 1.0)).astype(np.float32)
        losses[op] = CostModel.train(ctx, op, x[sel], y, mask.astype(np.uint8), steps=steps, lr=lr, seed=op)
    return cost, losses
