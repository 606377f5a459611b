"""lela_amd -- MI355X backend for LELA's exact level-3 path.

Host-side mirror of LELA's interface for the operations the HIP library
implements, so that calls read like LELA's own (lela/blas/level3.h,
lela/solutions/echelon-form.h, lela/algorithms/faugere-lachartre.h):

    F = Modular(65521, storage="uint16")        # Modular<uint16>
    ctx = Context(F)                            # Context<Ring, HIPModule>
    BLAS3.gemm(ctx, 1, A, B, 1, C)              # C <- A*B + C   (numpy arrays)
    BLAS3.trsm(ctx, 1, L, X, LowerTriangular, True)

Errors are raised as the LELA exceptions they stand for
(DiagonalEntryNotInvertible(idx), WrongMatrixForm(row), LELAError).
Every call goes through liblela_amd.so; there is no other compute path.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (ALGO_AUTO, ALGO_FP64, ALGO_INT8, ALGO_PLAIN, MLZ_F64_SYM, MLZ_U16,
                   MLZ_U32, PTR_DEVICE, PTR_HOST)

__all__ = [
    "Modular", "Context", "BLAS3", "DeviceMatrix", "LELAError", "DiagonalEntryNotInvertible",
    "WrongMatrixForm", "LowerTriangular", "UpperTriangular", "ALGO_AUTO", "ALGO_PLAIN",
    "ALGO_INT8", "ALGO_FP64", "EchelonForm", "Elimination", "GF2", "Dense01Matrix", "FaugereLachartre",
    "dense_to_csr", "dense01_to_hybrid", "DensePivotStrategy", "SparsePartialPivotStrategy",
]

LowerTriangular = "lower"
UpperTriangular = "upper"


class LELAError(RuntimeError):
    """LELA's generic error (lela/util/error.h)."""


class DiagonalEntryNotInvertible(LELAError):
    """Raised by trsm with the index of the offending diagonal entry
    (lela/util/error.h:83-95)."""

    def __init__(self, idx):
        super().__init__(f"Diagonal entry {idx} not invertible")
        self.idx = idx


class WrongMatrixForm(LELAError):
    """Raised by Faugere-Lachartre for a row whose head decreases
    (lela/algorithms/faugere-lachartre.tcc:32-41)."""

    def __init__(self, row):
        super().__init__(f"Bad input-matrix for Faugere-Lachartre at row-index {row}")
        self.row = row


_STORAGE = {
    "uint16": (MLZ_U16, np.uint16),
    "uint32": (MLZ_U32, np.uint32),
    "double": (MLZ_F64_SYM, np.float64),
}


class Modular:
    """Z/p with LELA's storage conventions: uint16/uint32 hold [0, p);
    double holds the symmetric range (lela/ring/modular.h:158-327)."""

    def __init__(self, p, storage="uint32"):
        if storage not in _STORAGE:
            raise ValueError(f"storage must be one of {sorted(_STORAGE)}")
        self.p = int(p)
        self.storage = storage
        self.elem, self.dtype = _STORAGE[storage]
        limits = {"uint16": 1 << 16, "uint32": 1 << 32, "double": 1 << 26}
        if not 2 <= self.p < limits[storage]:
            raise ValueError(f"modulus {p} not valid for Modular<{storage}>")

    def from_canonical(self, X):
        """Canonical residues [0, p) -> this ring's storage."""
        X = np.asarray(X)
        if self.storage == "double":
            Xs = X.astype(np.int64)
            Xs = np.where(2 * Xs > self.p, Xs - self.p, Xs)
            return Xs.astype(np.float64)
        return X.astype(self.dtype)

    def to_canonical(self, X):
        """This ring's storage -> canonical residues as uint32."""
        X = np.asarray(X)
        if self.storage == "double":
            Xi = X.astype(np.int64)
            return np.where(Xi < 0, Xi + self.p, Xi).astype(np.uint32)
        return X.astype(np.uint32)

    def __repr__(self):
        return f"Modular<{self.storage}>({self.p})"


class Context:
    """Context<Ring, HIPModule>: one HIP stream + device workspace
    (lela/blas/context.h:73-85).  Not thread-safe, like LELA's Context."""

    def __init__(self, F=None, device=0):
        self.F = F
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.mlz_ctx_create(ctypes.byref(h), device, None)
        if rc != 0:
            raise LELAError(f"mlz_ctx_create failed with code {rc} (no usable GPU?)")
        self.handle = h
        self._ptr_mode = PTR_HOST

    def close(self):
        if self.handle:
            self.lib.mlz_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def last_error(self):
        return self.lib.mlz_last_error(self.handle).decode()

    def check(self, rc, what, idx=None):
        if rc == 0:
            return
        if rc == _lib.MLZ_E_NOT_INVERTIBLE and idx is not None:
            raise DiagonalEntryNotInvertible(idx)
        if rc == _lib.MLZ_E_WRONG_FORM and idx is not None:
            raise WrongMatrixForm(idx)
        raise LELAError(f"{what}: code {rc}: {self.last_error()}")

    def set_option(self, key, value):
        """Tuning knob of the context (mlz_ctx_set_option); never changes results."""
        self.check(self.lib.mlz_ctx_set_option(self.handle, key.encode(), int(value)), "set_option")

    def set_ptr_mode(self, mode):
        self.check(self.lib.mlz_ctx_set_ptr_mode(self.handle, mode), "set_ptr_mode")
        self._ptr_mode = mode

    def synchronize(self):
        self.check(self.lib.mlz_ctx_synchronize(self.handle), "synchronize")

    # -- profiling -------------------------------------------------------
    def prof_enable(self, on=True):
        self.check(self.lib.mlz_prof_enable(self.handle, int(on)), "prof_enable")

    def prof_reset(self):
        self.check(self.lib.mlz_prof_reset(self.handle), "prof_reset")

    def prof_query(self, name):
        c = ctypes.c_uint64()
        ms = ctypes.c_double()
        self.check(self.lib.mlz_prof_query(self.handle, name.encode(), ctypes.byref(c),
                                           ctypes.byref(ms)), "prof_query")
        return c.value, ms.value

    def timer_start(self):
        self.check(self.lib.mlz_timer_start(self.handle), "timer_start")

    def timer_stop(self):
        ms = ctypes.c_double()
        self.check(self.lib.mlz_timer_stop(self.handle, ctypes.byref(ms)), "timer_stop")
        return ms.value


class DeviceMatrix:
    """A matrix resident in device memory (for callers that keep data in HBM)."""

    def __init__(self, ctx, rows, cols, dtype):
        self.ctx = ctx
        self.rows, self.cols = int(rows), int(cols)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.rows * self.cols * self.dtype.itemsize
        ptr = ctypes.c_void_p()
        ctx.check(ctx.lib.mlz_dev_alloc(ctx.handle, self.nbytes, ctypes.byref(ptr)), "dev_alloc")
        self.ptr = ptr

    @classmethod
    def from_host(cls, ctx, X):
        X = np.ascontiguousarray(X)
        d = cls(ctx, X.shape[0], X.shape[1] if X.ndim > 1 else 1, X.dtype)
        d.upload(X)
        return d

    def upload(self, X):
        X = np.ascontiguousarray(X, dtype=self.dtype)
        assert X.nbytes == self.nbytes
        self.ctx.check(self.ctx.lib.mlz_memcpy_htod(self.ctx.handle, self.ptr, X.ctypes.data,
                                                    self.nbytes), "memcpy_htod")

    def download(self):
        out = np.empty((self.rows, self.cols), dtype=self.dtype)
        self.ctx.check(self.ctx.lib.mlz_memcpy_dtoh(self.ctx.handle, out.ctypes.data, self.ptr,
                                                    self.nbytes), "memcpy_dtoh")
        return out

    def free(self):
        if self.ptr:
            self.ctx.lib.mlz_dev_free(self.ctx.handle, self.ptr)
            self.ptr = None

    @property
    def ld(self):
        return self.cols


def _host_ptr(X, dtype):
    if not isinstance(X, np.ndarray) or X.dtype != dtype:
        raise TypeError(f"expected numpy array of {dtype}")
    if X.ndim != 2:
        raise TypeError("expected a 2-D array")
    if X.size == 0:
        return X.ctypes.data, max(1, X.shape[1])
    if (X.shape[1] > 1 and X.strides[1] != X.itemsize) or (X.shape[0] > 1 and X.strides[0] % X.itemsize):
        raise TypeError("expected a row-major 2-D array (rows may be strided)")
    if X.shape[0] == 1:
        return X.ctypes.data, X.shape[1]
    return X.ctypes.data, X.strides[0] // X.itemsize


def _operand(ctx, X, dtype):
    """(pointer, leading dimension, rows, cols) of a host array or DeviceMatrix."""
    if isinstance(X, DeviceMatrix):
        if ctx._ptr_mode != PTR_DEVICE:
            ctx.set_ptr_mode(PTR_DEVICE)
        return X.ptr, X.ld, X.rows, X.cols
    if ctx._ptr_mode != PTR_HOST:
        ctx.set_ptr_mode(PTR_HOST)
    ptr, ld = _host_ptr(X, dtype)
    return ptr, ld, X.shape[0], X.shape[1]


class BLAS3:
    """BLAS3 front door (lela/blas/level3.h) for Z/p on the HIP module."""

    @staticmethod
    def gemm(ctx, a, A, B, b, C, algo=ALGO_AUTO):
        """C <- a*A*B + b*C (lela/blas/level3.h:81-83)."""
        F = ctx.F
        if isinstance(F, GF2):
            return _gf2_gemm(ctx, a, A, B, b, C)
        pa, lda, m, k = _operand(ctx, A, F.dtype)
        pb, ldb, kb, n = _operand(ctx, B, F.dtype)
        pc, ldc, mc, nc = _operand(ctx, C, F.dtype)
        if kb != k or mc != m or nc != n:
            raise LELAError("gemm: dimension mismatch")
        rc = ctx.lib.mlz_zp_gemm(ctx.handle, F.elem, F.p, m, n, k, int(a) % F.p, pa, lda, pb, ldb,
                                 int(b) % F.p, pc, ldc, algo)
        ctx.check(rc, "gemm")
        return C

    @staticmethod
    def trsm(ctx, a, A, B, type_, diagIsOne, algo=ALGO_AUTO):
        """B <- a*A^-1*B, A triangular (lela/blas/level3.h:120-122)."""
        F = ctx.F
        if isinstance(F, GF2):
            return _gf2_trsm(ctx, a, A, B, type_, diagIsOne)
        pa, lda, n, n2 = _operand(ctx, A, F.dtype)
        pb, ldb, nb, nrhs = _operand(ctx, B, F.dtype)
        if n != n2 or nb != n:
            raise LELAError("trsm: dimension mismatch")
        bad = ctypes.c_size_t(0)
        rc = ctx.lib.mlz_zp_trsm(ctx.handle, F.elem, F.p, int(type_ == UpperTriangular),
                                 int(bool(diagIsOne)), n, nrhs, int(a) % F.p, pa, lda, pb, ldb,
                                 ctypes.byref(bad), algo)
        ctx.check(rc, "trsm", idx=bad.value if rc == _lib.MLZ_E_NOT_INVERTIBLE else None)
        return B


class EchelonForm:
    """EchelonForm<Ring, HIPModule> (lela/solutions/echelon-form.h:31-164) for
    dense Z/p matrices.  echelonize(A, reduced=True) replaces A by its reduced
    row-echelon form and returns it; rank(A) gives the rank recorded for A."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._rank_table = {}
        self.pivots = None

    def echelonize(self, A, reduced=True, algo=ALGO_AUTO):
        if isinstance(self.ctx.F, GF2):
            if not reduced:
                raise LELAError("only the reduced form is computed on the GPU")
            return _gf2_echelonize(self, A)
        if not reduced:
            # echelon-form.h:113-131: both methods give Elimination's U
            _, rank, _ = Elimination(self.ctx).echelonize(A, compute_L=False, algo=algo)
            self._rank_table[id(A)] = rank
            self.pivots = None
            return A
        ctx, F = self.ctx, self.ctx.F
        pa, lda, m, n = _operand(ctx, A, F.dtype)
        rank = ctypes.c_size_t(0)
        piv = np.zeros(max(1, min(m, n)), dtype=np.uint32)
        rc = ctx.lib.mlz_zp_rref(ctx.handle, F.elem, F.p, m, n, pa, lda, ctypes.byref(rank),
                                 piv.ctypes.data, algo)
        ctx.check(rc, "rref")
        self._rank_table[id(A)] = rank.value
        self.pivots = piv[:rank.value].copy()
        return A

    def rank(self, A):
        return self._rank_table[id(A)]


class DensePivotStrategy:
    """LELA's pivot rule for dense rows: the first nonzero column, the first
    row in it (lela/algorithms/pivot-strategy.h:57-91)."""


class SparsePartialPivotStrategy:
    """LELA's pivot rule for sparse rows: the leftmost first nonzero, then the
    fewest nonzeros from there on, the first row on ties; words=True counts
    nonzero 64-column words as for GF(2) hybrid rows
    (lela/algorithms/pivot-strategy.h:93-131, pivot-strategy.tcc:69-188)."""

    def __init__(self, words=False):
        self.words = bool(words)


class Elimination:
    """Elimination<Ring, HIPModule> (lela/algorithms/elimination.h:34-238) for
    Z/p matrices, dense storage.  PS is the pivot strategy LELA would take
    from the row type (DefaultPivotStrategy, pivot-strategy.h:138-160):
    DensePivotStrategy() for dense rows (the default here),
    SparsePartialPivotStrategy() for sparse rows.  Permutations are LELA's:
    lists of transpositions (i, j)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def _run(self, A, mode, algo, PS=None):
        ctx, F = self.ctx, self.ctx.F
        pa, lda, m, n = _operand(ctx, A, F.dtype)
        rank = ctypes.c_size_t(0)
        det = ctypes.c_uint64(0)
        pr = np.zeros(max(1, min(m, n)), dtype=np.uint32)
        pc = np.zeros(max(1, min(m, n)), dtype=np.uint32)
        if isinstance(PS, SparsePartialPivotStrategy):
            rc = ctx.lib.mlz_zp_echelonize_sparse(ctx.handle, F.elem, F.p, m, n, pa, lda, mode, int(PS.words),
                                                  ctypes.byref(rank), ctypes.byref(det), pr.ctypes.data,
                                                  pc.ctypes.data)
        else:
            rc = ctx.lib.mlz_zp_echelonize(ctx.handle, F.elem, F.p, m, n, pa, lda, mode, ctypes.byref(rank),
                                           ctypes.byref(det), pr.ctypes.data, pc.ctypes.data, algo)
        if rc == _lib.MLZ_E_NOT_INVERTIBLE:
            raise LELAError("Could not invert pivot-element in the ring")
        ctx.check(rc, "echelonize")
        r = rank.value
        return pr[:r].copy(), pc[:r].copy(), r, int(det.value)

    def echelonize(self, A, compute_L=True, algo=ALGO_AUTO, PS=None):
        """A <- its row-echelon form U (elimination.tcc:47-139); with
        compute_L (LELA's default) the part left of the diagonal holds T with
        T P A = U.  Returns (P, rank, det)."""
        pr, _, r, det = self._run(A, _lib.ECH_TRANSFORM if compute_L else _lib.ECH_U, algo, PS)
        self.steps = pr
        return [(i, int(j)) for i, j in enumerate(pr) if j != i], r, det

    def pluq(self, A, algo=ALGO_AUTO, PS=None):
        """A <- L and U of P A Q = L U (elimination.tcc:266-370).  Returns
        (P, Q, rank, det) with LELA's lists (reversed step order)."""
        pr, pc, r, det = self._run(A, _lib.ECH_PLUQ, algo, PS)
        P = [(i, int(j)) for i, j in enumerate(pr) if j != i][::-1]
        Q = [(i, int(j)) for i, j in enumerate(pc) if j != i][::-1]
        return P, Q, r, det


class GF2:
    """The field with two elements (lela/ring/gf2.h:48-293); matrices are
    Dense01Matrix objects (64-bit words, LSB first)."""
    p = 2

    def __repr__(self):
        return "GF2"


class Dense01Matrix:
    """Bit-packed dense GF(2) matrix: `words` is a (rows x ceil(cols/64))
    uint64 array (or a DeviceMatrix of words); bit j of word w is column
    64*w + j (lela/matrix/dense-zero-one.h:132-360)."""

    def __init__(self, words, ncols):
        self.words = words
        self.ncols = int(ncols)

    def rowdim(self):
        return self.words.rows if isinstance(self.words, DeviceMatrix) else self.words.shape[0]

    def coldim(self):
        return self.ncols


def _gf2_operand(ctx, X):
    if not isinstance(X, Dense01Matrix):
        raise TypeError("GF(2) operands must be Dense01Matrix")
    ptr, ld, rows, _ = _operand(ctx, X.words, np.dtype(np.uint64))
    return ptr, ld, rows, X.ncols


def _gf2_gemm(ctx, a, A, B, b, C):
    pa, lda, m, k = _gf2_operand(ctx, A)
    pb, ldb, kb, n = _gf2_operand(ctx, B)
    pc, ldc, mc, nc = _gf2_operand(ctx, C)
    if kb != k or mc != m or nc != n:
        raise LELAError("gemm: dimension mismatch")
    ctx.check(ctx.lib.mlz_gf2_gemm(ctx.handle, m, n, k, int(a) & 1, pa, lda, pb, ldb, int(b) & 1,
                                   pc, ldc), "gf2 gemm")
    return C


def _gf2_trsm(ctx, a, A, B, type_, diagIsOne):
    pa, lda, n, n2 = _gf2_operand(ctx, A)
    pb, ldb, nb, nrhs = _gf2_operand(ctx, B)
    if n != n2 or nb != n:
        raise LELAError("trsm: dimension mismatch")
    bad = ctypes.c_size_t(0)
    rc = ctx.lib.mlz_gf2_trsm(ctx.handle, int(type_ == UpperTriangular), int(bool(diagIsOne)), n, nrhs,
                              int(a) & 1, pa, lda, pb, ldb, ctypes.byref(bad))
    ctx.check(rc, "gf2 trsm", idx=bad.value if rc == _lib.MLZ_E_NOT_INVERTIBLE else None)
    return B


def _gf2_echelonize(EF, A):
    ctx = EF.ctx
    pa, lda, m, n = _gf2_operand(ctx, A)
    rank = ctypes.c_size_t(0)
    piv = np.zeros(max(1, min(m, n)), dtype=np.uint32)
    ctx.check(ctx.lib.mlz_gf2_rref(ctx.handle, m, n, pa, lda, ctypes.byref(rank), piv.ctypes.data),
              "gf2 rref")
    EF._rank_table[id(A)] = rank.value
    EF.pivots = piv[:rank.value].copy()
    return A


def dense01_to_hybrid(X):
    """Hybrid rows (row_ptr uint64, word index uint32, word uint64) of a
    Dense01Matrix: the nonzero words of each row (lela/vector/hybrid.h:30-33)."""
    W = np.asarray(X.words, dtype=np.uint64)
    rows, blks = np.nonzero(W)
    row_ptr = np.zeros(W.shape[0] + 1, dtype=np.uint64)
    np.add.at(row_ptr, rows + 1, 1)
    return np.cumsum(row_ptr, dtype=np.uint64), blks.astype(np.uint32), W[rows, blks]


def dense_to_csr(X):
    """Row-compressed form (row_ptr uint64, col uint32, values) of a dense
    matrix, dropping zeros -- LELA's sparse rows store only nonzeros
    (lela/vector/sparse.h:311-491)."""
    X = np.asarray(X)
    rows, cols = np.nonzero(X)
    row_ptr = np.zeros(X.shape[0] + 1, dtype=np.uint64)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr, dtype=np.uint64)
    return row_ptr, cols.astype(np.uint32), X[rows, cols]


class FaugereLachartre:
    """FaugereLachartre<Ring, HIPModule> (lela/algorithms/faugere-lachartre.h:40-81)
    over Modular<T> (CSR or dense rows) and GF2 (hybrid rows)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.pivots = None

    def echelonize(self, R, X, reduced=True, algo=ALGO_AUTO, shape=None):
        """R <- reduced row-echelon form of X; returns (rank, det).

        X is a dense array or a CSR triple (row_ptr, col, val) with `shape`.
        det is the product of the heads of X's scanned rows
        (faugere-lachartre.tcc:119).  WrongMatrixForm is raised for a
        decreasing head; when a head is not invertible (composite modulus)
        LELA reports it and returns with R unchanged, and so does this call."""
        if not reduced:
            raise LELAError("only the reduced form is computed on the GPU")
        ctx, F = self.ctx, self.ctx.F
        if isinstance(F, GF2):
            return self._gf2(R, X, shape)
        if isinstance(X, tuple):
            row_ptr, col, val = X
            m, n = shape
        else:
            m, n = X.shape
            row_ptr, col, val = dense_to_csr(X)
        on_device = isinstance(R, DeviceMatrix) or isinstance(row_ptr, DeviceMatrix)
        if on_device:
            # resident inputs and output: every pointer must be device memory
            if not all(isinstance(v, DeviceMatrix) for v in (R, row_ptr, col, val)):
                raise TypeError("device FL needs R, row_ptr, col and val as DeviceMatrix")
            if ctx._ptr_mode != PTR_DEVICE:
                ctx.set_ptr_mode(PTR_DEVICE)
            return self._call(m, n, row_ptr.ptr, col.ptr, val.ptr, R.ptr, R.ld, algo)
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=F.dtype)
        if R is not None:
            pr, ldr, rm, rn = _operand(ctx, R, F.dtype)
            if (rm, rn) != (m, n):
                raise LELAError("R must have the shape of X")
        else:
            pr, ldr = None, n
            if ctx._ptr_mode != PTR_HOST:
                ctx.set_ptr_mode(PTR_HOST)
        return self._call(m, n, row_ptr.ctypes.data, col.ctypes.data, val.ctypes.data, pr, ldr, algo)

    def echelonize_compact(self, X, algo=ALGO_AUTO, shape=None, NP=None):
        """The reduced form of X in compact form (mlz_zp_fl_echelonize_compact):
        returns (rank, det, pivots, NP) where NP (rank x (n - rank)) holds the
        rows of R on the non-pivot columns in increasing order -- R is the
        identity on the pivot columns.  X as for echelonize (host arrays), or
        device CSR triples with NP a DeviceMatrix of at least rank x (n - rank)
        entries (then NP's buffer is returned as filled, row stride n - rank)."""
        ctx, F = self.ctx, self.ctx.F
        if isinstance(X, tuple):
            row_ptr, col, val = X
            m, n = shape
        else:
            m, n = X.shape
            row_ptr, col, val = dense_to_csr(X)
        rank = ctypes.c_size_t(0)
        det = ctypes.c_uint64(0)
        bad = ctypes.c_size_t(0)
        piv = np.zeros(max(1, min(m, n)), dtype=np.uint32)
        if isinstance(row_ptr, DeviceMatrix):
            if ctx._ptr_mode != PTR_DEVICE:
                ctx.set_ptr_mode(PTR_DEVICE)
            args = (row_ptr.ptr, col.ptr, val.ptr, NP.ptr, NP.rows * NP.cols)
            out = NP
        else:
            if ctx._ptr_mode != PTR_HOST:
                ctx.set_ptr_mode(PTR_HOST)
            row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
            col = np.ascontiguousarray(col, dtype=np.uint32)
            val = np.ascontiguousarray(val, dtype=F.dtype)
            cap = max(1, min(m, n) * n)        # rank * (n - rank) never exceeds it
            out = np.zeros(cap, dtype=F.dtype)
            args = (row_ptr.ctypes.data, col.ctypes.data, val.ctypes.data, out.ctypes.data, cap)
        rc = ctx.lib.mlz_zp_fl_echelonize_compact(ctx.handle, F.elem, F.p, m, n, args[0], args[1], args[2],
                                                  args[3], args[4], ctypes.byref(rank), piv.ctypes.data,
                                                  ctypes.byref(det), ctypes.byref(bad), algo)
        ctx.check(rc, "fl_echelonize_compact", idx=bad.value if rc == _lib.MLZ_E_WRONG_FORM else None)
        r = rank.value
        self.pivots = piv[:r].copy()
        if isinstance(out, np.ndarray):
            out = out[:r * (n - r)].reshape(r, n - r).copy()
        return r, det.value, self.pivots, out

    def _gf2(self, R, X, shape):
        """FaugereLachartre<GF2> on hybrid rows: X is a Dense01Matrix or a
        triple (row_ptr, blk, words) with `shape` (lela/vector/hybrid.h:30-33);
        R a Dense01Matrix or None.  Returns (rank, 1): every head is 1."""
        ctx = self.ctx
        if isinstance(X, Dense01Matrix):
            m, n = X.rowdim(), X.coldim()
            row_ptr, blk, words = dense01_to_hybrid(X)
        else:
            row_ptr, blk, words = X
            m, n = shape
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        blk = np.ascontiguousarray(blk, dtype=np.uint32)
        words = np.ascontiguousarray(words, dtype=np.uint64)
        if ctx._ptr_mode != PTR_HOST:
            ctx.set_ptr_mode(PTR_HOST)
        if R is not None:
            pr, ldr, rm, rn = _gf2_operand(ctx, R)
            if (rm, rn) != (m, n):
                raise LELAError("R must have the shape of X")
        else:
            pr, ldr = None, (n + 63) // 64
        rank = ctypes.c_size_t(0)
        bad = ctypes.c_size_t(0)
        piv = np.zeros(max(1, min(m, n)), dtype=np.uint32)
        rc = ctx.lib.mlz_gf2_fl_echelonize(ctx.handle, m, n, row_ptr.ctypes.data, blk.ctypes.data,
                                           words.ctypes.data, pr, ldr, ctypes.byref(rank), piv.ctypes.data,
                                           ctypes.byref(bad))
        ctx.check(rc, "gf2 fl_echelonize", idx=bad.value if rc == _lib.MLZ_E_WRONG_FORM else None)
        self.pivots = piv[:rank.value].copy()
        return rank.value, 1

    def _call(self, m, n, row_ptr, col, val, pr, ldr, algo):
        ctx, F = self.ctx, self.ctx.F
        rank = ctypes.c_size_t(0)
        det = ctypes.c_uint64(0)
        bad = ctypes.c_size_t(0)
        piv = np.zeros(max(1, min(m, n)), dtype=np.uint32)
        rc = ctx.lib.mlz_zp_fl_echelonize(ctx.handle, F.elem, F.p, m, n, row_ptr, col, val, pr, ldr,
                                          ctypes.byref(rank), piv.ctypes.data, ctypes.byref(det),
                                          ctypes.byref(bad), algo)
        if rc == _lib.MLZ_E_NOT_INVERTIBLE:
            return rank.value, det.value
        ctx.check(rc, "fl_echelonize", idx=bad.value if rc == _lib.MLZ_E_WRONG_FORM else None)
        self.pivots = piv[:rank.value].copy()
        return rank.value, det.value
