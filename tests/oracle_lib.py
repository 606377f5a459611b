"""ctypes wrapper of oracle/liboracle.so -- the CPU checker the tests compare against.

Test infrastructure only (never imported by the product package).
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

ORC_OK, ORC_E_ARG, ORC_E_NOT_INVERTIBLE, ORC_E_WRONG_FORM = 0, 1, 3, 4

_lib = None
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    lib = ctypes.CDLL(ORACLE_SO)
    u32, u16, i = ctypes.c_uint32, ctypes.c_uint16, ctypes.c_int
    lib.orc_zp_gemm_u32.argtypes = [u32, _sz, _sz, _sz, u32, _vp, _sz, _vp, _sz, u32, _vp, _sz]
    lib.orc_zp_gemm_u16.argtypes = [u16, _sz, _sz, _sz, u16, _vp, _sz, _vp, _sz, u16, _vp, _sz]
    lib.orc_zp_trsm.argtypes = [u32, i, i, _sz, _sz, u32, _vp, _sz, _vp, _sz, ctypes.POINTER(_sz)]
    lib.orc_zp_echelonize.argtypes = [u32, _sz, _sz, _vp, _sz, i, ctypes.POINTER(_sz), _vp,
                                      ctypes.POINTER(u32)]
    lib.orc_zp_elim.argtypes = [u32, _sz, _sz, _vp, _sz, i, ctypes.POINTER(_sz), ctypes.POINTER(u32), _vp]
    lib.orc_zp_pluq.argtypes = [u32, _sz, _sz, _vp, _sz, ctypes.POINTER(_sz), ctypes.POINTER(u32), _vp, _vp]
    lib.orc_zp_elim_sparse.argtypes = [u32, _sz, _sz, _vp, _sz, i, i, ctypes.POINTER(_sz), ctypes.POINTER(u32),
                                       _vp, _vp]
    lib.orc_gf2_gemm.argtypes = [_sz, _sz, _sz, i, _vp, _sz, _vp, _sz, i, _vp, _sz]
    lib.orc_gf2_echelonize.argtypes = [_sz, _sz, _vp, _sz, i, ctypes.POINTER(_sz), _vp]
    lib.orc_gf2_trsm.argtypes = [i, i, _sz, _sz, i, _vp, _sz, _vp, _sz, ctypes.POINTER(_sz)]
    lib.orc_fl_echelonize.argtypes = [u32, _sz, _sz, _vp, _sz, _vp, _sz, ctypes.POINTER(_sz),
                                      ctypes.POINTER(u32), ctypes.POINTER(_sz)]
    lib.orc_inv.argtypes = [u32, u32, ctypes.POINTER(u32)]
    lib.orc_mlp_forward.argtypes = [_vp, _vp, _sz, _vp]
    lib.orc_mlp_loss_grad.argtypes = [_vp, _vp, _vp, _vp, _sz, _vp, ctypes.POINTER(ctypes.c_double)]
    lib.orc_mlp_adam.argtypes = [_vp, _vp, _vp, _vp, _sz, i, ctypes.c_double, _vp]
    for name in ("orc_zp_gemm_u32", "orc_zp_gemm_u16", "orc_zp_trsm", "orc_zp_echelonize", "orc_zp_elim",
                 "orc_zp_pluq", "orc_zp_elim_sparse",
                 "orc_gf2_gemm", "orc_gf2_echelonize", "orc_gf2_trsm", "orc_fl_echelonize", "orc_inv",
                 "orc_mlp_forward", "orc_mlp_loss_grad", "orc_mlp_adam"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _c(X, dtype):
    return np.ascontiguousarray(X, dtype=dtype)


def zp_gemm(p, a, A, B, b, C, storage="uint32"):
    """Returns a*A*B + b*C mod p (LELA ZpModule semantics)."""
    lib = load()
    dt = np.uint16 if storage == "uint16" else np.uint32
    A, B, C = _c(A, dt), _c(B, dt), _c(C, dt).copy()
    m, k = A.shape
    n = B.shape[1]
    fn = lib.orc_zp_gemm_u16 if storage == "uint16" else lib.orc_zp_gemm_u32
    rc = fn(p, m, n, k, a, A.ctypes.data, k, B.ctypes.data, n, b, C.ctypes.data, n)
    assert rc == ORC_OK
    return C


def zp_trsm(p, upper, unit, a, A, B):
    """Returns (rc, bad_index, X) for X = a*A^-1*B."""
    lib = load()
    A, B = _c(A, np.uint32), _c(B, np.uint32).copy()
    n, nrhs = B.shape
    bad = _sz(0)
    rc = lib.orc_zp_trsm(p, int(upper), int(unit), n, nrhs, a, A.ctypes.data, n, B.ctypes.data,
                         nrhs, ctypes.byref(bad))
    return rc, bad.value, B


def zp_echelonize(p, A, reduced=True):
    """Returns (rc, R, rank, pivots, det) from LELA's Elimination."""
    lib = load()
    R = _c(A, np.uint32).copy()
    m, n = R.shape
    rank = _sz(0)
    piv = np.zeros(max(1, min(m, n)), dtype=np.uint32)
    det = ctypes.c_uint32(0)
    rc = lib.orc_zp_echelonize(p, m, n, R.ctypes.data, n, int(reduced), ctypes.byref(rank),
                               piv.ctypes.data, ctypes.byref(det))
    return rc, R, rank.value, piv[:rank.value].copy(), det.value


def zp_elim(p, A, compute_L=False):
    """Returns (rc, A_out, rank, det, perm_rows) of Elimination::echelonize."""
    lib = load()
    R = _c(A, np.uint32).copy()
    m, n = R.shape
    rank = _sz(0)
    det = ctypes.c_uint32(0)
    perm = np.zeros(max(1, min(m, n)), dtype=np.uint32)
    rc = lib.orc_zp_elim(p, m, n, R.ctypes.data, n, int(compute_L), ctypes.byref(rank), ctypes.byref(det),
                         perm.ctypes.data)
    return rc, R, rank.value, det.value, perm[:rank.value].copy()


def zp_pluq(p, A):
    """Returns (rc, A_out, rank, det, perm_rows, perm_cols) of Elimination::pluq (step order)."""
    lib = load()
    R = _c(A, np.uint32).copy()
    m, n = R.shape
    rank = _sz(0)
    det = ctypes.c_uint32(0)
    pr = np.zeros(max(1, min(m, n)), dtype=np.uint32)
    pc = np.zeros(max(1, min(m, n)), dtype=np.uint32)
    rc = lib.orc_zp_pluq(p, m, n, R.ctypes.data, n, ctypes.byref(rank), ctypes.byref(det), pr.ctypes.data,
                         pc.ctypes.data)
    return rc, R, rank.value, det.value, pr[:rank.value].copy(), pc[:rank.value].copy()


def zp_elim_sparse(p, A, mode, words=False):
    """Returns (rc, A_out, rank, det, perm_rows, perm_cols) of Elimination with
    SparsePartialPivotStrategy: mode 0 echelonize, 1 echelonize with
    compute_L, 2 pluq (step order)."""
    lib = load()
    R = _c(A, np.uint32).copy()
    m, n = R.shape
    rank = _sz(0)
    det = ctypes.c_uint32(0)
    pr = np.zeros(max(1, min(m, n)), dtype=np.uint32)
    pc = np.zeros(max(1, min(m, n)), dtype=np.uint32)
    rc = lib.orc_zp_elim_sparse(p, m, n, R.ctypes.data, n, int(mode), int(words), ctypes.byref(rank),
                                ctypes.byref(det), pr.ctypes.data, pc.ctypes.data)
    return rc, R, rank.value, det.value, pr[:rank.value].copy(), pc[:rank.value].copy()


def gf2_gemm(m, n, k, a, A, B, b, C):
    lib = load()
    A, B, C = _c(A, np.uint64), _c(B, np.uint64), _c(C, np.uint64).copy()
    rc = lib.orc_gf2_gemm(m, n, k, int(a), A.ctypes.data, A.shape[1], B.ctypes.data, B.shape[1],
                          int(b), C.ctypes.data, C.shape[1])
    assert rc == ORC_OK
    return C


def gf2_trsm(upper, unit, n, nrhs, a, A, B):
    """Returns (rc, bad_index, X) for X = a*A^-1*B over GF(2) (word rows)."""
    lib = load()
    A, B = _c(A, np.uint64), _c(B, np.uint64).copy()
    bad = _sz(0)
    rc = lib.orc_gf2_trsm(int(upper), int(unit), n, nrhs, int(a), A.ctypes.data, A.shape[1], B.ctypes.data,
                          B.shape[1], ctypes.byref(bad))
    return rc, bad.value, B


def gf2_echelonize(m, n, A, reduced=True):
    lib = load()
    R = _c(A, np.uint64).copy()
    rank = _sz(0)
    piv = np.zeros(max(1, min(m, n)), dtype=np.uint32)
    rc = lib.orc_gf2_echelonize(m, n, R.ctypes.data, R.shape[1], int(reduced), ctypes.byref(rank),
                                piv.ctypes.data)
    assert rc == ORC_OK
    return R, rank.value, piv[:rank.value].copy()


def fl_echelonize(p, X):
    """Returns (rc, R, rank, det, bad_row) of FaugereLachartre::echelonize(reduced=true)."""
    lib = load()
    X = _c(X, np.uint32)
    m, n = X.shape
    R = np.zeros((m, n), dtype=np.uint32)
    rank, bad = _sz(0), _sz(0)
    det = ctypes.c_uint32(0)
    rc = lib.orc_fl_echelonize(p, m, n, X.ctypes.data, n, R.ctypes.data, n, ctypes.byref(rank),
                               ctypes.byref(det), ctypes.byref(bad))
    return rc, R, rank.value, det.value, bad.value


def inv(p, y):
    lib = load()
    x = ctypes.c_uint32(0)
    ok = lib.orc_inv(p, y, ctypes.byref(x))
    return x.value if ok else None


MLP_PARAMS = 1380


def mlp_forward(params, x):
    """Predictions [n, 4] of the cost model (double precision)."""
    lib = load()
    params, x = _c(params, np.float32), _c(x, np.float32)
    y = np.zeros((x.shape[0], 4), dtype=np.float64)
    assert lib.orc_mlp_forward(params.ctypes.data, x.ctypes.data, x.shape[0], y.ctypes.data) == ORC_OK
    return y


def mlp_loss_grad(params, x, t, mask):
    """(loss, gradient[1380]) of the masked mean squared error."""
    lib = load()
    params, x, t = _c(params, np.float32), _c(x, np.float32), _c(t, np.float32)
    mask = _c(mask, np.uint8)
    g = np.zeros(MLP_PARAMS, dtype=np.float64)
    loss = ctypes.c_double(0.0)
    assert lib.orc_mlp_loss_grad(params.ctypes.data, x.ctypes.data, t.ctypes.data, mask.ctypes.data,
                                 x.shape[0], g.ctypes.data, ctypes.byref(loss)) == ORC_OK
    return loss.value, g


def mlp_adam(params, x, t, mask, steps, lr):
    """(params after `steps` Adam steps, loss before each step)."""
    lib = load()
    w = _c(params, np.float32).copy()
    x, t, mask = _c(x, np.float32), _c(t, np.float32), _c(mask, np.uint8)
    hist = np.zeros(max(1, steps), dtype=np.float64)
    assert lib.orc_mlp_adam(w.ctypes.data, x.ctypes.data, t.ctypes.data, mask.ctypes.data, x.shape[0],
                            int(steps), float(lr), hist.ctypes.data) == ORC_OK
    return w, hist[:steps]
