"""ctypes binding of liblela_amd.so, the C ABI declared in include/mlz.h.

The library is the only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblela_amd.so")

MLZ_U16, MLZ_U32, MLZ_F64_SYM = 1, 2, 3
ALGO_AUTO, ALGO_PLAIN, ALGO_INT8, ALGO_FP64 = 0, 1, 2, 3
PTR_HOST, PTR_DEVICE = 0, 1

MLZ_OK = 0
MLZ_E_ARG = 1
MLZ_E_HIP = 2
MLZ_E_NOT_INVERTIBLE = 3
MLZ_E_WRONG_FORM = 4
MLZ_E_NOMEM = 5
MLZ_E_UNSUPPORTED = 6
MLZ_E_FORMAT = 7
MLZ_E_INPUT = 8
MLZ_E_FILE = 9
MLZ_IO_GF2 = 4
NOALLOC_SCRATCH_MB = 64

_c_size = ctypes.c_size_t
_c_u64 = ctypes.c_uint64
_c_int = ctypes.c_int
_vp = ctypes.c_void_p
_ctx = ctypes.c_void_p

# name -> (restype, argtypes); every function declared in include/mlz.h.
SIGNATURES = {
    "mlz_ctx_create": (_c_int, [ctypes.POINTER(_ctx), _c_int, ctypes.c_char_p]),
    "mlz_ctx_destroy": (_c_int, [_ctx]),
    "mlz_ctx_set_ptr_mode": (_c_int, [_ctx, _c_int]),
    "mlz_ctx_set_option": (_c_int, [_ctx, ctypes.c_char_p, ctypes.c_int64]),
    "mlz_ctx_synchronize": (_c_int, [_ctx]),
    "mlz_last_error": (ctypes.c_char_p, [_ctx]),
    "mlz_version": (ctypes.c_char_p, []),
    "mlz_dev_alloc": (_c_int, [_ctx, _c_size, ctypes.POINTER(_vp)]),
    "mlz_dev_free": (_c_int, [_ctx, _vp]),
    "mlz_memcpy_htod": (_c_int, [_ctx, _vp, _vp, _c_size]),
    "mlz_memcpy_dtoh": (_c_int, [_ctx, _vp, _vp, _c_size]),
    "mlz_memcpy_dtod": (_c_int, [_ctx, _vp, _vp, _c_size]),
    "mlz_ctx_trim": (_c_int, [_ctx]),
    "mlz_prof_enable": (_c_int, [_ctx, _c_int]),
    "mlz_prof_reset": (_c_int, [_ctx]),
    "mlz_prof_query": (_c_int, [_ctx, ctypes.c_char_p, ctypes.POINTER(_c_u64),
                                ctypes.POINTER(ctypes.c_double)]),
    "mlz_timer_start": (_c_int, [_ctx]),
    "mlz_timer_stop": (_c_int, [_ctx, ctypes.POINTER(ctypes.c_double)]),
    "mlz_zp_gemm": (_c_int, [_ctx, _c_int, _c_u64, _c_size, _c_size, _c_size, _c_u64,
                             _vp, _c_size, _vp, _c_size, _c_u64, _vp, _c_size, _c_int]),
    "mlz_zp_gemm_noalloc": (_c_int, [_ctx, _c_int, _c_u64, _c_size, _c_size, _c_size, _c_u64,
                                     _vp, _c_size, _vp, _c_size, _c_u64, _vp, _c_size, _c_int]),
    "mlz_zp_trsm": (_c_int, [_ctx, _c_int, _c_u64, _c_int, _c_int, _c_size, _c_size,
                             _c_u64, _vp, _c_size, _vp, _c_size, ctypes.POINTER(_c_size), _c_int]),
    "mlz_zp_rref": (_c_int, [_ctx, _c_int, _c_u64, _c_size, _c_size, _vp, _c_size,
                             ctypes.POINTER(_c_size), _vp, _c_int]),
    "mlz_gf2_gemm": (_c_int, [_ctx, _c_size, _c_size, _c_size, _c_int, _vp, _c_size, _vp, _c_size,
                              _c_int, _vp, _c_size]),
    "mlz_gf2_rref": (_c_int, [_ctx, _c_size, _c_size, _vp, _c_size, ctypes.POINTER(_c_size), _vp]),
    "mlz_gf2_fl_echelonize": (_c_int, [_ctx, _c_size, _c_size, _vp, _vp, _vp, _vp, _c_size, ctypes.POINTER(_c_size),
                                       _vp, ctypes.POINTER(_c_size)]),
    "mlz_gf2_trsm": (_c_int, [_ctx, _c_int, _c_int, _c_size, _c_size, _c_int, _vp, _c_size, _vp, _c_size,
                              ctypes.POINTER(_c_size)]),
    "mlz_zp_echelonize": (_c_int, [_ctx, _c_int, _c_u64, _c_size, _c_size, _vp, _c_size, _c_int,
                                   ctypes.POINTER(_c_size), ctypes.POINTER(_c_u64), _vp, _vp, _c_int]),
    "mlz_zp_echelonize_sparse": (_c_int, [_ctx, _c_int, _c_u64, _c_size, _c_size, _vp, _c_size, _c_int, _c_int,
                                          ctypes.POINTER(_c_size), ctypes.POINTER(_c_u64), _vp, _vp]),
    "mlz_zp_fl_echelonize": (_c_int, [_ctx, _c_int, _c_u64, _c_size, _c_size, _vp, _vp, _vp, _vp,
                                      _c_size, ctypes.POINTER(_c_size), _vp, ctypes.POINTER(_c_u64),
                                      ctypes.POINTER(_c_size), _c_int]),
    "mlz_zp_fl_echelonize_compact": (_c_int, [_ctx, _c_int, _c_u64, _c_size, _c_size, _vp, _vp, _vp, _vp,
                                              _c_size, ctypes.POINTER(_c_size), _vp, ctypes.POINTER(_c_u64),
                                              ctypes.POINTER(_c_size), _c_int]),
    "mlz_batch_plan": (_c_int, [_ctx, _c_int, _c_u64, _vp, _c_size, _c_int, _c_int, ctypes.POINTER(_vp)]),
    "mlz_batch_launch": (_c_int, [_ctx, _vp]),
    "mlz_batch_finish": (_c_int, [_ctx, _vp, _vp, _vp, ctypes.POINTER(_c_size)]),
    "mlz_batch_destroy": (_c_int, [_ctx, _vp]),
    "mlz_batch_run": (_c_int, [_ctx, _c_int, _c_u64, _vp, _c_size, _c_int]),
    "mlz_tune": (_c_int, [_ctx, _c_int, _c_u64, _vp, _c_size, _c_int, _vp]),
    "mlz_tune_table_get": (_c_int, [_ctx, _c_int, _vp]),
    "mlz_tune_table_set": (_c_int, [_ctx, _c_int, _vp]),
    "mlz_task_features": (_c_int, [_vp, _c_u64, _vp]),
    "mlz_model_train": (_c_int, [_ctx, _c_int, _vp, _vp, _vp, _c_size, _c_int, ctypes.c_float, _c_u64,
                                 ctypes.POINTER(ctypes.c_float)]),
    "mlz_model_get": (_c_int, [_ctx, _c_int, _vp]),
    "mlz_model_set": (_c_int, [_ctx, _c_int, _vp]),
    "mlz_model_predict": (_c_int, [_ctx, _c_int, _vp, _c_size, _vp]),
    "mlz_model_grad": (_c_int, [_ctx, _vp, _vp, _vp, _vp, _c_size, _vp, ctypes.POINTER(ctypes.c_float)]),
    # include/mlz_io.h: matrix files (host code)
    "mlz_matrix_detect_format": (_c_int, [ctypes.c_char_p, ctypes.POINTER(_c_int)]),
    "mlz_matrix_read": (_c_int, [ctypes.c_char_p, _c_int, _c_int, _c_u64, _c_size, _c_size, ctypes.POINTER(_vp)]),
    "mlz_matrix_shape": (_c_int, [_vp, ctypes.POINTER(_c_size), ctypes.POINTER(_c_size), ctypes.POINTER(_c_size)]),
    "mlz_matrix_get_csr": (_c_int, [_vp, _vp, _vp, _vp]),
    "mlz_matrix_release": (_c_int, [_vp]),
    "mlz_matrix_write_csr": (_c_int, [ctypes.c_char_p, _c_int, _c_int, _c_u64, _c_size, _c_size, _vp, _vp, _vp]),
    "mlz_matrix_write_dense": (_c_int, [ctypes.c_char_p, _c_int, _c_int, _c_u64, _c_size, _c_size, _vp, _c_size]),
    "mlz_io_last_error": (ctypes.c_char_p, []),
}

TASK_GEMM, TASK_RREF = 1, 2
ECH_U, ECH_TRANSFORM, ECH_PLUQ = 0, 1, 2
MAX_VARIANTS = 4
MODEL_PARAMS = 1380
POLICY_FIXED, POLICY_TABLE, POLICY_MODEL = 0, 1, 2
BATCH_GRAPH, BATCH_TIMING, BATCH_SERIAL = 1, 2, 4
GEMM_PLAIN, GEMM_SPARSE, GEMM_I8S, GEMM_I8L = 0, 1, 2, 3
RREF_W64, RREF_W32, RREF_I8_W64, RREF_I8_W32 = 0, 1, 2, 3


class Task(ctypes.Structure):
    """struct mlz_task of include/mlz.h (same field order and types)."""
    _fields_ = [
        ("op", ctypes.c_int),
        ("variant", ctypes.c_int),
        ("m", ctypes.c_uint32),
        ("n", ctypes.c_uint32),
        ("k", ctypes.c_uint32),
        ("density", ctypes.c_float),
        ("a", ctypes.c_uint64),
        ("b", ctypes.c_uint64),
        ("A", ctypes.c_void_p),
        ("lda", ctypes.c_size_t),
        ("B", ctypes.c_void_p),
        ("ldb", ctypes.c_size_t),
        ("C", ctypes.c_void_p),
        ("ldc", ctypes.c_size_t),
        ("pivots", ctypes.c_void_p),
        ("rank", ctypes.c_size_t),
    ]

_lib = None


def load():
    """Loads liblela_amd.so once; raises RuntimeError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
