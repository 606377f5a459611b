"""Matrix files: LELA's MatrixReader / MatrixWriter (lela/matrix/io.h:36-223)
and BLAS3::read / BLAS3::write (lela/blas/level3.h:191-205), through the
host-side readers and writers of liblela_amd.so (include/mlz_io.h).

    F = Modular(65521, "uint16")
    A = MatrixReader(F).read("m.sms", FORMAT_DUMAS)      # dense numpy array
    MatrixWriter(F).write("out.txt", A, FORMAT_PRETTY)

Over GF2 dense matrices are Dense01Matrix objects.  Reading needs no GPU.
"""
import ctypes

import numpy as np

from . import _lib

(FORMAT_DETECT, FORMAT_UNKNOWN, FORMAT_TURNER, FORMAT_ONE_BASED, FORMAT_DUMAS, FORMAT_MAPLE, FORMAT_MATLAB,
 FORMAT_SAGE, FORMAT_PRETTY) = range(9)

FORMAT_NAMES = {FORMAT_DETECT: "detect", FORMAT_UNKNOWN: "unknown", FORMAT_TURNER: "Turner",
                FORMAT_ONE_BASED: "one-based", FORMAT_DUMAS: "Dumas", FORMAT_MAPLE: "Maple",
                FORMAT_MATLAB: "Matlab", FORMAT_SAGE: "Sage", FORMAT_PRETTY: "pretty"}


def _errors():
    from . import LELAError
    return LELAError


class UnrecognisedFormat(RuntimeError):
    """The format of an input file could not be detected (lela/matrix/io.h:46)."""


class InvalidMatrixInput(RuntimeError):
    """Malformed matrix input (lela/matrix/io.h:51)."""


class NotImplemented_(RuntimeError):
    """LELA's NotImplemented: reading one-based, Maple or Sage files."""


def _ring(F):
    from . import GF2
    if isinstance(F, GF2):
        return _lib.MLZ_IO_GF2, 2
    return F.elem, F.p


def _check(rc, what):
    if rc == 0:
        return
    msg = f"{what}: {_lib.load().mlz_io_last_error().decode()}"
    if rc == _lib.MLZ_E_FORMAT:
        raise UnrecognisedFormat(msg)
    if rc == _lib.MLZ_E_INPUT:
        raise InvalidMatrixInput(msg)
    if rc == _lib.MLZ_E_UNSUPPORTED:
        raise NotImplemented_(msg)
    raise _errors()(f"{msg} (code {rc})")


class SparseRows:
    """A matrix read from a file as rows of nonzeros: row_ptr (m+1), col, val
    (canonical residues); the form LELA's SparseMatrix keeps."""

    def __init__(self, m, n, row_ptr, col, val):
        self.m, self.n = m, n
        self.row_ptr, self.col, self.val = row_ptr, col, val

    @property
    def shape(self):
        return self.m, self.n

    def dense(self, dtype=np.uint64):
        X = np.zeros((self.m, self.n), dtype=dtype)
        rows = np.repeat(np.arange(self.m), np.diff(self.row_ptr).astype(np.int64))
        X[rows, self.col.astype(np.int64)] = self.val.astype(dtype)
        return X


class MatrixReader:
    """MatrixReader<Ring> (lela/matrix/io.h:57-175)."""

    def __init__(self, F):
        self.F = F

    @staticmethod
    def detectFormat(path):
        fmt = ctypes.c_int(0)
        _check(_lib.load().mlz_matrix_detect_format(str(path).encode(), ctypes.byref(fmt)), "detect")
        return fmt.value

    def read_rows(self, path, format=FORMAT_DETECT, rows=0, cols=0):
        """The file's matrix as SparseRows (rows/cols: the shape of the
        matrix LELA would read into; 0 takes it from the entries)."""
        lib = _lib.load()
        elem, p = _ring(self.F)
        h = ctypes.c_void_p()
        _check(lib.mlz_matrix_read(str(path).encode(), format, elem, p, rows, cols, ctypes.byref(h)), "read")
        try:
            m, n, nnz = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
            lib.mlz_matrix_shape(h, ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz))
            rp = np.zeros(m.value + 1, dtype=np.uint64)
            col = np.zeros(max(1, nnz.value), dtype=np.uint32)
            val = np.zeros(max(1, nnz.value), dtype=np.uint64)
            _check(lib.mlz_matrix_get_csr(h, rp.ctypes.data, col.ctypes.data, val.ctypes.data), "read")
        finally:
            lib.mlz_matrix_release(h)
        return SparseRows(m.value, n.value, rp, col[:nnz.value], val[:nnz.value])

    def read(self, path, format=FORMAT_DETECT, rows=0, cols=0):
        """The file's matrix, dense: a numpy array in the ring's storage, or a
        Dense01Matrix over GF2."""
        from . import GF2, Dense01Matrix
        S = self.read_rows(path, format, rows, cols)
        if isinstance(self.F, GF2):
            nw = (S.n + 63) // 64
            W = np.zeros((S.m, max(1, nw)), dtype=np.uint64)
            rows_ = np.repeat(np.arange(S.m), np.diff(S.row_ptr).astype(np.int64))
            c = S.col.astype(np.int64)
            np.bitwise_or.at(W, (rows_, c // 64), np.left_shift(np.uint64(1), (c % 64).astype(np.uint64)))
            return Dense01Matrix(W, S.n)
        return self.F.from_canonical(S.dense())


class MatrixWriter:
    """MatrixWriter<Ring> (lela/matrix/io.h:180-223)."""

    def __init__(self, F):
        self.F = F

    def write(self, path, A, format=FORMAT_PRETTY):
        """Writes A (dense array in the ring's storage, a Dense01Matrix over
        GF2, or SparseRows) with the text of MatrixWriter::write."""
        from . import GF2, Dense01Matrix
        lib = _lib.load()
        elem, p = _ring(self.F)
        if isinstance(A, SparseRows):
            rc = lib.mlz_matrix_write_csr(str(path).encode(), format, elem, p, A.m, A.n,
                                          np.ascontiguousarray(A.row_ptr, dtype=np.uint64).ctypes.data,
                                          np.ascontiguousarray(A.col, dtype=np.uint32).ctypes.data,
                                          np.ascontiguousarray(A.val, dtype=np.uint64).ctypes.data)
        elif isinstance(self.F, GF2):
            if not isinstance(A, Dense01Matrix):
                raise TypeError("GF(2) matrices are Dense01Matrix")
            W = np.ascontiguousarray(A.words, dtype=np.uint64)
            rc = lib.mlz_matrix_write_dense(str(path).encode(), format, elem, p, W.shape[0], A.ncols,
                                            W.ctypes.data, W.shape[1])
        else:
            X = np.ascontiguousarray(A, dtype=self.F.dtype)
            m, n = X.shape
            rc = lib.mlz_matrix_write_dense(str(path).encode(), format, elem, p, m, n, X.ctypes.data,
                                            max(1, n))
        _check(rc, "write")
        return path
