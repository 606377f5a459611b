/* include/mlz_io.h -- matrix files: C ABI of the readers and writers in
 * lela_amd/liblela_amd.so (host code only; no GPU is touched).
 *
 * Replaces LELA's MatrixReader / MatrixWriter (lela/matrix/io.h:36-223,
 * lela/matrix/io.tcc) as reached through BLAS3::read / BLAS3::write
 * (lela/blas/level3.h:191-205, lela/blas/level3-generic.h:340-355).  The
 * formats and their text are LELA's; the differences are listed per function.
 *
 * Entries: elem selects the ring as in mlz.h -- MLZ_U16 / MLZ_U32 /
 * MLZ_F64_SYM for Modular<uint16 / uint32 / double>(p), or MLZ_IO_GF2 for GF2.
 * Values cross this ABI as canonical residues in [0, p) (0 or 1 for GF(2)),
 * whatever the storage; the writers print MLZ_F64_SYM entries as their
 * symmetric representatives, as ModularTraits<double>::write does
 * (lela/ring/modular.h:305-306).
 *
 * Return codes are mlz.h's plus MLZ_E_FORMAT, MLZ_E_INPUT and MLZ_E_FILE;
 * mlz_io_last_error() describes the last failure of this thread.
 */
#ifndef MLZ_IO_H
#define MLZ_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LELA's FileFormatTag, same order (lela/matrix/io.h:36-41; PNG is not built). */
typedef enum {
	MLZ_FORMAT_DETECT = 0,
	MLZ_FORMAT_UNKNOWN = 1,
	MLZ_FORMAT_TURNER = 2,     /* "i j x" per nonzero, 0-based, then "-1" */
	MLZ_FORMAT_ONE_BASED = 3,  /* "i j x" per nonzero, 1-based (written only) */
	MLZ_FORMAT_DUMAS = 4,      /* SMS: "m n M", 1-based triples, "0 0 0" */
	MLZ_FORMAT_MAPLE = 5,      /* "[ [a, b], [c, d] ]" (written only) */
	MLZ_FORMAT_MATLAB = 6,     /* "[ a, b; c, d]" */
	MLZ_FORMAT_SAGE = 7,       /* "matrix([ [ a, b ], ... ])" (written only) */
	MLZ_FORMAT_PRETTY = 8      /* "  [ a . b ]" per row, '.' for zero */
} mlz_format;

#define MLZ_IO_GF2 4          /* elem value selecting GF2 for the functions below */

enum {
	MLZ_E_FORMAT = 7,          /* UnrecognisedFormat (lela/matrix/io.h:46) */
	MLZ_E_INPUT = 8,           /* InvalidMatrixInput (lela/matrix/io.h:51) */
	MLZ_E_FILE = 9             /* the file could not be opened, read or written */
};

typedef struct mlz_file_matrix mlz_file_matrix;

/* Format of the file at path: MLZ_FORMAT_DUMAS, _TURNER, _MAPLE, _MATLAB,
 * _SAGE, _PRETTY or _UNKNOWN, from its first non-blank line.  LELA's
 * MatrixReader::detectFormat (lela/matrix/io.tcc:79-187) tests the first 31
 * characters against regular expressions written with POSIX classes outside
 * brackets ("[:digit:]+" is the set {':', 'd', 'i', 'g', 't'}), so under
 * glibc it recognises none of its own formats and every FORMAT_DETECT read
 * throws UnrecognisedFormat; this function classifies by the shape the
 * writers produce instead (a superset of LELA's behaviour). */
int mlz_matrix_detect_format(const char *path, int *format);

/* Reads the matrix at path into a row-compressed host matrix owned by the
 * library (release with mlz_matrix_release).  Replaces MatrixReader::read
 * (lela/matrix/io.tcc:33-70) with readTurner, readDumas, readMatlab and
 * readPretty (:191-413); one-based, Maple and Sage input return
 * MLZ_E_UNSUPPORTED as LELA's NotImplemented readers do (:260-358), and
 * format MLZ_FORMAT_DETECT uses mlz_matrix_detect_format.  Entries are read
 * as Ring::read does (lela/ring/modular.h:405-413, lela/ring/gf2.h:110-111):
 * an integer of any length, taken as a signed 64-bit value the way GMP's
 * mpz_get_si does and reduced mod p; over GF(2) any nonzero integer is 1,
 * read as an int: a value outside int ends the read there as the failed
 * stream ends LELA's loops (Dumas, Turner; Matlab and pretty input return
 * MLZ_E_INPUT).
 * Zero entries are not stored and a repeated (i, j) keeps the last nonzero,
 * as setEntry does.  Dumas files carry their dimensions; for the other
 * formats LELA reads into a matrix the caller sized, which rows / cols give
 * here (entries outside them are MLZ_E_INPUT); rows = cols = 0 sizes the
 * matrix from the entries.  Malformed input is MLZ_E_INPUT (LELA's
 * InvalidMatrixInput; where LELA's pretty reader would loop forever on a
 * non-numeric token, this returns MLZ_E_INPUT). */
int mlz_matrix_read(const char *path, int format, int elem, uint64_t p, size_t rows, size_t cols,
		    mlz_file_matrix **out);
int mlz_matrix_shape(const mlz_file_matrix *A, size_t *m, size_t *n, size_t *nnz);
/* Copies the rows out: row_ptr [m+1], col [nnz] increasing in each row,
 * val [nnz] canonical residues. */
int mlz_matrix_get_csr(const mlz_file_matrix *A, uint64_t *row_ptr, uint32_t *col, uint64_t *val);
int mlz_matrix_release(mlz_file_matrix *A);

/* Writes an m x n matrix to path in format (not MLZ_FORMAT_DETECT / _UNKNOWN)
 * with the text of MatrixWriter::write (lela/matrix/io.tcc:443-651):
 * writeTurner, writeOneBased, writeDumas, writeMaple, writeMatlab, writeSage
 * and writePretty (column width Ring::elementWidth, lela/ring/modular.h:416-417,
 * lela/ring/gf2.h:120-121).  The _csr form takes canonical residues; the
 * _dense form takes A in the storage of elem (row-major, lda elements), or
 * for MLZ_IO_GF2 bit rows of uint64 words (lda in words, LSB first). */
int mlz_matrix_write_csr(const char *path, int format, int elem, uint64_t p, size_t m, size_t n,
			 const uint64_t *row_ptr, const uint32_t *col, const uint64_t *val);
int mlz_matrix_write_dense(const char *path, int format, int elem, uint64_t p, size_t m, size_t n,
			   const void *A, size_t lda);

/* Message of this thread's last failed call above (never NULL). */
const char *mlz_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MLZ_IO_H */
