/* include/mlz.h -- C ABI of the MI355X exact linear-algebra backend
 * (lela_amd/liblela_amd.so).
 *
 * This is the boundary a LELA module (HIPModule, include/lela_amd/hip-module.h) binds:
 * plain pointers and sizes, no C++ or HIP types.  Each entry point names the
 * LELA routine whose semantics it reproduces (paths relative to the LELA
 * source tree).
 *
 * Conventions
 *  - Matrices are row-major; leading dimensions count elements (words for GF(2)).
 *  - Z/p values are stored canonically in [0, p) for MLZ_U16 / MLZ_U32
 *    (ModularTraits<uint16/uint32>, lela/ring/modular.h:158-255) and in the
 *    symmetric range [-(p-1)/2, (p-1)/2] for MLZ_F64_SYM (ModularTraits<double>,
 *    lela/ring/modular.h:288-327).  Inputs must already be valid residues, as
 *    LELA assumes.
 *  - GF(2) rows are uint64 words, bit j of word w = column 64*w + j (LSB first,
 *    lela/vector/bit-iterator.h:249-311, lela/matrix/dense-zero-one.h:253-257).
 *  - Pointer mode: by default (MLZ_PTR_HOST) matrix pointers are host memory;
 *    the library stages them through device memory and returns when the result
 *    is back on the host.  With mlz_ctx_set_ptr_mode(ctx, MLZ_PTR_DEVICE) they
 *    are device pointers (from mlz_dev_alloc or any hipMalloc); calls are then
 *    queued on the context's stream and return without waiting unless they
 *    produce a host-side scalar (rank, pivots, error index).
 *  - Nothing is retained after a call returns; a context is not thread-safe
 *    (one per thread, like LELA's Context, lela/blas/context.h:62-71).
 *  - Return codes below.  The LELA-side shim converts them back into
 *    DiagonalEntryNotInvertible(idx), WrongMatrixForm(row) and LELAError.
 */
#ifndef MLZ_H
#define MLZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlz_ctx mlz_ctx;

/* Storage of Z/p elements. */
typedef enum {
	MLZ_U16 = 1,      /* Modular<uint16>, p < 2^16 */
	MLZ_U32 = 2,      /* Modular<uint32>, p < 2^32 */
	MLZ_F64_SYM = 3   /* Modular<double>, p < 2^26, symmetric representatives */
} mlz_elem;

/* gemm kernel families (all exact; they differ only in speed). */
enum {
	MLZ_ALGO_AUTO = 0,
	MLZ_ALGO_PLAIN = 1,  /* VALU uint64 multiply-accumulate, one reduction per entry */
	MLZ_ALGO_INT8 = 2,   /* 8-bit limbs on INT8 MFMA, int32 accumulation; p <= 2^16 */
	MLZ_ALGO_FP64 = 3    /* FP64 MFMA as exact integer carrier; p < 2^26 */
};

enum {
	MLZ_PTR_HOST = 0,
	MLZ_PTR_DEVICE = 1
};

/* Return codes. */
enum {
	MLZ_OK = 0,
	MLZ_E_ARG = 1,             /* bad argument (dimensions, modulus, null pointer) */
	MLZ_E_HIP = 2,             /* HIP runtime error; see mlz_last_error */
	MLZ_E_NOT_INVERTIBLE = 3,  /* DiagonalEntryNotInvertible / pivot not invertible */
	MLZ_E_WRONG_FORM = 4,      /* FaugereLachartre WrongMatrixForm */
	MLZ_E_NOMEM = 5,           /* device allocation failed */
	MLZ_E_UNSUPPORTED = 6      /* algorithm not available for this modulus/element */
};

/* ---- context ---------------------------------------------------------- */

/* One context per LELA Context: owns a HIP stream, device workspace and the
 * profiling state.  device = HIP device ordinal.  config may be NULL. */
int mlz_ctx_create(mlz_ctx **out, int device, const char *config);
int mlz_ctx_destroy(mlz_ctx *ctx);
int mlz_ctx_set_ptr_mode(mlz_ctx *ctx, int mode);
/* Tuning knobs (results never depend on them):
 *   "gemm.i8.variant"  INT8 gemm workgroup tile: 0 by shape, 1 = 128x128,
 *                      2 = 128x256, 3 = 256x128 (4 waves each), 4 = 128x128
 *                      with two workgroups per CU, 5 = 128x256 and
 *                      6 = 256x128 with 8 waves (two per SIMD); 7, 8 = 5
 *                      staged through 3 / 4 LDS buffers by direct loads,
 *                      9 = 7 with 64 k per step, 10 = 4 through 4 buffers,
 *                      11 = 64x64 (4 waves of 32x32) through 4 buffers.
 *   "rref.search.threads" threads of the echelon form's fused pivot search
 *                      and inverse (one workgroup per panel): 256, 512 or 1024.
 *   "rref.update.blocks" 32-column blocks per wave of the echelon form's
 *                      trailing update on INT8 limbs: 1 (default), 2 or 4.
 *   "update.i8.lds"    that update's full 32x32 tiles staged in LDS by direct
 *                      loads, 1, 2 or 4 tiles per wave; 0 = registers only;
 *                      -1 (default) = 2 when rows start on 128-byte
 *                      boundaries, else 1.
 *   "gf2.gemm.kernel"  GF(2) gemm: 0 = Four Russians tables in LDS (default),
 *                      1 = the plain AND/XOR-popcount dot form. */
int mlz_ctx_set_option(mlz_ctx *ctx, const char *key, int64_t value);
int mlz_ctx_synchronize(mlz_ctx *ctx);
/* Message for the last error on this context (static storage, never NULL). */
const char *mlz_last_error(mlz_ctx *ctx);
/* Library version string. */
const char *mlz_version(void);

/* ---- device memory helpers (for callers that keep data resident) ------- */
int mlz_dev_alloc(mlz_ctx *ctx, size_t bytes, void **ptr);
int mlz_dev_free(mlz_ctx *ctx, void *ptr);
int mlz_memcpy_htod(mlz_ctx *ctx, void *dst, const void *src, size_t bytes);
int mlz_memcpy_dtoh(mlz_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy on the context's stream (waits for completion). */
int mlz_memcpy_dtod(mlz_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Release cached workspace buffers. */
int mlz_ctx_trim(mlz_ctx *ctx);

/* ---- profiling: HIP events around every kernel launch, on the context's
 *      stream, accumulated per kernel name. ---------------------------------- */
int mlz_prof_enable(mlz_ctx *ctx, int on);
int mlz_prof_reset(mlz_ctx *ctx);
/* Synchronises, then returns launch count and summed milliseconds for the
 * kernels whose name contains `name`. */
int mlz_prof_query(mlz_ctx *ctx, const char *name, uint64_t *count, double *ms);
/* Events on the context's stream for timing a region. */
int mlz_timer_start(mlz_ctx *ctx);
int mlz_timer_stop(mlz_ctx *ctx, double *ms);

/* ---- Z/p level 3 -------------------------------------------------------- */

/* C <- a*A*B + b*C mod p; A m x k, B k x n, C m x n.
 * Replaces BLAS3::gemm (lela/blas/level3.h:81-83) for Modular<T> through
 * _gemm<Modular<T>, ZpModule<T>::Tag> -> GenericModule gemm_impl
 * (lela/blas/level3-generic.tcc:285-322) with the ZpModule dots
 * (lela/blas/level1-modular.tcc:25-63,135-158) and, for Modular<double>, the
 * BLAS-carrier gemm (lela/blas/level3-modular.tcc:25-123). */
int mlz_zp_gemm(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, size_t k,
		uint64_t a, const void *A, size_t lda, const void *B, size_t ldb,
		uint64_t b, void *C, size_t ldc, int algo);

/* The same product with no device allocation: for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in LELA.  A, B and C are device
 * pointers whatever the context's pointer mode; the call queues on the
 * context's stream and returns.  The INT8 and FP64 families pack their
 * operands into a static device array of MLZ_NOALLOC_SCRATCH_MB MiB that each
 * device holds until the process ends (the plain family needs none); a call
 * whose packed operands exceed it returns MLZ_E_ARG before any launch (INT8:
 * (m' + n') * k' * 2 bytes with m', n' padded to the 128- or 256-row tile and
 * k' = min(k, 65504) rounded up to 64; FP64: (m' * 32 + n' * 64) * ceil(k'/4)
 * bytes, m' = m rounded to 64, n' = n to 128, k' = min(k, the exact FP64
 * k-chunk, 65552 at p = 67108859)).  Every call on one device shares
 * that array: calls that overlap on a device (two contexts or threads) race.
 * No profiling events are recorded for it. */
#define MLZ_NOALLOC_SCRATCH_MB 64
int mlz_zp_gemm_noalloc(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, size_t k,
			uint64_t a, const void *A, size_t lda, const void *B, size_t ldb,
			uint64_t b, void *C, size_t ldc, int algo);

/* B <- a * A^-1 * B with A n x n triangular (upper != 0: upper), B n x nrhs.
 * Replaces BLAS3::trsm (lela/blas/level3.h:120-122) ->
 * _trsm<Ring, GenericModule::Tag>::trsm_impl (lela/blas/level3-generic.tcc:421-498).
 * Non-invertible diagonal: returns MLZ_E_NOT_INVERTIBLE, *bad_diag = the index
 * LELA's recursion reports (smallest for lower, largest for upper). */
int mlz_zp_trsm(mlz_ctx *ctx, mlz_elem elem, uint64_t p, int upper, int unit_diag,
		size_t n, size_t nrhs, uint64_t a, const void *A, size_t lda,
		void *B, size_t ldb, size_t *bad_diag, int algo);

/* Reduced row-echelon form of A (m x n) in place, standard layout: rows
 * 0..rank-1 hold the pivot rows (pivot 1, increasing pivot columns, zeros in
 * every other row of a pivot column), rows rank..m-1 are zero.  pivots[0..rank)
 * receives the pivot columns (array of min(m, n) entries, may be NULL).
 * Replaces EchelonForm::echelonize(A, reduced=true) (lela/solutions/echelon-form.h:103-154)
 * -> Elimination::echelonize_reduced (lela/algorithms/elimination.tcc:141-264) /
 * GaussJordan::echelonize_reduced (lela/algorithms/gauss-jordan.tcc:413-466).
 * algo selects the gemm family used for the block updates.
 * MLZ_E_NOT_INVERTIBLE: a pivot has no inverse (composite modulus). */
int mlz_zp_rref(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, void *A, size_t lda,
		size_t *rank, uint32_t *pivots, int algo);

/* Non-reduced row-echelon forms of A (m x n) in place, with LELA's
 * transposition pivoting: step i takes the first column (from the previous
 * pivot's on) in which a row at a position >= i is nonzero, the first such
 * row as pivot, swaps it with position i and clears the column below
 * (DensePivotStrategy, lela/algorithms/pivot-strategy.tcc:19-41).
 *   MLZ_ECH_U:         Elimination::echelonize (A, P, rank, det, compute_L =
 *                      false) (lela/algorithms/elimination.tcc:47-139): A
 *                      becomes U, rows past the rank zero.  This is also
 *                      EchelonForm::echelonize (A, false) for both its methods
 *                      (echelon-form.h:113-131; tests/test-gauss-jordan.C:96-99).
 *   MLZ_ECH_TRANSFORM: the same with compute_L = true (the method's default):
 *                      left of the diagonal, A[t][s] = T[t][s] with T P A = U
 *                      (T unit lower), as LELA leaves it (:121-122).
 *   MLZ_ECH_PLUQ:      Elimination::pluq (:266-370): P A Q = L U with L's
 *                      entries left of the diagonal in the first rank columns
 *                      and U from the diagonal on, columns in Q's order.
 * perm_rows[s] (s < rank, may be NULL): the position swapped with s at step s
 * (s itself when there was no swap), so LELA's P is the list of
 * (s, perm_rows[s]) with perm_rows[s] != s; perm_cols likewise for pluq's Q.
 * Both in step order (pluq's own lists are these reversed, :357-358).
 * det: the product of the pivots (:100).  MLZ_E_NOT_INVERTIBLE: a pivot has
 * no inverse (composite p; LELA throws LELAError there, :103-104). */
enum { MLZ_ECH_U = 0, MLZ_ECH_TRANSFORM = 1, MLZ_ECH_PLUQ = 2 };
int mlz_zp_echelonize(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, void *A, size_t lda,
		      int mode, size_t *rank, uint64_t *det, uint32_t *perm_rows, uint32_t *perm_cols, int algo);

/* The same three forms with SparsePartialPivotStrategy, the pivot rule LELA's
 * Elimination uses for matrices with sparse, sparse 0-1 and hybrid 0-1 rows
 * (DefaultPivotStrategy, lela/algorithms/pivot-strategy.h:150-160;
 * pivot-strategy.tcc:69-188): at step i, among the rows at positions >= i,
 * those whose first nonzero at or after the search column (the previous
 * pivot's column + 1; i for pluq) is leftmost, and of those the one with the
 * fewest nonzeros from that column on -- count_words != 0: the fewest nonzero
 * 64-column words from the pivot's word on, the hybrid-row count -- the first
 * such row on ties.  Each step then eliminates as mlz_zp_echelonize's modes
 * do (:112-123, :330-350).  A is dense here (LELA's sparse rows unpacked;
 * GF(2) rows as Z/2 entries); outputs as mlz_zp_echelonize.  m, n < 2^21. */
int mlz_zp_echelonize_sparse(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, void *A, size_t lda,
			     int mode, int count_words, size_t *rank, uint64_t *det, uint32_t *perm_rows,
			     uint32_t *perm_cols);

/* Faugere-Lachartre reduction of an F4-shaped matrix X (m x n, CSR: row_ptr
 * [m+1], col [nnz] strictly increasing within each row, val [nnz] in the
 * element storage) to its reduced row-echelon form R (m x n dense, ldr; may be
 * NULL to compute only rank/pivots).  pivots receives rank column indices
 * (array of min(m, n) entries, may be NULL).  det (may be NULL, which skips
 * its work) receives LELA's determinant output: the product of the heads of the
 * scanned rows of X (lela/algorithms/faugere-lachartre.tcc:119) and of the rows
 * of the non-reduced echelon form of D - C A^-1 B (:822, :833).
 * Replaces FaugereLachartre<Modular<T>>::echelonize(R, X, rank, det, reduced=true)
 * (lela/algorithms/faugere-lachartre.tcc:735-917).
 * MLZ_E_WRONG_FORM: a head decreases at row *bad_row (WrongMatrixForm, :80-81).
 * MLZ_E_NOT_INVERTIBLE: a head has no inverse mod p (LELA then returns with R
 * unchanged, :790-799); *rank is the number of pivot rows of X. */
int mlz_zp_fl_echelonize(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n,
			 const uint64_t *row_ptr, const uint32_t *col, const void *val, void *R,
			 size_t ldr, size_t *rank, uint32_t *pivots, uint64_t *det, size_t *bad_row,
			 int algo);

/* The same reduction with a compact result: pivots (rank column indices)
 * and NP, the rank rows of R on the n - rank non-pivot columns in increasing
 * column order, row-major with row stride n - rank.  R is the identity on the
 * pivot columns, so NP and the pivots determine it; NP is the block LELA's
 * splicer assembles R from (lela/algorithms/faugere-lachartre.tcc:363-504,
 * MatrixGrid3) and, for sparse rows, the nonzeros R holds.  On the
 * 20000 x 30000 F4-shaped matrix (rank 20000) it is a third of the dense R.  np_capacity
 * counts elements; a rank * (n - rank) above it returns MLZ_E_ARG with *rank
 * set (min(m, n) * n always suffices).  Other arguments, outputs and errors
 * as mlz_zp_fl_echelonize. */
int mlz_zp_fl_echelonize_compact(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n,
				 const uint64_t *row_ptr, const uint32_t *col, const void *val, void *NP,
				 size_t np_capacity, size_t *rank, uint32_t *pivots, uint64_t *det,
				 size_t *bad_row, int algo);

/* ---- GF(2) ------------------------------------------------------------- */

/* C <- a*A*B + b*C over GF(2) (a, b in {0, 1}); A m x k, B k x n, C m x n bits,
 * leading dimensions in 64-bit words.  Bits of C beyond column n are left as
 * they are; bits of A beyond k and of B beyond n are ignored.
 * Replaces BLAS3::gemm for GF2: _gemm<GF2, M4RIModule::Tag> -> mzd_mul /
 * mzd_addmul (lela/blas/level3-m4ri.h:89-108, lela/blas/level3-m4ri.tcc:61-86)
 * and the no-M4RI GenericModule path (lela/blas/level3-generic.tcc:243-263,
 * lela/blas/level2-gf2.tcc:28-55). */
int mlz_gf2_gemm(mlz_ctx *ctx, size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda_w,
		 const uint64_t *B, size_t ldb_w, int b, uint64_t *C, size_t ldc_w);

/* Faugere-Lachartre reduction over GF(2) of an F4-shaped matrix X (m x n
 * bits) given as hybrid rows (lela/vector/hybrid.h:30-33): row i holds the
 * pairs (blk[e], words[e]) for e in [row_ptr[i], row_ptr[i+1]), word indices
 * strictly increasing (LELA's are uint16, these uint32), bit j of a word =
 * column 64*blk + j.  R (m x n bits, ldr_w words; may be NULL for rank and
 * pivots only) receives the reduced row-echelon form; bits of R past column n
 * keep their values.  pivots: rank column indices (min(m, n) entries, may be
 * NULL).  Every head is 1 over GF(2), so LELA's det is 1 and is not returned.
 * Replaces FaugereLachartre<GF2>::echelonize(R, X, rank, det, reduced=true)
 * (lela/algorithms/faugere-lachartre.tcc:735-917) on
 * SparseMatrix<bool, HybridVector<..>> (tests/test-faugere-lachartre.C:278-285).
 * MLZ_E_WRONG_FORM: a head decreases at row *bad_row (WrongMatrixForm, :80-81). */
int mlz_gf2_fl_echelonize(mlz_ctx *ctx, size_t m, size_t n, const uint64_t *row_ptr, const uint32_t *blk,
			  const uint64_t *words, uint64_t *R, size_t ldr_w, size_t *rank, uint32_t *pivots,
			  size_t *bad_row);

/* B <- a*A^-1*B over GF(2) (a in {0, 1}); A n x n triangular (upper != 0:
 * upper), B n x nrhs bits, leading dimensions in words; bits of B beyond
 * column nrhs keep their values.  unit_diag != 0 takes A's diagonal as ones
 * (what M4RI's mzd_trsm_upper_left / _lower_left do); otherwise a zero on the
 * diagonal returns MLZ_E_NOT_INVERTIBLE with *bad_diag = the index LELA's
 * recursion reports (smallest for lower, largest for upper) and B unchanged.
 * Replaces BLAS3::trsm for GF2: _trsm<GF2, GenericModule::Tag>
 * (lela/blas/level3-generic.tcc:421-498) and _trsm<GF2, M4RIModule::Tag>
 * (lela/blas/level3-m4ri.tcc:88-109). */
int mlz_gf2_trsm(mlz_ctx *ctx, int upper, int unit_diag, size_t n, size_t nrhs, int a, const uint64_t *A,
		 size_t lda_w, uint64_t *B, size_t ldb_w, size_t *bad_diag);

/* Reduced row-echelon form over GF(2) in place (standard layout as
 * mlz_zp_rref; bits beyond column n are written as zero).
 * Replaces EchelonForm<GF2>::echelonize(A, reduced=true)
 * (lela/solutions/echelon-form-gf2.h:97-134: mzd_echelonize with M4RI,
 * GaussJordan / Elimination without it). */
int mlz_gf2_rref(mlz_ctx *ctx, size_t m, size_t n, uint64_t *A, size_t lda_w, size_t *rank,
		 uint32_t *pivots);

/* ---- batches of independent Z/p instances ------------------------------ */
/* Many small gemm and echelon-form instances of uneven shapes (the blocks a
 * Faugere-Lachartre splitting produces, or independent solves) run together
 * in persistent work-list launches, optionally captured once as a hipGraph
 * and replayed.  Each instance runs one of several exact kernel variants;
 * the choice changes speed only.  LELA has no batch interface: a caller makes
 * one BLAS3::gemm / EchelonForm::echelonize call per instance
 * (lela/blas/level3.h:81-83, lela/solutions/echelon-form.h:103-154); the
 * per-instance semantics here are those of mlz_zp_gemm and mlz_zp_rref.
 * Matrices are device pointers (data resident in HBM). */
enum { MLZ_TASK_GEMM = 1, MLZ_TASK_RREF = 2 };
#define MLZ_MAX_VARIANTS 4
/* gemm variants: 0 plain VALU 64x64, 1 sparse-A VALU 16x256, 2 INT8 MFMA
 * 64x64, 3 INT8 MFMA 128x128 (2 and 3 need p <= 2^16 and k <= 65504).
 * rref variants (m <= 4096, one workgroup per instance): 0 panels of 64
 * columns, 1 panels of 32 columns (VALU updates); 2 and 3 the same with the
 * trailing updates on INT8 MFMA (p <= 2^16). */
typedef struct mlz_task {
	int op;                /* MLZ_TASK_GEMM or MLZ_TASK_RREF */
	int variant;           /* in: -1 = chosen by the policy, else forced; out: the variant run */
	uint32_t m, n, k;      /* gemm: A m x k, B k x n, C m x n; rref: A m x n (k unused) */
	float density;         /* caller's estimate of the nonzero fraction (cost-model input) */
	uint64_t a, b;         /* gemm: C <- a*A*B + b*C */
	void *A;               /* gemm: A; rref: the matrix, reduced in place */
	size_t lda;
	const void *B;
	size_t ldb;
	void *C;
	size_t ldc;
	uint32_t *pivots;      /* rref: device array of min(m, n) pivot columns, may be NULL */
	size_t rank;           /* rref: out */
} mlz_task;

enum {
	MLZ_POLICY_FIXED = 0,  /* every task names its variant */
	MLZ_POLICY_TABLE = 1,  /* autotuner table (built by mlz_tune); a shape rule where it is empty */
	MLZ_POLICY_MODEL = 2   /* cost model (mlz_model_train); the table for an op without a model */
};
enum {
	MLZ_BATCH_GRAPH = 1,   /* capture the launches once as a hipGraph; mlz_batch_launch replays it */
	MLZ_BATCH_TIMING = 2,  /* record each task's kernel time (see mlz_batch_finish) */
	MLZ_BATCH_SERIAL = 4   /* one stream: kernels of different variants do not overlap */
};

typedef struct mlz_batch mlz_batch;

/* Chooses a variant per task (tasks[i].variant is set on return), builds the
 * work lists and device workspace, and with MLZ_BATCH_GRAPH captures the
 * launch sequence.  The task array must stay unchanged while the plan lives. */
int mlz_batch_plan(mlz_ctx *ctx, mlz_elem elem, uint64_t p, mlz_task *tasks, size_t ntasks, int policy,
		   int flags, mlz_batch **plan);
/* Queues one execution of the plan on the context's stream. */
int mlz_batch_launch(mlz_ctx *ctx, mlz_batch *plan);
/* Waits for the last execution; sets tasks[i].rank for echelon-form tasks and,
 * with MLZ_BATCH_TIMING and cost_us != NULL, cost_us[i] = the task's summed
 * workgroup time divided by the resident workgroups per CU of its kernels
 * (microseconds of one CU).  MLZ_E_NOT_INVERTIBLE: some pivot had no inverse
 * (composite p); *bad_task (may be NULL) = the first such task. */
int mlz_batch_finish(mlz_ctx *ctx, mlz_batch *plan, mlz_task *tasks, float *cost_us, size_t *bad_task);
int mlz_batch_destroy(mlz_ctx *ctx, mlz_batch *plan);
/* plan + launch + finish + destroy. */
int mlz_batch_run(mlz_ctx *ctx, mlz_elem elem, uint64_t p, mlz_task *tasks, size_t ntasks, int policy);

/* Autotuner: runs the tasks once per variant (serial, timed, reps times,
 * median kept) and writes cost_us[i * MLZ_MAX_VARIANTS + v] (NaN where v does
 * not apply); the context's table then maps each shape bucket
 * (floor(log2 m), floor(log2 n), floor(log2 k)) and op to the variant of
 * least total cost over the tasks in it.  The tasks' matrices are
 * overwritten (gemm C, rref A): pass scratch copies. */
int mlz_tune(mlz_ctx *ctx, mlz_elem elem, uint64_t p, mlz_task *tasks, size_t ntasks, int reps, float *cost_us);
/* Table entries (op 1 or 2): 16 x 16 x 16 bytes, index (lm*16 + ln)*16 + lk,
 * -1 = empty. */
int mlz_tune_table_get(mlz_ctx *ctx, int op, int8_t *table);
int mlz_tune_table_set(mlz_ctx *ctx, int op, const int8_t *table);

/* Cost model: per op an MLP 5 -> 32 -> 32 -> MLZ_MAX_VARIANTS (ReLU, FP32)
 * that predicts log2 of each variant's cost from the task features below.
 * Weights: W1[5][32] b1[32] W2[32][32] b2[32] W3[32][4] b3[4] (1380 floats). */
#define MLZ_MODEL_PARAMS 1380
/* Features of a task: (log2 m - 7.5)/1.5, (log2 n - 7.5)/1.5, (log2 k' - 7.5)/1.5
 * with k' = k for gemm and min(m, n) for rref, (log2 max(density, 2^-16) + 8)/8,
 * (log2 p - 16)/8. */
int mlz_task_features(const mlz_task *task, uint64_t p, float *x5);
/* Trains the op's model on the GPU (full-batch Adam, `steps` steps, learning
 * rate lr, weights initialised from seed) on x [n][5], y [n][MLZ_MAX_VARIANTS]
 * (log2 microseconds) and mask [n][MLZ_MAX_VARIANTS] (1 = measured); the loss
 * is the masked mean squared error.  *loss (may be NULL) = final loss. */
int mlz_model_train(mlz_ctx *ctx, int op, const float *x, const float *y, const uint8_t *mask, size_t n,
		    int steps, float lr, uint64_t seed, float *loss);
int mlz_model_get(mlz_ctx *ctx, int op, float *params);
int mlz_model_set(mlz_ctx *ctx, int op, const float *params);
/* Predictions y [n][MLZ_MAX_VARIANTS] of the op's model (GPU inference kernel). */
int mlz_model_predict(mlz_ctx *ctx, int op, const float *x, size_t n, float *y);
/* Loss and its gradient (MLZ_MODEL_PARAMS floats) at params on (x, y, mask),
 * computed by the training kernels. */
int mlz_model_grad(mlz_ctx *ctx, const float *params, const float *x, const float *y, const uint8_t *mask,
		   size_t n, float *grad, float *loss);

#ifdef __cplusplus
}
#endif

#endif /* MLZ_H */
