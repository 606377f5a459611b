/* oracle/oracle.h -- CPU restatement of LELA's exact level-3 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * backend (lela_amd/liblela_amd.so); only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product never links it.
 *
 * Every function restates the algorithm of the LELA routine it cites
 * (paths relative to the LELA source tree).  Storage is canonical: Z/p
 * values are uint32 in [0, p); GF(2) matrices are rows of uint64 words,
 * bit j of word w holding column 64*w + j (LSB first,
 * lela/vector/bit-iterator.h:249-311, lela/matrix/dense-zero-one.h:253-257).
 *
 * Pinning (what checks each function beyond retyping its formula) is listed
 * per function in tests/test_oracle.py; nothing here is "parity unpinned".
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	ORC_OK = 0,
	ORC_E_ARG = 1,
	ORC_E_NOT_INVERTIBLE = 3,  /* DiagonalEntryNotInvertible / LELAError("Could not invert pivot") */
	ORC_E_WRONG_FORM = 4       /* FaugereLachartre WrongMatrixForm */
};

/* C <- a*A*B + b*C over Z/p, uint32 storage (Modular<uint32> + ZpModule<uint32>).
 * Per-entry dot with the 2^64-wrap correction of lela/blas/level1-modular.tcc:135-158,
 * C update of lela/blas/level3-generic.tcc:301-319. */
int orc_zp_gemm_u32(uint32_t p, size_t m, size_t n, size_t k, uint32_t a,
		    const uint32_t *A, size_t lda, const uint32_t *B, size_t ldb,
		    uint32_t b, uint32_t *C, size_t ldc);

/* Same product for Modular<uint16> + ZpModule<uint16>: k-blocked delayed
 * reduction with block_size = floor((2^64-1)/(p-1)^2)
 * (lela/ring/modular.h:578-580, lela/blas/level1-modular.tcc:25-63). */
int orc_zp_gemm_u16(uint16_t p, size_t m, size_t n, size_t k, uint16_t a,
		    const uint16_t *A, size_t lda, const uint16_t *B, size_t ldb,
		    uint16_t b, uint16_t *C, size_t ldc);

/* B <- a * A^-1 * B, A n x n triangular, B n x nrhs (lela/blas/level3-generic.tcc:421-498).
 * On a non-invertible diagonal returns ORC_E_NOT_INVERTIBLE with *bad_diag set to
 * the index LELA's recursion reports (B is then partially updated, as in LELA). */
int orc_zp_trsm(uint32_t p, int upper, int unit_diag, size_t n, size_t nrhs, uint32_t a,
		const uint32_t *A, size_t lda, uint32_t *B, size_t ldb, size_t *bad_diag);

/* Elimination::echelonize / echelonize_reduced with DensePivotStrategy
 * (lela/algorithms/elimination.tcc:47-264, lela/algorithms/pivot-strategy.tcc:19-41).
 * pivots[0..rank) receives the pivot columns.  det as LELA accumulates it. */
int orc_zp_echelonize(uint32_t p, size_t m, size_t n, uint32_t *A, size_t lda, int reduced,
		      size_t *rank, uint32_t *pivots, uint32_t *det);

/* GF(2): C <- a*A*B + b*C (a, b in {0,1}); row/dot form of
 * lela/blas/level3-generic.tcc:243-263 + lela/blas/level2-gf2.tcc:28-55 +
 * lela/blas/level1-gf2.tcc:26-52.  Leading dimensions in words. */
int orc_gf2_gemm(size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda_w,
		 const uint64_t *B, size_t ldb_w, int b, uint64_t *C, size_t ldc_w);

/* GF(2) Elimination::echelonize[_reduced] with the Dense01 pivot search
 * (lela/algorithms/pivot-strategy.tcc:44-67) and word-XOR axpy
 * (lela/blas/level1-gf2.tcc:282-299). */
/* Elimination::echelonize with the transposition list and compute_L
 * (elimination.tcc:47-139); perm_rows[i] = pivot row of step i. */
int orc_zp_elim(uint32_t p, size_t m, size_t n, uint32_t *A, size_t lda, int compute_L, size_t *rank,
		uint32_t *det, uint32_t *perm_rows);
/* BLAS3::trsm over GF(2), B <- a A^-1 B (lela/blas/level3-generic.tcc:421-498
 * with Dense01Matrix's 64-column alignment).  unit_diag takes the diagonal as
 * ones; otherwise a zero diagonal returns ORC_E_NOT_INVERTIBLE with *bad_diag
 * the index the recursion reports, and B unchanged. */
int orc_gf2_trsm(int upper, int unit_diag, size_t n, size_t nrhs, int a, const uint64_t *A, size_t lda_w,
		 uint64_t *B, size_t ldb_w, size_t *bad_diag);

/* Elimination::pluq (elimination.tcc:266-370); per-step row and column swaps. */
int orc_zp_pluq(uint32_t p, size_t m, size_t n, uint32_t *A, size_t lda, size_t *rank, uint32_t *det,
		uint32_t *perm_rows, uint32_t *perm_cols);
/* Elimination::echelonize (mode 0: compute_L = false, 1: compute_L = true)
 * and Elimination::pluq (mode 2) with SparsePartialPivotStrategy
 * (lela/algorithms/pivot-strategy.tcc:69-188; words != 0 counts nonzero
 * 64-column words as for hybrid rows).  perm_rows / perm_cols as orc_zp_elim
 * and orc_zp_pluq. */
int orc_zp_elim_sparse(uint32_t p, size_t m, size_t n, uint32_t *A, size_t lda, int mode, int words,
		       size_t *rank, uint32_t *det, uint32_t *perm_rows, uint32_t *perm_cols);
int orc_gf2_echelonize(size_t m, size_t n, uint64_t *A, size_t lda_w, int reduced,
		       size_t *rank, uint32_t *pivots);

/* FaugereLachartre<Modular<uint32>>::echelonize(R, X, rank, det, reduced=true)
 * (lela/algorithms/faugere-lachartre.tcc:47-146,735-917) on dense storage.
 * R (m x n, ldr) receives the reduced row-echelon form.  On WrongMatrixForm
 * returns ORC_E_WRONG_FORM with *bad_row set.  det follows :119 for the X scan
 * and :833 for the scan of D's non-reduced echelon form, which is
 * orc_zp_elim's (LELA's transposition pivoting). */
int orc_fl_echelonize(uint32_t p, size_t m, size_t n, const uint32_t *X, size_t ldx,
		      uint32_t *R, size_t ldr, size_t *rank, uint32_t *det, size_t *bad_row);

/* Modular inverse by the extended Euclidean algorithm (lela/ring/modular.h:460-474, :536-557).
 * Returns 0 when gcd(y, p) != 1. */
int orc_inv(uint32_t p, uint32_t y, uint32_t *x);

/* ---- batch cost model (no LELA counterpart: restated from its definition) ---
 * MLP 5 -> 32 -> 32 -> 4 with ReLU; params laid out W1[5][32] b1[32]
 * W2[32][32] b2[32] W3[32][4] b3[4] (1380 floats).  Arithmetic in double.
 * Loss: masked mean squared error (1/M) sum mask*(y - t)^2, M = sum mask. */
#define ORC_MLP_PARAMS 1380
int orc_mlp_forward(const float *params, const float *x, size_t n, double *y);
int orc_mlp_loss_grad(const float *params, const float *x, const float *t, const uint8_t *mask, size_t n,
		      double *grad, double *loss);
/* `steps` full-batch Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) from params
 * (updated in place, kept in double inside); loss_hist[s] = loss before step s. */
int orc_mlp_adam(float *params, const float *x, const float *t, const uint8_t *mask, size_t n, int steps,
		 double lr, double *loss_hist);

#ifdef __cplusplus
}
#endif

#endif
