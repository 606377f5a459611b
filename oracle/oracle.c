/* oracle/oracle.c -- CPU restatement of LELA's exact level-3 path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP backend, never part of
 * the product (see oracle.h).  Citations are to the LELA source tree.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Modular<uint32> scalar arithmetic (lela/ring/modular.h:419-559).    */
/* ------------------------------------------------------------------ */

static inline uint32_t zp_mul(uint32_t p, uint32_t x, uint32_t y)
{
	/* Modular::mul: FatElement = uint64, reduce with % (modular.h:498-503). */
	return (uint32_t) (((uint64_t) x * (uint64_t) y) % p);
}

static inline uint32_t zp_axpy(uint32_t p, uint32_t a, uint32_t x, uint32_t y)
{
	/* Modular::axpy / axpyin: t = a*x + y in uint64, then % (modular.h:527-532).
	 * a, x, y < p < 2^32 so a*x + y < 2^64. */
	return (uint32_t) (((uint64_t) a * (uint64_t) x + (uint64_t) y) % p);
}

static inline uint32_t zp_neg(uint32_t p, uint32_t x)
{
	/* ModularTraits<uint32>::neg: v = p - v, then shift_down. */
	uint32_t v = p - x;
	return v >= p ? v - p : v;
}

int orc_inv(uint32_t p, uint32_t y, uint32_t *x)
{
	/* Extended Euclid on (y, p) as in Modular::eea (modular.h:540-557),
	 * iterative and in int64 so that p up to 2^32 cannot overflow. */
	int64_t r0 = y % p, r1 = p, s0 = 1, s1 = 0;
	while (r1 != 0) {
		int64_t q = r0 / r1, t;
		t = r0 - q * r1; r0 = r1; r1 = t;
		t = s0 - q * s1; s0 = s1; s1 = t;
	}
	if (r0 != 1)
		return 0;
	s0 %= (int64_t) p;
	if (s0 < 0)
		s0 += p;
	*x = (uint32_t) s0;
	return 1;
}

/* ------------------------------------------------------------------ */
/* gemm                                                                */
/* ------------------------------------------------------------------ */

/* 2^64 mod p, computed as ZpModule<uint32>::init_two_64 does (modular.h:616-625). */
static uint32_t two_64_mod(uint32_t p)
{
	uint64_t t = 2;
	for (int i = 0; i < 6; ++i)
		t = (t * t) % p;
	return (uint32_t) t;
}

int orc_zp_gemm_u32(uint32_t p, size_t m, size_t n, size_t k, uint32_t a,
		    const uint32_t *A, size_t lda, const uint32_t *B, size_t ldb,
		    uint32_t b, uint32_t *C, size_t ldc)
{
	if (p < 2)
		return ORC_E_ARG;
	const uint64_t two_64 = two_64_mod(p);
	for (size_t i = 0; i < m; ++i) {
		const uint32_t *Ai = A + i * lda;
		for (size_t j = 0; j < n; ++j) {
			/* _dot<Modular<uint32>,ZpModule<uint32>::Tag> dense/dense
			 * (level1-modular.tcc:135-158): uint64 sum, add 2^64 mod p on wrap. */
			uint64_t s = 0, t;
			const uint32_t *Bj = B + j;
			for (size_t l = 0; l < k; ++l) {
				t = (uint64_t) Ai[l] * (uint64_t) Bj[l * ldb];
				s += t;
				if (s < t)
					s += two_64;
			}
			uint32_t d = (uint32_t) (s % p);
			/* level3-generic.tcc:309-311: cij *= b; cij += a*d. */
			uint32_t cij = C[i * ldc + j];
			cij = zp_mul(p, cij, b);
			cij = zp_axpy(p, a, d, cij);
			C[i * ldc + j] = cij;
		}
	}
	return ORC_OK;
}

int orc_zp_gemm_u16(uint16_t p, size_t m, size_t n, size_t k, uint16_t a,
		    const uint16_t *A, size_t lda, const uint16_t *B, size_t ldb,
		    uint16_t b, uint16_t *C, size_t ldc)
{
	if (p < 2)
		return ORC_E_ARG;
	/* ZpModule<uint16>::block_size (modular.h:578-580); DoubleFatElement = uint64. */
	const uint64_t pm1 = (uint64_t) p - 1;
	size_t block = (pm1 == 0) ? 0 : (size_t) (UINT64_MAX / (pm1 * pm1));
	for (size_t i = 0; i < m; ++i) {
		const uint16_t *Ai = A + i * lda;
		for (size_t j = 0; j < n; ++j) {
			/* level1-modular.tcc:25-63: the first (k mod block) products, reduce,
			 * then whole blocks, reducing after each. */
			size_t bs = block == 0 ? k : block;
			size_t first = bs ? k % bs : 0;
			uint64_t t = 0, s;
			size_t l = 0;
			for (; l < first; ++l)
				t += (uint64_t) Ai[l] * (uint64_t) B[l * ldb + j];
			t %= p;
			while (l < k) {
				s = 0;
				for (size_t e = l + bs; l < e; ++l)
					s += (uint64_t) Ai[l] * (uint64_t) B[l * ldb + j];
				t += s;
				t %= p;
			}
			/* Modular<uint16>: FatElement uint32 (modular.h:198-202). */
			uint32_t cij = C[i * ldc + j];
			cij = (uint32_t) (((uint32_t) cij * (uint32_t) b) % p);
			cij = (uint32_t) (((uint32_t) a * (uint32_t) t + cij) % p);
			C[i * ldc + j] = (uint16_t) cij;
		}
	}
	return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* trsm (lela/blas/level3-generic.tcc:421-498)                         */
/* ------------------------------------------------------------------ */

static void scal_rows(uint32_t p, uint32_t a, uint32_t *B, size_t ldb, size_t rows, size_t cols)
{
	for (size_t i = 0; i < rows; ++i)
		for (size_t j = 0; j < cols; ++j)
			B[i * ldb + j] = zp_mul(p, B[i * ldb + j], a);
}

static int trsm_rec(uint32_t p, int upper, int unit, size_t n, size_t nrhs, uint32_t a,
		    const uint32_t *A, size_t lda, uint32_t *B, size_t ldb, size_t *bad)
{
	if (n == 0)
		return ORC_OK;
	if (a == 0) {                               /* :434-435 */
		scal_rows(p, 0, B, ldb, n, nrhs);
		return ORC_OK;
	}
	if (n == 1) {                               /* :436-448 */
		if (unit) {
			scal_rows(p, a, B, ldb, 1, nrhs);
			return ORC_OK;
		}
		uint32_t ai;
		if (!orc_inv(p, A[0], &ai)) {
			*bad = 0;
			return ORC_E_NOT_INVERTIBLE;
		}
		ai = zp_mul(p, ai, a);
		scal_rows(p, ai, B, ldb, 1, nrhs);
		return ORC_OK;
	}
	/* Dense matrices have align 1, so the split is l = n/2 (:449-497). */
	size_t l = n / 2;
	const uint32_t *A11 = A, *A22 = A + l * lda + l;
	uint32_t *B1 = B, *B2 = B + l * ldb;
	int rc;
	if (!upper) {
		const uint32_t *A21 = A + l * lda;
		rc = trsm_rec(p, upper, unit, l, nrhs, a, A11, lda, B1, ldb, bad);
		if (rc)
			return rc;
		/* gemm(-1, A21, B1, a, B2) */
		orc_zp_gemm_u32(p, n - l, nrhs, l, p - 1, A21, lda, B1, ldb, a, B2, ldb);
		rc = trsm_rec(p, upper, unit, n - l, nrhs, 1, A22, lda, B2, ldb, bad);
		if (rc) {
			*bad += l;
			return rc;
		}
	} else {
		const uint32_t *A12 = A + l;
		rc = trsm_rec(p, upper, unit, n - l, nrhs, a, A22, lda, B2, ldb, bad);
		if (rc) {
			*bad += l;
			return rc;
		}
		orc_zp_gemm_u32(p, l, nrhs, n - l, p - 1, A12, lda, B2, ldb, a, B1, ldb);
		rc = trsm_rec(p, upper, unit, l, nrhs, 1, A11, lda, B1, ldb, bad);
		if (rc)
			return rc;
	}
	return ORC_OK;
}

int orc_zp_trsm(uint32_t p, int upper, int unit_diag, size_t n, size_t nrhs, uint32_t a,
		const uint32_t *A, size_t lda, uint32_t *B, size_t ldb, size_t *bad_diag)
{
	if (p < 2)
		return ORC_E_ARG;
	size_t bad = 0;
	int rc = trsm_rec(p, upper, unit_diag, n, nrhs, a, A, lda, B, ldb, &bad);
	if (bad_diag)
		*bad_diag = bad;
	return rc;
}

/* ------------------------------------------------------------------ */
/* Elimination over Z/p (lela/algorithms/elimination.tcc:47-264)       */
/* ------------------------------------------------------------------ */

/* DensePivotStrategy::getPivot_spec, dense (pivot-strategy.tcc:19-41):
 * scan columns from col, rows from row, take the first nonzero. */
static int dense_pivot(uint32_t *A, size_t lda, size_t m, size_t n, size_t *row, size_t *col)
{
	for (size_t c = *col; c < n; ++c)
		for (size_t r = *row; r < m; ++r)
			if (A[r * lda + c] != 0) {
				*row = r;
				*col = c;
				return 1;
			}
	return 0;
}

static void swap_rows_u32(uint32_t *A, size_t lda, size_t n, size_t i, size_t j)
{
	for (size_t c = 0; c < n; ++c) {
		uint32_t t = A[i * lda + c];
		A[i * lda + c] = A[j * lda + c];
		A[j * lda + c] = t;
	}
}

int orc_zp_echelonize(uint32_t p, size_t m, size_t n, uint32_t *A, size_t lda, int reduced,
		      size_t *rank, uint32_t *pivots, uint32_t *det)
{
	if (p < 2)
		return ORC_E_ARG;
	size_t r = 0;
	uint32_t d = 1 % p;
	size_t i = 0, pivot_col = 0;
	for (; i < m && pivot_col < n; ++i, ++pivot_col) {
		size_t pivot_row = i;
		if (!dense_pivot(A, lda, m, n, &pivot_row, &pivot_col))
			break;
		if (i != pivot_row)                          /* :92-97 / :197-207 */
			swap_rows_u32(A, lda, n, i, pivot_row);
		uint32_t x = A[i * lda + pivot_col];
		d = zp_mul(p, d, x);
		uint32_t xinv;
		if (!orc_inv(p, x, &xinv))                   /* :103-108 / :217-220 */
			return ORC_E_NOT_INVERTIBLE;
		uint32_t *Pi = A + i * lda;
		if (!reduced) {
			uint32_t negxinv = zp_neg(p, xinv);
			for (size_t j = i + 1; j < m; ++j) {   /* :112-123 */
				uint32_t *Rj = A + j * lda;
				uint32_t aj = Rj[pivot_col];
				if (aj != 0) {
					uint32_t c = zp_mul(p, aj, negxinv);
					for (size_t col = 0; col < n; ++col)
						Rj[col] = zp_axpy(p, c, Pi[col], Rj[col]);
				}
			}
		} else {
			for (size_t col = 0; col < n; ++col)  /* :223 scal(xinv, pivot row) */
				Pi[col] = zp_mul(p, Pi[col], xinv);
			for (size_t j = 0; j < m; ++j) {       /* :230-244 */
				if (j == i)
					continue;
				uint32_t *Rj = A + j * lda;
				uint32_t aj = Rj[pivot_col];
				if (aj != 0) {
					uint32_t c = zp_neg(p, aj);
					for (size_t col = 0; col < n; ++col)
						Rj[col] = zp_axpy(p, c, Pi[col], Rj[col]);
				}
			}
		}
		if (pivots)
			pivots[r] = (uint32_t) pivot_col;
		++r;
	}
	if (rank)
		*rank = r;
	if (det)
		*det = d;
	return ORC_OK;
}

/* Elimination::echelonize (lela/algorithms/elimination.tcc:47-139) with the
 * row transpositions: step i takes the pivot of DensePivotStrategy
 * (pivot-strategy.tcc:19-41), swaps rows i and pivot_row (:92-97), and for
 * every row j below with a nonzero a in the pivot column adds -a/x times the
 * pivot row to it over the whole row (:112-120).  compute_L (the method's
 * default): A.setEntry (j, i, -a/x) after that axpy (:121-122), so the part
 * left of each row's pivot accumulates the transform T with T P A = U.
 * perm_rows[i] = pivot_row of step i (i itself when no swap); det = the
 * product of the pivots (:100). */
int orc_zp_elim(uint32_t p, size_t m, size_t n, uint32_t *A, size_t lda, int compute_L, size_t *rank,
		uint32_t *det, uint32_t *perm_rows)
{
	if (p < 2)
		return ORC_E_ARG;
	size_t r = 0, i = 0, pivot_col = 0;
	uint32_t d = 1 % p;
	for (; i < m && pivot_col < n; ++i, ++pivot_col) {
		size_t pivot_row = i;
		if (!dense_pivot(A, lda, m, n, &pivot_row, &pivot_col))
			break;
		if (i != pivot_row)
			swap_rows_u32(A, lda, n, i, pivot_row);
		if (perm_rows)
			perm_rows[i] = (uint32_t) pivot_row;
		const uint32_t x = A[i * lda + pivot_col];
		d = zp_mul(p, d, x);
		uint32_t xinv;
		if (!orc_inv(p, x, &xinv))
			return ORC_E_NOT_INVERTIBLE;
		const uint32_t negxinv = zp_neg(p, xinv);
		const uint32_t *Pi = A + i * lda;
		for (size_t j = i + 1; j < m; ++j) {
			uint32_t *Rj = A + j * lda;
			const uint32_t aj = Rj[pivot_col];
			if (aj == 0)
				continue;
			const uint32_t c = zp_mul(p, aj, negxinv);
			for (size_t col = 0; col < n; ++col)
				Rj[col] = zp_axpy(p, c, Pi[col], Rj[col]);
			if (compute_L)
				Rj[i] = c;
		}
		++r;
	}
	if (rank)
		*rank = r;
	if (det)
		*det = d;
	return ORC_OK;
}

static void swap_cols_u32(uint32_t *A, size_t lda, size_t m, size_t i, size_t j)
{
	for (size_t r = 0; r < m; ++r) {
		uint32_t t = A[r * lda + i];
		A[r * lda + i] = A[r * lda + j];
		A[r * lda + j] = t;
	}
}

/* Elimination::pluq (lela/algorithms/elimination.tcc:266-370): step i takes
 * the pivot of DensePivotStrategy from (i, i), swaps rows i and pivot_row and
 * columns i and pivot_col (whole rows and columns, :298-309), and for every
 * row j below with a nonzero a in column i adds -a/x times the pivot row to it
 * on the columns right of i (:321-331), then stores L's entry a/x in column i
 * (:333-335).  perm_rows[i] / perm_cols[i] = the row / column swapped with i
 * at step i, in step order (LELA reverses both lists at the end, :357-358). */
int orc_zp_pluq(uint32_t p, size_t m, size_t n, uint32_t *A, size_t lda, size_t *rank, uint32_t *det,
		uint32_t *perm_rows, uint32_t *perm_cols)
{
	if (p < 2)
		return ORC_E_ARG;
	size_t r = 0;
	uint32_t d = 1 % p;
	for (size_t i = 0; i < m && i < n; ++i) {
		size_t pivot_row = i, pivot_col = i;
		if (!dense_pivot(A, lda, m, n, &pivot_row, &pivot_col))
			break;
		if (i != pivot_row)
			swap_rows_u32(A, lda, n, i, pivot_row);
		if (i != pivot_col)
			swap_cols_u32(A, lda, m, i, pivot_col);
		if (perm_rows)
			perm_rows[i] = (uint32_t) pivot_row;
		if (perm_cols)
			perm_cols[i] = (uint32_t) pivot_col;
		const uint32_t x = A[i * lda + i];
		d = zp_mul(p, d, x);
		uint32_t xinv;
		if (!orc_inv(p, x, &xinv))
			return ORC_E_NOT_INVERTIBLE;
		const uint32_t negxinv = zp_neg(p, xinv);
		const uint32_t *Pi = A + i * lda;
		for (size_t j = i + 1; j < m; ++j) {
			uint32_t *Rj = A + j * lda;
			const uint32_t aj = Rj[i];
			if (aj == 0)
				continue;
			const uint32_t c = zp_mul(p, aj, negxinv);
			for (size_t col = i + 1; col < n; ++col)
				Rj[col] = zp_axpy(p, c, Pi[col], Rj[col]);
			Rj[i] = zp_neg(p, c);
		}
		++r;
	}
	if (rank)
		*rank = r;
	if (det)
		*det = d;
	return ORC_OK;
}

/* SparsePartialPivotStrategy::getPivot (lela/algorithms/pivot-strategy.tcc:69-99
 * for sparse rows, :132-183 for hybrid 0-1 rows) on dense storage: scanning the
 * rows from *row down, the row whose first nonzero at or after start = *col is
 * leftmost, and of those the one with the fewest nonzeros in [start, n) --
 * words != 0: the fewest nonzero 64-column words from the pivot's word on
 * (the hybrid row's stored blocks, i->end () - block) -- the first on ties
 * (strict <).  Returns 0 when no row has a nonzero at or after start. */
static int sparse_pivot(const uint32_t *A, size_t lda, size_t m, size_t n, int words, size_t *row, size_t *col)
{
	const size_t start = *col;
	size_t best_col = n, best_cnt = (size_t) -1, best_row = 0;
	for (size_t r = *row; r < m; ++r) {
		const uint32_t *R = A + r * lda;
		size_t lead = start;
		while (lead < n && R[lead] == 0)
			++lead;
		if (lead == n)
			continue;
		size_t cnt = 0;
		if (!words) {
			for (size_t c = lead; c < n; ++c)
				cnt += R[c] != 0;
		} else {
			for (size_t w = lead / 64; w * 64 < n; ++w) {
				int any = 0;
				for (size_t c = w * 64; c < n && c < w * 64 + 64; ++c)
					any |= R[c] != 0;
				cnt += any;
			}
		}
		if (lead < best_col) {
			best_col = lead;
			best_cnt = cnt;
			best_row = r;
		} else if (lead == best_col && cnt < best_cnt) {
			best_cnt = cnt;
			best_row = r;
		}
	}
	if (best_cnt == (size_t) -1)
		return 0;
	*row = best_row;
	*col = best_col;
	return 1;
}

int orc_zp_elim_sparse(uint32_t p, size_t m, size_t n, uint32_t *A, size_t lda, int mode, int words,
		       size_t *rank, uint32_t *det, uint32_t *perm_rows, uint32_t *perm_cols)
{
	if (p < 2 || mode < 0 || mode > 2)
		return ORC_E_ARG;
	const int pluq = mode == 2;
	size_t r = 0;
	uint32_t d = 1 % p;
	size_t i = 0, pivot_col = 0;
	/* elimination.tcc:73-131 (echelonize), :296-355 (pluq) */
	for (; i < m && (pluq ? i : pivot_col) < n; ++i, ++pivot_col) {
		size_t pivot_row = i;
		if (pluq)
			pivot_col = i;
		if (!sparse_pivot(A, lda, m, n, words, &pivot_row, &pivot_col))
			break;
		if (i != pivot_row)
			swap_rows_u32(A, lda, n, i, pivot_row);
		if (pluq && i != pivot_col)
			swap_cols_u32(A, lda, m, i, pivot_col);
		if (perm_rows)
			perm_rows[i] = (uint32_t) pivot_row;
		if (perm_cols)
			perm_cols[i] = (uint32_t) (pluq ? pivot_col : i);
		const size_t pc = pluq ? i : pivot_col;
		const uint32_t x = A[i * lda + pc];
		d = zp_mul(p, d, x);
		uint32_t xinv;
		if (!orc_inv(p, x, &xinv))
			return ORC_E_NOT_INVERTIBLE;
		const uint32_t negxinv = zp_neg(p, xinv);
		const uint32_t *Pi = A + i * lda;
		for (size_t j = i + 1; j < m; ++j) {
			uint32_t *Rj = A + j * lda;
			const uint32_t aj = Rj[pc];
			if (aj == 0)
				continue;
			const uint32_t c = zp_mul(p, aj, negxinv);
			const size_t c0 = mode == 0 ? pc : mode == 1 ? 0 : i + 1;
			for (size_t col = c0; col < n; ++col)
				Rj[col] = zp_axpy(p, c, Pi[col], Rj[col]);
			if (mode == 1)
				Rj[i] = c;
			else if (mode == 2)
				Rj[i] = zp_neg(p, c);
		}
		++r;
	}
	if (rank)
		*rank = r;
	if (det)
		*det = d;
	return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* GF(2)                                                               */
/* ------------------------------------------------------------------ */

static inline int gf2_get(const uint64_t *row, size_t col)
{
	return (int) ((row[col >> 6] >> (col & 63)) & 1u);
}

static inline void gf2_flip(uint64_t *row, size_t col)
{
	row[col >> 6] ^= (uint64_t) 1 << (col & 63);
}

static inline uint64_t tail_mask(size_t ncols)
{
	size_t r = ncols & 63;
	return r ? (((uint64_t) 1 << r) - 1) : ~(uint64_t) 0;
}

int orc_gf2_gemm(size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda_w,
		 const uint64_t *B, size_t ldb_w, int b, uint64_t *C, size_t ldc_w)
{
	size_t kw = (k + 63) / 64, nw = (n + 63) / 64;
	/* B columns as rows of B^T (the TransposeMatrix of level3-generic.tcc:257). */
	uint64_t *BT = (uint64_t *) calloc(n * (kw ? kw : 1), sizeof(uint64_t));
	if (!BT && n != 0 && kw != 0)
		return ORC_E_ARG;
	for (size_t l = 0; l < k; ++l)
		for (size_t j = 0; j < n; ++j)
			if (gf2_get(B + l * ldb_w, j))
				gf2_flip(BT + j * kw, l);
	uint64_t kmask = tail_mask(k);
	for (size_t i = 0; i < m; ++i) {
		uint64_t *Ci = C + i * ldc_w;
		/* gemv (level2-gf2.tcc:28-55): y <- 0 if !b; y[j] += dot(BT_j, x) if a. */
		if (!b)
			for (size_t w = 0; w < nw; ++w)
				Ci[w] &= ~(w + 1 == nw ? tail_mask(n) : ~(uint64_t) 0);
		if (!a)
			continue;
		const uint64_t *Ai = A + i * lda_w;
		for (size_t j = 0; j < n; ++j) {
			/* _dot Dense01 (level1-gf2.tcc:26-52): XOR of word ANDs, parity. */
			uint64_t t = 0;
			const uint64_t *Bj = BT + j * kw;
			for (size_t w = 0; w < kw; ++w)
				t ^= Ai[w] & Bj[w] & (w + 1 == kw ? kmask : ~(uint64_t) 0);
			if (__builtin_parityll(t))
				gf2_flip(Ci, j);
		}
	}
	free(BT);
	return ORC_OK;
}

/* BLAS3::trsm over GF(2) on unpacked bits: the generic recursion
 * (lela/blas/level3-generic.tcc:421-498) with Dense01Matrix's alignment of 64
 * columns: l = n/2 below 128 columns (:449-450), else a multiple of 64
 * (:472-474).  Over GF(2) -1 = 1, so gemm(-1, A21, B1, a, B2) is B2 ^= A21 B1
 * once a = 1. */
static int gf2_trsm_rec(int upper, int unit, size_t n, size_t nrhs, int a, const unsigned char *A, size_t lda,
			unsigned char *B, size_t ldb, size_t *bad)
{
	if (n == 0)
		return ORC_OK;
	if (!a) {                                   /* :434-435 */
		for (size_t i = 0; i < n; ++i)
			memset(B + i * ldb, 0, nrhs);
		return ORC_OK;
	}
	if (n == 1) {                               /* :436-448: the only invertible value is 1 */
		if (!unit && !A[0]) {
			*bad = 0;
			return ORC_E_NOT_INVERTIBLE;
		}
		return ORC_OK;
	}
	size_t l = n < 128 ? n / 2 : (n / 128) * 64;
	unsigned char *B1 = B, *B2 = B + l * ldb;
	int rc;
	if (!upper) {
		rc = gf2_trsm_rec(upper, unit, l, nrhs, a, A, lda, B1, ldb, bad);
		if (rc)
			return rc;
		for (size_t i = l; i < n; ++i)
			for (size_t j = 0; j < l; ++j)
				if (A[i * lda + j])
					for (size_t c = 0; c < nrhs; ++c)
						B[i * ldb + c] ^= B1[j * ldb + c];
		rc = gf2_trsm_rec(upper, unit, n - l, nrhs, 1, A + l * lda + l, lda, B2, ldb, bad);
		if (rc) {
			*bad += l;
			return rc;
		}
	} else {
		rc = gf2_trsm_rec(upper, unit, n - l, nrhs, a, A + l * lda + l, lda, B2, ldb, bad);
		if (rc) {
			*bad += l;
			return rc;
		}
		for (size_t i = 0; i < l; ++i)
			for (size_t j = l; j < n; ++j)
				if (A[i * lda + j])
					for (size_t c = 0; c < nrhs; ++c)
						B[i * ldb + c] ^= B[j * ldb + c];
		rc = gf2_trsm_rec(upper, unit, l, nrhs, 1, A, lda, B1, ldb, bad);
		if (rc)
			return rc;
	}
	return ORC_OK;
}

int orc_gf2_trsm(int upper, int unit_diag, size_t n, size_t nrhs, int a, const uint64_t *A, size_t lda_w,
		 uint64_t *B, size_t ldb_w, size_t *bad_diag)
{
	unsigned char *Au = (unsigned char *) calloc(n * n + 1, 1), *Bu = (unsigned char *) calloc(n * nrhs + 1, 1);
	if (!Au || !Bu) {
		free(Au);
		free(Bu);
		return ORC_E_ARG;
	}
	for (size_t i = 0; i < n; ++i) {
		for (size_t j = 0; j < n; ++j)
			Au[i * n + j] = (unsigned char) gf2_get(A + i * lda_w, j);
		for (size_t c = 0; c < nrhs; ++c)
			Bu[i * nrhs + c] = (unsigned char) gf2_get(B + i * ldb_w, c);
	}
	size_t bad = 0;
	int rc = gf2_trsm_rec(upper, unit_diag, n, nrhs, a & 1, Au, n, Bu, nrhs, &bad);
	if (bad_diag)
		*bad_diag = bad;
	if (rc == ORC_OK)
		for (size_t i = 0; i < n; ++i)
			for (size_t c = 0; c < nrhs; ++c)
				if ((unsigned char) gf2_get(B + i * ldb_w, c) != Bu[i * nrhs + c])
					gf2_flip(B + i * ldb_w, c);
	free(Au);
	free(Bu);
	return rc;
}

int orc_gf2_echelonize(size_t m, size_t n, uint64_t *A, size_t lda_w, int reduced,
		       size_t *rank, uint32_t *pivots)
{
	size_t nw = (n + 63) / 64;
	size_t r = 0, i = 0, pivot_col = 0;
	for (; i < m && pivot_col < n; ++i, ++pivot_col) {
		/* pivot-strategy.tcc:44-67 */
		size_t pivot_row = i;
		int found = 0;
		for (size_t c = pivot_col; c < n && !found; ++c)
			for (size_t rr = i; rr < m; ++rr)
				if (gf2_get(A + rr * lda_w, c)) {
					pivot_row = rr;
					pivot_col = c;
					found = 1;
					break;
				}
		if (!found)
			break;
		if (pivot_row != i)
			for (size_t w = 0; w < nw; ++w) {
				uint64_t t = A[i * lda_w + w];
				A[i * lda_w + w] = A[pivot_row * lda_w + w];
				A[pivot_row * lda_w + w] = t;
			}
		const uint64_t *Pi = A + i * lda_w;
		size_t j0 = reduced ? 0 : i + 1;
		for (size_t j = j0; j < m; ++j) {
			if (j == i)
				continue;
			uint64_t *Rj = A + j * lda_w;
			if (gf2_get(Rj, pivot_col))          /* axpy = row XOR (level1-gf2.tcc:282-299) */
				for (size_t w = 0; w < nw; ++w)
					Rj[w] ^= Pi[w];
		}
		if (pivots)
			pivots[r] = (uint32_t) pivot_col;
		++r;
	}
	if (rank)
		*rank = r;
	return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Faugere-Lachartre (lela/algorithms/faugere-lachartre.tcc)           */
/* ------------------------------------------------------------------ */

/* BLAS1::head: index of the first nonzero, -1 for a zero row. */
static long row_head(const uint32_t *row, size_t n)
{
	for (size_t c = 0; c < n; ++c)
		if (row[c] != 0)
			return (long) c;
	return -1;
}

/* Row/column classification of setup_splicer (:47-146).  is_piv_row[i] marks
 * the A/B rows; piv_col_of_row lists their heads.  Returns ORC_E_WRONG_FORM
 * on a decreasing head among scanned rows.  det is multiplied by every
 * scanned nonzero row's head (:119). */
static int fl_classify(uint32_t p, const uint32_t *X, size_t ldx, size_t m, size_t n,
		       unsigned char *is_piv_row, unsigned char *is_piv_col, size_t *npiv,
		       uint32_t *det, size_t *bad_row)
{
	long last_col = -1;
	size_t np = 0;
	memset(is_piv_row, 0, m);
	memset(is_piv_col, 0, n);
	for (size_t row = 0; row < m; ++row) {
		long col = row_head(X + row * ldx, n);
		if (col == -1)
			break;                               /* :64-70: first zero row ends the scan */
		if (col < last_col) {                        /* :80-81 */
			if (bad_row)
				*bad_row = row;
			return ORC_E_WRONG_FORM;
		}
		if (col > last_col) {                        /* new pivot row */
			is_piv_row[row] = 1;
			is_piv_col[col] = 1;
			last_col = col;
			++np;
		}
		*det = zp_mul(p, *det, X[row * ldx + (size_t) col]);
	}
	*npiv = np;
	return ORC_OK;
}

int orc_fl_echelonize(uint32_t p, size_t m, size_t n, const uint32_t *X, size_t ldx,
		      uint32_t *R, size_t ldr, size_t *rank, uint32_t *det, size_t *bad_row)
{
	if (p < 2)
		return ORC_E_ARG;
	int rc = ORC_OK;
	uint32_t d = 1 % p;
	unsigned char *prow = (unsigned char *) malloc(m + 1), *pcol = (unsigned char *) malloc(n + 1);
	size_t *rows_tr = (size_t *) malloc((m + 1) * sizeof(size_t));
	size_t *rows_res = (size_t *) malloc((m + 1) * sizeof(size_t));
	size_t *cols_tr = (size_t *) malloc((n + 1) * sizeof(size_t));
	size_t *cols_res = (size_t *) malloc((n + 1) * sizeof(size_t));
	uint32_t *Am = NULL, *Bm = NULL, *Cm = NULL, *Dm = NULL, *D1 = NULL, *D2 = NULL,
		 *B1 = NULL, *B2 = NULL, *dpiv = NULL;
	size_t r = 0;
	rc = fl_classify(p, X, ldx, m, n, prow, pcol, &r, &d, bad_row);
	if (rc)
		goto out;
	{
		size_t nt = 0, nr = 0, ct = 0, cr = 0;
		for (size_t i = 0; i < m; ++i)
			if (prow[i]) rows_tr[nt++] = i; else rows_res[nr++] = i;
		for (size_t c = 0; c < n; ++c)
			if (pcol[c]) cols_tr[ct++] = c; else cols_res[cr++] = c;
		/* MatrixGrid1 split (:148-275): A sparse in LELA, dense here; same values. */
		size_t mr = nr, nc = cr;
		Am = (uint32_t *) calloc(r * r + 1, 4);
		Bm = (uint32_t *) calloc(r * nc + 1, 4);
		Cm = (uint32_t *) calloc(mr * r + 1, 4);
		Dm = (uint32_t *) calloc(mr * nc + 1, 4);
		for (size_t i = 0; i < r; ++i) {
			const uint32_t *Xi = X + rows_tr[i] * ldx;
			for (size_t j = 0; j < r; ++j) Am[i * r + j] = Xi[cols_tr[j]];
			for (size_t j = 0; j < nc; ++j) Bm[i * nc + j] = Xi[cols_res[j]];
		}
		for (size_t i = 0; i < mr; ++i) {
			const uint32_t *Xi = X + rows_res[i] * ldx;
			for (size_t j = 0; j < r; ++j) Cm[i * r + j] = Xi[cols_tr[j]];
			for (size_t j = 0; j < nc; ++j) Dm[i * nc + j] = Xi[cols_res[j]];
		}
		/* :790-799: B <- A^-1 B; a failure returns with R untouched. */
		size_t bad = 0;
		rc = orc_zp_trsm(p, 1, 0, r, nc, 1, Am, r, Bm, nc, &bad);
		if (rc)
			goto out;
		/* :805 D <- D - C B */
		orc_zp_gemm_u32(p, mr, nc, r, p - 1, Cm, r, Bm, nc, 1, Dm, nc);
		/* :822 EF.echelonize(D): for a dense matrix GaussJordan::echelonize and
		 * move_L (solutions/echelon-form.h:124-131), whose result LELA's test
		 * asserts equal to Elimination::echelonize's (tests/test-gauss-jordan.C:
		 * 92-99): orc_zp_elim, LELA's transposition pivoting. */
		size_t rd = 0;
		uint32_t ddet = 1 % p;
		dpiv = (uint32_t *) malloc((nc + 1) * 4);
		rc = orc_zp_elim(p, mr, nc, Dm, nc, 0, &rd, &ddet, NULL);
		if (rc)
			goto out;
		/* :833 setup_splicer(D): the scan of the echelon rows; every head enters det. */
		for (size_t i = 0; i < rd; ++i) {
			long h = row_head(Dm + i * nc, nc);
			dpiv[i] = (uint32_t) h;
			d = zp_mul(p, d, Dm[i * nc + (size_t) h]);
		}
		/* :852-861 D1^-1 D2 over D's pivot / non-pivot columns. */
		unsigned char *dcol = (unsigned char *) calloc(nc + 1, 1);
		for (size_t i = 0; i < rd; ++i) dcol[dpiv[i]] = 1;
		size_t nd2 = nc - rd;
		size_t *d2cols = (size_t *) malloc((nd2 + 1) * sizeof(size_t));
		for (size_t c = 0, t = 0; c < nc; ++c) if (!dcol[c]) d2cols[t++] = c;
		D1 = (uint32_t *) calloc(rd * rd + 1, 4);
		D2 = (uint32_t *) calloc(rd * nd2 + 1, 4);
		for (size_t i = 0; i < rd; ++i) {
			for (size_t j = 0; j < rd; ++j) D1[i * rd + j] = Dm[i * nc + dpiv[j]];
			for (size_t j = 0; j < nd2; ++j) D2[i * nd2 + j] = Dm[i * nc + d2cols[j]];
		}
		rc = orc_zp_trsm(p, 1, 0, rd, nd2, 1, D1, rd, D2, nd2, &bad);
		if (rc) {
			free(dcol); free(d2cols);
			goto out;
		}
		/* :883-892 B2 <- B2 - B1 D2 */
		B1 = (uint32_t *) calloc(r * rd + 1, 4);
		B2 = (uint32_t *) calloc(r * nd2 + 1, 4);
		for (size_t i = 0; i < r; ++i) {
			for (size_t j = 0; j < rd; ++j) B1[i * rd + j] = Bm[i * nc + dpiv[j]];
			for (size_t j = 0; j < nd2; ++j) B2[i * nd2 + j] = Bm[i * nc + d2cols[j]];
		}
		orc_zp_gemm_u32(p, r, nd2, rd, p - 1, B1, rd, D2, nd2, 1, B2, nd2);
		/* :901 MatrixGrid3 assembly: standard-layout RREF.  Rows are the X pivot
		 * rows (identity on cols_tr, B2 on the final non-pivot columns) and the
		 * D pivot rows (identity on D's pivot columns, D2), merged by pivot column. */
		for (size_t i = 0; i < m; ++i)
			memset(R + i * ldr, 0, n * sizeof(uint32_t));
		size_t ia = 0, id = 0, out_row = 0;
		while (ia < r || id < rd) {
			size_t ca = ia < r ? cols_tr[ia] : (size_t) -1;
			size_t cd = id < rd ? cols_res[dpiv[id]] : (size_t) -1;
			uint32_t *Ro = R + out_row * ldr;
			if (ca < cd) {
				Ro[ca] = 1 % p;
				for (size_t j = 0; j < nd2; ++j) Ro[cols_res[d2cols[j]]] = B2[ia * nd2 + j];
				++ia;
			} else {
				Ro[cd] = 1 % p;
				for (size_t j = 0; j < nd2; ++j) Ro[cols_res[d2cols[j]]] = D2[id * nd2 + j];
				++id;
			}
			++out_row;
		}
		r += rd;
		free(dcol);
		free(d2cols);
	}
out:
	if (rank) *rank = r;
	if (det) *det = d;
	free(prow); free(pcol); free(rows_tr); free(rows_res); free(cols_tr); free(cols_res);
	free(Am); free(Bm); free(Cm); free(Dm); free(D1); free(D2); free(B1); free(B2); free(dpiv);
	return rc;
}
