// Elimination with LELA's SparsePartialPivotStrategy on the GPU:
// Elimination::echelonize (lela/algorithms/elimination.tcc:47-139) and
// Elimination::pluq (:266-370) with the pivot rule of
// lela/algorithms/pivot-strategy.tcc:69-188 -- among the rows at or below the
// current position, the one whose first nonzero at or after the search column
// is leftmost, and of those the one with the fewest nonzeros from there on
// (for GF(2) hybrid rows: the fewest nonzero 64-bit words from the pivot's
// word on), the first such row on ties.
//
// The rule reads every candidate's current nonzero count, so the trailing
// matrix cannot be updated in delayed blocks the way the dense-pivoting forms
// are: each step updates the affected rows in full.  One step is two
// launches on the context's stream:
//   sp_swap_kernel  takes the step's pivot key (leftmost column, count,
//                   position: one 64-bit word, the minimum over rows), swaps
//                   the pivot row into place (and for pluq the pivot column),
//                   and records the pivot and its inverse;
//   sp_elim_kernel  one workgroup per row below: a row with a nonzero in the
//                   pivot column gets LELA's axpy, then its (leftmost column,
//                   count) for the next search; a row without one keeps its
//                   cached pair, which the step cannot change (its entries in
//                   the swapped columns and left of its first nonzero are zero).
//                   Every live row folds its key into the next step's minimum
//                   with one atomicMin.
// Rows live in a canonical uint32 working copy (m x n, rows swapped in place).

#include "common.hpp"

namespace mlz {

namespace {

constexpr int SP_NT = 256;               // threads per row workgroup
constexpr unsigned long long SP_NONE = ~0ull;
constexpr int SP_BITS = 21;              // key fields: column, count, position
constexpr unsigned long long SP_MASK = (1ull << SP_BITS) - 1;

__device__ __host__ inline unsigned long long sp_key(unsigned col, unsigned cnt, unsigned pos)
{
	return ((unsigned long long) col << (2 * SP_BITS)) | ((unsigned long long) cnt << SP_BITS) | pos;
}

// Per-step record: pivot position before the swap, pivot column, value, inverse.
struct SpStep {
	unsigned pos, col, x, xinv;
};

// Leftmost nonzero column >= start and the count of row X[0..n) (entries, or
// nonzero 64-column words from the leftmost one's word on), by one workgroup.
// Returns through LDS; every thread sees the result.
__device__ void sp_row_scan(const uint32_t *__restrict__ row, int n, int start, bool words, unsigned *s_lead,
			    unsigned *s_cnt)
{
	const int t = threadIdx.x;
	if (t == 0) {
		*s_lead = (unsigned) n;
		*s_cnt = 0;
	}
	__syncthreads();
	unsigned lead = (unsigned) n, cnt = 0;
	// columns in chunks of SP_NT so that each wave covers one aligned 64-column word
	for (int base = start & ~63; base < n; base += SP_NT) {
		const int col = base + t;
		const bool nz = col >= start && col < n && row[col] != 0u;
		if (nz && (unsigned) col < lead)
			lead = (unsigned) col;
		if (!words)
			cnt += nz;
	}
	if (lead < (unsigned) n)
		atomicMin(s_lead, lead);
	if (!words && cnt)
		atomicAdd(s_cnt, cnt);
	__syncthreads();
	if (words && *s_lead < (unsigned) n) {
		const int w0 = (int) (*s_lead & ~63u);
		for (int base = w0; base < n; base += SP_NT) {
			const int col = base + t;
			const bool nz = col < n && row[col] != 0u;
			const unsigned long long b = __ballot(nz);
			if ((t & 63) == 0 && b)
				atomicAdd(s_cnt, 1u);
		}
		__syncthreads();
	}
}

// Initial (leftmost column, count) of every row from column 0, and step 0's key.
__global__ __launch_bounds__(SP_NT) void sp_init_kernel(const uint32_t *__restrict__ X, size_t ldx, int n, int words,
							unsigned *__restrict__ lead, unsigned *__restrict__ cnt,
							unsigned long long *__restrict__ key)
{
	__shared__ unsigned s_lead, s_cnt;
	const int j = blockIdx.x;
	sp_row_scan(X + (size_t) j * ldx, n, 0, words != 0, &s_lead, &s_cnt);
	if (threadIdx.x == 0) {
		lead[j] = s_lead;
		cnt[j] = s_cnt;
		if (s_lead < (unsigned) n)
			atomicMin(key, sp_key(s_lead, s_cnt, (unsigned) j));
	}
}

__global__ void sp_fill_kernel(unsigned long long *__restrict__ key, int count, unsigned *__restrict__ err)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < count)
		key[i] = SP_NONE;
	if (i == 0)
		*err = 0;
}

// Step s: swap the pivot row into position s (all n columns; thread 0 also
// swaps the rows' cached search pairs), or with cols != 0 -- pluq's second
// launch, after the row swap -- the pivot column into column s (all m rows).
__global__ void sp_swap_kernel(uint32_t *__restrict__ X, size_t ldx, int m, int n, int s, int cols,
			       const unsigned long long *__restrict__ key, unsigned *__restrict__ lead,
			       unsigned *__restrict__ cnt, const unsigned *__restrict__ err)
{
	const unsigned long long k = key[s];
	if (k == SP_NONE || *err)
		return;
	const int pos = (int) (k & SP_MASK), col = (int) (k >> (2 * SP_BITS));
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (!cols) {
		if (pos == s)
			return;
		if (i < n) {
			const uint32_t a = X[(size_t) s * ldx + i];
			X[(size_t) s * ldx + i] = X[(size_t) pos * ldx + i];
			X[(size_t) pos * ldx + i] = a;
		}
		if (i == 0) {
			const unsigned l = lead[s], c = cnt[s];
			lead[s] = lead[pos];
			cnt[s] = cnt[pos];
			lead[pos] = l;
			cnt[pos] = c;
		}
	} else if (col != s && i < m) {
		const uint32_t a = X[(size_t) i * ldx + s];
		X[(size_t) i * ldx + s] = X[(size_t) i * ldx + col];
		X[(size_t) i * ldx + col] = a;
	}
}

// Records step s (after the swaps): pivot value and inverse; a pivot without
// an inverse (composite p) stops the elimination with the error flag.
__global__ void sp_record_kernel(const uint32_t *__restrict__ X, size_t ldx, int s, int pluq,
				 const unsigned long long *__restrict__ key, SpStep *__restrict__ steps, ModP mp,
				 unsigned *__restrict__ err)
{
	const unsigned long long k = key[s];
	if (k == SP_NONE || *err)
		return;
	const unsigned pos = (unsigned) (k & SP_MASK), col = (unsigned) (k >> (2 * SP_BITS));
	const uint32_t x = X[(size_t) s * ldx + (pluq ? (unsigned) s : col)];
	const uint32_t xinv = mp.inv(x);
	steps[s] = SpStep{pos, col, x, xinv};
	if (xinv == 0u)
		*err = 1;
}

// Step s, one workgroup per row position j > s.  Modes as MLZ_ECH_*.
__global__ __launch_bounds__(SP_NT) void sp_elim_kernel(uint32_t *__restrict__ X, size_t ldx, int m, int n, int s,
							int mode, int words,
							const unsigned long long *__restrict__ key,
							const SpStep *__restrict__ steps, unsigned *__restrict__ lead,
							unsigned *__restrict__ cnt, ModP mp,
							const unsigned *__restrict__ err)
{
	__shared__ unsigned s_lead, s_cnt, s_a;
	const int j = s + 1 + blockIdx.x;
	if (j >= m || key[s] == SP_NONE || *err)
		return;
	const SpStep st = steps[s];
	const bool pluq = mode == MLZ_ECH_PLUQ;
	const int pcol = pluq ? s : (int) st.col;
	uint32_t *__restrict__ Rj = X + (size_t) j * ldx;
	const uint32_t *__restrict__ Rs = X + (size_t) s * ldx;
	// the multiplier is read once, before any thread updates the row
	if (threadIdx.x == 0)
		s_a = Rj[pcol];
	__syncthreads();
	const uint32_t a = s_a;
	const int next_start = pluq ? s + 1 : (int) st.col + 1;
	if (a != 0u) {
		// negaxinv = -a / x; row_j += negaxinv * row_s over LELA's range
		const uint32_t c = mp.mul(a, mp.neg(st.xinv));
		const int c0 = mode == MLZ_ECH_U ? pcol : mode == MLZ_ECH_TRANSFORM ? 0 : s + 1;
		for (int col = c0 + (int) threadIdx.x; col < n; col += SP_NT)
			Rj[col] = mp.axpy(c, Rs[col], Rj[col]);
		__syncthreads();
		if (threadIdx.x == 0) {
			if (mode == MLZ_ECH_TRANSFORM)
				Rj[s] = c;                       // elimination.tcc:121-122
			else if (mode == MLZ_ECH_PLUQ)
				Rj[s] = mp.neg(c);               // a / x (:347-348)
		}
		__syncthreads();
		sp_row_scan(Rj, n, next_start, words != 0, &s_lead, &s_cnt);
		if (threadIdx.x == 0) {
			lead[j] = s_lead;
			cnt[j] = s_cnt;
		}
	} else if (threadIdx.x == 0) {
		s_lead = lead[j];
		s_cnt = cnt[j];
	}
	if (threadIdx.x == 0 && s_lead < (unsigned) n)
		atomicMin((unsigned long long *) &key[s + 1], sp_key(s_lead, s_cnt, (unsigned) j));
}

template <int E>
__global__ void sp_load_kernel(const typename Elem<E>::T *__restrict__ A, size_t lda, int n, uint32_t p,
			       uint32_t *__restrict__ X, size_t ldx)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, row = blockIdx.y;
	if (col < n)
		X[(size_t) row * ldx + col] = Elem<E>::load(A + (size_t) row * lda + col, p);
}

template <int E>
__global__ void sp_store_kernel(const uint32_t *__restrict__ X, size_t ldx, int n, uint32_t p,
				typename Elem<E>::T *__restrict__ A, size_t lda)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, row = blockIdx.y;
	if (col < n)
		Elem<E>::store(A + (size_t) row * lda + col, X[(size_t) row * ldx + col], p);
}

}  // namespace

int zp_sparse_elim_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, void *A, size_t lda, int mode,
		       int words, size_t *rank_out, uint64_t *det_out, uint32_t *perm_rows, uint32_t *perm_cols)
{
	const size_t rmax = std::min(m, n);
	if (rmax == 0) {
		if (rank_out)
			*rank_out = 0;
		if (det_out)
			*det_out = 1 % p;
		return MLZ_OK;
	}
	if (m > SP_MASK || n > SP_MASK)
		return set_error(ctx, MLZ_E_ARG, "sparse pivoting supports up to %llu rows and columns",
				 (unsigned long long) SP_MASK);
	const ModP mp = make_modp(p);
	const uint32_t pp = (uint32_t) p;
	const size_t ldx = round_up(n, 64);
	void *w0, *w1;
	MLZ_TRY(ws_get(ctx, WS_RREF_0, m * ldx * sizeof(uint32_t), &w0));
	const size_t sKey = round_up((rmax + 1) * 8, 256), sStep = round_up(rmax * sizeof(SpStep), 256),
		     sRow = round_up(m * 4, 256);
	MLZ_TRY(ws_get(ctx, WS_RREF_1, sKey + sStep + 2 * sRow + 256, &w1));
	uint32_t *X = (uint32_t *) w0;
	unsigned long long *key = (unsigned long long *) w1;
	SpStep *steps = (SpStep *) ((char *) w1 + sKey);
	unsigned *lead = (unsigned *) ((char *) w1 + sKey + sStep);
	unsigned *cnt = (unsigned *) ((char *) w1 + sKey + sStep + sRow);
	unsigned *err = (unsigned *) ((char *) w1 + sKey + sStep + 2 * sRow);
	hipStream_t st = ctx->stream;
	{
		ProfScope ps(ctx, "sp_load");
		dim3 g(ceil_div(n, 256), (unsigned) m);
		if (elem == MLZ_U16)
			sp_load_kernel<MLZ_U16><<<g, 256, 0, st>>>((const uint16_t *) A, lda, (int) n, pp, X, ldx);
		else if (elem == MLZ_U32)
			sp_load_kernel<MLZ_U32><<<g, 256, 0, st>>>((const uint32_t *) A, lda, (int) n, pp, X, ldx);
		else
			sp_load_kernel<MLZ_F64_SYM><<<g, 256, 0, st>>>((const double *) A, lda, (int) n, pp, X, ldx);
		MLZ_TRY(check_launch(ctx, "sp_load_kernel"));
	}
	sp_fill_kernel<<<ceil_div(rmax + 1, 256), 256, 0, st>>>(key, (int) (rmax + 1), err);
	MLZ_TRY(check_launch(ctx, "sp_fill_kernel"));
	{
		ProfScope ps(ctx, "sp_init");
		sp_init_kernel<<<(unsigned) m, SP_NT, 0, st>>>(X, ldx, (int) n, words, lead, cnt, key);
		MLZ_TRY(check_launch(ctx, "sp_init_kernel"));
	}
	const bool pluq = mode == MLZ_ECH_PLUQ;
	// steps in chunks; the host looks at the next key between chunks and stops
	// once no row is left (the kernels of a finished elimination return at once)
	const size_t chunk = 64;
	size_t s = 0;
	bool more = true;
	while (more && s < rmax) {
		const size_t end = std::min(rmax, s + chunk);
		for (; s < end; ++s) {
			ProfScope ps(ctx, "sp_step");
			sp_swap_kernel<<<ceil_div(n, 256), 256, 0, st>>>(X, ldx, (int) m, (int) n, (int) s, 0, key, lead,
									 cnt, err);
			if (pluq)
				sp_swap_kernel<<<ceil_div(m, 256), 256, 0, st>>>(X, ldx, (int) m, (int) n, (int) s, 1, key,
										 lead, cnt, err);
			sp_record_kernel<<<1, 1, 0, st>>>(X, ldx, (int) s, pluq, key, steps, mp, err);
			if (s + 1 < m)
				sp_elim_kernel<<<(unsigned) (m - s - 1), SP_NT, 0, st>>>(X, ldx, (int) m, (int) n, (int) s,
											 mode, words, key, steps, lead,
											 cnt, mp, err);
			MLZ_TRY(check_launch(ctx, "sp_elim_kernel"));
		}
		unsigned long long next = 0;
		MLZ_HIP(ctx, hipMemcpyAsync(&next, key + s, 8, hipMemcpyDeviceToHost, st));
		MLZ_HIP(ctx, hipStreamSynchronize(st));
		more = next != SP_NONE;
	}
	std::vector<unsigned long long> keys(rmax + 1);
	std::vector<SpStep> rec(rmax);
	unsigned herr = 0;
	MLZ_HIP(ctx, hipMemcpyAsync(keys.data(), key, (rmax + 1) * 8, hipMemcpyDeviceToHost, st));
	MLZ_HIP(ctx, hipMemcpyAsync(rec.data(), steps, rmax * sizeof(SpStep), hipMemcpyDeviceToHost, st));
	MLZ_HIP(ctx, hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
	MLZ_HIP(ctx, hipStreamSynchronize(st));
	size_t rank = 0;
	while (rank < rmax && keys[rank] != SP_NONE)
		++rank;
	if (herr)
		return set_error(ctx, MLZ_E_NOT_INVERTIBLE, "Could not invert pivot-element in the ring");
	uint64_t det = 1 % p;
	for (size_t t = 0; t < rank; ++t) {
		det = (det * rec[t].x) % p;
		if (perm_rows)
			perm_rows[t] = rec[t].pos;
		if (perm_cols)
			perm_cols[t] = pluq ? rec[t].col : (uint32_t) t;
	}
	{
		ProfScope ps(ctx, "sp_store");
		dim3 g(ceil_div(n, 256), (unsigned) m);
		if (elem == MLZ_U16)
			sp_store_kernel<MLZ_U16><<<g, 256, 0, st>>>(X, ldx, (int) n, pp, (uint16_t *) A, lda);
		else if (elem == MLZ_U32)
			sp_store_kernel<MLZ_U32><<<g, 256, 0, st>>>(X, ldx, (int) n, pp, (uint32_t *) A, lda);
		else
			sp_store_kernel<MLZ_F64_SYM><<<g, 256, 0, st>>>(X, ldx, (int) n, pp, (double *) A, lda);
		MLZ_TRY(check_launch(ctx, "sp_store_kernel"));
	}
	if (rank_out)
		*rank_out = rank;
	if (det_out)
		*det_out = det;
	return MLZ_OK;
}

}  // namespace mlz

using namespace mlz;

extern "C" int mlz_zp_echelonize_sparse(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, void *A,
					size_t lda, int mode, int count_words, size_t *rank, uint64_t *det,
					uint32_t *perm_rows, uint32_t *perm_cols)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_TRY(check_modulus(ctx, elem, p));
	if (mode != MLZ_ECH_U && mode != MLZ_ECH_TRANSFORM && mode != MLZ_ECH_PLUQ)
		return set_error(ctx, MLZ_E_ARG, "unknown echelon mode %d", mode);
	if (m && n && !A)
		return set_error(ctx, MLZ_E_ARG, "null matrix pointer");
	if (m && n && lda < n)
		return set_error(ctx, MLZ_E_ARG, "leading dimension smaller than row length");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (ctx->ptr_mode == MLZ_PTR_DEVICE || m == 0 || n == 0)
		return zp_sparse_elim_dev(ctx, elem, p, m, n, A, lda, mode, count_words != 0, rank, det, perm_rows,
					  perm_cols);
	const size_t es = elem_size(elem);
	void *dA = nullptr;
	MLZ_TRY(stage_in(ctx, WS_STAGE_A, A, m, n, lda, es, &dA));
	MLZ_TRY(zp_sparse_elim_dev(ctx, elem, p, m, n, dA, n, mode, count_words != 0, rank, det, perm_rows,
				   perm_cols));
	return stage_out(ctx, A, dA, m, n, lda, es);
}
