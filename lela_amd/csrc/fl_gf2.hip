// lela_amd/csrc/fl_gf2.hip -- Faugere-Lachartre reduction over GF(2) of an
// F4-shaped matrix given as hybrid rows, to reduced row-echelon form on MI355X.
//
// Follows FaugereLachartre<GF2>::echelonize(R, X, rank, det, reduced=true)
// (lela/algorithms/faugere-lachartre.tcc:735-917) on
// SparseMatrix<bool, HybridVector<.., uint16, uint64>> rows
// (lela/vector/hybrid.h:30-33: sorted (word index, 64-bit word) pairs, bit j
// of word w = column 64 w + j), the configuration LELA's own test runs
// (tests/test-faugere-lachartre.C:278-285):
//   1. the row scan of setup_splicer (:47-146) on the rows' heads (first set
//      bit): a head larger than the last starts a pivot row (A|B), a repeated
//      head marks a C|D row, a smaller one is WrongMatrixForm(row), the first
//      zero row ends the scan and the rows after it join C|D (:64-70,
//      :137-142).  Sequential and O(m): on the host, from a kernel's heads.
//   2. the MatrixGrid1 split (:148-275) into bit-packed dense A (pivot rows x
//      pivot columns: upper triangular, ones on the diagonal), B, C, D.
//   3. B <- A^-1 B (:791) by the GF(2) trsm (unit upper).
//   4. D <- D + C B (:805; -1 = 1) by the GF(2) gemm.
//   5. D <- RREF(D) (:822-861: LELA's echelon form of D reduced by D1^-1 D2,
//      whose result is the unique RREF).
//   6. B <- B + B[:, J_D] RREF(D) (:883-892).
//   7. the standard-layout RREF, rows ordered by pivot column (:901).
// det: every head is 1 over GF(2), so LELA's det is 1 (true) and nothing is
// computed for it.
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace mlz {

// head[i]: first set bit of row i (column index), -1 for a zero row.
__global__ void gf2fl_heads_kernel(const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ blk,
				   const uint64_t *__restrict__ words, int m, int *__restrict__ head)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= m)
		return;
	int h = -1;
	for (uint64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
		if (words[e]) {
			h = 64 * (int) blk[e] + __builtin_ctzll(words[e]);
			break;
		}
	head[i] = h;
}

__global__ void gf2fl_zero_kernel(uint64_t *__restrict__ X, size_t count)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i < count)
		X[i] = 0ull;
}

// Every set bit of row i (workgroup i) to its block: rowmap[i] >= 0 is a
// pivot row (A|B), -1-t the C|D row t; colmap[c] >= 0 a pivot column (A|C),
// -1-f the other column f (B|D).  Bits past column n are ignored.
__global__ void gf2fl_split_kernel(const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ blk,
				   const uint64_t *__restrict__ words, int n, const int *__restrict__ rowmap,
				   const int *__restrict__ colmap, uint64_t *__restrict__ A, size_t lda,
				   uint64_t *__restrict__ B, size_t ldb, uint64_t *__restrict__ C, size_t ldc,
				   uint64_t *__restrict__ D, size_t ldd)
{
	const int i = blockIdx.x;
	const int rm = rowmap[i];
	const uint64_t e0 = row_ptr[i], e1 = row_ptr[i + 1];
	for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
		uint64_t w = words[e];
		const int c0 = 64 * (int) blk[e];
		while (w) {
			const int b = __builtin_ctzll(w);
			w &= w - 1;
			const int c = c0 + b;
			if (c >= n)
				break;
			const int cm = colmap[c];
			uint64_t *dst;
			int r, cc;
			if (rm >= 0) {
				r = rm;
				if (cm >= 0) { dst = A + (size_t) r * lda; cc = cm; }
				else { dst = B + (size_t) r * ldb; cc = -1 - cm; }
			} else {
				r = -1 - rm;
				if (cm >= 0) { dst = C + (size_t) r * ldc; cc = cm; }
				else { dst = D + (size_t) r * ldd; cc = -1 - cm; }
			}
			atomicOr((unsigned long long *) (dst + (cc >> 6)), (unsigned long long) (1ull << (cc & 63)));
		}
	}
}

// G (r x rd bits) = the columns J of B (r x nr bits); workgroup i: row i.
__global__ void gf2fl_gather_cols_kernel(const uint64_t *__restrict__ B, size_t ldb, const int *__restrict__ J, int rd,
					 uint64_t *__restrict__ G, size_t ldg)
{
	const int i = blockIdx.x;
	const uint64_t *Bi = B + (size_t) i * ldb;
	for (int w = threadIdx.x; w < (rd + 63) / 64; w += blockDim.x) {
		uint64_t out = 0;
		for (int b = 0; b < 64 && 64 * w + b < rd; ++b) {
			const int c = J[64 * w + b];
			out |= ((Bi[c >> 6] >> (c & 63)) & 1ull) << b;
		}
		G[(size_t) i * ldg + w] = out;
	}
}

// Output row t of R: src[t] >= 0: pivot row src of X (its pivot column pc[t]
// plus B's row on the other columns), -1-j: row j of RREF(D); rows past the
// rank are zero.  resmap[c] = f >= 0 for the non-pivot column f of X, -1 for
// X's pivot columns.  Workgroup t: row t; word w of it gathers its 64 columns.
__global__ void gf2fl_assemble_kernel(const int *__restrict__ src, const int *__restrict__ pc, int rank,
				      const int *__restrict__ resmap, int n, const uint64_t *__restrict__ B, size_t ldb,
				      const uint64_t *__restrict__ D, size_t ldd, uint64_t *__restrict__ R, size_t ldr)
{
	const int t = blockIdx.x;
	const int nw = (n + 63) / 64;
	const int s = t < rank ? src[t] : 0;
	const uint64_t *row = s >= 0 ? B + (size_t) s * ldb : D + (size_t) (-1 - s) * ldd;
	const int pcol = s >= 0 ? pc[t] : -1;
	for (int w = threadIdx.x; w < nw; w += blockDim.x) {
		uint64_t out = 0;
		if (t < rank)
			for (int b = 0; b < 64; ++b) {
				const int c = 64 * w + b;
				if (c >= n)
					break;
				const int f = resmap[c];
				const uint64_t bit = f >= 0 ? (row[f >> 6] >> (f & 63)) & 1ull : (uint64_t) (c == pcol);
				out |= bit << b;
			}
		// bits past column n keep their values
		const uint64_t mask = w == nw - 1 ? (n % 64 ? (1ull << (n % 64)) - 1ull : ~0ull) : ~0ull;
		uint64_t *dst = R + (size_t) t * ldr + w;
		*dst = (out & mask) | (*dst & ~mask);
	}
}

int gf2_fl_dev(mlz_ctx *ctx, size_t m, size_t n, const uint64_t *row_ptr, const uint32_t *blk, const uint64_t *words,
	       uint64_t *R, size_t ldr, size_t *rank_out, uint32_t *pivots_out, size_t *bad_row)
{
	// 1. heads -> the row scan
	void *w0;
	MLZ_TRY(ws_get(ctx, WS_FL_0, (2 * m + 2 * n + 16) * sizeof(int), &w0));
	int *d_head = (int *) w0, *d_rowmap = d_head + m, *d_colmap = d_rowmap + m, *d_resmap = d_colmap + n;
	gf2fl_heads_kernel<<<ceil_div(m, 256), 256, 0, ctx->stream>>>(row_ptr, blk, words, (int) m, d_head);
	MLZ_TRY(check_launch(ctx, "gf2fl_heads_kernel"));
	std::vector<int> head(m);
	MLZ_HIP(ctx, hipMemcpyAsync(head.data(), d_head, m * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	std::vector<char> is_piv_row(m, 0), is_piv_col(n, 0);
	long last = -1;
	size_t r = 0;
	for (size_t i = 0; i < m; ++i) {
		const long h = head[i];
		if (h < 0)
			break;                       // :64-70
		if (h < last) {                      // :80-81
			if (bad_row)
				*bad_row = i;
			return set_error(ctx, MLZ_E_WRONG_FORM, "Bad input-matrix for Faugere-Lachartre at row-index %zu", i);
		}
		if (h > last) {
			is_piv_row[i] = 1;
			is_piv_col[h] = 1;
			last = h;
			++r;
		}
	}
	const size_t mr = m - r, nr = n - r;
	std::vector<int> rowmap(m), colmap(n), resmap(n), piv_x, res_cols;
	piv_x.reserve(r);
	res_cols.reserve(nr);
	for (size_t i = 0, a = 0, b = 0; i < m; ++i)
		rowmap[i] = is_piv_row[i] ? (int) a++ : -1 - (int) b++;
	for (size_t j = 0, a = 0, b = 0; j < n; ++j) {
		if (is_piv_col[j]) {
			colmap[j] = (int) a++;
			resmap[j] = -1;
			piv_x.push_back((int) j);
		} else {
			resmap[j] = (int) b;
			colmap[j] = -1 - (int) b++;
			res_cols.push_back((int) j);
		}
	}
	MLZ_HIP(ctx, hipMemcpyAsync(d_rowmap, rowmap.data(), m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(d_colmap, colmap.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(d_resmap, resmap.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));

	// 2. the split into bit-packed A (r x r), B (r x nr), C (mr x r), D (mr x nr)
	const size_t wr = std::max((size_t) 1, (size_t) ceil_div(r, 64)), wn = std::max((size_t) 1, (size_t) ceil_div(nr, 64));
	const size_t sA = r * wr, sB = r * wn, sC = mr * wr, sD = mr * wn;
	void *wa, *wb, *wc, *wd;
	MLZ_TRY(ws_get(ctx, WS_FL_1, sA * 8 + 256, &wa));
	MLZ_TRY(ws_get(ctx, WS_FL_2, sB * 8 + 256, &wb));
	MLZ_TRY(ws_get(ctx, WS_FL_3, sC * 8 + 256, &wc));
	MLZ_TRY(ws_get(ctx, WS_FL_4, sD * 8 + 256, &wd));
	uint64_t *dA = (uint64_t *) wa, *dB = (uint64_t *) wb, *dC = (uint64_t *) wc, *dD = (uint64_t *) wd;
	{
		ProfScope ps(ctx, "gf2fl_split");
		uint64_t *bufs[4] = {dA, dB, dC, dD};
		const size_t cnt[4] = {sA, sB, sC, sD};
		for (int q = 0; q < 4; ++q)
			if (cnt[q])
				gf2fl_zero_kernel<<<ceil_div(cnt[q], 256), 256, 0, ctx->stream>>>(bufs[q], cnt[q]);
		gf2fl_split_kernel<<<(unsigned) m, 64, 0, ctx->stream>>>(row_ptr, blk, words, (int) n, d_rowmap, d_colmap,
									  dA, wr, dB, wn, dC, wr, dD, wn);
		MLZ_TRY(check_launch(ctx, "gf2fl_split_kernel"));
	}
	// 3. B <- A^-1 B (A upper triangular with ones on the diagonal)
	if (r && nr) {
		size_t bad = 0;
		MLZ_TRY(gf2_trsm_dev(ctx, 1, 1, r, nr, 1, dA, wr, dB, wn, &bad));
	}
	// 4. D <- D + C B
	if (mr && nr && r)
		MLZ_TRY(gf2_gemm_dev(ctx, mr, nr, r, 1, dC, wr, dB, wn, 1, dD, wn));
	// 5. RREF(D)
	size_t rd = 0;
	std::vector<uint32_t> jd(std::min(mr, nr) + 1);
	if (mr && nr)
		MLZ_TRY(gf2_rref_dev(ctx, mr, nr, dD, wn, &rd, jd.data()));
	// 6. B <- B + B[:, J_D] RREF(D)[0 .. rd)
	if (r && rd) {
		void *wg, *wj;
		const size_t wgd = (size_t) ceil_div(rd, 64);
		MLZ_TRY(ws_get(ctx, WS_FL_5, r * wgd * 8 + 256, &wg));
		MLZ_TRY(ws_get(ctx, WS_FL_6, rd * sizeof(int) + 256, &wj));
		std::vector<int> jdi(jd.begin(), jd.begin() + rd);
		MLZ_HIP(ctx, hipMemcpyAsync(wj, jdi.data(), rd * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
		gf2fl_gather_cols_kernel<<<(unsigned) r, 64, 0, ctx->stream>>>(dB, wn, (const int *) wj, (int) rd,
										(uint64_t *) wg, wgd);
		MLZ_TRY(check_launch(ctx, "gf2fl_gather_cols_kernel"));
		MLZ_TRY(gf2_gemm_dev(ctx, r, nr, rd, 1, (const uint64_t *) wg, wgd, dD, wn, 1, dB, wn));
	}
	// 7. the rows of R by pivot column: X's pivot rows and D's
	const size_t rank = r + rd;
	std::vector<int> src(rank + 1), pcv(rank + 1);
	std::vector<uint32_t> piv(rank + 1);
	{
		size_t ia = 0, id = 0, t = 0;
		while (ia < r || id < rd) {
			const long ca = ia < r ? piv_x[ia] : -1, cd = id < rd ? res_cols[jd[id]] : -1;
			if (id >= rd || (ia < r && ca < cd)) {
				src[t] = (int) ia;
				pcv[t] = (int) ca;
				piv[t] = (uint32_t) ca;
				++ia;
			} else {
				src[t] = -1 - (int) id;
				pcv[t] = (int) cd;
				piv[t] = (uint32_t) cd;
				++id;
			}
			++t;
		}
	}
	if (R) {
		void *wm;
		MLZ_TRY(ws_get(ctx, WS_FL_7, 2 * (rank + 1) * sizeof(int), &wm));
		int *d_src = (int *) wm, *d_pc = d_src + rank + 1;
		MLZ_HIP(ctx, hipMemcpyAsync(d_src, src.data(), (rank + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
		MLZ_HIP(ctx, hipMemcpyAsync(d_pc, pcv.data(), (rank + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
		ProfScope ps(ctx, "gf2fl_assemble");
		gf2fl_assemble_kernel<<<(unsigned) m, 64, 0, ctx->stream>>>(d_src, d_pc, (int) rank, d_resmap, (int) n, dB,
									     wn, dD, wn, R, ldr);
		MLZ_TRY(check_launch(ctx, "gf2fl_assemble_kernel"));
	}
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));   // the host vectors leave scope
	if (pivots_out)
		for (size_t t = 0; t < rank; ++t)
			pivots_out[t] = piv[t];
	if (rank_out)
		*rank_out = rank;
	return MLZ_OK;
}

}  // namespace mlz

using namespace mlz;

extern "C" int mlz_gf2_fl_echelonize(mlz_ctx *ctx, size_t m, size_t n, const uint64_t *row_ptr, const uint32_t *blk,
				     const uint64_t *words, uint64_t *R, size_t ldr_w, size_t *rank, uint32_t *pivots,
				     size_t *bad_row)
{
	if (!ctx)
		return MLZ_E_ARG;
	const size_t nw = (n + 63) / 64;
	if (!row_ptr || (m && n && R && ldr_w < nw))
		return set_error(ctx, MLZ_E_ARG, "bad hybrid rows or output pointer");
	if (m > (size_t) INT32_MAX / 2 || n > (size_t) INT32_MAX / 2)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (m == 0 || n == 0) {
		if (rank)
			*rank = 0;
		return MLZ_OK;
	}
	if (ctx->ptr_mode == MLZ_PTR_DEVICE)
		return gf2_fl_dev(ctx, m, n, row_ptr, blk, words, R, ldr_w, rank, pivots, bad_row);
	// host rows: validate (word indices increasing and in range) and stage
	const uint64_t nnz = row_ptr[m];
	for (size_t i = 0; i < m; ++i) {
		if (row_ptr[i + 1] < row_ptr[i])
			return set_error(ctx, MLZ_E_ARG, "row_ptr not monotone at row %zu", i);
		for (uint64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
			if (blk[e] >= nw || (e > row_ptr[i] && blk[e] <= blk[e - 1]))
				return set_error(ctx, MLZ_E_ARG, "row %zu: word indices not increasing or out of range", i);
	}
	void *a, *b, *c, *dR = nullptr;
	MLZ_TRY(ws_get(ctx, WS_STAGE_A, (m + 1) * 8, &a));
	MLZ_TRY(ws_get(ctx, WS_STAGE_B, nnz * 4 + 8, &b));
	MLZ_TRY(ws_get(ctx, WS_STAGE_C, nnz * 8 + 8, &c));
	MLZ_HIP(ctx, hipMemcpyAsync(a, row_ptr, (m + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
	if (nnz) {
		MLZ_HIP(ctx, hipMemcpyAsync(b, blk, nnz * 4, hipMemcpyHostToDevice, ctx->stream));
		MLZ_HIP(ctx, hipMemcpyAsync(c, words, nnz * 8, hipMemcpyHostToDevice, ctx->stream));
	}
	if (R)
		MLZ_TRY(stage_in(ctx, WS_BATCH_0, R, m, nw, ldr_w, 8, &dR));
	MLZ_TRY(gf2_fl_dev(ctx, m, n, (const uint64_t *) a, (const uint32_t *) b, (const uint64_t *) c, (uint64_t *) dR,
			   nw, rank, pivots, bad_row));
	if (R)
		return stage_out(ctx, R, dR, m, nw, ldr_w, 8);
	return MLZ_OK;
}
