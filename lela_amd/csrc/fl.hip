// lela_amd/csrc/fl.hip -- Faugere-Lachartre reduction of an F4-shaped matrix
// to reduced row-echelon form on MI355X.
//
// Follows FaugereLachartre::echelonize(R, X, rank, det, reduced=true)
// (lela/algorithms/faugere-lachartre.tcc:735-917):
//   1. row scan of setup_splicer (:47-146): each row's head (first nonzero
//      column); a head larger than the last starts a pivot row (A|B), a
//      repeated head marks a C|D row, a smaller head is WrongMatrixForm(row),
//      the first zero row ends the scan; det *= every scanned head (:119).
//      This scan is sequential and O(m), so it runs on the host from the
//      heads a kernel computes.
//   2. MatrixGrid1 split (:148-275) into dense A (pivot rows x pivot columns,
//      upper triangular), B, C, D by one scatter kernel over the nonzeros.
//   3. B <- A^-1 B   (:791, upper, non-unit): blocked back-substitution that
//      skips the zero columns right of each block-row's last nonzero (the
//      pivot block of an F4 matrix is sparse; LELA hands trsm a sparse A);
//      diagonal blocks by zp_trsm.hip, off-diagonal slabs by zp_gemm.hip.
//   4. D <- D - C B  (:805, Schur complement; zp_gemm.hip)
//   5. D <- RREF(D)  (:822-852; zp_rref.hip) -- LELA takes the non-reduced
//      echelon form and reduces it with D1^-1 D2 (:861); RREF(D) is that
//      result directly.  When det is asked for, the non-reduced form's heads
//      (setup_splicer on D, :833) come from zp_lu.hip's elimination of D with
//      LELA's transposition pivoting: for a dense D, EchelonForm::echelonize
//      runs GaussJordan::echelonize and move_L (solutions/echelon-form.h:
//      124-131), whose result LELA's own test asserts equal to
//      Elimination::echelonize's (tests/test-gauss-jordan.C:92-99).
//   6. B <- B - B[:, J_D] RREF(D)  (:883-892, B2 - B1 D1^-1 D2)
//   7. assembly of the standard-layout RREF (MatrixGrid3, :901; :363-504).
// det (when asked for): the product of the heads of the X scan (:119) and of
// the rows of D's non-reduced echelon form (:833).
#include "common.hpp"

#include <algorithm>

namespace mlz {

// head[i] = first column whose value is nonzero mod p, -1 for a zero row;
// hval[i] = that value (canonical).
template <int E>
__global__ void fl_heads_kernel(const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col,
				const typename Elem<E>::T *__restrict__ val, int m, uint32_t p, int *head,
				uint32_t *hval)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= m)
		return;
	int h = -1;
	uint32_t hv = 0;
	for (uint64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
		uint32_t v = Elem<E>::load(val + e, p) % p;
		if (v != 0) {
			h = (int) col[e];
			hv = v;
			break;
		}
	}
	head[i] = h;
	hval[i] = hv;
}

// X[0 .. count) = 0, four entries per thread (X is 16-byte aligned).
__global__ void fl_zero_kernel(uint32_t *__restrict__ X, size_t count)
{
	const size_t i = 4 * ((size_t) blockIdx.x * blockDim.x + threadIdx.x);
	if (i + 4 <= count)
		*(uint4 *) (X + i) = make_uint4(0u, 0u, 0u, 0u);
	else
		for (size_t q = i; q < count; ++q)
			X[q] = 0;
}

// rowmap[i] = index among pivot rows (>= 0) or -(1 + index among the others);
// colmap[j] likewise for columns.  A: r x r, B: r x nr, C: mr x r, D: mr x nr.
template <int E>
__global__ void fl_scatter_kernel(const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col,
				  const typename Elem<E>::T *__restrict__ val, int m, uint32_t p,
				  const int *__restrict__ rowmap, const int *__restrict__ colmap, int r, int nr,
				  uint32_t *__restrict__ A, uint32_t *__restrict__ B, uint32_t *__restrict__ C,
				  uint32_t *__restrict__ D)
{
	const int i = blockIdx.x;
	if (i >= m)
		return;
	const int ri = rowmap[i];
	for (uint64_t e = row_ptr[i] + threadIdx.x; e < row_ptr[i + 1]; e += blockDim.x) {
		const uint32_t v = Elem<E>::load(val + e, p) % p;
		const int cj = colmap[col[e]];
		if (ri >= 0) {
			if (cj >= 0)
				A[(size_t) ri * r + cj] = v;
			else
				B[(size_t) ri * nr + (-1 - cj)] = v;
		} else {
			if (cj >= 0)
				C[(size_t) (-1 - ri) * r + cj] = v;
			else
				D[(size_t) (-1 - ri) * nr + (-1 - cj)] = v;
		}
	}
}

// last[ri] = largest A-column index with a nonzero entry in pivot row ri (its
// head at least, since A is upper triangular with the heads on the diagonal).
// One wave per row (FL_LASTCOL_ROWS rows per workgroup): the lanes stride
// over the row's entries and a shuffle reduction takes the maximum.
constexpr int FL_LASTCOL_ROWS = 4;
template <int E>
__global__ __launch_bounds__(64 * FL_LASTCOL_ROWS) void fl_lastcol_kernel(
	const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col, const typename Elem<E>::T *__restrict__ val,
	int m, uint32_t p, const int *__restrict__ rowmap, const int *__restrict__ colmap, int *__restrict__ last)
{
	const int i = blockIdx.x * FL_LASTCOL_ROWS + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	if (i >= m)
		return;   // i and ri are uniform across the wave
	const int ri = rowmap[i];
	if (ri < 0)
		return;
	int mx = ri;
	for (uint64_t e = row_ptr[i] + lane; e < row_ptr[i + 1]; e += 64) {
		const int cj = colmap[col[e]];
		if (cj > mx && Elem<E>::load(val + e, p) % p != 0)
			mx = cj;
	}
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1)
		mx = max(mx, __shfl_xor(mx, off, 64));
	if (lane == 0)
		last[ri] = mx;
}

// B[i][j0 .. j0 + w) = 0 for every row i < r (grid.y strides over the rows).
__global__ void fl_zero_cols_kernel(uint32_t *__restrict__ B, int r, int nr, int j0, int w)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= w)
		return;
	for (int i = blockIdx.y; i < r; i += gridDim.y)
		B[(size_t) i * nr + j0 + j] = 0u;
}

// G[i][s] = B[i][jd[s]]  (r x rd); x covers the columns s, grid.y strides
// over the rows i.
__global__ void fl_gather_cols_kernel(const uint32_t *__restrict__ B, int r, int nr, const int *__restrict__ jd,
				      int rd, uint32_t *__restrict__ G)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= rd)
		return;
	const int js = jd[s];
	for (int i = blockIdx.y; i < r; i += gridDim.y)
		G[(size_t) i * rd + s] = B[(size_t) i * nr + js];
}

// Output row t: src[t] >= 0 -> X pivot row src[t] (B row), pivot column
// pcol[t]; src[t] < 0 -> D pivot row -1-src[t] (RREF(D) row); t >= rank -> 0.
template <int E, bool VEC>
__global__ void fl_assemble_kernel(const uint32_t *__restrict__ B, const uint32_t *__restrict__ D, int nr,
				   const int *__restrict__ colmap, const int *__restrict__ src,
				   const int *__restrict__ pcol, int rank, int m, int n, uint32_t p,
				   typename Elem<E>::T *__restrict__ R, size_t ldr)
{
	// four consecutive columns per thread; VEC (R and ldr aligned to four
	// entries) writes them as one vector store
	typedef typename Elem<E>::T T;
	typedef T T4 __attribute__((ext_vector_type(4)));
	const int j0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
	if (j0 >= n)
		return;
	int cj[4];
#pragma unroll
	for (int q = 0; q < 4; ++q)
		cj[q] = j0 + q < n ? colmap[j0 + q] : 0;
	for (int t = blockIdx.y; t < m; t += gridDim.y) {
		uint32_t v[4] = {0u, 0u, 0u, 0u};
		if (t < rank) {
			const int s = src[t], pc = pcol[t];
			const uint32_t *row = s >= 0 ? B + (size_t) s * nr : D + (size_t) (-1 - s) * nr;
#pragma unroll
			for (int q = 0; q < 4; ++q)
				if (j0 + q < n)
					v[q] = cj[q] >= 0 ? ((s >= 0 && j0 + q == pc) ? 1u : 0u) : row[-1 - cj[q]];
		}
		T *out = R + (size_t) t * ldr + j0;
		if (VEC && j0 + 4 <= n) {
			T o[4];
#pragma unroll
			for (int q = 0; q < 4; ++q)
				Elem<E>::store(&o[q], v[q], p);   // B and D hold residues below p
			*(T4 *) out = T4{o[0], o[1], o[2], o[3]};
		} else {
#pragma unroll
			for (int q = 0; q < 4; ++q)
				if (j0 + q < n)
					Elem<E>::store(out + q, v[q], p);
		}
	}
}

// The compact result: row t of R on the non-pivot columns (B/D columns
// bdcol[0 .. nf)), rows 0 .. rank-1, four columns per thread.
template <int E>
__global__ void fl_compact_kernel(const uint32_t *__restrict__ B, const uint32_t *__restrict__ D, int nr,
				  const int *__restrict__ bdcol, int nf, const int *__restrict__ src, int rank, uint32_t p,
				  typename Elem<E>::T *__restrict__ NP, size_t ldnp)
{
	const int f0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
	if (f0 >= nf)
		return;
	int cb[4];
#pragma unroll
	for (int q = 0; q < 4; ++q)
		cb[q] = f0 + q < nf ? bdcol[f0 + q] : 0;
	for (int t = blockIdx.y; t < rank; t += gridDim.y) {
		const int s = src[t];
		const uint32_t *row = s >= 0 ? B + (size_t) s * nr : D + (size_t) (-1 - s) * nr;
#pragma unroll
		for (int q = 0; q < 4; ++q)
			if (f0 + q < nf)
				Elem<E>::store(NP + (size_t) t * ldnp + f0 + q, row[cb[q]], p);
	}
}

// Row groups of the assembly launch: each workgroup walks m / FL_ASM_ROWS
// output rows of its 1024 columns (fewer, longer workgroups than one per row).
constexpr size_t FL_ASM_ROWS = 1024;

constexpr size_t FL_TRSM_BLOCK = 512;   // rows per block of the pivot-block solve (multiple of TRSM_LEAF_ROWS)
static_assert(FL_TRSM_BLOCK % TRSM_LEAF_ROWS == 0, "solve blocks must align with the inverted diagonal blocks");

struct FlResult {
	size_t rank;
	uint64_t det;
	size_t bad_row;
	size_t nf;          // n - rank: columns of the compact block
};

// D's echelon rows 0 .. rd-1 split by a column map: cmap[j] = s >= 0 puts
// column j at D1[:, s] (a pivot column), -1 - f at D2[:, f].
__global__ void fl_split_d_kernel(const uint32_t *__restrict__ D, size_t ldd, int nr, const int *__restrict__ cmap,
				  uint32_t *__restrict__ D1, size_t ld1, uint32_t *__restrict__ D2, size_t ld2)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
	if (j >= nr)
		return;
	const uint32_t v = D[(size_t) t * ldd + j];
	const int s = cmap[j];
	if (s >= 0)
		D1[(size_t) t * ld1 + s] = v;
	else
		D2[(size_t) t * ld2 + (-1 - s)] = v;
}

// Row t of the reduced form: the unit vector on pivot column t, D1^-1 D2 on the others.
__global__ void fl_join_d_kernel(uint32_t *__restrict__ D, size_t ldd, int nr, const int *__restrict__ cmap,
				 const uint32_t *__restrict__ D2, size_t ld2)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
	if (j >= nr)
		return;
	const int s = cmap[j];
	D[(size_t) t * ldd + j] = s >= 0 ? (s == t ? 1u : 0u) : D2[(size_t) t * ld2 + (-1 - s)];
}

// D (rd echelon rows of nr columns, pivot columns J) <- its reduced form:
// D2 <- D1^-1 D2 with D1 = D[:, J] upper triangular (:852-861).  The
// non-reduced rows past rd are zero already.
static int fl_reduce_d(mlz_ctx *ctx, uint32_t p, size_t rd, size_t nr, uint32_t *D, const uint32_t *J, int algo)
{
	const size_t nf = nr - rd;
	std::vector<int> cmap(nr);
	for (size_t j = 0, f = 0, s = 0; j < nr; ++j) {
		if (s < rd && J[s] == j)
			cmap[j] = (int) s++;
		else
			cmap[j] = -1 - (int) f++;
	}
	void *wa, *wm;
	MLZ_TRY(ws_get(ctx, WS_FL_3, (rd * rd + rd * nf) * 4 + 512, &wa));
	MLZ_TRY(ws_get(ctx, WS_FL_6, nr * sizeof(int) + 256, &wm));
	uint32_t *D1 = (uint32_t *) wa, *D2 = D1 + round_up(rd * rd, 64);
	MLZ_HIP(ctx, hipMemcpyAsync(wm, cmap.data(), nr * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	dim3 g(ceil_div(nr, 256), (unsigned) rd);
	fl_split_d_kernel<<<g, 256, 0, ctx->stream>>>(D, nr, (int) nr, (const int *) wm, D1, rd, D2, nf ? nf : 1);
	MLZ_TRY(check_launch(ctx, "fl_split_d_kernel"));
	if (nf) {
		size_t bad = 0;
		MLZ_TRY(zp_trsm_dev(ctx, MLZ_U32, p, 1, 0, rd, nf, 1, D1, rd, D2, nf, &bad, algo));
	}
	fl_join_d_kernel<<<g, 256, 0, ctx->stream>>>(D, nr, (int) nr, (const int *) wm, D2, nf ? nf : 1);
	MLZ_TRY(check_launch(ctx, "fl_join_d_kernel"));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));   // cmap leaves scope
	return MLZ_OK;
}

template <int E>
static int fl_run(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, const uint64_t *d_rowptr, const uint32_t *d_col,
		  const void *d_val, void *R, size_t ldr, uint32_t *pivots_out, FlResult *res, int algo, bool want_det,
		  void *NP = nullptr, size_t np_cap = 0, void **np_dev = nullptr)
{
	typedef typename Elem<E>::T T;
	const ModP mp = make_modp(p);
	// 1. heads (device) -> row scan (host)
	void *w;
	MLZ_TRY(ws_get(ctx, WS_FL_0, (2 * m + n + 16) * sizeof(int) + m * sizeof(uint32_t), &w));
	int *d_head = (int *) w;
	uint32_t *d_hval = (uint32_t *) (d_head + m);
	int *d_rowmap = (int *) (d_hval + m);
	int *d_colmap = d_rowmap + m;
	fl_heads_kernel<E><<<ceil_div(m, 256), 256, 0, ctx->stream>>>(d_rowptr, d_col, (const T *) d_val, (int) m, p,
								       d_head, d_hval);
	MLZ_TRY(check_launch(ctx, "fl_heads_kernel"));
	std::vector<int> head(m);
	std::vector<uint32_t> hval(m);
	MLZ_HIP(ctx, hipMemcpyAsync(head.data(), d_head, m * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(hval.data(), d_hval, m * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));

	std::vector<char> is_piv_row(m, 0), is_piv_col(n, 0);
	uint64_t det = 1 % p;
	long last = -1;
	size_t r = 0;
	for (size_t i = 0; i < m; ++i) {
		const long h = head[i];
		if (h < 0)
			break;
		if (h < last) {
			res->bad_row = i;
			return set_error(ctx, MLZ_E_WRONG_FORM, "Bad input-matrix for Faugere-Lachartre at row-index %zu", i);
		}
		if (h > last) {
			is_piv_row[i] = 1;
			is_piv_col[h] = 1;
			last = h;
			++r;
		}
		det = mp.mul((uint32_t) det, hval[i]);
	}
	res->det = det;
	const size_t mr = m - r, nr = n - r;
	std::vector<int> rowmap(m), colmap(n), piv_cols_x;
	piv_cols_x.reserve(r);
	std::vector<int> res_cols;
	res_cols.reserve(nr);
	for (size_t i = 0, a = 0, b = 0; i < m; ++i)
		rowmap[i] = is_piv_row[i] ? (int) a++ : -1 - (int) b++;
	for (size_t j = 0, a = 0, b = 0; j < n; ++j) {
		if (is_piv_col[j]) {
			colmap[j] = (int) a++;
			piv_cols_x.push_back((int) j);
		} else {
			colmap[j] = -1 - (int) b++;
			res_cols.push_back((int) j);
		}
	}
	MLZ_HIP(ctx, hipMemcpyAsync(d_rowmap, rowmap.data(), m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(d_colmap, colmap.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));

	// 2. dense blocks
	void *wa, *wb, *wc, *wd;
	MLZ_TRY(ws_get(ctx, WS_FL_1, r * r * 4 + 256, &wa));
	MLZ_TRY(ws_get(ctx, WS_FL_2, r * nr * 4 + 256, &wb));
	MLZ_TRY(ws_get(ctx, WS_FL_3, mr * r * 4 + 256, &wc));
	MLZ_TRY(ws_get(ctx, WS_FL_4, mr * nr * 4 + 256, &wd));
	uint32_t *dA = (uint32_t *) wa, *dB = (uint32_t *) wb, *dC = (uint32_t *) wc, *dD = (uint32_t *) wd;
	{
		ProfScope ps(ctx, "fl_split");
		const size_t cnt[4] = {r * r, r * nr, mr * r, mr * nr};
		uint32_t *bufs[4] = {dA, dB, dC, dD};
		for (int q = 0; q < 4; ++q)
			if (cnt[q])
				fl_zero_kernel<<<ceil_div(ceil_div(cnt[q], 4), 256), 256, 0, ctx->stream>>>(bufs[q], cnt[q]);
		fl_scatter_kernel<E><<<(unsigned) m, 128, 0, ctx->stream>>>(d_rowptr, d_col, (const T *) d_val, (int) m, p,
									   d_rowmap, d_colmap, (int) r, (int) nr, dA, dB,
									   dC, dD);
		MLZ_TRY(check_launch(ctx, "fl_scatter_kernel"));
	}
	// 3. B <- A^-1 B, as blocked back-substitution over FL_TRSM_BLOCK-row
	//    blocks from the bottom: B_I <- B_I - A[I, I+1..c1) X[I+1..c1), then
	//    X_I = A_II^-1 B_I, where c1 is one past the last nonzero column of
	//    block-row I.  F4 pivot blocks are sparse and banded, so the columns
	//    right of c1 (known zeros) are skipped; for a dense A this is the
	//    ordinary blocked solve.
	if (r && nr) {
		void *wl;
		MLZ_TRY(ws_get(ctx, WS_FL_5, r * sizeof(int) + 256, &wl));
		fl_lastcol_kernel<E><<<ceil_div(m, (size_t) FL_LASTCOL_ROWS), 64 * FL_LASTCOL_ROWS, 0, ctx->stream>>>(
			d_rowptr, d_col, (const T *) d_val, (int) m, p, d_rowmap, d_colmap, (int *) wl);
		MLZ_TRY(check_launch(ctx, "fl_lastcol_kernel"));
		std::vector<int> last(r);
		MLZ_HIP(ctx, hipMemcpyAsync(last.data(), wl, r * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
		MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
		// every diagonal block's inverse in one launch, and LELA's
		// non-invertible-head check over all of A
		const uint32_t *Winv = nullptr;
		{
			size_t bad = 0;
			const int rc = zp_trsm_prepare(ctx, MLZ_U32, p, 1, 0, r, dA, r, &bad, &Winv);
			if (rc != MLZ_OK) {
				res->rank = r;
				return rc;
			}
		}
		const size_t nb = (r + FL_TRSM_BLOCK - 1) / FL_TRSM_BLOCK;
		const bool i8 = p <= 65536 && (algo == MLZ_ALGO_AUTO || algo == MLZ_ALGO_INT8) && r <= 65504;
		if (i8) {
			// 64-row blocks on INT8 MFMA, each block's band ending one past
			// its rows' last nonzero column
			const size_t nb64 = (r + 63) / 64;
			std::vector<int> c1(nb64);
			for (size_t I = 0; I < nb64; ++I) {
				const size_t i0 = 64 * I, i1 = std::min(r, i0 + 64);
				size_t c = i1;
				for (size_t i = i0; i < i1; ++i)
					c = std::max(c, (size_t) last[i] + 1);
				c1[I] = (int) c;
			}
			ProfScope ps(ctx, "fl_solve");
			// from two 1024-row blocks on: their inverses first, then one band
			// gemm and one inverse gemm per block (15 dependent steps instead
			// of 235 at 15000 rows)
			if (r >= 2048)
				MLZ_TRY(zp_trsm_banded_inv_i8(ctx, p, r, nr, dA, r, dB, nr, Winv, c1.data()));
			else
				MLZ_TRY(zp_trsm_blocked_i8(ctx, MLZ_U32, p, 1, r, nr, 1, dA, r, dB, nr, Winv, c1.data()));
		}
		for (size_t I = i8 ? 0 : nb; I-- > 0;) {
			const size_t i0 = I * FL_TRSM_BLOCK, i1 = std::min(r, i0 + FL_TRSM_BLOCK);
			size_t c1 = i1;
			for (size_t i = i0; i < i1; ++i)
				c1 = std::max(c1, (size_t) last[i] + 1);
			if (c1 > i1)
				MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, i1 - i0, nr, c1 - i1, p - 1, dA + i0 * r + i1, r,
						    dB + i1 * nr, nr, 1, dB + i0 * nr, nr, algo));
			MLZ_TRY(zp_trsm_apply(ctx, MLZ_U32, p, 1, 0, i1 - i0, nr, 1, dA + i0 * r + i0, r, dB + i0 * nr, nr,
					      Winv + i0 * TRSM_LEAF_ROWS, algo));
		}
	}
	// 4. D <- D - C B
	if (mr && nr && r)
		MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, mr, nr, r, p - 1, dC, r, dB, nr, 1, dD, nr, algo));
	// 5. RREF(D).  With det: LELA's way, the non-reduced echelon form (its
	// heads enter det, :833) reduced by D1^-1 D2 (:852-861); else directly.
	size_t rd = 0;
	std::vector<uint32_t> jd(std::min(mr, nr) + 1);
	if (mr && nr && want_det) {
		uint64_t dl = 1;
		MLZ_TRY(zp_echelon_dev(ctx, MLZ_U32, p, mr, nr, dD, nr, MLZ_ECH_U, &rd, &dl, nullptr, nullptr, algo, true,
				       jd.data()));
		res->det = mp.mul((uint32_t) res->det, (uint32_t) dl);
		if (rd)
			MLZ_TRY(fl_reduce_d(ctx, p, rd, nr, dD, jd.data(), algo));
	} else if (mr && nr) {
		MLZ_TRY(zp_rref_dev(ctx, MLZ_U32, p, mr, nr, dD, nr, &rd, jd.data(), algo));
	}
	// 6. B <- B - B[:, J_D] RREF(D)[0:rd]
	if (r && rd) {
		void *wg, *wj;
		MLZ_TRY(ws_get(ctx, WS_FL_5, r * rd * 4 + 256, &wg));
		MLZ_TRY(ws_get(ctx, WS_FL_6, rd * sizeof(int) + 256, &wj));
		std::vector<int> jdi(jd.begin(), jd.begin() + rd);
		MLZ_HIP(ctx, hipMemcpyAsync(wj, jdi.data(), rd * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
		fl_gather_cols_kernel<<<dim3(ceil_div(rd, 256), (unsigned) std::min(r, FL_ASM_ROWS)), 256, 0, ctx->stream>>>(
			dB, (int) r, (int) nr, (const int *) wj, (int) rd, (uint32_t *) wg);
		MLZ_TRY(check_launch(ctx, "fl_gather_cols_kernel"));
		// On D's pivot columns the product is B's own column (RREF(D) has the
		// identity there), so those columns of B become zero.  When they are
		// one contiguous range (D of full row rank has them first), only the
		// other columns go through the gemm and the range is cleared.
		const size_t j0 = jd[0];
		if (jd[rd - 1] - j0 == rd - 1) {
			if (j0)
				MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, r, j0, rd, p - 1, wg, rd, dD, nr, 1, dB, nr, algo));
			if (j0 + rd < nr)
				MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, r, nr - j0 - rd, rd, p - 1, wg, rd, dD + j0 + rd, nr, 1,
						    dB + j0 + rd, nr, algo));
			fl_zero_cols_kernel<<<dim3(ceil_div(rd, 256), std::min(r, (size_t) 65535)), 256, 0, ctx->stream>>>(
				dB, (int) r, (int) nr, (int) j0, (int) rd);
			MLZ_TRY(check_launch(ctx, "fl_zero_cols_kernel"));
		} else {
			MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, r, nr, rd, p - 1, wg, rd, dD, nr, 1, dB, nr, algo));
		}
		MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));   // jdi leaves scope
	}
	// 7. assembly: merge the X pivots and the D pivots by column
	const size_t rank = r + rd;
	std::vector<int> src(rank + 1), pcol(rank + 1);
	{
		size_t ia = 0, id = 0, t = 0;
		while (ia < r || id < rd) {
			const long ca = ia < r ? piv_cols_x[ia] : (long) n + 1;
			const long cd = id < rd ? res_cols[jd[id]] : (long) n + 1;
			if (ca < cd) {
				src[t] = (int) ia;
				pcol[t] = (int) ca;
				++ia;
			} else {
				src[t] = -1 - (int) id;
				pcol[t] = (int) cd;
				++id;
			}
			++t;
		}
	}
	void *wm;
	MLZ_TRY(ws_get(ctx, WS_FL_7, 2 * (rank + 1) * sizeof(int), &wm));
	int *d_src = (int *) wm, *d_pcol = d_src + rank + 1;
	MLZ_HIP(ctx, hipMemcpyAsync(d_src, src.data(), (rank + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(d_pcol, pcol.data(), (rank + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	if (R) {
		ProfScope ps(ctx, "fl_assemble");
		const dim3 grid(ceil_div(ceil_div(n, 4), 256), (unsigned) std::min<size_t>(m, FL_ASM_ROWS));
		if ((uintptr_t) R % (4 * sizeof(T)) == 0 && ldr % 4 == 0)
			fl_assemble_kernel<E, true><<<grid, 256, 0, ctx->stream>>>(dB, dD, (int) nr, d_colmap, d_src, d_pcol,
										 (int) rank, (int) m, (int) n, p, (T *) R, ldr);
		else
			fl_assemble_kernel<E, false><<<grid, 256, 0, ctx->stream>>>(dB, dD, (int) nr, d_colmap, d_src, d_pcol,
										  (int) rank, (int) m, (int) n, p, (T *) R, ldr);
		MLZ_TRY(check_launch(ctx, "fl_assemble_kernel"));
	}
	res->nf = n - rank;
	if (NP || np_dev) {
		// the non-pivot columns in B/D's numbering: all but D's pivot columns
		const size_t nf = n - rank;
		if (rank * nf > np_cap) {
			MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
			res->rank = rank;
			return set_error(ctx, MLZ_E_ARG, "the compact block needs %zu x %zu entries, %zu given", rank, nf,
					 np_cap);
		}
		std::vector<int> bdcol;
		bdcol.reserve(nf + 1);
		for (size_t b = 0, q = 0; b < nr; ++b) {
			if (q < rd && jd[q] == b)
				++q;
			else
				bdcol.push_back((int) b);
		}
		if (np_dev)
			MLZ_TRY(ws_get(ctx, WS_BATCH_1, rank * nf * sizeof(T) + 256, np_dev));
		T *dst = np_dev ? (T *) *np_dev : (T *) NP;
		void *wb;
		MLZ_TRY(ws_get(ctx, WS_FL_6, (nf + 1) * sizeof(int), &wb));
		MLZ_HIP(ctx, hipMemcpyAsync(wb, bdcol.data(), nf * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
		if (nf && rank) {
			ProfScope ps(ctx, "fl_compact");
			const dim3 grid(ceil_div(ceil_div(nf, 4), 256), (unsigned) std::min<size_t>(rank, FL_ASM_ROWS));
			fl_compact_kernel<E><<<grid, 256, 0, ctx->stream>>>(dB, dD, (int) nr, (const int *) wb, (int) nf, d_src,
									   (int) rank, p, dst, nf);
			MLZ_TRY(check_launch(ctx, "fl_compact_kernel"));
		}
	}
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	if (pivots_out)
		for (size_t t = 0; t < rank; ++t)
			pivots_out[t] = (uint32_t) pcol[t];
	res->rank = rank;
	return MLZ_OK;
}

}  // namespace mlz

using namespace mlz;

extern "C" int mlz_zp_fl_echelonize(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n,
				    const uint64_t *row_ptr, const uint32_t *col, const void *val, void *R,
				    size_t ldr, size_t *rank, uint32_t *pivots, uint64_t *det, size_t *bad_row,
				    int algo)
{
	if (!ctx)
		return MLZ_E_ARG;
	if (p < 2 || (elem == MLZ_U16 && p > 65535) || (elem == MLZ_U32 && p > 0xffffffffull) ||
	    (elem == MLZ_F64_SYM && p >= 67108864ull) || elem_size(elem) == 0)
		return set_error(ctx, MLZ_E_ARG, "bad modulus or element type");
	if (!row_ptr || (m && n && R && ldr < n))
		return set_error(ctx, MLZ_E_ARG, "bad CSR or output pointer");
	if (m > (size_t) INT32_MAX / 2 || n > (size_t) INT32_MAX / 2)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	FlResult res = {0, 1 % p, 0, 0};
	int rc;
	const size_t es = elem_size(elem);
	if (m == 0 || n == 0) {
		if (rank) *rank = 0;
		if (det) *det = 1 % p;
		return MLZ_OK;
	}
	const uint64_t *d_rowptr;
	const uint32_t *d_col;
	const void *d_val;
	void *dR = R;
	if (ctx->ptr_mode == MLZ_PTR_DEVICE) {
		d_rowptr = row_ptr;
		d_col = col;
		d_val = val;
	} else {
		// host CSR: validate (sorted, in-range columns) and stage
		const uint64_t nnz = row_ptr[m];
		for (size_t i = 0; i < m; ++i) {
			if (row_ptr[i + 1] < row_ptr[i])
				return set_error(ctx, MLZ_E_ARG, "row_ptr not monotone at row %zu", i);
			for (uint64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
				if (col[e] >= n || (e > row_ptr[i] && col[e] <= col[e - 1]))
					return set_error(ctx, MLZ_E_ARG, "row %zu: columns not sorted or out of range", i);
		}
		void *a, *b, *c;
		MLZ_TRY(ws_get(ctx, WS_STAGE_A, (m + 1) * 8, &a));
		MLZ_TRY(ws_get(ctx, WS_STAGE_B, nnz * 4 + 8, &b));
		MLZ_TRY(ws_get(ctx, WS_STAGE_C, nnz * es + 8, &c));
		MLZ_HIP(ctx, hipMemcpyAsync(a, row_ptr, (m + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
		if (nnz) {
			MLZ_HIP(ctx, hipMemcpyAsync(b, col, nnz * 4, hipMemcpyHostToDevice, ctx->stream));
			MLZ_HIP(ctx, hipMemcpyAsync(c, val, nnz * es, hipMemcpyHostToDevice, ctx->stream));
		}
		d_rowptr = (const uint64_t *) a;
		d_col = (const uint32_t *) b;
		d_val = c;
		dR = nullptr;
		if (R)
			MLZ_TRY(ws_get(ctx, WS_BATCH_0, m * n * es, &dR));
	}
	const uint32_t pp = (uint32_t) p;
	const size_t ldr_dev = ctx->ptr_mode == MLZ_PTR_DEVICE ? ldr : n;
	if (elem == MLZ_U16)
		rc = fl_run<MLZ_U16>(ctx, pp, m, n, d_rowptr, d_col, d_val, dR, ldr_dev, pivots, &res, algo, det != nullptr);
	else if (elem == MLZ_U32)
		rc = fl_run<MLZ_U32>(ctx, pp, m, n, d_rowptr, d_col, d_val, dR, ldr_dev, pivots, &res, algo, det != nullptr);
	else
		rc = fl_run<MLZ_F64_SYM>(ctx, pp, m, n, d_rowptr, d_col, d_val, dR, ldr_dev, pivots, &res, algo, det != nullptr);
	if (rank) *rank = res.rank;
	if (det) *det = res.det;
	if (bad_row) *bad_row = res.bad_row;
	if (rc != MLZ_OK)
		return rc;
	if (ctx->ptr_mode != MLZ_PTR_DEVICE && R)
		return stage_out(ctx, R, dR, m, n, ldr, es);
	return MLZ_OK;
}

extern "C" int mlz_zp_fl_echelonize_compact(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n,
					    const uint64_t *row_ptr, const uint32_t *col, const void *val, void *NP,
					    size_t np_capacity, size_t *rank, uint32_t *pivots, uint64_t *det,
					    size_t *bad_row, int algo)
{
	if (!ctx)
		return MLZ_E_ARG;
	if (p < 2 || (elem == MLZ_U16 && p > 65535) || (elem == MLZ_U32 && p > 0xffffffffull) ||
	    (elem == MLZ_F64_SYM && p >= 67108864ull) || elem_size(elem) == 0)
		return set_error(ctx, MLZ_E_ARG, "bad modulus or element type");
	if (!row_ptr || !NP)
		return set_error(ctx, MLZ_E_ARG, "bad CSR or output pointer");
	if (m > (size_t) INT32_MAX / 2 || n > (size_t) INT32_MAX / 2)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	FlResult res = {0, 1 % p, 0, 0};
	if (m == 0 || n == 0) {
		if (rank) *rank = 0;
		if (det) *det = 1 % p;
		return MLZ_OK;
	}
	const size_t es = elem_size(elem);
	const uint64_t *d_rowptr;
	const uint32_t *d_col;
	const void *d_val;
	const bool dev = ctx->ptr_mode == MLZ_PTR_DEVICE;
	if (dev) {
		d_rowptr = row_ptr;
		d_col = col;
		d_val = val;
	} else {
		const uint64_t nnz = row_ptr[m];
		for (size_t i = 0; i < m; ++i) {
			if (row_ptr[i + 1] < row_ptr[i])
				return set_error(ctx, MLZ_E_ARG, "row_ptr not monotone at row %zu", i);
			for (uint64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
				if (col[e] >= n || (e > row_ptr[i] && col[e] <= col[e - 1]))
					return set_error(ctx, MLZ_E_ARG, "row %zu: columns not sorted or out of range", i);
		}
		void *a, *b, *c;
		MLZ_TRY(ws_get(ctx, WS_STAGE_A, (m + 1) * 8, &a));
		MLZ_TRY(ws_get(ctx, WS_STAGE_B, nnz * 4 + 8, &b));
		MLZ_TRY(ws_get(ctx, WS_STAGE_C, nnz * es + 8, &c));
		MLZ_HIP(ctx, hipMemcpyAsync(a, row_ptr, (m + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
		if (nnz) {
			MLZ_HIP(ctx, hipMemcpyAsync(b, col, nnz * 4, hipMemcpyHostToDevice, ctx->stream));
			MLZ_HIP(ctx, hipMemcpyAsync(c, val, nnz * es, hipMemcpyHostToDevice, ctx->stream));
		}
		d_rowptr = (const uint64_t *) a;
		d_col = (const uint32_t *) b;
		d_val = c;
	}
	const uint32_t pp = (uint32_t) p;
	void *npd = nullptr;
	void **np_dev = dev ? nullptr : &npd;
	void *np = dev ? NP : nullptr;
	int rc;
	if (elem == MLZ_U16)
		rc = fl_run<MLZ_U16>(ctx, pp, m, n, d_rowptr, d_col, d_val, nullptr, 0, pivots, &res, algo, det != nullptr,
				     np, np_capacity, np_dev);
	else if (elem == MLZ_U32)
		rc = fl_run<MLZ_U32>(ctx, pp, m, n, d_rowptr, d_col, d_val, nullptr, 0, pivots, &res, algo, det != nullptr,
				     np, np_capacity, np_dev);
	else
		rc = fl_run<MLZ_F64_SYM>(ctx, pp, m, n, d_rowptr, d_col, d_val, nullptr, 0, pivots, &res, algo,
					 det != nullptr, np, np_capacity, np_dev);
	if (rank) *rank = res.rank;
	if (det) *det = res.det;
	if (bad_row) *bad_row = res.bad_row;
	if (rc != MLZ_OK)
		return rc;
	if (!dev && res.rank * res.nf)
		MLZ_HIP(ctx, hipMemcpy(NP, npd, res.rank * res.nf * es, hipMemcpyDeviceToHost));
	return MLZ_OK;
}
