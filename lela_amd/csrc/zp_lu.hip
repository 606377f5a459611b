// lela_amd/csrc/zp_lu.hip -- non-reduced row-echelon forms over Z/p with
// LELA's transposition pivoting, on MI355X.
//
// Replaces Elimination::echelonize (lela/algorithms/elimination.tcc:47-139,
// with and without compute_L) and Elimination::pluq (:266-370) for dense
// matrices with DensePivotStrategy (lela/algorithms/pivot-strategy.tcc:19-41):
// step i takes the first column (from the previous pivot's on) in which some
// row at a position >= i is nonzero, the first such row as pivot, and swaps
// it with position i.  Unlike the reduced form, the result depends on this
// order, so it is reproduced exactly; the arithmetic is blocked:
//
// Rows stay where they are in a working copy; perm[pos] names the physical
// row at each position, and a transposition swaps two entries of perm.  Per
// panel of up to LU_W columns from column c at position i:
//  1. one workgroup loads the panel part of the LU_K rows at positions
//     i .. i+LU_K-1 (the window), augmented with the identity, and runs the
//     elimination above on them column by column (the first nonzero row of
//     the window is the first nonzero row overall, because the window holds
//     the first positions in order).  The identity part becomes T, the
//     transform of the window rows: the new pivot rows are T_piv X_w.  A
//     backward pass over the pivot rows turns T_piv into M = U11^-1 T_piv, so
//     that W = M X_w = P_J^-1 P (P: the pivot rows on entry, J: their
//     columns) and V = U11^-1.
//  2. [T_piv; M] X_w is one gemm over the columns from c on; every row below
//     the new pivots (window or not) becomes X_j - X_j[J] W, one rank-LU_W
//     update of all rows (zero multipliers elsewhere); the pivot rows receive
//     T_piv X_w.  This is the unique result of the row operations of the
//     column-by-column elimination: every row below changes by a combination
//     of pivot rows that clears the pivot columns, and that combination is
//     X_j[J] P_J^-1.
//  3. If a column of the panel has no nonzero in the window (and rows below
//     it exist), the panel ends there; after its update a search over every
//     remaining row finds the next pivot (first column, then first row), and
//     its transposition is made before the next panel.
// compute_L: the working copy carries [A | I_m]; the augmented part of row
// perm[t] at column perm[s] is T[t][s] of T P A = U, the value LELA leaves in
// A[t][s] left of the pivots.  pluq: the multipliers of the rows below are
// X_j[J] V (one more product per panel), the pivot rows' own from the window
// elimination, stored per physical row; the column order comes from the
// pivot columns.
#include "common.hpp"

#include <climits>

namespace mlz {

constexpr int LU_W = 64;     // panel width (columns)
constexpr int LU_K = 128;    // window rows
constexpr int LU_NT = 1024;  // threads of the window kernel
constexpr int LU_SEARCH_COLS = 2048;   // columns per launch of the pivot search

// Device outputs of one window kernel.
struct LuPanel {
	int rp;            // pivots found
	int stop;          // 1: a column of the panel needs the rows beyond the window
	int stop_col;      // that column (absolute)
	int bad;           // a pivot had no inverse (composite modulus)
	int steps[LU_W];   // window position of step s's pivot row (before its swap)
	int J[LU_W];       // pivot column of step s (absolute)
	uint32_t x[LU_W];  // pivot value of step s
	int wperm[LU_K];   // window order on exit: original window index at each position
};

// One panel: window rows i0 .. i0+kw-1 (kw = min(LU_K, m - i0)), columns
// c .. c+wc-1 (pivots searched in them only).  Outputs: TM rows 0..63 =
// T_piv, rows 64..127 = M (LU_K columns each, zero past rp), V = U11^-1
// (LU_W x LU_W), F[orig window index][s] = the multiplier a/x of step s.
__global__ __launch_bounds__(LU_NT) void lu_window_kernel(const uint32_t *__restrict__ X, size_t ldx, int m, int c,
							  int wc, int i0, const int *__restrict__ perm, ModP mp,
							  const uint16_t *__restrict__ invtab, LuPanel *__restrict__ out,
							  uint32_t *__restrict__ TM, uint32_t *__restrict__ V,
							  uint32_t *__restrict__ F)
{
	__shared__ uint32_t Xl[LU_K][LU_W];
	__shared__ uint32_t Tl[LU_K][LU_K];
	__shared__ uint32_t Il[LU_W][LU_W];
	__shared__ int ord[LU_K];
	__shared__ uint32_t fv[LU_K];
	__shared__ uint32_t sinv[LU_W];
	__shared__ int jl[LU_W];
	__shared__ int s_first, s_bad;
	const int tid = threadIdx.x;
	const int kw = min(LU_K, m - i0);
	const bool covers = i0 + LU_K >= m;
	for (int idx = tid; idx < LU_K * LU_W; idx += LU_NT) {
		const int q = idx / LU_W, cc = idx % LU_W;
		Xl[q][cc] = (q < kw && cc < wc) ? X[(size_t) perm[i0 + q] * ldx + c + cc] : 0u;
	}
	for (int idx = tid; idx < LU_K * LU_K; idx += LU_NT) {
		const int q = idx / LU_K, r = idx % LU_K;
		Tl[q][r] = (q == r && q < kw) ? 1u : 0u;
	}
	if (tid < LU_K)
		ord[tid] = tid;
	if (tid == 0)
		s_bad = 0;
	__syncthreads();
	int cur = 0, stop = 0, stop_col = 0;
	for (int cc = 0; cc < wc; ++cc) {
		if (cur == kw) {             // every window row is a pivot row
			if (!covers) {
				stop = 1;
				stop_col = c + cc;
			}
			break;
		}
		if (tid == 0)
			s_first = INT_MAX;
		__syncthreads();
		if (tid >= cur && tid < kw && Xl[tid][cc] != 0u)
			atomicMin(&s_first, tid);
		__syncthreads();
		const int first = s_first;
		if (first == INT_MAX) {
			if (!covers) {       // rows beyond the window may have one
				stop = 1;
				stop_col = c + cc;
				break;
			}
			continue;            // a zero column: no pivot (pivot-strategy.tcc:31-39)
		}
		if (first != cur) {
			for (int col = tid; col < LU_W + LU_K; col += LU_NT) {
				if (col < LU_W) {
					const uint32_t t = Xl[cur][col];
					Xl[cur][col] = Xl[first][col];
					Xl[first][col] = t;
				} else {
					const int tc = col - LU_W;
					const uint32_t t = Tl[cur][tc];
					Tl[cur][tc] = Tl[first][tc];
					Tl[first][tc] = t;
				}
			}
			if (tid == 0) {
				const int t = ord[cur];
				ord[cur] = ord[first];
				ord[first] = t;
			}
			__syncthreads();
		}
		if (tid == 0) {
			const uint32_t x = Xl[cur][cc];
			const uint32_t inv = invtab ? (uint32_t) invtab[x] : mp.inv(x);
			if (inv == 0u)
				s_bad = 1;
			sinv[cur] = inv;
			jl[cur] = cc;
			out->steps[cur] = first;
			out->J[cur] = c + cc;
			out->x[cur] = x;
		}
		__syncthreads();
		if (tid > cur && tid < kw) {
			const uint32_t f = mp.mul(Xl[tid][cc], sinv[cur]);
			fv[tid] = f;
			F[(size_t) ord[tid] * LU_W + cur] = f;
			Xl[tid][cc] = 0u;
		}
		__syncthreads();
		// rows below: row_r -= f_r * pivot row on the panel's later columns and on T
		const int pcols = wc - cc - 1, cols = pcols + LU_K, rows = kw - cur - 1;
		for (int idx = tid; idx < rows * cols; idx += LU_NT) {
			const int r = cur + 1 + idx / cols, ci = idx % cols;
			const uint32_t f = fv[r];
			if (f == 0u)
				continue;
			if (ci < pcols) {
				const int col = cc + 1 + ci;
				Xl[r][col] = mp.sub(Xl[r][col], mp.mul(f, Xl[cur][col]));
			} else {
				const int tc = ci - pcols;
				Tl[r][tc] = mp.sub(Tl[r][tc], mp.mul(f, Tl[cur][tc]));
			}
		}
		__syncthreads();
		++cur;
	}
	const int rp = cur;
	if (s_bad) {
		if (tid == 0) {
			out->bad = 1;
			out->rp = rp;
		}
		return;
	}
	for (int idx = tid; idx < LU_W * LU_K; idx += LU_NT) {
		const int t = idx / LU_K, q = idx % LU_K;
		TM[idx] = t < rp ? Tl[t][q] : 0u;
	}
	for (int idx = tid; idx < LU_W * LU_W; idx += LU_NT) {
		const int t = idx / LU_W, u = idx % LU_W;
		Il[t][u] = (t == u && t < rp) ? 1u : 0u;
	}
	__syncthreads();
	// backward pass over the pivot rows [U | T | I]: U11^-1 from the right
	for (int t = rp - 1; t >= 0; --t) {
		const uint32_t inv = sinv[t];
		for (int col = tid; col < LU_W + LU_K + LU_W; col += LU_NT) {
			if (col < LU_W)
				Xl[t][col] = mp.mul(Xl[t][col], inv);
			else if (col < LU_W + LU_K)
				Tl[t][col - LU_W] = mp.mul(Tl[t][col - LU_W], inv);
			else
				Il[t][col - LU_W - LU_K] = mp.mul(Il[t][col - LU_W - LU_K], inv);
		}
		__syncthreads();
		if (tid < t)
			fv[tid] = Xl[tid][jl[t]];
		__syncthreads();
		const int cols = LU_W + LU_K + LU_W;
		for (int idx = tid; idx < t * cols; idx += LU_NT) {
			const int s = idx / cols, col = idx % cols;
			const uint32_t g = fv[s];
			if (g == 0u)
				continue;
			if (col < LU_W)
				Xl[s][col] = mp.sub(Xl[s][col], mp.mul(g, Xl[t][col]));
			else if (col < LU_W + LU_K)
				Tl[s][col - LU_W] = mp.sub(Tl[s][col - LU_W], mp.mul(g, Tl[t][col - LU_W]));
			else
				Il[s][col - LU_W - LU_K] = mp.sub(Il[s][col - LU_W - LU_K], mp.mul(g, Il[t][col - LU_W - LU_K]));
		}
		__syncthreads();
	}
	for (int idx = tid; idx < LU_W * LU_K; idx += LU_NT) {
		const int t = idx / LU_K, q = idx % LU_K;
		TM[LU_W * LU_K + idx] = t < rp ? Tl[t][q] : 0u;
	}
	for (int idx = tid; idx < LU_W * LU_W; idx += LU_NT)
		V[idx] = Il[idx / LU_W][idx % LU_W];
	if (tid < LU_K)
		out->wperm[tid] = ord[tid];
	if (tid == 0) {
		out->rp = rp;
		out->stop = stop;
		out->stop_col = stop_col;
		out->bad = 0;
	}
}

// Xw[q][col] = X[perm[i0 + q]][c + col] for q < kw, zero rows up to LU_K.
__global__ void lu_gather_window_kernel(const uint32_t *__restrict__ X, size_t ldx, int kw, int c, int nc, int i0,
					const int *__restrict__ perm, uint32_t *__restrict__ Xw)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
	if (col >= nc)
		return;
	Xw[(size_t) q * nc + col] = q < kw ? X[(size_t) perm[i0 + q] * ldx + c + col] : 0u;
}

// perm[i0 + q] <- perm[i0 + wperm[q]] for q < kw (one workgroup).
__global__ void lu_permute_window_kernel(int *__restrict__ perm, int i0, int kw, const LuPanel *__restrict__ st)
{
	__shared__ int old[LU_K];
	const int q = threadIdx.x;
	if (q < kw)
		old[q] = perm[i0 + q];
	__syncthreads();
	if (q < kw)
		perm[i0 + q] = old[st->wperm[q]];
}

// G[perm[pos]][t] = X[perm[pos]][J_t] for positions pos >= i1 and t < rp,
// zero for every other row and for t >= rp (G is m x LU_W).
__global__ void lu_gather_mult_kernel(const uint32_t *__restrict__ X, size_t ldx, int m, int i1,
				      const int *__restrict__ perm, const LuPanel *__restrict__ st,
				      uint32_t *__restrict__ G)
{
	const int pos = blockIdx.x * (blockDim.x / LU_W) + threadIdx.x / LU_W, t = threadIdx.x % LU_W;
	if (pos >= m)
		return;
	const int row = perm[pos];
	const int rp = st->rp;
	G[(size_t) row * LU_W + t] = (pos >= i1 && t < rp) ? X[(size_t) row * ldx + st->J[t]] : 0u;
}

// The pivot rows: X[perm[i0 + t]][c + col] = U[t][col] for t < rp.
__global__ void lu_scatter_pivots_kernel(uint32_t *__restrict__ X, size_t ldx, int c, int nc, int i0,
					 const int *__restrict__ perm, const LuPanel *__restrict__ st,
					 const uint32_t *__restrict__ U)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
	if (col >= nc || t >= st->rp)
		return;
	X[(size_t) perm[i0 + t] * ldx + c + col] = U[(size_t) t * nc + col];
}

// pluq multipliers of this panel's steps i0 .. i0+rp-1: rows below take
// GV[row][t] (= X_j[J] U11^-1), the pivot rows their window multipliers.
__global__ void lu_store_mult_kernel(uint32_t *__restrict__ L, size_t ldl, int m, int i0,
				     const int *__restrict__ perm, const LuPanel *__restrict__ st,
				     const uint32_t *__restrict__ GV, const uint32_t *__restrict__ F)
{
	const int pos = blockIdx.x * (blockDim.x / LU_W) + threadIdx.x / LU_W, t = threadIdx.x % LU_W;
	const int rp = st->rp;
	if (pos < i0 || pos >= m || t >= rp)
		return;
	const int row = perm[pos];
	uint32_t v;
	if (pos >= i0 + rp)
		v = GV[(size_t) row * LU_W + t];
	else {
		const int k = pos - i0;           // pivot row k of the panel: its multipliers of steps t < k
		v = t < k ? F[(size_t) st->wperm[k] * LU_W + t] : 0u;
		if (t >= k)
			return;
	}
	L[(size_t) row * ldl + i0 + t] = v;
}

// First nonzero at positions >= i0, columns [c0, c1): key = col << 32 | pos.
__global__ void lu_search_kernel(const uint32_t *__restrict__ X, size_t ldx, int m, int i0, int c0, int c1,
				 const int *__restrict__ perm, unsigned long long *__restrict__ key)
{
	const int col = c0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (col >= c1)
		return;
	for (int pos = i0 + blockIdx.y; pos < m; pos += gridDim.y) {
		if (X[(size_t) perm[pos] * ldx + col] != 0u) {
			atomicMin(key, ((unsigned long long) col << 32) | (unsigned) pos);
			return;
		}
	}
}

__global__ void lu_key_init_kernel(unsigned long long *key) { *key = ~0ull; }

__global__ void lu_swap_kernel(int *perm, int a, int b)
{
	const int t = perm[a];
	perm[a] = perm[b];
	perm[b] = t;
}

__global__ void lu_iota_kernel(int *perm, int m)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < m)
		perm[i] = i;
}

// Working copy: canonical residues of A in columns 0 .. n-1, for compute_L the
// identity in columns n .. n+m-1.
template <int E>
__global__ void lu_load_kernel(const typename Elem<E>::T *__restrict__ A, size_t lda, int m, int n, int ntot, uint32_t p,
			       uint32_t *__restrict__ X)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, row = blockIdx.y;
	if (col >= ntot)
		return;
	X[(size_t) row * ntot + col] = col < n ? Elem<E>::load(A + (size_t) row * lda + col, p)
					       : (col - n == row ? 1u : 0u);
}

// Output row t (position order).  mode 0: U.  mode 1: U right of the
// diagonal, T[t][s] = X[perm t][n + perm s] left of it.  mode 2 (pluq): L
// left of the diagonal (first `rank` columns), U in column order corder.
template <int E>
__global__ void lu_store_kernel(const uint32_t *__restrict__ X, size_t ldx, int m, int n, int rank, int mode,
				const int *__restrict__ perm, const int *__restrict__ corder,
				const uint32_t *__restrict__ L, size_t ldl, uint32_t p, typename Elem<E>::T *__restrict__ A,
				size_t lda)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
	if (col >= n)
		return;
	const size_t row = (size_t) perm[t];
	uint32_t v;
	if (mode == 0)
		v = X[row * ldx + col];
	else if (mode == 1)
		v = col < t ? X[row * ldx + n + perm[col]] : X[row * ldx + col];
	else
		v = col < t ? (col < rank ? L[row * ldl + col] : 0u) : X[row * ldx + corder[col]];
	Elem<E>::store(A + (size_t) t * lda + col, v, p);
}

int zp_echelon_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, void *A, size_t lda, int mode,
		   size_t *rank_out, uint64_t *det_out, uint32_t *perm_rows, uint32_t *perm_cols, int algo)
{
	if (m == 0 || n == 0) {
		if (rank_out)
			*rank_out = 0;
		if (det_out)
			*det_out = 1 % p;
		return MLZ_OK;
	}
	if (m >= (size_t) INT32_MAX || n >= (size_t) INT32_MAX / 2)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	const ModP mp = make_modp(p);
	const uint32_t pp = (uint32_t) p;
	const size_t ntot = mode == MLZ_ECH_TRANSFORM ? n + m : n;
	const size_t rmax = std::min(m, n);
	// workspace: X (m x ntot), Xw (LU_K x ntot), UW (128 x ntot), G and GV (m x LU_W),
	// TM (128 x LU_K), V (LU_W^2), F (LU_K x LU_W), L (m x rmax for pluq), perm, corder, state
	void *w0, *w1, *w2, *w3;
	MLZ_TRY(ws_get(ctx, WS_RREF_0, m * ntot * sizeof(uint32_t), &w0));
	const size_t sXw = round_up(LU_K * ntot * 4, 256), sUW = round_up(2 * LU_W * ntot * 4, 256);
	MLZ_TRY(ws_get(ctx, WS_RREF_1, sXw + sUW, &w1));
	const size_t sG = round_up(m * LU_W * 4, 256);
	MLZ_TRY(ws_get(ctx, WS_RREF_2, 2 * sG + (mode == MLZ_ECH_PLUQ ? m * rmax * 4 : 0), &w2));
	const size_t small = round_up(2 * LU_W * LU_K * 4, 256) + round_up(LU_W * LU_W * 4, 256) +
			     round_up(LU_K * LU_W * 4, 256) + round_up(m * 4, 256) + round_up(n * 4, 256) +
			     round_up(sizeof(LuPanel), 256) + 256;
	MLZ_TRY(ws_get(ctx, WS_RREF_3, small, &w3));
	uint32_t *X = (uint32_t *) w0;
	uint32_t *Xw = (uint32_t *) w1, *UW = (uint32_t *) ((char *) w1 + sXw);
	uint32_t *G = (uint32_t *) w2, *GV = (uint32_t *) ((char *) w2 + sG);
	uint32_t *L = (uint32_t *) ((char *) w2 + 2 * sG);
	char *q = (char *) w3;
	uint32_t *TM = (uint32_t *) q; q += round_up(2 * LU_W * LU_K * 4, 256);
	uint32_t *V = (uint32_t *) q; q += round_up(LU_W * LU_W * 4, 256);
	uint32_t *F = (uint32_t *) q; q += round_up(LU_K * LU_W * 4, 256);
	int *perm = (int *) q; q += round_up(m * 4, 256);
	int *corder = (int *) q; q += round_up(n * 4, 256);
	LuPanel *st = (LuPanel *) q; q += round_up(sizeof(LuPanel), 256);
	unsigned long long *key = (unsigned long long *) q;

	const uint16_t *invtab = nullptr;
	if (p <= 65536)
		MLZ_TRY(zp_inverse_table(ctx, mp, &invtab));
	{
		ProfScope ps(ctx, "lu_load");
		lu_iota_kernel<<<ceil_div(m, 256), 256, 0, ctx->stream>>>(perm, (int) m);
		dim3 g(ceil_div(ntot, 256), (unsigned) m);
		if (elem == MLZ_U16)
			lu_load_kernel<MLZ_U16><<<g, 256, 0, ctx->stream>>>((const uint16_t *) A, lda, (int) m, (int) n,
									     (int) ntot, pp, X);
		else if (elem == MLZ_U32)
			lu_load_kernel<MLZ_U32><<<g, 256, 0, ctx->stream>>>((const uint32_t *) A, lda, (int) m, (int) n,
									     (int) ntot, pp, X);
		else
			lu_load_kernel<MLZ_F64_SYM><<<g, 256, 0, ctx->stream>>>((const double *) A, lda, (int) m, (int) n,
										 (int) ntot, pp, X);
		MLZ_TRY(check_launch(ctx, "lu_load_kernel"));
	}
	const bool upd_i8 = p <= 65536 && (algo == MLZ_ALGO_AUTO || algo == MLZ_ALGO_INT8);
	std::vector<int> steps, pcols;     // per step: pivot position, pivot column
	steps.reserve(rmax);
	pcols.reserve(rmax);
	uint64_t det = 1 % p;
	size_t i = 0, c = 0;
	// A transposition made after a search: the next window then finds its
	// first pivot at position i, and its step records this position instead.
	long pending = -1;
	LuPanel h;
	while (i < m && c < n) {
		const int wc = (int) std::min((size_t) LU_W, n - c);
		const int kw = (int) std::min((size_t) LU_K, m - i);
		{
			ProfScope ps(ctx, "lu_window");
			lu_window_kernel<<<1, LU_NT, 0, ctx->stream>>>(X, ntot, (int) m, (int) c, wc, (int) i, perm, mp, invtab,
									st, TM, V, F);
		}
		MLZ_TRY(check_launch(ctx, "lu_window_kernel"));
		MLZ_HIP(ctx, hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
		MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
		if (h.bad)
			return set_error(ctx, MLZ_E_NOT_INVERTIBLE, "Could not invert pivot-element in the ring");
		const int rp = h.rp;
		for (int s = 0; s < rp; ++s) {
			steps.push_back(s == 0 && pending >= 0 ? (int) pending : (int) i + h.steps[s]);
			pcols.push_back(h.J[s]);
			det = mp.mul((uint32_t) det, h.x[s]);
		}
		if (rp > 0)
			pending = -1;
		if (rp > 0) {
			const size_t nc = ntot - c;
			ProfScope ps(ctx, "lu_update");
			lu_gather_window_kernel<<<dim3(ceil_div(nc, 256), LU_K), 256, 0, ctx->stream>>>(X, ntot, kw, (int) c,
												  (int) nc, (int) i, perm, Xw);
			MLZ_TRY(check_launch(ctx, "lu_gather_window_kernel"));
			// [T_piv; M] X_w: the new pivot rows and W
			MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, 2 * LU_W, nc, LU_K, 1, TM, LU_K, Xw, nc, 0, UW, nc, algo));
			lu_permute_window_kernel<<<1, LU_K, 0, ctx->stream>>>(perm, (int) i, kw, st);
			lu_gather_mult_kernel<<<ceil_div(m, 256 / LU_W), 256, 0, ctx->stream>>>(X, ntot, (int) m, (int) i + rp,
												  perm, st, G);
			MLZ_TRY(check_launch(ctx, "lu_gather_mult_kernel"));
			if (mode == MLZ_ECH_PLUQ) {
				// multipliers X_j[J] U11^-1 (before X changes)
				MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, m, LU_W, LU_W, 1, G, LU_W, V, LU_W, 0, GV, LU_W, algo));
				lu_store_mult_kernel<<<ceil_div(m, 256 / LU_W), 256, 0, ctx->stream>>>(L, rmax, (int) m, (int) i,
													 perm, st, GV, F);
				MLZ_TRY(check_launch(ctx, "lu_store_mult_kernel"));
			}
			const uint32_t *Wp = UW + (size_t) LU_W * nc;
			if (upd_i8)
				MLZ_TRY(zp_rank_update_i8_dev(ctx, pp, m, nc, LU_W, G, LU_W, Wp, nc, X + c, ntot));
			else
				MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, m, nc, LU_W, p - 1, G, LU_W, Wp, nc, 1, X + c, ntot, algo));
			lu_scatter_pivots_kernel<<<dim3(ceil_div(nc, 256), LU_W), 256, 0, ctx->stream>>>(X, ntot, (int) c,
													(int) nc, (int) i, perm,
													st, UW);
			MLZ_TRY(check_launch(ctx, "lu_scatter_pivots_kernel"));
		}
		i += rp;
		if (!h.stop) {
			c += wc;
			continue;
		}
		// the window ran out: the next pivot (first column, then first
		// position) among all remaining rows, searched in column chunks
		c = h.stop_col;
		bool found = false;
		while (i < m && c < n) {
			const size_t c1 = std::min(n, c + LU_SEARCH_COLS);
			unsigned long long k = ~0ull;
			lu_key_init_kernel<<<1, 1, 0, ctx->stream>>>(key);
			lu_search_kernel<<<dim3(ceil_div(c1 - c, 256), (unsigned) std::min<size_t>(m - i, 64)), 256, 0,
					   ctx->stream>>>(X, ntot, (int) m, (int) i, (int) c, (int) c1, perm, key);
			MLZ_TRY(check_launch(ctx, "lu_search_kernel"));
			MLZ_HIP(ctx, hipMemcpyAsync(&k, key, sizeof k, hipMemcpyDeviceToHost, ctx->stream));
			MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
			if (k == ~0ull) {
				c = c1;
				continue;
			}
			found = true;
			c = (size_t) (k >> 32);
			const size_t pos = (size_t) (k & 0xffffffffull);
			if (pos != i) {
				lu_swap_kernel<<<1, 1, 0, ctx->stream>>>(perm, (int) i, (int) pos);
				MLZ_TRY(check_launch(ctx, "lu_swap_kernel"));
				pending = (long) pos;
			}
			break;
		}
		if (!found)
			break;
	}
	const std::vector<int> &st_final = steps, &pc_final = pcols;
	const size_t rank = st_final.size();
	if (perm_rows)
		for (size_t s = 0; s < rank; ++s)
			perm_rows[s] = (uint32_t) st_final[s];
	// pluq's column order: the pivot columns swapped into place step by step
	std::vector<int> co(n), where(n);
	for (size_t j = 0; j < n; ++j)
		co[j] = where[j] = (int) j;
	for (size_t s = 0; s < rank; ++s) {
		const int cpos = where[pc_final[s]];
		if (perm_cols)
			perm_cols[s] = (uint32_t) cpos;
		const int a = co[s], b = co[cpos];
		co[s] = b;
		co[cpos] = a;
		where[b] = (int) s;
		where[a] = cpos;
	}
	if (mode == MLZ_ECH_PLUQ)
		MLZ_HIP(ctx, hipMemcpyAsync(corder, co.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	{
		ProfScope ps(ctx, "lu_store");
		dim3 g(ceil_div(n, 256), (unsigned) m);
#define STORE(EE, TT) lu_store_kernel<EE><<<g, 256, 0, ctx->stream>>>(X, ntot, (int) m, (int) n, (int) rank, mode, \
								     perm, corder, L, rmax, pp, (TT *) A, lda)
		if (elem == MLZ_U16)
			STORE(MLZ_U16, uint16_t);
		else if (elem == MLZ_U32)
			STORE(MLZ_U32, uint32_t);
		else
			STORE(MLZ_F64_SYM, double);
#undef STORE
		MLZ_TRY(check_launch(ctx, "lu_store_kernel"));
	}
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));   // co leaves scope
	if (rank_out)
		*rank_out = rank;
	if (det_out)
		*det_out = det;
	return MLZ_OK;
}

}  // namespace mlz

using namespace mlz;

extern "C" int mlz_zp_echelonize(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, void *A, size_t lda,
				 int mode, size_t *rank, uint64_t *det, uint32_t *perm_rows, uint32_t *perm_cols,
				 int algo)
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
	if (algo < MLZ_ALGO_AUTO || algo > MLZ_ALGO_FP64)
		return set_error(ctx, MLZ_E_ARG, "unknown gemm algo %d", algo);
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (ctx->ptr_mode == MLZ_PTR_DEVICE || m == 0 || n == 0)
		return zp_echelon_dev(ctx, elem, p, m, n, A, lda, mode, rank, det, perm_rows, perm_cols, algo);
	const size_t es = elem_size(elem);
	void *dA = nullptr;
	MLZ_TRY(stage_in(ctx, WS_STAGE_A, A, m, n, lda, es, &dA));
	MLZ_TRY(zp_echelon_dev(ctx, elem, p, m, n, dA, n, mode, rank, det, perm_rows, perm_cols, algo));
	return stage_out(ctx, A, dA, m, n, lda, es);
}

