// lela_amd/csrc/gf2.hip -- GF(2) level 3 on bit-packed rows.
//
// Layout: rows of uint64 words, bit j of word w = column 64*w + j (LSB first,
// lela/vector/bit-iterator.h:249-311), row stride in words
// (lela/matrix/dense-zero-one.h:253-257).
//
// gemm (C <- a*A*B + b*C): "Method of Four Russians" multiplication.  A
// workgroup owns G2_TM rows x G2_TW words of C in registers.  For each 64-row
// slab of B it builds, in LDS, eight 256-entry tables T_b[i] = XOR of the
// rows 8b+j of the slab for the set bits j of i (one XOR per entry, built by
// doubling), then every C row XORs T_b[byte b of its A word] for b = 0..7.
// This replaces the word-XOR axpy gemm of the no-M4RI build
// (lela/blas/level3-generic.tcc:243-263, lela/blas/level2-gf2.tcc:28-55,
// lela/blas/level1-gf2.tcc:26-52,282-299) and libm4ri's mzd_mul/mzd_addmul
// (lela/blas/level3-m4ri.tcc:61-86).
//
// RREF: blocked Gauss-Jordan over 256-bit panels, the GF(2) counterpart of
// zp_rref.hip (pivot tournament in one kernel, P_J^-1 * P, one M4RM update of
// all rows), replacing EchelonForm<GF2> -> GaussJordan / mzd_echelonize
// (lela/solutions/echelon-form-gf2.h:97-134).
#include "common.hpp"

namespace mlz {

constexpr int G2_TW = 4;                    // C words per workgroup tile (256 columns)
constexpr int G2_RPT = 4;                   // C rows per thread
constexpr int G2_TM = 256 * G2_RPT;         // C rows per workgroup

__device__ inline uint64_t tail_mask64(int bits)
{
	return bits >= 64 ? ~0ull : ((1ull << bits) - 1);
}

// k_bits: inner dimension in bits; nbits: C columns.
__global__ __launch_bounds__(256) void gf2_gemm_m4rm_kernel(const uint64_t *__restrict__ A, size_t lda,
							    const uint64_t *__restrict__ B, size_t ldb,
							    uint64_t *__restrict__ C, size_t ldc, int m,
							    int nbits, int k_bits, int a, int b)
{
	__shared__ uint64_t T[8][256][G2_TW];
	__shared__ uint64_t Bs[64][G2_TW];
	const int tid = threadIdx.x;
	const int nw = (nbits + 63) / 64;
	const int kw = (k_bits + 63) / 64;
	const int w0 = blockIdx.x * G2_TW;
	const int row0 = blockIdx.y * G2_TM;
	uint64_t acc[G2_RPT][G2_TW];
#pragma unroll
	for (int q = 0; q < G2_RPT; ++q)
#pragma unroll
		for (int w = 0; w < G2_TW; ++w)
			acc[q][w] = 0;
	if (tid < 8 * G2_TW)
		T[tid / G2_TW][0][tid % G2_TW] = 0;
	if (a) {
		for (int kwi = 0; kwi < kw; ++kwi) {
			{
				const int r = tid / G2_TW, w = tid % G2_TW;   // 64 rows x 4 words
				const int kk = kwi * 64 + r;
				Bs[r][w] = (kk < k_bits && w0 + w < nw) ? B[(size_t) kk * ldb + w0 + w] : 0ull;
			}
			__syncthreads();
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				const int items = (8 << j) * G2_TW;
				for (int idx = tid; idx < items; idx += 256) {
					const int w = idx % G2_TW, rest = idx / G2_TW;
					const int bb = rest >> j, i = (1 << j) + (rest & ((1 << j) - 1));
					T[bb][i][w] = T[bb][i - (1 << j)][w] ^ Bs[bb * 8 + j][w];
				}
				__syncthreads();
			}
			const uint64_t km = (kwi == kw - 1) ? tail_mask64(k_bits - kwi * 64) : ~0ull;
#pragma unroll
			for (int q = 0; q < G2_RPT; ++q) {
				const int row = row0 + tid + 256 * q;
				uint64_t av = row < m ? (A[(size_t) row * lda + kwi] & km) : 0ull;
#pragma unroll
				for (int bb = 0; bb < 8; ++bb) {
					const int idx = (int) ((av >> (8 * bb)) & 255);
#pragma unroll
					for (int w = 0; w < G2_TW; ++w)
						acc[q][w] ^= T[bb][idx][w];
				}
			}
			__syncthreads();
		}
	}
#pragma unroll
	for (int q = 0; q < G2_RPT; ++q) {
		const int row = row0 + tid + 256 * q;
		if (row >= m)
			continue;
#pragma unroll
		for (int w = 0; w < G2_TW; ++w) {
			const int gw = w0 + w;
			if (gw >= nw)
				continue;
			uint64_t *cp = C + (size_t) row * ldc + gw;
			const uint64_t mask = (gw == nw - 1) ? tail_mask64(nbits - gw * 64) : ~0ull;
			const uint64_t old = *cp;
			const uint64_t v = (b ? old : 0ull) ^ acc[q][w];
			*cp = (v & mask) | (old & ~mask);
		}
	}
}

int gf2_gemm_dev(mlz_ctx *ctx, size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda,
		 const uint64_t *B, size_t ldb, int b, uint64_t *C, size_t ldc)
{
	if (m == 0 || n == 0)
		return MLZ_OK;
	const size_t nw = (n + 63) / 64;
	dim3 g(ceil_div(nw, G2_TW), ceil_div(m, G2_TM));
	ProfScope ps(ctx, "gf2_gemm_m4rm");
	gf2_gemm_m4rm_kernel<<<g, 256, 0, ctx->stream>>>(A, lda, B, ldb, C, ldc, (int) m, (int) n, (int) k,
							  (a & 1) && k > 0, b & 1);
	return check_launch(ctx, "gf2_gemm_m4rm_kernel");
}

// ===========================================================================
// RREF
// ===========================================================================

constexpr int G2P = 4;              // panel width in words (256 columns)
constexpr int G2_PB = G2P * 64;     // panel width in bits
constexpr int G2_CHUNK = 1024;      // rows per tournament leaf
constexpr int G2_MAXLVL = 24;

struct Gf2State {
	int *is_pivot, *rank, *piv_rows, *piv_cols, *k, *J, *prow, *lists_rows, *lists_cnt, *counters;
};

__device__ inline int bit_of(const uint64_t *row, int col)
{
	return (int) ((row[col >> 6] >> (col & 63)) & 1ull);
}

// Echelon form of nr rows x wb bits (rows of G2P words in LDS); returns pivots.
__device__ int gf2_lds_echelon(uint64_t (*M)[G2P], const int *rowid, int *alive, int nr, int wb, int *out_rows,
			       int *out_cols, int *s_min, int *flip)
{
	const int tid = threadIdx.x;
	int cnt = 0;
	for (int col = 0; col < wb; ++col) {
		if (tid == 0)
			*s_min = 0x7fffffff;
		__syncthreads();
		for (int r = tid; r < nr; r += blockDim.x)
			if (alive[r] && bit_of(M[r], col))
				atomicMin(s_min, r);
		__syncthreads();
		const int pr = *s_min;
		__syncthreads();
		if (pr == 0x7fffffff)
			continue;
		for (int r = tid; r < nr; r += blockDim.x)
			flip[r] = (r != pr && alive[r] && bit_of(M[r], col));
		if (tid == 0) {
			alive[pr] = 0;
			out_rows[cnt] = rowid[pr];
			out_cols[cnt] = col;
		}
		__syncthreads();
		for (int idx = tid; idx < nr * G2P; idx += blockDim.x) {
			const int r = idx / G2P, w = idx % G2P;
			if (flip[r])
				M[r][w] ^= M[pr][w];
		}
		__syncthreads();
		++cnt;
	}
	return cnt;
}

__device__ inline int g2_nodes_at(int nodes0, int level)
{
	int n = nodes0;
	for (int l = 0; l < level; ++l)
		n = (n + 1) / 2;
	return n;
}

__global__ __launch_bounds__(256) void gf2_tourney_kernel(const uint64_t *__restrict__ A, size_t lda, int m,
							  int cw, int nw, int wb, Gf2State st, int nodes0)
{
	__shared__ uint64_t M[G2_CHUNK][G2P];
	__shared__ int rowid[G2_CHUNK], alive[G2_CHUNK], flip[G2_CHUNK];
	__shared__ int outr[G2_PB], outc[G2_PB];
	__shared__ int s_min, s_go, s_nr, s_n0;
	const int tid = threadIdx.x;
	int node = blockIdx.x, level = 0;
	const int r0 = node * G2_CHUNK;
	const int nr0 = min(G2_CHUNK, m - r0);
	for (int idx = tid; idx < nr0 * G2P; idx += blockDim.x) {
		const int r = idx / G2P, w = idx % G2P;
		M[r][w] = cw + w < nw ? A[(size_t) (r0 + r) * lda + cw + w] : 0ull;
	}
	for (int r = tid; r < nr0; r += blockDim.x) {
		rowid[r] = r0 + r;
		alive[r] = st.is_pivot[r0 + r] == 0;
	}
	__syncthreads();
	int cnt = gf2_lds_echelon(M, rowid, alive, nr0, wb, outr, outc, &s_min, flip);
	for (;;) {
		const int here = g2_nodes_at(nodes0, level);
		if (here == 1) {
			for (int t = tid; t < G2_PB; t += blockDim.x) {
				st.J[t] = t < cnt ? outc[t] : 0;
				st.prow[t] = t < cnt ? outr[t] : 0;
			}
			if (tid == 0)
				*st.k = cnt;
			for (int t = tid; t < G2_MAXLVL * nodes0; t += blockDim.x)
				st.counters[t] = 0;
			return;
		}
		int *lr = st.lists_rows + ((size_t) level * nodes0 + node) * G2_PB;
		for (int t = tid; t < cnt; t += blockDim.x)
			lr[t] = outr[t];
		if (tid == 0)
			st.lists_cnt[(size_t) level * nodes0 + node] = cnt;
		__threadfence();
		__syncthreads();
		const int parent = node >> 1, sibling = node ^ 1;
		if (tid == 0) {
			int go = 1;
			if (sibling < here)
				go = atomicAdd(&st.counters[(size_t) level * nodes0 + parent], 1) == 1;
			s_go = go;
		}
		__syncthreads();
		if (!s_go)
			return;
		__threadfence();
		const int c0 = parent * 2, c1 = parent * 2 + 1;
		if (tid == 0) {
			const int n0 = ((volatile int *) st.lists_cnt)[(size_t) level * nodes0 + c0];
			const int n1 = c1 < here ? ((volatile int *) st.lists_cnt)[(size_t) level * nodes0 + c1] : 0;
			s_nr = n0 + n1;
			s_n0 = n0;
		}
		__syncthreads();
		const int nr = s_nr, n0 = s_n0;
		for (int r = tid; r < nr; r += blockDim.x) {
			const int child = r < n0 ? c0 : c1;
			const int pos = r < n0 ? r : r - n0;
			rowid[r] = ((volatile int *) st.lists_rows)[((size_t) level * nodes0 + child) * G2_PB + pos];
			alive[r] = 1;
		}
		__syncthreads();
		for (int idx = tid; idx < nr * G2P; idx += blockDim.x) {
			const int r = idx / G2P, w = idx % G2P;
			M[r][w] = cw + w < nw ? A[(size_t) rowid[r] * lda + cw + w] : 0ull;
		}
		__syncthreads();
		cnt = gf2_lds_echelon(M, rowid, alive, nr, wb, outr, outc, &s_min, flip);
		++level;
		node = parent;
	}
}

// P[t] = pivot row t from word cw on (t < k; zero rows up to G2_PB).
__global__ void gf2_gather_kernel(const uint64_t *__restrict__ A, size_t lda, int nwc, int cw, int panel,
				  Gf2State st, uint64_t *__restrict__ P)
{
	const int t = blockIdx.y;
	const int k = *st.k;
	const int w = blockIdx.x * blockDim.x + threadIdx.x;
	if (w < nwc)
		P[(size_t) t * nwc + w] = t < k ? A[(size_t) st.prow[t] * lda + cw + w] : 0ull;
	if (blockIdx.x == 0 && threadIdx.x == 0 && t < k) {
		const int r = *st.rank;
		st.is_pivot[st.prow[t]] = panel + 1;
		st.piv_rows[r + t] = st.prow[t];
		st.piv_cols[r + t] = cw * 64 + st.J[t];
	}
}

// Inverse of P_J (k x k bits, identity-padded to G2_PB) by Gauss-Jordan; rows
// of [P_J | I] are 2*G2P words.
__global__ __launch_bounds__(256) void gf2_inv_kernel(const uint64_t *__restrict__ P, int nwc, Gf2State st,
						      uint64_t *__restrict__ Pinv)
{
	__shared__ uint64_t X[G2_PB][2 * G2P];
	__shared__ int flip[G2_PB];
	__shared__ int s_piv;
	const int tid = threadIdx.x;
	const int k = *st.k;
	for (int r = tid; r < G2_PB; r += blockDim.x) {
		uint64_t row[2 * G2P];
		for (int w = 0; w < 2 * G2P; ++w)
			row[w] = 0;
		if (r < k) {
			const uint64_t *src = P + (size_t) r * nwc;
			for (int s = 0; s < k; ++s)
				if (bit_of(src, st.J[s]))
					row[s >> 6] |= 1ull << (s & 63);
		} else {
			row[r >> 6] |= 1ull << (r & 63);
		}
		row[G2P + (r >> 6)] |= 1ull << (r & 63);
		for (int w = 0; w < 2 * G2P; ++w)
			X[r][w] = row[w];
	}
	__syncthreads();
	for (int s = 0; s < k; ++s) {
		if (tid == 0) {
			int pr = -1;
			for (int r = s; r < k; ++r)
				if ((X[r][s >> 6] >> (s & 63)) & 1ull) { pr = r; break; }
			s_piv = pr;
		}
		__syncthreads();
		const int pr = s_piv;
		if (pr < 0)
			break;
		if (pr != s && tid < 2 * G2P) {
			uint64_t t = X[s][tid];
			X[s][tid] = X[pr][tid];
			X[pr][tid] = t;
		}
		__syncthreads();
		for (int r = tid; r < G2_PB; r += blockDim.x)
			flip[r] = r != s && ((X[r][s >> 6] >> (s & 63)) & 1ull);
		__syncthreads();
		for (int idx = tid; idx < G2_PB * 2 * G2P; idx += blockDim.x) {
			const int r = idx / (2 * G2P), w = idx % (2 * G2P);
			if (flip[r])
				X[r][w] ^= X[s][w];
		}
		__syncthreads();
	}
	for (int idx = tid; idx < G2_PB * G2P; idx += blockDim.x) {
		const int r = idx / G2P, w = idx % G2P;
		Pinv[idx] = X[r][G2P + w];
	}
}

// G[i] = bits J of row i (compressed to k bits), zero on this panel's pivots.
__global__ void gf2_gatherG_kernel(const uint64_t *__restrict__ A, size_t lda, int m, int cw, int nw, int panel,
				   Gf2State st, uint64_t *__restrict__ G)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= m)
		return;
	const int k = *st.k;
	uint64_t row[G2P], g[G2P];
	for (int w = 0; w < G2P; ++w) {
		row[w] = cw + w < nw ? A[(size_t) i * lda + cw + w] : 0ull;
		g[w] = 0;
	}
	if (st.is_pivot[i] != panel + 1)
		for (int s = 0; s < k; ++s) {
			const int j = st.J[s];
			if ((row[j >> 6] >> (j & 63)) & 1ull)
				g[s >> 6] |= 1ull << (s & 63);
		}
	for (int w = 0; w < G2P; ++w)
		G[(size_t) i * G2P + w] = g[w];
}

__global__ void gf2_scatter_kernel(uint64_t *__restrict__ A, size_t lda, int nwc, int cw, Gf2State st,
				   const uint64_t *__restrict__ W)
{
	const int t = blockIdx.y;
	const int k = *st.k;
	const int w = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < k && w < nwc)
		A[(size_t) st.prow[t] * lda + cw + w] = W[(size_t) t * nwc + w];
}

__global__ void gf2_advance_kernel(Gf2State st)
{
	*st.rank += *st.k;
}

__global__ void gf2_init_kernel(Gf2State st, int m, int cnt_counters)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < m)
		st.is_pivot[i] = 0;
	if (i < cnt_counters)
		st.counters[i] = 0;
	if (i == 0) {
		*st.rank = 0;
		*st.k = 0;
	}
}

// Copies A into the working matrix with the bits beyond n cleared.
__global__ void gf2_load_kernel(const uint64_t *__restrict__ A, size_t lda, int m, int n, uint64_t *__restrict__ W)
{
	const int nw = (n + 63) / 64;
	const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (size_t) m * nw)
		return;
	const int r = (int) (idx / nw), w = (int) (idx % nw);
	uint64_t v = A[(size_t) r * lda + w];
	if (w == nw - 1)
		v &= tail_mask64(n - w * 64);
	W[idx] = v;
}

__global__ void gf2_store_kernel(const uint64_t *__restrict__ W, int m, int nw, Gf2State st, uint64_t *__restrict__ A,
				 size_t lda)
{
	const size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (size_t) m * nw)
		return;
	const int t = (int) (idx / nw), w = (int) (idx % nw);
	A[(size_t) t * lda + w] = t < *st.rank ? W[(size_t) st.piv_rows[t] * nw + w] : 0ull;
}

int gf2_rref_dev(mlz_ctx *ctx, size_t m, size_t n, uint64_t *A, size_t lda, size_t *rank_out, uint32_t *pivots_out)
{
	if (m == 0 || n == 0) {
		if (rank_out)
			*rank_out = 0;
		return MLZ_OK;
	}
	const size_t nw = (n + 63) / 64;
	const int nodes0 = (int) ceil_div(m, G2_CHUNK);
	const size_t rmax = std::min(m, n);
	void *w0, *w1, *w2, *w3;
	MLZ_TRY(ws_get(ctx, WS_GF2_0, m * nw * 8, &w0));
	MLZ_TRY(ws_get(ctx, WS_GF2_1, 2 * (size_t) G2_PB * nw * 8 + (size_t) G2_PB * G2P * 8, &w1));
	MLZ_TRY(ws_get(ctx, WS_GF2_2, m * G2P * 8, &w2));
	const size_t ints = m + 1 + 2 * rmax + 1 + 2 * G2_PB + (size_t) G2_MAXLVL * nodes0 * G2_PB +
			    2 * (size_t) G2_MAXLVL * nodes0;
	MLZ_TRY(ws_get(ctx, WS_SMALL, ints * sizeof(int), &w3));
	uint64_t *Wm = (uint64_t *) w0;
	uint64_t *P = (uint64_t *) w1, *Wb = P + (size_t) G2_PB * nw, *Pinv = Wb + (size_t) G2_PB * nw;
	uint64_t *G = (uint64_t *) w2;
	Gf2State st;
	int *q = (int *) w3;
	st.is_pivot = q; q += m;
	st.rank = q; q += 1;
	st.piv_rows = q; q += rmax;
	st.piv_cols = q; q += rmax;
	st.k = q; q += 1;
	st.J = q; q += G2_PB;
	st.prow = q; q += G2_PB;
	st.lists_rows = q; q += (size_t) G2_MAXLVL * nodes0 * G2_PB;
	st.lists_cnt = q; q += (size_t) G2_MAXLVL * nodes0;
	st.counters = q;
	{
		const int cc = G2_MAXLVL * nodes0;
		gf2_init_kernel<<<ceil_div(std::max(m, (size_t) cc), 256), 256, 0, ctx->stream>>>(st, (int) m, cc);
		gf2_load_kernel<<<ceil_div(m * nw, 256), 256, 0, ctx->stream>>>(A, lda, (int) m, (int) n, Wm);
		MLZ_TRY(check_launch(ctx, "gf2_load_kernel"));
	}
	int panel = 0;
	for (size_t cw = 0; cw < nw; cw += G2P, ++panel) {
		const int wb = (int) std::min((size_t) G2_PB, n - cw * 64);
		const int nwc = (int) (nw - cw);
		{
			ProfScope ps(ctx, "gf2_tourney");
			gf2_tourney_kernel<<<nodes0, 256, 0, ctx->stream>>>(Wm, nw, (int) m, (int) cw, (int) nw, wb, st, nodes0);
		}
		MLZ_TRY(check_launch(ctx, "gf2_tourney_kernel"));
		gf2_gather_kernel<<<dim3(ceil_div(nwc, 256), G2_PB), 256, 0, ctx->stream>>>(Wm, nw, nwc, (int) cw, panel, st, P);
		gf2_inv_kernel<<<1, 256, 0, ctx->stream>>>(P, nwc, st, Pinv);
		MLZ_TRY(check_launch(ctx, "gf2_inv_kernel"));
		MLZ_TRY(gf2_gemm_dev(ctx, G2_PB, (size_t) nwc * 64, G2_PB, 1, Pinv, G2P, P, nwc, 0, Wb, nwc));
		gf2_gatherG_kernel<<<ceil_div(m, 256), 256, 0, ctx->stream>>>(Wm, nw, (int) m, (int) cw, (int) nw, panel, st, G);
		MLZ_TRY(gf2_gemm_dev(ctx, m, (size_t) nwc * 64, G2_PB, 1, G, G2P, Wb, nwc, 1, Wm + cw, nw));
		gf2_scatter_kernel<<<dim3(ceil_div(nwc, 256), G2_PB), 256, 0, ctx->stream>>>(Wm, nw, nwc, (int) cw, st, Wb);
		gf2_advance_kernel<<<1, 1, 0, ctx->stream>>>(st);
		MLZ_TRY(check_launch(ctx, "gf2_scatter_kernel"));
		if ((panel & 3) == 3) {
			int r = 0;
			MLZ_HIP(ctx, hipMemcpyAsync(&r, st.rank, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
			MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
			if ((size_t) r >= m)
				break;
		}
	}
	int r = 0;
	MLZ_HIP(ctx, hipMemcpyAsync(&r, st.rank, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	{
		ProfScope ps(ctx, "gf2_store");
		gf2_store_kernel<<<ceil_div(m * nw, 256), 256, 0, ctx->stream>>>(Wm, (int) m, (int) nw, st, A, lda);
		MLZ_TRY(check_launch(ctx, "gf2_store_kernel"));
	}
	if (pivots_out && r > 0) {
		std::vector<int> pc(r);
		MLZ_HIP(ctx, hipMemcpyAsync(pc.data(), st.piv_cols, r * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
		MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
		for (int i = 0; i < r; ++i)
			pivots_out[i] = (uint32_t) pc[i];
	}
	if (rank_out)
		*rank_out = (size_t) r;
	return MLZ_OK;
}

}  // namespace mlz

using namespace mlz;

extern "C" int mlz_gf2_gemm(mlz_ctx *ctx, size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda_w,
			    const uint64_t *B, size_t ldb_w, int b, uint64_t *C, size_t ldc_w)
{
	if (!ctx)
		return MLZ_E_ARG;
	const size_t kw = (k + 63) / 64, nw = (n + 63) / 64;
	if ((m && k && (!A || lda_w < kw)) || (k && n && (!B || ldb_w < nw)) || (m && n && (!C || ldc_w < nw)))
		return set_error(ctx, MLZ_E_ARG, "bad GF(2) matrix pointer or leading dimension");
	if (m > (size_t) INT32_MAX || n > (size_t) INT32_MAX || k > (size_t) INT32_MAX)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (ctx->ptr_mode == MLZ_PTR_DEVICE)
		return gf2_gemm_dev(ctx, m, n, k, a, A, lda_w, B, ldb_w, b, C, ldc_w);
	if (m == 0 || n == 0)
		return MLZ_OK;
	void *dA = nullptr, *dB = nullptr, *dC = nullptr;
	if (k) {
		MLZ_TRY(stage_in(ctx, WS_STAGE_A, A, m, kw, lda_w, 8, &dA));
		MLZ_TRY(stage_in(ctx, WS_STAGE_B, B, k, nw, ldb_w, 8, &dB));
	}
	MLZ_TRY(stage_in(ctx, WS_STAGE_C, C, m, nw, ldc_w, 8, &dC));
	MLZ_TRY(gf2_gemm_dev(ctx, m, n, k, a, (const uint64_t *) dA, kw, (const uint64_t *) dB, nw, b,
			     (uint64_t *) dC, nw));
	return stage_out(ctx, C, dC, m, nw, ldc_w, 8);
}

extern "C" int mlz_gf2_rref(mlz_ctx *ctx, size_t m, size_t n, uint64_t *A, size_t lda_w, size_t *rank,
			    uint32_t *pivots)
{
	if (!ctx)
		return MLZ_E_ARG;
	const size_t nw = (n + 63) / 64;
	if (m && n && (!A || lda_w < nw))
		return set_error(ctx, MLZ_E_ARG, "bad GF(2) matrix pointer or leading dimension");
	if (m > (size_t) INT32_MAX / 2 || n > (size_t) INT32_MAX / 2)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (ctx->ptr_mode == MLZ_PTR_DEVICE)
		return gf2_rref_dev(ctx, m, n, A, lda_w, rank, pivots);
	if (m == 0 || n == 0) {
		if (rank)
			*rank = 0;
		return MLZ_OK;
	}
	void *dA = nullptr;
	MLZ_TRY(stage_in(ctx, WS_STAGE_A, A, m, nw, lda_w, 8, &dA));
	MLZ_TRY(gf2_rref_dev(ctx, m, n, (uint64_t *) dA, nw, rank, pivots));
	return stage_out(ctx, A, dA, m, nw, lda_w, 8);
}
