// lela_amd/csrc/gf2.hip -- GF(2) level 3 on bit-packed rows.
//
// Layout: rows of uint64 words, bit j of word w = column 64*w + j (LSB first,
// lela/vector/bit-iterator.h:249-311), row stride in words
// (lela/matrix/dense-zero-one.h:253-257).
//
// gemm (C <- a*A*B + b*C): "Method of Four Russians" multiplication.  A
// workgroup owns 1024 or 2048 rows x G2_TW words of C in registers.  For each 64-row
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

#include <climits>

namespace mlz {

constexpr int G2_TW = 4;                    // C words per workgroup tile (256 columns)

__device__ inline uint64_t tail_mask64(int bits)
{
	return bits >= 64 ? ~0ull : ((1ull << bits) - 1);
}

// k_bits: inner dimension in bits; nbits: C columns.  G2_RPT rows per thread
// (256 * G2_RPT per workgroup).
template <int G2_RPT>
__global__ __launch_bounds__(256) void gf2_gemm_m4rm_kernel(const uint64_t *__restrict__ A, size_t lda,
							    const uint64_t *__restrict__ B, size_t ldb,
							    uint64_t *__restrict__ C, size_t ldc, int m,
							    int nbits, int k_bits, int a, int b, const int *__restrict__ kz,
							    const int *__restrict__ is_pivot, int rest_w, int *__restrict__ rest)
{
	if (kz && *kz == 0)
		return;                 // an echelon-form panel without pivots: nothing to add
	// T[h][b][e]: entry e of byte b's table, words 2h and 2h+1 (one 16-byte
	// read each).  Entries 16 bytes apart spread a wave's data-dependent
	// reads over twice the bank positions of 32-byte entries (those lost 28%
	// of the kernel's cycles to conflicts); one 8-byte array per word doubles
	// the reads and measured slower (5.40 against 4.64 ms at 16384^3).
	static_assert(G2_TW == 4, "two 16-byte halves per table entry");
	__shared__ ulonglong2 T[2][8][256];
	__shared__ uint64_t Bs[64][G2_TW];
	const int tid = threadIdx.x;
	const int nw = (nbits + 63) / 64;
	const int kw = (k_bits + 63) / 64;
	const int w0 = blockIdx.x * G2_TW;
	const int row0 = blockIdx.y * (256 * G2_RPT);
	uint64_t acc[G2_RPT][G2_TW];
#pragma unroll
	for (int q = 0; q < G2_RPT; ++q)
#pragma unroll
		for (int w = 0; w < G2_TW; ++w)
			acc[q][w] = 0;
	if (tid < 16)
		T[tid / 8][tid % 8][0] = make_ulonglong2(0ull, 0ull);
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
				const int per_h = 8 << j, items = per_h * 2;
				for (int idx = tid; idx < items; idx += 256) {
					const int h = idx / per_h, rest = idx % per_h;
					const int bb = rest >> j, i = (1 << j) + (rest & ((1 << j) - 1));
					const ulonglong2 prev = T[h][bb][i - (1 << j)];
					T[h][bb][i] = make_ulonglong2(prev.x ^ Bs[bb * 8 + j][2 * h], prev.y ^ Bs[bb * 8 + j][2 * h + 1]);
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
					const ulonglong2 t0 = T[0][bb][idx], t1 = T[1][bb][idx];
					acc[q][0] ^= t0.x;
					acc[q][1] ^= t0.y;
					acc[q][2] ^= t1.x;
					acc[q][3] ^= t1.y;
				}
			}
			__syncthreads();
		}
	}
	uint64_t live_or = 0ull;
#pragma unroll
	for (int q = 0; q < G2_RPT; ++q) {
		const int row = row0 + tid + 256 * q;
		if (row >= m)
			continue;
		const bool live = rest && is_pivot[row] == 0;
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
			if (live && gw >= rest_w)
				live_or |= v & mask;
		}
	}
	// echelon-form update: does a live row keep a bit right of the panel?
	if (rest && __ballot(live_or != 0ull) != 0ull && (threadIdx.x & 63) == 0)
		atomicOr(rest, 1);
}

// ---------------------------------------------------------------------------
// The plain form: every entry of C is a dot product over GF(2), the parity
// of popcount(A_i AND B^T_j) (lela/blas/level1-gf2.tcc:26-52 on the rows of A
// and of B's transpose, lela/blas/level3-generic.tcc:243-263).
// B is first transposed in 64 x 64-bit blocks into word-major columns,
// Bw[w][j] = bits 64w .. 64w+63 of column j (one wave per block: lane t holds
// row 64w + t, and the ballot of bit j over the lanes is column j's word).
// The product kernel gives each lane one column of C: a wave covers 64
// columns and 64 rows, and accumulates acc[r] ^= A[r][w] & Bw[w][col] in
// 32-bit halves (two ANDs and one three-way XOR per 64-bit word); A's rows
// are LDS broadcasts.  At the end, bit r of the output word of row r is the
// parity of acc[r], and the ballot over the wave's lanes is that word.
// ---------------------------------------------------------------------------
constexpr int G2D_KC = 16;    // k-words per LDS stage of the dot kernel
constexpr int G2D_TR = 128;   // rows per workgroup (2 waves of 64)
constexpr int G2D_TC = 128;   // columns per workgroup (2 waves of 64)

__global__ __launch_bounds__(64) void gf2_transpose64_kernel(const uint64_t *__restrict__ B, size_t ldb, int k_bits,
							     int nw, uint64_t *__restrict__ Bw, size_t ldw)
{
	const int c = blockIdx.x, w = blockIdx.y, t = threadIdx.x;
	const int row = w * 64 + t;
	const uint64_t v = (row < k_bits && c < nw) ? B[(size_t) row * ldb + c] : 0ull;
	uint64_t mine = 0;
#pragma unroll 8
	for (int j = 0; j < 64; ++j) {
		const uint64_t col = __ballot((v >> j) & 1ull);
		if (t == j)
			mine = col;
	}
	Bw[(size_t) w * ldw + (size_t) c * 64 + t] = mine;
}

__global__ __launch_bounds__(256) void gf2_gemm_dot_kernel(const uint64_t *__restrict__ A, size_t lda,
							   const uint64_t *__restrict__ Bw, size_t ldw,
							   uint64_t *__restrict__ C, size_t ldc, int m, int nbits,
							   int kw, int a, int b)
{
	__shared__ uint4 As[G2D_TR][G2D_KC / 2];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wr = wave >> 1, wc = wave & 1;
	const int r0 = blockIdx.y * G2D_TR, c0 = blockIdx.x * G2D_TC;
	const int col = c0 + wc * 64 + lane;
	const int nw = (nbits + 63) / 64;
	uint32_t acc[64];
#pragma unroll
	for (int r = 0; r < 64; ++r)
		acc[r] = 0u;
	if (a) {
		for (int k0 = 0; k0 < kw; k0 += G2D_KC) {
			// A tile: 128 rows x KC words, thread t loads row t/2, 8 words
			{
				const int r = tid >> 1, h = (tid & 1) * (G2D_KC / 2);
				const int row = r0 + r;
				uint64_t v[G2D_KC / 2];
#pragma unroll
				for (int q = 0; q < G2D_KC / 2; ++q) {
					const int w = k0 + h + q;
					v[q] = (row < m && w < kw) ? A[(size_t) row * lda + w] : 0ull;
				}
#pragma unroll
				for (int q = 0; q < G2D_KC / 2; q += 2)
					As[r][(h + q) / 2] = make_uint4((uint32_t) v[q], (uint32_t) (v[q] >> 32),
									(uint32_t) v[q + 1], (uint32_t) (v[q + 1] >> 32));
			}
			uint2 bv[G2D_KC];
#pragma unroll
			for (int q = 0; q < G2D_KC; ++q) {
				const uint64_t x = (k0 + q < kw) ? Bw[(size_t) (k0 + q) * ldw + col] : 0ull;
				bv[q] = make_uint2((uint32_t) x, (uint32_t) (x >> 32));
			}
			__syncthreads();
#pragma unroll
			for (int q = 0; q < G2D_KC; q += 2) {
#pragma unroll
				for (int r = 0; r < 64; ++r) {
					const uint4 av = As[wr * 64 + r][q / 2];
					acc[r] ^= (av.x & bv[q].x) ^ (av.y & bv[q].y) ^ (av.z & bv[q + 1].x) ^ (av.w & bv[q + 1].y);
				}
			}
			__syncthreads();
		}
	}
	// output word of row r: the ballot of the parities; lane r keeps row r's
	uint64_t mine = 0;
#pragma unroll
	for (int r = 0; r < 64; ++r) {
		const uint64_t word = __ballot(__popc(acc[r]) & 1u);
		if (lane == r)
			mine = word;
	}
	const int row = r0 + wr * 64 + lane;
	const int gw = (c0 >> 6) + wc;
	if (row < m && gw < nw) {
		uint64_t *cp = C + (size_t) row * ldc + gw;
		const uint64_t mask = (gw == nw - 1) ? tail_mask64(nbits - gw * 64) : ~0ull;
		const uint64_t old = *cp;
		const uint64_t v = (b ? old : 0ull) ^ mine;
		*cp = (v & mask) | (old & ~mask);
	}
}

static int gf2_gemm_dot(mlz_ctx *ctx, size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda,
			const uint64_t *B, size_t ldb, int b, uint64_t *C, size_t ldc)
{
	const size_t nw = (n + 63) / 64, kw = (k + 63) / 64;
	const size_t ldw = round_up(nw * 64, G2D_TC);
	void *w = nullptr;
	if (a && kw) {
		MLZ_TRY(ws_get(ctx, WS_GF2_3, std::max((size_t) 1, kw) * ldw * sizeof(uint64_t), &w));
		ProfScope ps(ctx, "gf2_transpose64");
		gf2_transpose64_kernel<<<dim3((unsigned) (ldw / 64), (unsigned) kw), 64, 0, ctx->stream>>>(
			B, ldb, (int) k, (int) nw, (uint64_t *) w, ldw);
		MLZ_TRY(check_launch(ctx, "gf2_transpose64_kernel"));
	}
	ProfScope ps(ctx, "gf2_gemm_dot");
	dim3 g(ceil_div(n, G2D_TC), ceil_div(m, G2D_TR));
	gf2_gemm_dot_kernel<<<g, 256, 0, ctx->stream>>>(A, lda, (const uint64_t *) w, ldw, C, ldc, (int) m, (int) n,
							 (int) kw, (a & 1) && kw > 0, b & 1);
	return check_launch(ctx, "gf2_gemm_dot_kernel");
}

// The method of four Russians; with kz, the launch does nothing when *kz is 0
// (read on the device, so that the host need not wait for it).
// With rest, a live row (is_pivot[row] == 0) with a bit in word rest_w or
// later of its result sets *rest.
static int gf2_gemm_m4rm_dev(mlz_ctx *ctx, size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda,
			     const uint64_t *B, size_t ldb, int b, uint64_t *C, size_t ldc, const int *kz,
			     const int *is_pivot = nullptr, int rest_w = 0, int *rest = nullptr,
			     hipStream_t stream = nullptr)
{
	if (m == 0 || n == 0)
		return MLZ_OK;
	if (!stream)
		stream = ctx->stream;
	const size_t nw = (n + 63) / 64;
	ProfScope ps(ctx, "gf2_gemm_m4rm");
	// 4 rows per thread (1024 per workgroup): 16384^3 in 2.72 ms, against
	// 3.87 ms with 8 (which spills registers) and 3.71 ms with 2 (twice the
	// table builds per row)
	gf2_gemm_m4rm_kernel<4><<<dim3(ceil_div(nw, G2_TW), ceil_div(m, 256 * 4)), 256, 0, stream>>>(
		A, lda, B, ldb, C, ldc, (int) m, (int) n, (int) k, (a & 1) && k > 0, b & 1, kz, is_pivot, rest_w, rest);
	return check_launch(ctx, "gf2_gemm_m4rm_kernel");
}

int gf2_gemm_dev(mlz_ctx *ctx, size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda,
		 const uint64_t *B, size_t ldb, int b, uint64_t *C, size_t ldc)
{
	if (m == 0 || n == 0)
		return MLZ_OK;
	if (ctx->gf2_kernel == 1)
		return gf2_gemm_dot(ctx, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	return gf2_gemm_m4rm_dev(ctx, m, n, k, a, A, lda, B, ldb, b, C, ldc, nullptr);
}

// ===========================================================================
// RREF
// ===========================================================================

constexpr int G2P = 4;              // panel width in words (256 columns)
constexpr int G2_PB = G2P * 64;     // panel width in bits
constexpr int G2_SRPT = 2;           // candidate rows per thread in the pivot search
constexpr int G2_CHUNK = 256 * G2_SRPT;  // rows per fast-path set / tournament leaf
constexpr int G2_MAXLVL = 24;

struct Gf2State {
	int *is_pivot, *rank, *piv_rows, *piv_cols, *k, *J, *prow, *lists_rows, *lists_cnt, *counters, *need_tree;
	int *inv_done;   // the fast path also produced P_J^-1
	int *finished;   // no live row has a bit left: every later panel is empty
	int *rest;       // [panel & 3]: the panel's update left a bit in a live row right of it
	int *kh;         // [panel & 3]: the panel's pivot count (st.k is overwritten by the next panel)
	int *prow_h;     // [panel & 1][G2_PB]: the panel's pivot rows, likewise
};

__device__ inline int bit_of(const uint64_t *row, int col)
{
	return (int) ((row[col >> 6] >> (col & 63)) & 1ull);
}

// Gaussian elimination over GF(2) with the rows in registers: thread t holds
// rows t + 256*q (q < RPT), WPR words each.  Bit columns 0..nsearch-1 are
// processed in order; the pivot of a column is the live row of lowest index
// with that bit set.  reduced = false clears the bit in live rows only
// (echelon form); reduced = true clears it in every other row (Gauss-Jordan).
// Two barriers per column: the arrival slot s_min alternates between two
// words so it can be reset while the other is in use.
// Optionally a second block of words per row (x, XW words) is carried along
// by every row operation without being searched (the identity part of a
// fast-path elimination); keeping it apart leaves the searched words
// addressable by selects rather than scratch memory.
template <int RPT, int WPR, int XW = 0>
__device__ int gf2_reg_eliminate(uint64_t (&v)[RPT][WPR], bool (&alive)[RPT], int nrows, int nsearch, bool reduced,
				 uint64_t *pivot, int *s_min, int *out_rows, int *out_cols,
				 uint64_t (*x)[XW > 0 ? XW : 1] = nullptr)
{
	// Columns where no live row has a bit cannot hold a pivot (XORs of live
	// rows keep them zero there): an OR over the live rows skips them, which
	// makes rank-exhausted panels and zero blocks nearly free.
	__shared__ unsigned long long s_or[WPR];
	const int tid = threadIdx.x;
	if (tid == 0) {
		s_min[0] = INT_MAX;
		s_min[1] = INT_MAX;
	}
	if (tid < WPR)
		s_or[tid] = 0ull;
	__syncthreads();
	{
		uint64_t orv[WPR];
#pragma unroll
		for (int w = 0; w < WPR; ++w)
			orv[w] = 0ull;
#pragma unroll
		for (int q = 0; q < RPT; ++q)
			if (tid + 256 * q < nrows && alive[q]) {
#pragma unroll
				for (int w = 0; w < WPR; ++w)
					orv[w] |= v[q][w];
			}
#pragma unroll
		for (int w = 0; w < WPR; ++w) {
			uint32_t lo = (uint32_t) orv[w], hi = (uint32_t) (orv[w] >> 32);
#pragma unroll
			for (int off = 32; off >= 1; off >>= 1) {
				lo |= (uint32_t) __shfl_xor((int) lo, off, 64);
				hi |= (uint32_t) __shfl_xor((int) hi, off, 64);
			}
			if ((tid & 63) == 0 && (lo | hi))
				atomicOr(&s_or[w], ((unsigned long long) hi << 32) | lo);
		}
	}
	__syncthreads();
	uint64_t live_cols[WPR];
#pragma unroll
	for (int w = 0; w < WPR; ++w)
		live_cols[w] = s_or[w];
	// s_min is two words used in turn by the columns searched (not by column
	// parity: skipped columns would leave the last pivot row in the word the
	// next search reads); the one not in use is cleared for the next search.
	int cnt = 0, ph = 0;
	for (int col = 0; col < nsearch; ++col) {
		if (!((live_cols[col >> 6] >> (col & 63)) & 1ull))
			continue;
		// the wave's lowest candidate row: rows tid + 256q grow with q and,
		// within q, with the lane, so the first q with a candidate decides
		int cand = INT_MAX;
#pragma unroll
		for (int q = 0; q < RPT; ++q) {
			const int r = tid + 256 * q;
			const unsigned long long bal =
				__ballot(r < nrows && alive[q] && ((v[q][col >> 6] >> (col & 63)) & 1ull));
			if (cand == INT_MAX && bal)
				cand = 256 * q + (tid & ~63) + __ffsll((long long) bal) - 1;
		}
		if ((tid & 63) == 0 && cand != INT_MAX)
			atomicMin(&s_min[ph], cand);
		__syncthreads();
		const int pr = s_min[ph];
		if (tid == 0)
			s_min[ph ^ 1] = INT_MAX;
		ph ^= 1;
		if (pr != INT_MAX) {
#pragma unroll
			for (int q = 0; q < RPT; ++q)
				if (tid + 256 * q == pr) {
#pragma unroll
					for (int w = 0; w < WPR; ++w)
						pivot[w] = v[q][w];
#pragma unroll
					for (int w = 0; w < XW; ++w)
						pivot[WPR + w] = x[q][w];
					alive[q] = false;
				}
			if (tid == 0) {
				out_rows[cnt] = pr;
				out_cols[cnt] = col;
			}
		}
		__syncthreads();
		if (pr == INT_MAX)
			continue;
#pragma unroll
		for (int q = 0; q < RPT; ++q) {
			const int r = tid + 256 * q;
			if (r != pr && (reduced || alive[q]) && ((v[q][col >> 6] >> (col & 63)) & 1ull)) {
#pragma unroll
				for (int w = 0; w < WPR; ++w)
					v[q][w] ^= pivot[w];
#pragma unroll
				for (int w = 0; w < XW; ++w)
					x[q][w] ^= pivot[WPR + w];
			}
		}
		++cnt;
	}
	__syncthreads();
	return cnt;
}

constexpr int G2_IW = G2_CHUNK / 64;   // identity words per row (fast path)
constexpr int G2_SCAN = 8;            // 256-row groups per pass of the live-row scan

__device__ inline uint64_t g2_readlane64(uint64_t x, int l)
{
	const uint32_t lo = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) x, l);
	const uint32_t hi = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) (x >> 32), l);
	return ((uint64_t) hi << 32) | lo;
}

// Gauss-Jordan on the fast path's G2_CHUNK rows, eight columns at a time
// (the method of four Russians applied to elimination).  Per strip of eight
// columns: one wave eliminates the rows' 8-bit strip patterns alone (lane l
// holds rows 8l .. 8l+7 as the bytes of one word; a row operation on all
// eight is one multiply and XOR), choosing as gf2_reg_eliminate does the
// lowest live row with the column's bit, and records for every row the
// combination of the strip's pivot rows -- as they stood at the strip's
// start -- that its operations amount to; the 2^k combinations of the k
// pivot rows (panel and carried words) are tabulated in LDS and every row
// XORs its entry once.  Four barriers per strip instead of two per column.
// The result is gf2_reg_eliminate's with reduced = true: every row is
// cleared in every pivot column, and the same pivot rows are chosen.
// IW carried words per row (x): the identity part on the fast path, none in
// the tournament.
template <int IW>
__device__ int gf2_strip_eliminate(uint64_t (&v)[G2_SRPT][G2P], bool (&alive)[G2_SRPT], int nrows, int nsearch,
				   int *out_rows, int *out_cols, uint64_t (&x)[G2_SRPT][IW > 0 ? IW : 1])
{
	constexpr int RW = G2P + IW;   // words of a row: panel, then carried
	static_assert(G2_CHUNK == 512 && G2_SRPT == 2, "one wave holds the strip patterns of 512 rows");
	__shared__ __align__(8) uint8_t sb[G2_CHUNK], sa[G2_CHUNK], sc[G2_CHUNK], sp[G2_CHUNK];
	__shared__ uint64_t prow[8][RW];
	__shared__ uint64_t T[2][16][RW];   // combinations of pivots 0-3 and of 4-7
	__shared__ int sk;
	__shared__ unsigned long long s_or[G2P];
	const int tid = threadIdx.x, lane = tid & 63;
	if (tid < G2P)
		s_or[tid] = 0ull;
	__syncthreads();
	{
		uint64_t orv[G2P] = {0ull, 0ull, 0ull, 0ull};
#pragma unroll
		for (int q = 0; q < G2_SRPT; ++q)
			if (tid + 256 * q < nrows && alive[q])
#pragma unroll
				for (int w = 0; w < G2P; ++w)
					orv[w] |= v[q][w];
#pragma unroll
		for (int w = 0; w < G2P; ++w) {
			uint32_t lo = (uint32_t) orv[w], hi = (uint32_t) (orv[w] >> 32);
#pragma unroll
			for (int off = 32; off >= 1; off >>= 1) {
				lo |= (uint32_t) __shfl_xor((int) lo, off, 64);
				hi |= (uint32_t) __shfl_xor((int) hi, off, 64);
			}
			if (lane == 0 && (lo | hi))
				atomicOr(&s_or[w], ((unsigned long long) hi << 32) | lo);
		}
	}
	__syncthreads();
	uint64_t live_cols[G2P];
#pragma unroll
	for (int w = 0; w < G2P; ++w)
		live_cols[w] = s_or[w];
	int cnt = 0;
	for (int c0 = 0; c0 < nsearch; c0 += 8) {
		const int w = c0 >> 6, sh = c0 & 63;
		const int nbits = min(8, nsearch - c0);
		const uint64_t smask = (1ull << nbits) - 1;
		uint64_t lcw = 0;
#pragma unroll
		for (int ww = 0; ww < G2P; ++ww)
			lcw = ww == w ? live_cols[ww] : lcw;
		if (((lcw >> sh) & smask) == 0)
			continue;                       // no live bit: no pivot in the strip
		// 1. the rows' strip patterns and live flags
#pragma unroll
		for (int q = 0; q < G2_SRPT; ++q) {
			const int r = tid + 256 * q;
			uint64_t vw = 0;
#pragma unroll
			for (int ww = 0; ww < G2P; ++ww)
				vw = ww == w ? v[q][ww] : vw;
			sb[r] = r < nrows ? (uint8_t) ((vw >> sh) & smask) : 0;
			sa[r] = (r < nrows && alive[q]) ? 1 : 0;
		}
		__syncthreads();
		// 2. one wave eliminates the patterns; each word of eight rows is
		// handled as two 32-bit halves (four rows each)
		if (tid < 64) {
			const uint2 b2 = reinterpret_cast<const uint2 *>(sb)[lane];
			const uint2 l2 = reinterpret_cast<const uint2 *>(sa)[lane];
			uint32_t Bl = b2.x, Bh = b2.y, Ll = l2.x, Lh = l2.y, Cl = 0, Ch = 0, Pl = 0, Ph = 0;
			int k = 0;
			for (int j = 0; j < nbits; ++j) {
				const uint32_t ml = (Bl >> j) & 0x01010101u, mh = (Bh >> j) & 0x01010101u;
				const uint32_t al = ml & Ll, ah = mh & Lh;
				const uint64_t bal = __ballot((al | ah) != 0);
				if (bal == 0)
					continue;
				const int pl = __ffsll((long long) bal) - 1;
				// this lane's lowest candidate (8 * row in lane), its pattern and combination
				const int mine = al ? __ffs(al) - 1 : 32 + __ffs(ah) - 1;
				const uint32_t mb = mine < 32 ? (Bl >> mine) & 0xffu : (Bh >> (mine - 32)) & 0xffu;
				const uint32_t mc = mine < 32 ? (Cl >> mine) & 0xffu : (Ch >> (mine - 32)) & 0xffu;
				const int pb = __builtin_amdgcn_readlane(mine, pl);
				const uint32_t pbyte = (uint32_t) __builtin_amdgcn_readlane((int) mb, pl);
				const uint32_t pc = (uint32_t) __builtin_amdgcn_readlane((int) mc, pl) ^ (1u << k);
				uint32_t xl = ml, xh = mh;
				if (lane == pl) {
					if (pb < 32) {
						xl &= ~(0xffu << pb);
						Ll &= ~(0xffu << pb);
						Pl |= (uint32_t) (k + 1) << pb;
					} else {
						xh &= ~(0xffu << (pb - 32));
						Lh &= ~(0xffu << (pb - 32));
						Ph |= (uint32_t) (k + 1) << (pb - 32);
					}
				}
				Bl ^= xl * pbyte;       // one byte per row: no carries
				Bh ^= xh * pbyte;
				Cl ^= xl * pc;
				Ch ^= xh * pc;
				if (lane == 0) {
					out_rows[cnt + k] = 8 * pl + (pb >> 3);
					out_cols[cnt + k] = c0 + j;
				}
				++k;
			}
			reinterpret_cast<uint2 *>(sc)[lane] = make_uint2(Cl, Ch);
			reinterpret_cast<uint2 *>(sp)[lane] = make_uint2(Pl, Ph);   // pivot slot + 1, 0 for other rows
			if (lane == 0)
				sk = k;
		}
		__syncthreads();
		const int k = sk;
		// 3. the strip's pivot rows as they stood, then their combinations
#pragma unroll
		for (int q = 0; q < G2_SRPT; ++q) {
			const int r = tid + 256 * q;
			const int t = (int) sp[r] - 1;
			if (t >= 0) {
#pragma unroll
				for (int ww = 0; ww < G2P; ++ww)
					prow[t][ww] = v[q][ww];
#pragma unroll
				for (int ww = 0; ww < IW; ++ww)
					prow[t][G2P + ww] = x[q][ww];
				alive[q] = false;
			}
		}
		__syncthreads();
		// two tables of 16: a combination is the XOR of one entry of each
		for (int idx = tid; idx < 32 * RW; idx += blockDim.x) {
			const int hb = idx / (16 * RW), e = (idx / RW) % 16, ww = idx % RW;
			uint64_t acc = 0ull;
#pragma unroll
			for (int t = 0; t < 4; ++t)
				if (((e >> t) & 1) && 4 * hb + t < k)
					acc ^= prow[4 * hb + t][ww];
			T[hb][e][ww] = acc;
		}
		__syncthreads();
		// 4. every row applies its combination
#pragma unroll
		for (int q = 0; q < G2_SRPT; ++q) {
			const int r = tid + 256 * q;
			const int cmb = sc[r];
			if (cmb) {
				const int lo = cmb & 15, hi = cmb >> 4;
#pragma unroll
				for (int ww = 0; ww < G2P; ++ww)
					v[q][ww] ^= T[0][lo][ww] ^ T[1][hi][ww];
#pragma unroll
				for (int ww = 0; ww < IW; ++ww)
					x[q][ww] ^= T[0][lo][G2P + ww] ^ T[1][hi][G2P + ww];
			}
		}
		cnt += k;
	}
	__syncthreads();
	return cnt;
}

__device__ inline int g2_nodes_at(int nodes0, int level)
{
	int n = nodes0;
	for (int l = 0; l < level; ++l)
		n = (n + 1) / 2;
	return n;
}

// Fast path: the first G2_CHUNK rows S that are not pivots yet, reduced
// (Gauss-Jordan) together with the identity, [S | I].  If S has full panel
// rank its pivot rows are the panel's, and P_J^-1 comes out of the same
// elimination: pivot rows only ever receive other pivot rows, so the
// identity part of the pivot row of column J[t] has bits on pivot rows only,
// and bit s of it is entry (t, pos of row s) of P_J^-1.
__global__ __launch_bounds__(256) void gf2_fast_kernel(const uint64_t *__restrict__ A, size_t lda, int m, int cw,
						       int nw, int wb, Gf2State st, uint64_t *__restrict__ Pinv)
{
	__shared__ int rowid[G2_CHUNK], outr[G2_PB], outc[G2_PB], pos[G2_CHUNK];
	__shared__ int s_cnt, s_scan[G2_SCAN][4];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	// A panel with pivots records in its update whether a live row kept a
	// bit right of it.  The previous panel's update may still be running
	// (its columns right of this panel's), so the one before it is used.
	const int panel = cw / G2P;
	const int prev_k = panel > 0 ? st.kh[(panel - 1) & 3] : 0;
	if (*st.finished || (panel >= 2 && st.kh[(panel - 2) & 3] > 0 && st.rest[(panel - 2) & 3] == 0)) {
		// every live row is zero to the end of the rows: no pivots are left
		if (tid == 0) {
			*st.finished = 1;
			*st.k = 0;
			*st.need_tree = 0;
			*st.inv_done = 1;
		}
		return;
	}
	if (tid == 0) {
		s_cnt = 0;
		st.rest[panel & 3] = 0;   // for this panel's update
	}
	__syncthreads();
	// the first G2_CHUNK live rows, G2_SCAN groups of 256 rows per pass (their
	// flags loaded together, one prefix count per pass)
	int scanned_all = 1;
	for (int base = 0; base < m; base += 256 * G2_SCAN) {
		bool live[G2_SCAN];
#pragma unroll
		for (int u = 0; u < G2_SCAN; ++u) {
			const int row = base + 256 * u + tid;
			live[u] = row < m && st.is_pivot[row] == 0;
		}
		unsigned long long bal[G2_SCAN];
#pragma unroll
		for (int u = 0; u < G2_SCAN; ++u) {
			bal[u] = __ballot(live[u]);
			if (lane == 0)
				s_scan[u][wave] = __popcll(bal[u]);
		}
		__syncthreads();
		int before = s_cnt;
#pragma unroll
		for (int u = 0; u < G2_SCAN; ++u) {
			int b = before;
			for (int w = 0; w < wave; ++w)
				b += s_scan[u][w];
			const int slot = b + __popcll(bal[u] & ((1ull << lane) - 1));
			if (live[u] && slot < G2_CHUNK)
				rowid[slot] = base + 256 * u + tid;
			before += s_scan[u][0] + s_scan[u][1] + s_scan[u][2] + s_scan[u][3];
		}
		__syncthreads();
		if (tid == 0)
			s_cnt = before;
		__syncthreads();
		if (s_cnt >= G2_CHUNK) {
			// every row was seen and exactly G2_CHUNK are live
			scanned_all = (s_cnt == G2_CHUNK && base + 256 * G2_SCAN >= m);
			break;
		}
	}
	const int nr = min(s_cnt, G2_CHUNK);
	uint64_t v[G2_SRPT][G2P], xi[G2_SRPT][G2_IW];
	bool alive[G2_SRPT];
#pragma unroll
	for (int q = 0; q < G2_SRPT; ++q) {
		const int r = tid + 256 * q;
		alive[q] = r < nr;
#pragma unroll
		for (int w = 0; w < G2P; ++w)
			v[q][w] = (r < nr && cw + w < nw) ? A[(size_t) rowid[r] * lda + cw + w] : 0ull;
#pragma unroll
		for (int w = 0; w < G2_IW; ++w)
			xi[q][w] = (r >> 6) == w ? 1ull << (r & 63) : 0ull;
	}
	const int cnt = gf2_strip_eliminate<G2_IW>(v, alive, nr, wb, outr, outc, xi);
	const bool full = cnt == wb;
	bool all_zero = false;
	if (cnt == 0 && !scanned_all) {
		// no bit in the first rows: if no live row has a bit in the panel
		// the panel has no pivot, and the tournament is not needed
		__shared__ int s_any;
		if (tid == 0)
			s_any = 0;
		__syncthreads();
		// G2_SCAN rows per thread per pass, their loads issued together
		int any = 0;
		for (int base = 0; base < m && !any; base += 256 * G2_SCAN) {
			uint64_t acc = 0ull;
#pragma unroll
			for (int u = 0; u < G2_SCAN; ++u) {
				const int row = base + 256 * u + tid;
				if (row < m && st.is_pivot[row] == 0)
#pragma unroll
					for (int w = 0; w < G2P; ++w)
						if (cw + w < nw)
							acc |= A[(size_t) row * lda + cw + w];
			}
			any |= acc != 0ull;
		}
		if (any)
			s_any = 1;
		__syncthreads();
		all_zero = s_any == 0;
	}
	const bool done = full || scanned_all || all_zero;
	if (cnt == 0 && done && prev_k == 0) {
		// the panel has no pivot and the previous one none either (so no
		// update runs now, and none will tell): if no live row has a bit
		// right of it, the rank is complete and the remaining panels are
		// empty
		__shared__ int s_rest;
		if (tid == 0)
			s_rest = 0;
		__syncthreads();
		const int w0 = cw + G2P;
		int any = 0;
		for (int base = 0; base < m && !any && w0 < nw; base += 256) {
			const int row = base + tid;
			uint64_t acc = 0ull;
			if (row < m && st.is_pivot[row] == 0)
				for (int w = w0; w < nw; ++w)
					acc |= A[(size_t) row * lda + w];
			any = __syncthreads_or(acc != 0ull);
		}
		if (tid == 0 && !any)
			*st.finished = 1;
	}
	if (full) {
		// P_J^-1 row t (pivot of column outc[t]) from the identity part of row outr[t]
		for (int s = tid; s < G2_CHUNK; s += blockDim.x)
			pos[s] = -1;
		__syncthreads();
		for (int t = tid; t < cnt; t += blockDim.x)
			pos[outr[t]] = t;
		__syncthreads();
#pragma unroll
		for (int q = 0; q < G2_SRPT; ++q) {
			const int r = tid + 256 * q;
			const int t = pos[r];
			if (t >= 0) {
				uint64_t out[G2P] = {0, 0, 0, 0};
#pragma unroll
				for (int w = 0; w < G2_IW; ++w) {
					uint64_t bits = xi[q][w];
					while (bits) {
						const int s = 64 * w + __ffsll((long long) bits) - 1;
						bits &= bits - 1;
						const int ts = pos[s];
#pragma unroll
						for (int w2 = 0; w2 < G2P; ++w2)
							out[w2] |= (ts >> 6) == w2 ? 1ull << (ts & 63) : 0ull;
					}
				}
#pragma unroll
				for (int w = 0; w < G2P; ++w)
					Pinv[(size_t) t * G2P + w] = out[w];
			}
		}
		// rows beyond the panel width (a narrow last panel): identity
		for (int t = cnt + tid; t < G2_PB; t += blockDim.x)
#pragma unroll
			for (int w = 0; w < G2P; ++w)
				Pinv[(size_t) t * G2P + w] = (t >> 6) == w ? 1ull << (t & 63) : 0ull;
	}
	if (done)
		for (int t = tid; t < G2_PB; t += blockDim.x) {
			st.J[t] = t < cnt ? outc[t] : 0;
			st.prow[t] = t < cnt ? rowid[outr[t]] : 0;
		}
	if (tid == 0) {
		if (done)
			*st.k = cnt;
		*st.need_tree = done ? 0 : 1;
		*st.inv_done = full ? 1 : 0;
	}
}

__global__ __launch_bounds__(256) void gf2_tourney_kernel(const uint64_t *__restrict__ A, size_t lda, int m,
							  int cw, int nw, int wb, Gf2State st, int nodes0)
{
	__shared__ int rowid[G2_CHUNK], outr[G2_PB], outc[G2_PB];
	__shared__ int s_go, s_nr, s_n0;
	if (*st.need_tree == 0)
		return;
	const int tid = threadIdx.x;
	int node = blockIdx.x, level = 0;
	uint64_t v[G2_SRPT][G2P];
	bool alive[G2_SRPT];
	const int r0 = node * G2_CHUNK;
	const int nr0 = min(G2_CHUNK, m - r0);
	for (int r = tid; r < G2_CHUNK; r += blockDim.x)
		rowid[r] = r0 + r;
#pragma unroll
	for (int q = 0; q < G2_SRPT; ++q) {
		const int r = tid + 256 * q;
		alive[q] = r < nr0 && st.is_pivot[r0 + r] == 0;
#pragma unroll
		for (int w = 0; w < G2P; ++w)
			v[q][w] = (r < nr0 && cw + w < nw) ? A[(size_t) (r0 + r) * lda + cw + w] : 0ull;
	}
	__syncthreads();
	uint64_t none[G2_SRPT][1];
	int cnt = gf2_strip_eliminate<0>(v, alive, nr0, wb, outr, outc, none);
	for (int t = tid; t < cnt; t += blockDim.x)
		outr[t] = rowid[outr[t]];
	__syncthreads();
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
		}
		__syncthreads();
#pragma unroll
		for (int q = 0; q < G2_SRPT; ++q) {
			const int r = tid + 256 * q;
			alive[q] = r < nr;
#pragma unroll
			for (int w = 0; w < G2P; ++w)
				v[q][w] = (r < nr && cw + w < nw) ? A[(size_t) rowid[r] * lda + cw + w] : 0ull;
		}
		cnt = gf2_strip_eliminate<0>(v, alive, nr, wb, outr, outc, none);
		for (int t = tid; t < cnt; t += blockDim.x)
			outr[t] = rowid[outr[t]];
		__syncthreads();
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
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		// the panel's count and pivot rows, for work still running when the
		// next panel's search overwrites st.k and st.prow
		if (t == 0)
			st.kh[panel & 3] = k;
		st.prow_h[(panel & 1) * G2_PB + t] = t < k ? st.prow[t] : 0;
	}
	if (k == 0)
		return;                 // no pivots: P is not read
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

// Inverse of P_J (k x k bits, identity-padded to G2_PB) by Gauss-Jordan on
// [P_J | I] (strip elimination, the identity part carried): row t of thread
// t; the pivot row of column c ends as row c of the inverse.
__global__ __launch_bounds__(256) void gf2_inv_kernel(const uint64_t *__restrict__ P, int nwc, Gf2State st,
						      uint64_t *__restrict__ Pinv)
{
	__shared__ int outr[G2_PB], outc[G2_PB], pos[G2_PB], sJ[G2_PB];
	if (*st.inv_done)
		return;
	const int tid = threadIdx.x;
	const int k = *st.k;
	uint64_t v[G2_SRPT][G2P], x[G2_SRPT][G2P];
	bool alive[G2_SRPT];
#pragma unroll
	for (int q = 0; q < G2_SRPT; ++q) {
		alive[q] = q == 0;
#pragma unroll
		for (int w = 0; w < G2P; ++w)
			v[q][w] = x[q][w] = 0ull;
	}
	sJ[tid] = tid < k ? st.J[tid] : 0;
	__syncthreads();
	{
		const int r = tid;
		if (r < k) {
			// the row's panel words, then its bits on the pivot columns
			uint64_t pw[G2P];
#pragma unroll
			for (int w = 0; w < G2P; ++w)
				pw[w] = w < nwc ? P[(size_t) r * nwc + w] : 0ull;
			for (int s = 0; s < k; ++s) {
				const int j = sJ[s];
				uint64_t word = pw[0];
#pragma unroll
				for (int w = 1; w < G2P; ++w)
					word = (j >> 6) == w ? pw[w] : word;
				const uint64_t bit = ((word >> (j & 63)) & 1ull) << (s & 63);
#pragma unroll
				for (int w = 0; w < G2P; ++w)
					v[0][w] |= (s >> 6) == w ? bit : 0ull;
			}
		} else {
			v[0][r >> 6] |= 1ull << (r & 63);
		}
		x[0][r >> 6] |= 1ull << (r & 63);
		pos[r] = r;
	}
	const int cnt = gf2_strip_eliminate<G2P>(v, alive, G2_PB, k, outr, outc, x);
	for (int s = tid; s < cnt; s += blockDim.x)
		pos[outr[s]] = outc[s];
	__syncthreads();
#pragma unroll
	for (int w = 0; w < G2P; ++w)
		Pinv[(size_t) pos[tid] * G2P + w] = x[0][w];
}

// G[i] = bits J of row i (compressed to k bits), zero on this panel's pivots.
// J is strictly increasing, so J[k-1] = k-1 means J[s] = s (a full-rank
// panel, the usual case): the compressed bits are then the row's first k
// panel bits as they stand.
__global__ void gf2_gatherG_kernel(const uint64_t *__restrict__ A, size_t lda, int m, int cw, int nw, int panel,
				   Gf2State st, uint64_t *__restrict__ G)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int k = *st.k;
	if (i >= m || k == 0)
		return;                 // no pivots: G is not read
	uint64_t row[G2P], g[G2P];
	for (int w = 0; w < G2P; ++w) {
		row[w] = cw + w < nw ? A[(size_t) i * lda + cw + w] : 0ull;
		g[w] = 0;
	}
	if (st.is_pivot[i] != panel + 1) {
		if (k > 0 && st.J[k - 1] == k - 1) {
			for (int w = 0; w < G2P; ++w) {
				const int nb = min(64, max(0, k - 64 * w));
				g[w] = nb == 64 ? row[w] : row[w] & ((1ull << nb) - 1ull);
			}
		} else {
			for (int s = 0; s < k; ++s) {
				const int j = st.J[s];
				if ((row[j >> 6] >> (j & 63)) & 1ull)
					g[s >> 6] |= 1ull << (s & 63);
			}
		}
	}
	for (int w = 0; w < G2P; ++w)
		G[(size_t) i * G2P + w] = g[w];
}

// The panel's pivot rows <- W on words [w0, w1) of the nwc from cw; kp and
// prow: the panel's count and pivot rows.  count: the pivots join the rank.
__global__ void gf2_scatter_kernel(uint64_t *__restrict__ A, size_t lda, int nwc, int cw, int w0, int w1,
				   const int *__restrict__ kp, const int *__restrict__ prow, int *__restrict__ rank,
				   int count, const uint64_t *__restrict__ W)
{
	const int t = blockIdx.y;
	const int k = *kp;
	const int w = w0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (t < k && w < w1)
		A[(size_t) prow[t] * lda + cw + w] = W[(size_t) t * nwc + w];
	// (no workgroup of this launch reads the rank)
	if (count && blockIdx.x == 0 && t == 0 && threadIdx.x == 0)
		*rank += k;
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
		*st.need_tree = 0;
		*st.inv_done = 0;
		*st.finished = 0;
		for (int j = 0; j < 4; ++j) {
			st.rest[j] = 1;
			st.kh[j] = 0;
		}
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
	MLZ_TRY(ws_get(ctx, WS_GF2_1, 2 * (size_t) G2_PB * nw * 8 + 2 * (size_t) G2_PB * G2P * 8, &w1));
	MLZ_TRY(ws_get(ctx, WS_GF2_2, m * G2P * 8, &w2));
	const size_t ints = m + 1 + 2 * rmax + 1 + 2 * G2_PB + (size_t) G2_MAXLVL * nodes0 * G2_PB +
			    2 * (size_t) G2_MAXLVL * nodes0 + 11 + 2 * G2_PB;
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
	st.need_tree = q; q += 1;
	st.inv_done = q; q += 1;
	st.finished = q; q += 1;
	st.rest = q; q += 4;
	st.kh = q; q += 4;
	st.prow_h = q; q += 2 * G2_PB;
	st.counters = q;
	{
		const int cc = G2_MAXLVL * nodes0;
		gf2_init_kernel<<<ceil_div(std::max(m, (size_t) cc), 256), 256, 0, ctx->stream>>>(st, (int) m, cc);
		gf2_load_kernel<<<ceil_div(m * nw, 256), 256, 0, ctx->stream>>>(A, lda, (int) m, (int) n, Wm);
		MLZ_TRY(check_launch(ctx, "gf2_load_kernel"));
	}
	// Look-ahead: a panel's update of its own and the next panel's words
	// (the narrow part) runs on the context's stream, the rest (the wide
	// part: Wb's remaining words, the update and the pivot rows' scatter
	// there) on the second stream, beside the next panel's search.  The next
	// panel's gather waits for the wide part; Pinv alternates between two
	// buffers (the search writes it), k and the pivot rows are read from the
	// panel's snapshots.
	struct Overlap {
		hipStream_t s1 = nullptr;
		hipEvent_t en = nullptr, ew = nullptr;
		~Overlap()
		{
			if (s1)
				(void) hipStreamSynchronize(s1);
			if (en)
				(void) hipEventDestroy(en);
			if (ew)
				(void) hipEventDestroy(ew);
		}
	} ov;
	if (nw > 2 * (size_t) G2P) {
		MLZ_TRY(aux_stream(ctx, &ov.s1));
		MLZ_HIP(ctx, hipEventCreateWithFlags(&ov.en, hipEventDisableTiming));
		MLZ_HIP(ctx, hipEventCreateWithFlags(&ov.ew, hipEventDisableTiming));
	}
	bool wide_live = false;
	int panel = 0;
	for (size_t cw = 0; cw < nw; cw += G2P, ++panel) {
		const int wb = (int) std::min((size_t) G2_PB, n - cw * 64);
		const int nwc = (int) (nw - cw);
		const int nn = std::min(nwc, 2 * G2P);
		uint64_t *Pi = Pinv + (size_t) (panel & 1) * G2_PB * G2P;
		const int *kp = st.kh + (panel & 3);
		const int *prow = st.prow_h + (panel & 1) * G2_PB;
		{
			ProfScope ps(ctx, "gf2_pivots");
			gf2_fast_kernel<<<1, 256, 0, ctx->stream>>>(Wm, nw, (int) m, (int) cw, (int) nw, wb, st, Pi);
			gf2_tourney_kernel<<<nodes0, 256, 0, ctx->stream>>>(Wm, nw, (int) m, (int) cw, (int) nw, wb, st, nodes0);
		}
		MLZ_TRY(check_launch(ctx, "gf2_tourney_kernel"));
		if (wide_live) {
			MLZ_HIP(ctx, hipStreamWaitEvent(ctx->stream, ov.ew, 0));
			wide_live = false;
		}
		gf2_gather_kernel<<<dim3(ceil_div(nwc, 256), G2_PB), 256, 0, ctx->stream>>>(Wm, nw, nwc, (int) cw, panel, st, P);
		gf2_inv_kernel<<<1, 256, 0, ctx->stream>>>(P, nwc, st, Pi);
		MLZ_TRY(check_launch(ctx, "gf2_inv_kernel"));
		MLZ_TRY(gf2_gemm_m4rm_dev(ctx, G2_PB, (size_t) nn * 64, G2_PB, 1, Pi, G2P, P, nwc, 0, Wb, nwc, kp));
		gf2_gatherG_kernel<<<ceil_div(m, 256), 256, 0, ctx->stream>>>(Wm, nw, (int) m, (int) cw, (int) nw, panel, st, G);
		MLZ_TRY(gf2_gemm_m4rm_dev(ctx, m, (size_t) nn * 64, G2_PB, 1, G, G2P, Wb, nwc, 1, Wm + cw, nw, kp,
					  st.is_pivot, G2P, st.rest + (panel & 3)));
		gf2_scatter_kernel<<<dim3(ceil_div(nn, 256), G2_PB), 256, 0, ctx->stream>>>(Wm, nw, nwc, (int) cw, 0, nn, kp,
												prow, st.rank, 1, Wb);
		MLZ_TRY(check_launch(ctx, "gf2_scatter_kernel"));
		if (nwc > nn) {
			MLZ_HIP(ctx, hipEventRecord(ov.en, ctx->stream));
			MLZ_HIP(ctx, hipStreamWaitEvent(ov.s1, ov.en, 0));
			MLZ_TRY(gf2_gemm_m4rm_dev(ctx, G2_PB, (size_t) (nwc - nn) * 64, G2_PB, 1, Pi, G2P, P + nn, nwc, 0,
						  Wb + nn, nwc, kp, nullptr, 0, nullptr, ov.s1));
			MLZ_TRY(gf2_gemm_m4rm_dev(ctx, m, (size_t) (nwc - nn) * 64, G2_PB, 1, G, G2P, Wb + nn, nwc, 1,
						  Wm + cw + nn, nw, kp, st.is_pivot, 0, st.rest + (panel & 3), ov.s1));
			gf2_scatter_kernel<<<dim3(ceil_div(nwc - nn, 256), G2_PB), 256, 0, ov.s1>>>(
				Wm, nw, nwc, (int) cw, nn, nwc, kp, prow, st.rank, 0, Wb);
			MLZ_TRY(check_launch(ctx, "gf2_scatter_kernel"));
			MLZ_HIP(ctx, hipEventRecord(ov.ew, ov.s1));
			wide_live = true;
		}
		if ((panel & 3) == 3) {
			int r = 0, fin = 0;
			MLZ_HIP(ctx, hipMemcpyAsync(&r, st.rank, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
			MLZ_HIP(ctx, hipMemcpyAsync(&fin, st.finished, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
			MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
			if ((size_t) r >= m || fin)
				break;
		}
	}
	if (wide_live)
		MLZ_HIP(ctx, hipStreamWaitEvent(ctx->stream, ov.ew, 0));
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

// ---------------------------------------------------------------------------
// trsm over GF(2): B <- a A^-1 B, A triangular (lela/blas/level3-generic.tcc:
// 421-498; the M4RI module's mzd_trsm_*_left, level3-m4ri.tcc:88-109, takes
// the diagonal as ones).  The 64 x 64 diagonal blocks are inverted up front;
// the recursion splits at multiples of 64 and updates with the GF(2) gemm;
// a leaf applies its block's inverse.
// ---------------------------------------------------------------------------

// Inverse of diagonal block b (rows and columns 64b .. 64b+63; a partial last
// block is completed by the identity), one wave per block: lane i holds row i
// of the block (its word b) and of the inverse.  Lower: x_i = e_i + sum over
// j < i with A_ij = 1 of x_j, so column j is added to the rows below it once
// x_j is final (j ascending); upper likewise with j descending.  Without
// unit_diag a zero diagonal bit is recorded (bad[0] = the smallest index,
// bad[1] = the largest).
__global__ __launch_bounds__(64) void gf2_trsm_inv_kernel(const uint64_t *__restrict__ A, size_t lda, int n, int upper,
							   int unit, uint64_t *__restrict__ W,
							   unsigned long long *__restrict__ bad)
{
	const int b = blockIdx.x, i = threadIdx.x;
	const int r0 = 64 * b, rows = min(64, n - r0);
	uint64_t a = 0;
	if (i < rows) {
		a = A[(size_t) (r0 + i) * lda + b] & tail_mask64(rows);   // columns of the block inside A
		if (!unit && !((a >> i) & 1ull)) {
			atomicMin(&bad[0], (unsigned long long) (r0 + i));
			atomicMax(&bad[1], (unsigned long long) (r0 + i));
		}
	}
	// the strict triangle only
	a &= upper ? (i == 63 ? 0ull : ~0ull << (i + 1)) : ((1ull << i) - 1ull);
	uint64_t x = 1ull << i;
	for (int s = 0; s < 64; ++s) {
		const int j = upper ? 63 - s : s;
		const uint64_t xj = __shfl(x, j, 64);
		if ((a >> j) & 1ull)
			x ^= xj;
	}
	W[(size_t) b * 64 + i] = x;
}

// The block's rows of B (Bb: its first row, `rows` rows) <- X_b times them
// (X_b = W's block b): output (i, w) is the XOR of the rows j whose bit is set
// in X_b[i].  Workgroup: G2T_CW words of the rows, staged in LDS.  Bits past
// nrhs in the last word keep their values.
constexpr int G2T_CW = 32;
__global__ __launch_bounds__(256) void gf2_trsm_leaf_kernel(const uint64_t *__restrict__ W, int b, int rows,
							    uint64_t *__restrict__ Bb, size_t ldb, int nrhs)
{
	__shared__ uint64_t Bt[64][G2T_CW + 1];
	__shared__ uint64_t Xs[64];
	const int tid = threadIdx.x;
	const int nw = (nrhs + 63) / 64, w0 = blockIdx.x * G2T_CW;
	if (tid < 64)
		Xs[tid] = W[(size_t) b * 64 + tid];
	for (int e = tid; e < 64 * G2T_CW; e += 256) {
		const int r = e / G2T_CW, w = e % G2T_CW;
		Bt[r][w] = (r < rows && w0 + w < nw) ? Bb[(size_t) r * ldb + w0 + w] : 0ull;
	}
	__syncthreads();
	for (int e = tid; e < 64 * G2T_CW; e += 256) {
		const int r = e / G2T_CW, w = e % G2T_CW, gw = w0 + w;
		if (r >= rows || gw >= nw)
			continue;
		uint64_t acc = 0, sel = Xs[r];
		while (sel) {
			const int j = __builtin_ctzll(sel);
			acc ^= Bt[j][w];
			sel &= sel - 1;
		}
		const uint64_t mask = gw == nw - 1 ? tail_mask64(nrhs - gw * 64) : ~0ull;
		Bb[(size_t) r * ldb + gw] = (acc & mask) | (Bt[r][w] & ~mask);
	}
}

static int gf2_trsm_rec(mlz_ctx *ctx, int upper, size_t n, size_t nrhs, const uint64_t *A, size_t lda, uint64_t *B,
			size_t ldb, const uint64_t *W, size_t b0)
{
	if (n <= 64) {
		gf2_trsm_leaf_kernel<<<ceil_div(ceil_div(nrhs, 64), G2T_CW), 256, 0, ctx->stream>>>(W, (int) b0, (int) n, B,
												   ldb, (int) nrhs);
		return check_launch(ctx, "gf2_trsm_leaf_kernel");
	}
	// split at a multiple of 64, as LELA's aligned case (level3-generic.tcc:474)
	const size_t l = std::max((size_t) 64, (n / 128) * 64), lw = l / 64;
	uint64_t *B2 = B + l * ldb;
	if (!upper) {
		MLZ_TRY(gf2_trsm_rec(ctx, upper, l, nrhs, A, lda, B, ldb, W, b0));
		MLZ_TRY(gf2_gemm_dev(ctx, n - l, nrhs, l, 1, A + l * lda, lda, B, ldb, 1, B2, ldb));   // B2 += A21 B1
		return gf2_trsm_rec(ctx, upper, n - l, nrhs, A + l * lda + lw, lda, B2, ldb, W, b0 + lw);
	}
	MLZ_TRY(gf2_trsm_rec(ctx, upper, n - l, nrhs, A + l * lda + lw, lda, B2, ldb, W, b0 + lw));
	MLZ_TRY(gf2_gemm_dev(ctx, l, nrhs, n - l, 1, A + lw, lda, B2, ldb, 1, B, ldb));         // B1 += A12 B2
	return gf2_trsm_rec(ctx, upper, l, nrhs, A, lda, B, ldb, W, b0);
}

// B (rows x nbits) <- 0, bits past nbits in the last word kept.
__global__ void gf2_clear_kernel(uint64_t *__restrict__ B, size_t ldb, int nbits)
{
	const int nw = (nbits + 63) / 64, w = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
	if (w >= nw)
		return;
	const uint64_t mask = w == nw - 1 ? tail_mask64(nbits - w * 64) : ~0ull;
	B[(size_t) r * ldb + w] &= ~mask;
}

__global__ void gf2_trsm_bad_init_kernel(unsigned long long *bad)
{
	bad[0] = ~0ull;
	bad[1] = 0ull;
}

int gf2_trsm_dev(mlz_ctx *ctx, int upper, int unit, size_t n, size_t nrhs, int a, const uint64_t *A, size_t lda,
		 uint64_t *B, size_t ldb, size_t *bad_diag)
{
	if (n == 0 || nrhs == 0)
		return MLZ_OK;
	if (!(a & 1)) {              // a = 0: B <- 0 (level3-generic.tcc:434-435)
		gf2_clear_kernel<<<dim3(ceil_div(ceil_div(nrhs, 64), 256), (unsigned) n), 256, 0, ctx->stream>>>(B, ldb,
													     (int) nrhs);
		return check_launch(ctx, "gf2_clear_kernel");
	}
	const size_t nb = ceil_div(n, 64);
	void *w;
	MLZ_TRY(ws_get(ctx, WS_TRSM_0, nb * 64 * 8 + 256, &w));
	uint64_t *W = (uint64_t *) w;
	unsigned long long *bad = (unsigned long long *) (W + nb * 64);
	gf2_trsm_bad_init_kernel<<<1, 1, 0, ctx->stream>>>(bad);
	gf2_trsm_inv_kernel<<<(unsigned) nb, 64, 0, ctx->stream>>>(A, lda, (int) n, upper, unit, W, bad);
	MLZ_TRY(check_launch(ctx, "gf2_trsm_inv_kernel"));
	if (!unit) {
		unsigned long long hb[2];
		MLZ_HIP(ctx, hipMemcpyAsync(hb, bad, sizeof hb, hipMemcpyDeviceToHost, ctx->stream));
		MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
		if (hb[0] != ~0ull) {
			// the recursion meets lower blocks top-down and upper ones bottom-up
			if (bad_diag)
				*bad_diag = (size_t) (upper ? hb[1] : hb[0]);
			return set_error(ctx, MLZ_E_NOT_INVERTIBLE, "Diagonal entry not invertible");
		}
	}
	return gf2_trsm_rec(ctx, upper, n, nrhs, A, lda, B, ldb, W, 0);
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

extern "C" int mlz_gf2_trsm(mlz_ctx *ctx, int upper, int unit_diag, size_t n, size_t nrhs, int a, const uint64_t *A,
			    size_t lda_w, uint64_t *B, size_t ldb_w, size_t *bad_diag)
{
	if (!ctx)
		return MLZ_E_ARG;
	const size_t nw = (n + 63) / 64, rw = (nrhs + 63) / 64;
	if ((n && (!A || lda_w < nw)) || (n && nrhs && (!B || ldb_w < rw)))
		return set_error(ctx, MLZ_E_ARG, "bad GF(2) matrix pointer or leading dimension");
	if (n > (size_t) INT32_MAX / 2 || nrhs > (size_t) INT32_MAX / 2)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (ctx->ptr_mode == MLZ_PTR_DEVICE)
		return gf2_trsm_dev(ctx, upper, unit_diag, n, nrhs, a, A, lda_w, B, ldb_w, bad_diag);
	if (n == 0 || nrhs == 0)
		return MLZ_OK;
	void *dA = nullptr, *dB = nullptr;
	MLZ_TRY(stage_in(ctx, WS_STAGE_A, A, n, nw, lda_w, 8, &dA));
	MLZ_TRY(stage_in(ctx, WS_STAGE_B, B, n, rw, ldb_w, 8, &dB));
	const int rc = gf2_trsm_dev(ctx, upper, unit_diag, n, nrhs, a, (const uint64_t *) dA, nw, (uint64_t *) dB, rw,
				    bad_diag);
	if (rc != MLZ_OK)
		return rc;
	return stage_out(ctx, B, dB, n, rw, ldb_w, 8);
}
