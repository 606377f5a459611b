// lela_amd/csrc/zp_trsm.hip -- B <- a * A^-1 * B over Z/p, A triangular.
//
// Follows the recursion of _trsm<Ring, GenericModule::Tag>::trsm_impl
// (lela/blas/level3-generic.tcc:421-498): split at l, solve one half, update
// the other with gemm(-1, A21, X1, a, B2) (or A12 for upper), solve the rest.
// The gemm updates run on the INT8 / FP64 / plain kernels of zp_gemm.hip.
// The recursion splits at multiples of TRSM_LEAF; all diagonal blocks of that
// size are inverted up front by one launch, and each leaf is then the small
// product B_b <- a * inv(A_bb) * B_b (on INT8 MFMA for p <= 2^16 unless the plain or FP64 path
// was asked for).
//
// Non-invertible diagonal entries are found up front by one kernel (modular
// inverse of every diagonal entry); the reported index is the one LELA's
// recursion throws first: the smallest for lower, the largest for upper
// (level3-generic.tcc:441-442,461-467,485-491).  B is left untouched then.
#include "common.hpp"

#include <vector>

namespace mlz {

constexpr int TRSM_LEAF = (int) TRSM_LEAF_ROWS;

// Modular inverse by the extended Euclidean algorithm; 0 if gcd(y, p) != 1.
__device__ inline int inv_mod(uint32_t y, uint32_t p, uint32_t *out)
{
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
	*out = (uint32_t) s0;
	return 1;
}

template <int E>
__global__ void diag_inv_kernel(const typename Elem<E>::T *A, size_t lda, int n, uint32_t p,
				uint32_t *dinv, unsigned long long *bad /* [min, max] */)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint32_t v = Elem<E>::load(A + (size_t) i * lda + i, p), x = 0;
	if (!inv_mod(v, p, &x)) {
		atomicMin(&bad[0], (unsigned long long) i);
		atomicMax(&bad[1], (unsigned long long) i);
		x = 0;
	}
	dinv[i] = x;
}

__global__ void init_bad_kernel(unsigned long long *bad)
{
	bad[0] = ~0ull;
	bad[1] = 0ull;
}

// Inverts every diagonal block of TRSM_LEAF rows in one launch, one
// workgroup of 256 threads per block (W receives block b's inverse in rows
// b*L .. b*L+nb-1, L columns each), in log depth: the eight 8 x 8 diagonal
// blocks first (64 threads, one column each, 8 substitution steps:
// x_i = (e_ij - sum_k A_ik x_k) / A_ii), then three doubling levels, each
// joining pairs of inverted s x s blocks:
//   lower [[P, 0], [C, Q]]^-1 = [[P^-1, 0], [-Q^-1 C P^-1, Q^-1]],
//   upper [[P, B], [0, Q]]^-1 = [[P^-1, -P^-1 B Q^-1], [0, Q^-1]],
// as two products T = C P^-1 (B Q^-1), then -Q^-1 T (-P^-1 T), all pairs of
// a level at once, one barrier after each.  A 64-step substitution is one
// dependent chain of reductions; here the longest chain is 8 steps plus six
// sums of at most 32 terms.  Sums are delayed, one reduction per entry: raw
// 64-bit products for p < 2^26 (32 terms below 2^57), reduced products
// otherwise (below 2^37).  A partial block is completed by the identity.
template <int E, bool UPPER, bool UNIT, bool SMALL>
__global__ __launch_bounds__(256) void trsm_diag_inv_blk_kernel(const typename Elem<E>::T *__restrict__ A, size_t lda,
								int n, const uint32_t *__restrict__ dinv,
								uint32_t *__restrict__ W, ModP mp)
{
	constexpr int L = TRSM_LEAF;
	__shared__ uint32_t As[L][L + 1];      // the block's strict triangle
	__shared__ uint32_t X[L][L + 1];       // its inverse as it is formed
	__shared__ uint32_t Tm[L * L / 4];     // the first products of a level
	__shared__ uint32_t Ds[L];             // inverses of the diagonal entries
	const uint32_t p = (uint32_t) mp.p;
	const int tid = threadIdx.x;
	const int b0 = blockIdx.x * L;
	const int nb = min(L, n - b0);
	auto prod = [&](uint32_t x, uint32_t y) -> uint64_t {
		return SMALL ? (uint64_t) x * y : (uint64_t) mp.mul(x, y);
	};
	{
		// row tid/4, columns 16 (tid%4) .. +15; all loads before the stores
		const int r = tid >> 2, c0 = 16 * (tid & 3);
		uint32_t v[16];
#pragma unroll
		for (int q = 0; q < 16; ++q) {
			const int c = c0 + q;
			const bool tri = UPPER ? (c > r) : (c < r);
			v[q] = (r < nb && c < nb && tri) ? Elem<E>::load(A + (size_t) (b0 + r) * lda + b0 + c, p) : 0u;
		}
#pragma unroll
		for (int q = 0; q < 16; ++q) {
			As[r][c0 + q] = v[q];
			X[r][c0 + q] = 0u;
		}
		if (tid < L)
			Ds[tid] = (UNIT || tid >= nb) ? 1u : dinv[b0 + tid];
	}
	__syncthreads();
	if (tid < L) {           // column c of the inverse of diagonal block d (8 x 8)
		const int o = 8 * (tid >> 3), c = tid & 7;
		uint32_t x[8];
#pragma unroll
		for (int s_ = 0; s_ < 8; ++s_) {
			const int i = UPPER ? 7 - s_ : s_;
			uint64_t acc = 0;
#pragma unroll
			for (int t = 0; t < s_; ++t) {
				const int k = UPPER ? 7 - t : t;     // the rows solved before i
				acc += prod(As[o + i][o + k], x[k]);
			}
			uint32_t xi = mp.sub(i == c ? 1u : 0u, (uint32_t) mp.reduce(acc));
			if (!UNIT)
				xi = mp.mul(xi, Ds[o + i]);
			x[i] = xi;
		}
#pragma unroll
		for (int i = 0; i < 8; ++i)
			X[o + i][o + c] = x[i];
	}
	__syncthreads();
#pragma unroll
	for (int s = 8; s < L; s *= 2) {
		constexpr int NT = 256;
		const int per = (L / (2 * s)) * s * s;     // outputs per product: 32 s
		// T = C P^-1 (lower) or B Q^-1 (upper) for every pair
		for (int idx = tid; idx < per; idx += NT) {
			const int pr = idx / (s * s), rem = idx % (s * s), r = rem / s, c = rem % s, o = 2 * s * pr;
			uint64_t acc = 0;
			for (int k = 0; k < s; ++k)
				acc += UPPER ? prod(As[o + r][o + s + k], X[o + s + k][o + s + c])
					     : prod(As[o + s + r][o + k], X[o + k][o + c]);
			Tm[idx] = (uint32_t) mp.reduce(acc);
		}
		__syncthreads();
		// the off-diagonal block of the pair's inverse: -Q^-1 T (lower) or -P^-1 T (upper)
		for (int idx = tid; idx < per; idx += NT) {
			const int pr = idx / (s * s), rem = idx % (s * s), r = rem / s, c = rem % s, o = 2 * s * pr;
			const uint32_t *Tp = Tm + pr * s * s;
			uint64_t acc = 0;
			for (int k = 0; k < s; ++k)
				acc += UPPER ? prod(X[o + r][o + k], Tp[k * s + c]) : prod(X[o + s + r][o + s + k], Tp[k * s + c]);
			const uint32_t v = mp.neg((uint32_t) mp.reduce(acc));
			if (UPPER)
				X[o + r][o + s + c] = v;
			else
				X[o + s + r][o + c] = v;
		}
		__syncthreads();
	}
	{
		const int r = tid >> 2, c0 = 16 * (tid & 3);
		if (r < nb)
#pragma unroll
			for (int q = 0; q < 16; ++q)
				W[(size_t) (b0 + r) * L + c0 + q] = c0 + q < nb ? X[r][c0 + q] : 0u;
	}
}

// Four canonical residues (p <= 2^16) as INT8 limb bytes: hi holds the high
// limbs of v0..v3 in bytes 0..3, lo the low ones.  The limb pair of v is the
// 16-bit value v + 32896 (less p above 32639) with both bytes offset by 0x80:
// the signed split of i8_limbs.  Two byte permutes gather the bytes of the
// four 16-bit values.
__device__ __forceinline__ void limb_pack4(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3, uint32_t p,
					   uint32_t &hi, uint32_t &lo)
{
	auto w = [&](uint32_t v) -> uint32_t { return v > 32639u ? v + 32896u - p : v + 32896u; };
	const uint32_t x01 = w(v0) | (w(v1) << 16), x23 = w(v2) | (w(v3) << 16);
	lo = __builtin_amdgcn_perm(x23, x01, 0x06040200u) ^ 0x80808080u;
	hi = __builtin_amdgcn_perm(x23, x01, 0x07050301u) ^ 0x80808080u;
}

// One product of the doubling steps of trsm_block_inverses:
// C_z = a * A_z * B_z for the s x s blocks of batch member z (s a multiple of
// 64), A in the caller's storage, B and C canonical uint32 (p <= 2^16), on
// INT8 MFMA with the limb split done in LDS, so that no separate packing
// launch is needed.  A workgroup computes one 64 x 64 tile of C (4 waves,
// 32 x 32 each) over k-chunks of 64.  The limb split is the kernel's main
// VALU work (every workgroup splits the entries it reads), so the loads are
// 16-byte ones where they can be: B as 4 x 4 blocks (four rows of four
// columns; a thread then holds four k-consecutive entries of each column), A
// as runs of four k-consecutive entries when AV (rows aligned to four
// entries); limb_pack4 forms each 32-bit fragment word.  k <= 512 keeps every
// limb sum far below 2^31 and the combined value an exact double.
template <int EA, bool AV>
__global__ __launch_bounds__(256) void trsm_dbl_kernel(const typename Elem<EA>::T *__restrict__ A, size_t lda,
						       const uint32_t *__restrict__ B, size_t ldb,
						       uint32_t *__restrict__ C, size_t ldc, int s, uint32_t a,
						       I8Batch bt, ModP mp)
{
	typedef typename Elem<EA>::T TA;
	__shared__ __attribute__((aligned(16))) int8_t Af[2][2][2][64][16];   // [row block][k block][limb]
	__shared__ __attribute__((aligned(16))) int8_t Bf[2][2][2][64][16];   // [column block][k block][limb]
	const uint32_t p = (uint32_t) mp.p;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int nt = s >> 6, z = blockIdx.y;
	const int r0 = (blockIdx.x / nt) * 64, c0 = (blockIdx.x % nt) * 64;
	const TA *Az = A + bt.offset(z, 0) + (size_t) r0 * lda;
	const uint32_t *Bz = B + bt.offset(z, 1) + c0;
	uint32_t *Cz = C + bt.offset(z, 2) + (size_t) r0 * ldc + c0;
	const int rb = wave >> 1, cb = wave & 1;
	v16i acc0 = v16i{0}, acc1 = v16i{0}, acc2 = v16i{0};
	// A: groups g = tid + 256 q (row g >> 4, k 4 (g & 15) ..); B: the 4 x 4
	// block at k 4 (tid >> 4) .., columns 4 (tid & 15) ..  The next chunk's
	// entries are loaded into registers while this one is multiplied (two
	// chunks ahead, in two LDS buffers, measured slower: 120 against 90 us for
	// the eight launches of an 8192-row solve; chunks of 128 measured equal)
	const int kq = 4 * (tid >> 4), cq = 4 * (tid & 15);
	uint32_t ra[4][4], rbv[4][4];
	auto load = [&](int k0) {
#pragma unroll
		for (int q = 0; q < 4; ++q) {
			const int g = tid + 256 * q;
			const TA *src = Az + (size_t) (g >> 4) * lda + k0 + 4 * (g & 15);
			if constexpr (AV && sizeof(TA) == 2) {
				const uint2 u = *(const uint2 *) src;
				ra[q][0] = u.x & 0xffffu;
				ra[q][1] = u.x >> 16;
				ra[q][2] = u.y & 0xffffu;
				ra[q][3] = u.y >> 16;
			} else if constexpr (AV && sizeof(TA) == 4) {
				const uint4 u = *(const uint4 *) src;
				ra[q][0] = u.x;
				ra[q][1] = u.y;
				ra[q][2] = u.z;
				ra[q][3] = u.w;
			} else {
#pragma unroll
				for (int t = 0; t < 4; ++t)
					ra[q][t] = Elem<EA>::load(src + t, p);
			}
		}
#pragma unroll
		for (int t = 0; t < 4; ++t) {
			const uint4 u = *(const uint4 *) (Bz + (size_t) (k0 + kq + t) * ldb + cq);
			rbv[t][0] = u.x;
			rbv[t][1] = u.y;
			rbv[t][2] = u.z;
			rbv[t][3] = u.w;
		}
	};
	auto put = [&]() {
#pragma unroll
		for (int q = 0; q < 4; ++q) {
			const int g = tid + 256 * q;
			const int r = g >> 4, ka = 4 * (g & 15);
			uint32_t hi, lo;
			limb_pack4(ra[q][0], ra[q][1], ra[q][2], ra[q][3], p, hi, lo);
			const int lna = (r & 31) + 32 * ((ka & 31) >> 4);
			*(uint32_t *) &Af[r >> 5][ka >> 5][0][lna][ka & 15] = hi;
			*(uint32_t *) &Af[r >> 5][ka >> 5][1][lna][ka & 15] = lo;
		}
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const int c = cq + j;
			uint32_t hi, lo;
			limb_pack4(rbv[0][j], rbv[1][j], rbv[2][j], rbv[3][j], p, hi, lo);
			const int lnb = (c & 31) + 32 * ((kq & 31) >> 4);
			*(uint32_t *) &Bf[c >> 5][kq >> 5][0][lnb][kq & 15] = hi;
			*(uint32_t *) &Bf[c >> 5][kq >> 5][1][lnb][kq & 15] = lo;
		}
	};
	load(0);
	for (int k0 = 0; k0 < s; k0 += 64) {
		put();
		__syncthreads();
		if (k0 + 64 < s)
			load(k0 + 64);
#pragma unroll
		for (int kb = 0; kb < 2; ++kb) {
			const v4i ah = *(const v4i *) &Af[rb][kb][0][lane][0], al = *(const v4i *) &Af[rb][kb][1][lane][0];
			const v4i bh = *(const v4i *) &Bf[cb][kb][0][lane][0], bl = *(const v4i *) &Bf[cb][kb][1][lane][0];
			acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, acc0, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bl, acc1, 0, 0, 0);
			acc2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bl, acc2, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh, acc1, 0, 0, 0);
		}
		__syncthreads();
	}
	const int col = cb * 32 + (lane & 31);
#pragma unroll
	for (int r = 0; r < 16; ++r) {
		const int row = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
		const double v = fma((double) acc0[r], 65536.0, fma((double) acc1[r], 256.0, (double) acc2[r]));
		Cz[(size_t) row * ldc + col] = mp.mod_f64((double) a * (double) mp.mod_f64(v));
	}
}

// Leaf solve B_b <- a * W_b * B_b for one diagonal block (nb <= L rows), in
// place: each workgroup owns 64 right-hand-side columns.
template <int E, bool WIDE>
__global__ __launch_bounds__(256) void trsm_leaf_apply_kernel(const uint32_t *__restrict__ W,
							      typename Elem<E>::T *__restrict__ B, size_t ldb,
							      int nb, int nrhs, uint32_t a, ModP mp,
							      uint64_t two_64)
{
	constexpr int L = TRSM_LEAF;
	__shared__ uint32_t Ws[L][L + 1];
	__shared__ uint32_t Bs[L][64 + 1];
	const uint32_t p = (uint32_t) mp.p;
	const int j0 = blockIdx.x * 64;
	for (int idx = threadIdx.x; idx < L * L; idx += blockDim.x) {
		int r = idx / L, c = idx % L;
		Ws[r][c] = (r < nb && c < nb) ? W[(size_t) r * L + c] : 0u;
	}
	for (int idx = threadIdx.x; idx < L * 64; idx += blockDim.x) {
		int r = idx / 64, c = idx % 64;
		Bs[r][c] = (r < nb && j0 + c < nrhs) ? Elem<E>::load(B + (size_t) r * ldb + j0 + c, p) : 0u;
	}
	__syncthreads();
	const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
	for (int q = 0; q < L / 4; ++q) {
		const int r = r0 + 4 * q;
		uint64_t acc = 0;
		for (int t = 0; t < nb; ++t) {
			uint64_t prod = (uint64_t) Ws[r][t] * Bs[t][c];
			acc += prod;
			if (WIDE && acc < prod)
				acc += two_64;
		}
		if (r < nb && j0 + c < nrhs)
			Elem<E>::store(B + (size_t) r * ldb + j0 + c, mp.mul((uint32_t) mp.reduce(acc), a), p);
	}
}

// The same leaf on INT8 MFMA (p <= 2^16): a workgroup of 4 waves owns 64
// right-hand-side columns.  W_b and the 64 x 64 block of B are turned into
// limb fragments in LDS (A operand: element (r, k) is byte k%16 of lane
// r%32 + 32*((k%32)/16) of fragment (r/32, k/32); B operand likewise with the
// column in place of r), each wave computes one 32 x 32 block of W_b * B_b with
// eight MFMAs, and the block is written back after the whole strip was read.
template <int E>
__global__ __launch_bounds__(256) void trsm_leaf_i8_kernel(const uint32_t *__restrict__ W,
							   typename Elem<E>::T *__restrict__ B, size_t ldb, int nb,
							   int nrhs, uint32_t a, ModP mp)
{
	constexpr int L = TRSM_LEAF;
	__shared__ __attribute__((aligned(16))) int8_t Wf[2][2][2][64][16];   // [row block][k block][limb]
	__shared__ __attribute__((aligned(16))) int8_t Bf[2][2][2][64][16];   // [column block][k block][limb]
	const uint32_t p = (uint32_t) mp.p;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int j0 = blockIdx.x * 64;
	for (int e = tid; e < L * L; e += blockDim.x) {
		const int r = e / L, k = e % L;
		const uint32_t wv = (r < nb && k < nb) ? W[(size_t) r * L + k] : 0u;
		const uint32_t lw = i8_limbs(wv, p);
		const int ln = (r & 31) + 32 * ((k & 31) >> 4), by = k & 15;
		Wf[r >> 5][k >> 5][0][ln][by] = (int8_t) (lw >> 8);
		Wf[r >> 5][k >> 5][1][ln][by] = (int8_t) lw;
		// B element (k = e / L, column e % L): rows of B are contiguous in the columns
		const int kb = e / L, c = e % L;
		const uint32_t bv = (kb < nb && j0 + c < nrhs) ? Elem<E>::load(B + (size_t) kb * ldb + j0 + c, p) : 0u;
		const uint32_t lb = i8_limbs(bv, p);
		const int lnb = (c & 31) + 32 * ((kb & 31) >> 4), byb = kb & 15;
		Bf[c >> 5][kb >> 5][0][lnb][byb] = (int8_t) (lb >> 8);
		Bf[c >> 5][kb >> 5][1][lnb][byb] = (int8_t) lb;
	}
	__syncthreads();
	const int rb = wave >> 1, cb = wave & 1;
	v16i acc0 = v16i{0}, acc1 = v16i{0}, acc2 = v16i{0};
#pragma unroll
	for (int kb = 0; kb < 2; ++kb) {
		const v4i ah = *(const v4i *) &Wf[rb][kb][0][lane][0], al = *(const v4i *) &Wf[rb][kb][1][lane][0];
		const v4i bh = *(const v4i *) &Bf[cb][kb][0][lane][0], bl = *(const v4i *) &Bf[cb][kb][1][lane][0];
		acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, acc0, 0, 0, 0);
		acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bl, acc1, 0, 0, 0);
		acc2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bl, acc2, 0, 0, 0);
		acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh, acc1, 0, 0, 0);
	}
	const int col = j0 + cb * 32 + (lane & 31);
	if (col >= nrhs)
		return;
#pragma unroll
	for (int r = 0; r < 16; ++r) {
		const int row = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
		if (row < nb) {
			const double v = fma((double) acc0[r], 65536.0, fma((double) acc1[r], 256.0, (double) acc2[r]));
			const uint32_t d = mp.mod_f64(v);
			Elem<E>::store(B + (size_t) row * ldb + col, mp.mod_f64((double) a * (double) d), p);
		}
	}
}

template <int E>
static int trsm_diag_inverses(mlz_ctx *ctx, uint32_t p, int upper, int unit, size_t n, const void *A,
			      size_t lda, const uint32_t *dinv, uint32_t *W)
{
	typedef typename Elem<E>::T T;
	const ModP mp = make_modp(p);
	unsigned g = ceil_div(n, TRSM_LEAF);
	ProfScope ps(ctx, "trsm_diag_block_inv");
#define INV(U, N, S) trsm_diag_inv_blk_kernel<E, U, N, S><<<g, 256, 0, ctx->stream>>>((const T *) A, lda, (int) n, dinv, W, mp)
	if (mp.small) {
		if (upper) { if (unit) INV(true, true, true); else INV(true, false, true); }
		else       { if (unit) INV(false, true, true); else INV(false, false, true); }
	} else {
		if (upper) { if (unit) INV(true, true, false); else INV(true, false, false); }
		else       { if (unit) INV(false, true, false); else INV(false, false, false); }
	}
#undef INV
	return check_launch(ctx, "trsm_diag_inv_blk_kernel");
}

template <int E>
static int trsm_leaf(mlz_ctx *ctx, uint32_t p, size_t nb, size_t nrhs, uint32_t a, const uint32_t *W,
		     void *B, size_t ldb, int algo)
{
	typedef typename Elem<E>::T T;
	const ModP mp = make_modp(p);
	uint64_t t = 2;
	for (int i = 0; i < 6; ++i)
		t = (t * t) % p;  // 2^64 mod p
	const bool wide = (unsigned __int128) (p - 1) * (p - 1) * TRSM_LEAF >= ((unsigned __int128) 1 << 64);
	dim3 g(ceil_div(nrhs, 64));
	ProfScope ps(ctx, "trsm_leaf");
	if (p <= 65536 && (algo == MLZ_ALGO_AUTO || algo == MLZ_ALGO_INT8))
		trsm_leaf_i8_kernel<E><<<g, 256, 0, ctx->stream>>>(W, (T *) B, ldb, (int) nb, (int) nrhs, a, mp);
	else if (wide)
		trsm_leaf_apply_kernel<E, true><<<g, 256, 0, ctx->stream>>>(W, (T *) B, ldb, (int) nb, (int) nrhs, a, mp, t);
	else
		trsm_leaf_apply_kernel<E, false><<<g, 256, 0, ctx->stream>>>(W, (T *) B, ldb, (int) nb, (int) nrhs, a, mp, t);
	return check_launch(ctx, "trsm_leaf_apply_kernel");
}

// ---------------------------------------------------------------------------
// Blocked solve on INT8 MFMA (p <= 2^16) with one launch per 64-row block:
// block I's rows of X = a * A^-1 * B are W_I (a B_I - A[I, band] X[band]),
// with W_I the inverse of its diagonal block and the band the solved rows it
// depends on (upper: the rows below it up to the block's last nonzero
// column; lower: all rows above it).  Packed operands ([32-index block]
// [k block][limb][64 lanes][16 B], as zp_gemm.hip's INT8 kernels): every
// block's band, the diagonal inverses, and X as it is solved (B operand:
// column blocks, then row blocks).

// 16-byte piece (hi and lo limb vectors) of the 16 residues v[0..15].
__device__ inline void i8_piece(const uint32_t (&v)[16], uint32_t p, int8_t *dst)
{
	uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
	for (int b = 0; b < 16; ++b) {
		const uint32_t l = i8_limbs(v[b], p);
		hw[b >> 2] |= ((l >> 8) & 0xffu) << (8 * (b & 3));
		lw[b >> 2] |= (l & 0xffu) << (8 * (b & 3));
	}
	*(uint4 *) dst = make_uint4(hw[0], hw[1], hw[2], hw[3]);
	*(uint4 *) (dst + 1024) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
}

// Block I's band of a dense triangular A: k blocks, first column, end
// column, and the byte offset of its fragments (blocks packed in order).
__host__ __device__ inline void dense_band(int upper, int n, int nb64, int I, int &kbn, int &kc0, int &kend,
					   size_t &off)
{
	const int i0 = 64 * I, i1 = i0 + 64 < n ? i0 + 64 : n;
	if (upper) {
		kc0 = i1;
		kend = n;
		kbn = 2 * (nb64 - 1 - I);
		off = (size_t) 8192 * ((size_t) I * (nb64 - 1) - (size_t) I * (I - 1) / 2);
	} else {
		kc0 = 0;
		kend = i0;
		kbn = 2 * I;
		off = (size_t) 4096 * I * (I - 1);
	}
}

// Band fragments: workgroup (q, I) packs k blocks 2q, 2q+1 of block I's band
// (columns kc0 .., zero from kend on; from meta/off, or dense_band when meta
// is null); thread = piece (row block, k block, lane).
template <int E>
__global__ __launch_bounds__(256) void trsm_band_pack_kernel(const typename Elem<E>::T *__restrict__ A, size_t lda,
							     int n, const int *__restrict__ meta, const size_t *__restrict__ moff,
							     int upper, int8_t *__restrict__ band, uint32_t p)
{
	const int I = blockIdx.y, kb = 2 * blockIdx.x + ((threadIdx.x >> 6) & 1);
	int n_kb, kc0, kend;
	size_t off;
	if (meta) {
		n_kb = meta[3 * I];
		kc0 = meta[3 * I + 1];
		kend = meta[3 * I + 2];
		off = moff[I];
	} else {
		dense_band(upper, n, (int) gridDim.y, I, n_kb, kc0, kend, off);
	}
	if (kb >= n_kb)
		return;
	const int lane = threadIdx.x & 63, rb = threadIdx.x >> 7;
	const int i = 64 * I + 32 * rb + (lane & 31);
	const int k0 = kc0 + 32 * kb + 16 * (lane >> 5);
	uint32_t v[16];
#pragma unroll
	for (int b = 0; b < 16; ++b)
		v[b] = (i < n && k0 + b < kend) ? Elem<E>::load(A + (size_t) i * lda + k0 + b, p) : 0u;
	i8_piece(v, p, band + off + ((size_t) rb * n_kb + kb) * 2048 + lane * 16);
}

// Diagonal-block inverse fragments (W from zp_trsm_prepare: block I's inverse
// in rows 64I .. of a 64-column array, valid for its first nb rows/columns).
__global__ __launch_bounds__(256) void trsm_winv_pack_kernel(const uint32_t *__restrict__ W, int n,
							     int8_t *__restrict__ wp, uint32_t p)
{
	const int I = blockIdx.x, nb = min(64, n - 64 * I);
	const int lane = threadIdx.x & 63, kb = (threadIdx.x >> 6) & 1, rb = threadIdx.x >> 7;
	const int t = 32 * rb + (lane & 31), s0 = 32 * kb + 16 * (lane >> 5);
	uint32_t v[16];
#pragma unroll
	for (int b = 0; b < 16; ++b)
		v[b] = (t < nb && s0 + b < nb) ? W[(size_t) (64 * I + t) * 64 + s0 + b] : 0u;
	i8_piece(v, p, wp + (size_t) I * 8192 + (rb * 2 + kb) * 2048 + lane * 16);
}

// Block rows [i0, i0 + nbr) of X; workgroup j owns columns [64j, 64j + 64).
// Waves 0-3 (rb, cb) own one 32 x 32 block each; the band product is split
// over k, waves 4-7 taking the upper half of the k blocks and handing their
// sums to waves 0-3 through LDS (the product is a chain of dependent loads and
// MFMAs, and the launch has only nrhs/64 workgroups).  T = a B_I - band *
// X[band rows] (kbn k blocks of packed X from row block kb0), then X_I = W_I T,
// written to B in place and, packed, to Xp (rkb row blocks per column block).
template <int E>
__global__ __launch_bounds__(512) void trsm_block_solve_kernel(typename Elem<E>::T *__restrict__ B, size_t ldb,
							       int nrhs, int i0, int nbr, int kbn, int kb0,
							       const int8_t *__restrict__ band,
							       const int8_t *__restrict__ winv, int8_t *__restrict__ Xp,
							       int rkb, uint32_t a, ModP mp)
{
	__shared__ uint32_t Ts[64][65];
	__shared__ int Pk[4][3][16][64];   // upper k half's sums
	const uint32_t p = (uint32_t) mp.p;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, half = wave >> 2, rb = (wave & 3) >> 1,
		  cb = wave & 1;
	const int cbg = 2 * blockIdx.x + cb, col = 32 * cbg + (lane & 31);
	// B_I's entries are loaded first: their latency overlaps the band product
	uint32_t bv[16];
	if (half) {
#pragma unroll
		for (int r = 0; r < 16; ++r)
			bv[r] = 0u;
	} else if (nbr == 64 && 64 * (int) blockIdx.x + 64 <= nrhs) {
#pragma unroll
		for (int r = 0; r < 16; ++r)
			bv[r] = Elem<E>::load(B + (size_t) (i0 + 32 * rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * ldb + col, p);
	} else {
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			const int row = 32 * rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
			bv[r] = (row < nbr && col < nrhs) ? Elem<E>::load(B + (size_t) (i0 + row) * ldb + col, p) : 0u;
		}
	}
	v16i a0 = v16i{0}, a1 = v16i{0}, a2 = v16i{0};
	{
		const int kh = kbn / 2, kend = half ? kbn : kh;
		const int8_t *ga = band + (size_t) rb * kbn * 2048 + lane * 16;
		const int8_t *gx = Xp + ((size_t) cbg * rkb + kb0) * 2048 + lane * 16;
		// four k blocks' fragments are loaded before their MFMAs so that the
		// loads' latency is paid once per group
		int kb = half ? kh : 0;
		for (; kb + 4 <= kend; kb += 4) {
			v4i ah[4], al[4], bh[4], bl[4];
#pragma unroll
			for (int q = 0; q < 4; ++q) {
				const size_t o = (size_t) (kb + q) * 2048;
				ah[q] = *(const v4i *) (ga + o);
				al[q] = *(const v4i *) (ga + o + 1024);
				bh[q] = *(const v4i *) (gx + o);
				bl[q] = *(const v4i *) (gx + o + 1024);
			}
#pragma unroll
			for (int q = 0; q < 4; ++q) {
				a0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[q], bh[q], a0, 0, 0, 0);
				a1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah[q], bl[q], a1, 0, 0, 0);
				a2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al[q], bl[q], a2, 0, 0, 0);
				a1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al[q], bh[q], a1, 0, 0, 0);
			}
		}
		for (; kb < kend; ++kb) {
			const size_t o = (size_t) kb * 2048;
			const v4i ah = *(const v4i *) (ga + o), al = *(const v4i *) (ga + o + 1024);
			const v4i bh = *(const v4i *) (gx + o), bl = *(const v4i *) (gx + o + 1024);
			a0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, a0, 0, 0, 0);
			a1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bl, a1, 0, 0, 0);
			a2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bl, a2, 0, 0, 0);
			a1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh, a1, 0, 0, 0);
		}
	}
	if (half) {
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			Pk[wave & 3][0][r][lane] = a0[r];
			Pk[wave & 3][1][r][lane] = a1[r];
			Pk[wave & 3][2][r][lane] = a2[r];
		}
	}
	__syncthreads();
	// T = a B_I - band product; rows >= nbr and columns >= nrhs are zero
	if (!half) {
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			const int row = 32 * rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
			uint32_t tv = 0;
			if (row < nbr && col < nrhs) {
				const double h = (double) (a0[r] + Pk[wave][0][r][lane]);
				const double m = (double) (a1[r] + Pk[wave][1][r][lane]);
				const double l = (double) (a2[r] + Pk[wave][2][r][lane]);
				const uint32_t d = mp.mod_f64(fma(h, 65536.0, fma(m, 256.0, l)));
				const uint32_t ab = a == 1 ? bv[r] : mp.mod_f64((double) a * (double) bv[r]);
				tv = ab >= d ? ab - d : ab + p - d;
			}
			Ts[row][32 * cb + (lane & 31)] = tv;
		}
	}
	__syncthreads();
	// X_I = W_I T (waves 0-3): B fragments of T built from LDS, W_I's from winv
	a0 = v16i{0};
	a1 = v16i{0};
	a2 = v16i{0};
#pragma unroll
	for (int kb = 0; kb < 2 && !half; ++kb) {
		uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
		const int s0 = 32 * kb + 16 * (lane >> 5), cc = 32 * cb + (lane & 31);
#pragma unroll
		for (int b = 0; b < 16; ++b) {
			const uint32_t l = i8_limbs(Ts[s0 + b][cc], p);
			hw[b >> 2] |= ((l >> 8) & 0xffu) << (8 * (b & 3));
			lw[b >> 2] |= (l & 0xffu) << (8 * (b & 3));
		}
		const v4i bh = v4i{(int) hw[0], (int) hw[1], (int) hw[2], (int) hw[3]};
		const v4i bl = v4i{(int) lw[0], (int) lw[1], (int) lw[2], (int) lw[3]};
		const int8_t *gw = winv + (rb * 2 + kb) * 2048 + lane * 16;
		const v4i ah = *(const v4i *) gw, al = *(const v4i *) (gw + 1024);
		a0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, a0, 0, 0, 0);
		a1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bl, a1, 0, 0, 0);
		a2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bl, a2, 0, 0, 0);
		a1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh, a1, 0, 0, 0);
	}
	__syncthreads();
	if (!half) {
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			const int row = 32 * rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
			const uint32_t x = mp.mod_f64(fma((double) a0[r], 65536.0, fma((double) a1[r], 256.0, (double) a2[r])));
			if (row < nbr && col < nrhs)
				Elem<E>::store(B + (size_t) (i0 + row) * ldb + col, x, p);
			Ts[row][32 * cb + (lane & 31)] = x;
		}
	}
	__syncthreads();
	// X_I packed: piece (column block, row block, lane) of this workgroup
	if (tid < 256) {
		const int cb2 = tid >> 7, kb2 = (tid >> 6) & 1;
		const int cc = 32 * cb2 + (lane & 31), s0 = 32 * kb2 + 16 * (lane >> 5);
		uint32_t v[16];
#pragma unroll
		for (int b = 0; b < 16; ++b)
			v[b] = Ts[s0 + b][cc];
		i8_piece(v, p, Xp + ((size_t) (2 * blockIdx.x + cb2) * rkb + i0 / 32 + kb2) * 2048 + lane * 16);
	}
}

// X = a A^-1 B by 64-row blocks on INT8 MFMA (p <= 2^16, n <= 65504), W from
// zp_trsm_prepare for this A.  upper: blocks from the bottom, block I's band
// from its diagonal block's end to bend[I] (default n); lower: from the top,
// band = all rows above the block.
template <int E>
static int trsm_blocked_i8_t(mlz_ctx *ctx, uint32_t p, int upper, size_t n, size_t nrhs, uint32_t a,
			     const typename Elem<E>::T *A, size_t lda, typename Elem<E>::T *B, size_t ldb,
			     const uint32_t *W, const int *bend)
{
	if (n == 0 || nrhs == 0)
		return MLZ_OK;
	const size_t nb64 = (n + 63) / 64, rpad = nb64 * 64, npad = round_up(nrhs, 64);
	std::vector<int> meta(3 * nb64);
	std::vector<size_t> off(nb64 + 1, 0);
	int kbmax = 0;
	for (size_t I = 0; I < nb64; ++I) {
		int kbn, kc0, kend;
		size_t o;
		dense_band(upper, (int) n, (int) nb64, (int) I, kbn, kc0, kend, o);
		if (bend) {   // banded: the band ends at bend[I]
			kend = std::max(bend[I], kc0);
			kbn = 2 * ((kend - kc0 + 63) / 64);
			o = off[I];
		}
		meta[3 * I] = kbn;
		meta[3 * I + 1] = kc0;
		meta[3 * I + 2] = kend;
		off[I] = o;
		off[I + 1] = o + (size_t) 2 * kbn * 2048;
		kbmax = std::max(kbmax, kbn);
	}
	void *wx, *wbp;
	MLZ_TRY(ws_get(ctx, WS_TRSM_1, rpad * npad * 2, &wx));
	const size_t mbytes = round_up(3 * nb64 * sizeof(int), 256) + round_up((nb64 + 1) * sizeof(size_t), 256);
	MLZ_TRY(ws_get(ctx, WS_TRSM_2, mbytes + nb64 * 8192 + off[nb64] + 256, &wbp));
	int *d_meta = (int *) wbp;
	size_t *d_off = (size_t *) ((char *) wbp + round_up(3 * nb64 * sizeof(int), 256));
	int8_t *winvp = (int8_t *) wbp + mbytes, *band = winvp + nb64 * 8192;
	if (bend) {
		// host tables (one call per FL): uploaded, and waited for below
		MLZ_HIP(ctx, hipMemcpyAsync(d_meta, meta.data(), 3 * nb64 * sizeof(int), hipMemcpyHostToDevice,
					    ctx->stream));
		MLZ_HIP(ctx, hipMemcpyAsync(d_off, off.data(), (nb64 + 1) * sizeof(size_t), hipMemcpyHostToDevice,
					    ctx->stream));
	}
	const ModP mp = make_modp(p);
	if (kbmax)
		trsm_band_pack_kernel<E><<<dim3(kbmax / 2, nb64), 256, 0, ctx->stream>>>(
			A, lda, (int) n, bend ? d_meta : nullptr, bend ? d_off : nullptr, upper, band, p);
	trsm_winv_pack_kernel<<<nb64, 256, 0, ctx->stream>>>(W, (int) n, winvp, p);
	MLZ_TRY(check_launch(ctx, "trsm_winv_pack_kernel"));
	for (size_t s = 0; s < nb64; ++s) {
		const size_t I = upper ? nb64 - 1 - s : s;
		const size_t i0 = 64 * I;
		trsm_block_solve_kernel<E><<<npad / 64, 512, 0, ctx->stream>>>(
			B, ldb, (int) nrhs, (int) i0, (int) std::min((size_t) 64, n - i0), meta[3 * I], meta[3 * I + 1] / 32,
			band + off[I], winvp + I * 8192, (int8_t *) wx, (int) (rpad / 32), a, mp);
	}
	MLZ_TRY(check_launch(ctx, "trsm_block_solve_kernel"));
	if (bend)
		MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));   // meta and off leave scope
	return MLZ_OK;
}

int zp_trsm_blocked_i8(mlz_ctx *ctx, int elem, uint64_t p, int upper, size_t n, size_t nrhs, uint64_t a,
		       const void *A, size_t lda, void *B, size_t ldb, const uint32_t *W, const int *bend)
{
	const uint32_t pp = (uint32_t) p, aa = (uint32_t) (a % p);
	if (elem == MLZ_U16)
		return trsm_blocked_i8_t<MLZ_U16>(ctx, pp, upper, n, nrhs, aa, (const uint16_t *) A, lda, (uint16_t *) B, ldb,
						  W, bend);
	if (elem == MLZ_U32)
		return trsm_blocked_i8_t<MLZ_U32>(ctx, pp, upper, n, nrhs, aa, (const uint32_t *) A, lda, (uint32_t *) B, ldb,
						  W, bend);
	return trsm_blocked_i8_t<MLZ_F64_SYM>(ctx, pp, upper, n, nrhs, aa, (const double *) A, lda, (double *) B, ldb, W,
					      bend);
}

constexpr size_t TRSM_BLOCK_I8 = 1024;   // 512: same speed at 8192 x 1024; 2048: 20% slower

// ---------------------------------------------------------------------------
// Inverses of the full TRSM_INV_BLOCK-row diagonal blocks (p <= 2^16).  With
// them the solve is a chain of n / TRSM_INV_BLOCK gemm pairs instead of
// n / 64 dependent block solves.  They do not depend on B, so all blocks are
// inverted together, by doubling from the 64-row inverses of
// zp_trsm_prepare: a lower block [L11 0; L21 L22] has the inverse
// [L11^-1 0; -L22^-1 L21 L11^-1 L22^-1] (upper: the off-diagonal block is
// -U11^-1 U12 U22^-1), so each doubling step is two batched gemms over every
// pair of diagonal blocks of the current size.  Storage: block I's inverse
// is the dense S x S array Z + I*S*S (S = TRSM_INV_BLOCK), canonical uint32.
// ---------------------------------------------------------------------------
constexpr size_t TRSM_INV_BLOCK = 1024;
constexpr size_t TRSM_ZP_BYTES = TRSM_INV_BLOCK * TRSM_INV_BLOCK * 2;   // one inverse as INT8 limbs

// Z <- the 64-row diagonal inverses of W in place, zero elsewhere (one
// thread per entry of the nfull S x S arrays).  The 64-blocks on the
// triangle's side of the diagonal are left alone: the doubling steps write
// each of them before anything reads it.
template <bool UPPER>
__global__ __launch_bounds__(256) void trsm_inv_init_kernel(const uint32_t *__restrict__ W, size_t total,
							    uint32_t *__restrict__ Z)
{
	constexpr size_t S = TRSM_INV_BLOCK;
	const size_t idx = (size_t) blockIdx.x * 256 + threadIdx.x;
	if (idx >= total)
		return;
	const size_t I = idx / (S * S), rem = idx % (S * S);
	const int r = (int) (rem / S), c = (int) (rem % S);
	if (UPPER ? (c >> 6) > (r >> 6) : (c >> 6) < (r >> 6))
		return;
	const bool diag = (r >> 6) == (c >> 6) && (UPPER ? c >= r : c <= r);
	Z[idx] = diag ? W[(I * S + r) * TRSM_LEAF + (c & (TRSM_LEAF - 1))] : 0u;
}

template <int E>
static int trsm_block_inverses(mlz_ctx *ctx, uint32_t p, int upper, size_t nfull, const typename Elem<E>::T *A,
			       size_t lda, const uint32_t *W, uint32_t *Z, uint32_t *Tb)
{
	constexpr size_t S = TRSM_INV_BLOCK;
	const size_t total = nfull * S * S;
	const ModP mp = make_modp(p);
	{
		ProfScope ps(ctx, "trsm_inv_init");
		if (upper)
			trsm_inv_init_kernel<true><<<ceil_div(total, 256), 256, 0, ctx->stream>>>(W, total, Z);
		else
			trsm_inv_init_kernel<false><<<ceil_div(total, 256), 256, 0, ctx->stream>>>(W, total, Z);
	}
	MLZ_TRY(check_launch(ctx, "trsm_inv_init_kernel"));
	// A's rows as runs of four entries: every block offset below is a multiple
	// of 64 entries, so aligned rows of A suffice (uint16 / uint32 storage)
	typedef typename Elem<E>::T TE;
	const bool av = sizeof(TE) <= 4 && lda % 4 == 0 && (uintptr_t) A % (4 * sizeof(TE)) == 0;
	for (size_t s = TRSM_LEAF; 2 * s <= S; s *= 2) {
		// pair j of block I: rows and columns o .. o + 2s of the block, o = 2sj.
		// Entry (o + r, o + c) of block I is at I*S*S + o*(S + 1) + r*S + c in
		// Z and at (I*S + o + r)*lda + I*S + o + c in A.
		const int inner = (int) (S / (2 * s));
		I8Batch bt;
		bt.count = (int) nfull * inner;
		bt.inner = inner;
		const long long zo = (long long) (S * S), zi = (long long) (2 * s * (S + 1));
		const long long ao = (long long) (S * (lda + 1)), ai = (long long) (2 * s * (lda + 1));
		const long long to = (long long) (inner * s * s), ti = (long long) (s * s);
		// T = off-diagonal block of A (U12 / L21) times one diagonal inverse (U22^-1 / L11^-1)
		bt.outer[0] = ao; bt.in[0] = ai;
		bt.outer[1] = zo; bt.in[1] = zi;
		bt.outer[2] = to; bt.in[2] = ti;
		const typename Elem<E>::T *offA = upper ? A + s : A + s * lda;
		const uint32_t *inv_r = upper ? Z + s * S + s : Z;
		const dim3 grid((unsigned) ((s / 64) * (s / 64)), (unsigned) bt.count);
		if (av)
			trsm_dbl_kernel<E, true><<<grid, 256, 0, ctx->stream>>>(offA, lda, inv_r, S, Tb, s, (int) s, 1u, bt, mp);
		else
			trsm_dbl_kernel<E, false><<<grid, 256, 0, ctx->stream>>>(offA, lda, inv_r, S, Tb, s, (int) s, 1u, bt, mp);
		MLZ_TRY(check_launch(ctx, "trsm_dbl_kernel"));
		// off-diagonal block of the inverse = -(other diagonal inverse) * T
		bt.outer[0] = zo; bt.in[0] = zi;
		bt.outer[1] = to; bt.in[1] = ti;
		bt.outer[2] = zo; bt.in[2] = zi;
		const uint32_t *inv_l = upper ? Z : Z + s * S + s;             // U11^-1 / L22^-1
		uint32_t *out = upper ? Z + s : Z + s * S;
		trsm_dbl_kernel<MLZ_U32, true><<<grid, 256, 0, ctx->stream>>>(inv_l, S, Tb, s, out, S, (int) s, p - 1, bt, mp);
		MLZ_TRY(check_launch(ctx, "trsm_dbl_kernel"));
	}
	return MLZ_OK;
}

// The off-diagonal blocks of A that trsm_rec's updates read when it runs with
// the TRSM_INV_BLOCK inverses (Zp), in the order it reads them: rows r0 ..,
// columns c0 .. of A, and the offset of the block's packed copy.
struct TrsmUpd {
	size_t r0, c0, rows, cols, off;
};

static void trsm_updates(size_t n, size_t base, int upper, std::vector<TrsmUpd> &out)
{
	// the cases in which trsm_rec does not split (p <= 2^16 here)
	if (n == TRSM_INV_BLOCK || n <= TRSM_BLOCK_I8)
		return;
	const size_t l = (n / 2 + TRSM_INV_BLOCK - 1) / TRSM_INV_BLOCK * TRSM_INV_BLOCK;
	if (!upper) {
		trsm_updates(l, base, upper, out);
		out.push_back({base + l, base, n - l, l, 0});
		trsm_updates(n - l, base + l, upper, out);
	} else {
		trsm_updates(n - l, base + l, upper, out);
		out.push_back({base, base + l, l, n - l, 0});
		trsm_updates(l, base, upper, out);
	}
}

// The right-hand sides kept as packed INT8 B operands between the products
// of the recursion (every block a full TRSM_INV_BLOCK leaf): PB holds B's
// current rows, PX the solved rows X; both in the packed B layout over all n
// rows (32-column blocks `panel` bytes apart, 2048 bytes per 32 rows).  A
// leaf reads B_I from PB and writes X_I also into PX; an update reads X from
// PX and writes its rows also into PB.  Each product thus finds its B operand
// packed by the one before it, with no packing launch between them.
struct TrsmPk {
	int8_t *PB, *PX;
	size_t panel;
	I8PackIO io(const int8_t *in, size_t in_row, int8_t *out, size_t out_row) const
	{
		I8PackIO r;
		r.Bpre = in + in_row / 32 * 2048;
		r.bpanel = panel;
		r.Cpk = out + out_row / 32 * 2048;
		r.cpanel = panel;
		return r;
	}
};

// Upk: the packed update blocks in trsm_updates' order (with Zp), *ui the
// next one to use; pk (with Upk): the packed right-hand sides, row0 the
// first row of B in them.  (A right-looking form with look-ahead -- per block
// the leaf and the next block's update on one stream, the update of all later
// blocks on a second -- measured slower for 8192 x 1024, 0.49-0.50 against
// 0.39-0.40 ms: its updates have k = 1024 where the recursion's reach 4096,
// and each step still waits for the second stream's larger update.)
template <int E>
static int trsm_rec(mlz_ctx *ctx, uint32_t p, int upper, int unit, size_t n, size_t nrhs, uint32_t a,
		    const typename Elem<E>::T *A, size_t lda, typename Elem<E>::T *B, size_t ldb,
		    const uint32_t *W, int algo, const int8_t *Zp = nullptr, const int8_t *const *Upk = nullptr,
		    size_t *ui = nullptr, const TrsmPk *pk = nullptr, size_t row0 = 0)
{
	if (n == 0)
		return MLZ_OK;
	if (Zp && n == TRSM_INV_BLOCK) {
		// a full diagonal block with its inverse (packed): X = a Z_I B_I.  The
		// product is written over B_I, which the gemm reads only through its
		// packed copy.
		return zp_gemm_i8_batched_dev(ctx, E, MLZ_U32, p, I8Batch::single(), n, nrhs, n, a, nullptr, n, B, ldb, 0, B,
					      ldb, Zp, pk ? pk->io(pk->PB, row0, pk->PX, row0) : I8PackIO());
	}
	if (n <= (size_t) TRSM_LEAF)
		return trsm_leaf<E>(ctx, p, n, nrhs, a, W, B, ldb, algo);
	// diagonal blocks up to TRSM_BLOCK_I8 rows: one launch per 64-row block
	// instead of the recursion's small gemms and leaves
	if (n <= TRSM_BLOCK_I8 && p <= 65536 && (algo == MLZ_ALGO_AUTO || algo == MLZ_ALGO_INT8))
		return trsm_blocked_i8_t<E>(ctx, p, upper, n, nrhs, a, A, lda, B, ldb, W, nullptr);
	// Split at a multiple of the leaf size near n/2, so that every leaf is
	// one of the precomputed diagonal blocks (any split gives the same unique
	// solution; LELA splits at n/2).
	// With block inverses the split is at a multiple of their size, so
	// that every full block is a leaf; a last partial block (at the end of
	// the matrix) runs the 64-row chain.
	const size_t unit_split = Zp ? TRSM_INV_BLOCK : (size_t) TRSM_LEAF;
	size_t l = (n / 2 + unit_split - 1) / unit_split * unit_split;
	const typename Elem<E>::T *A11 = A, *A22 = A + l * lda + l;
	typename Elem<E>::T *B1 = B, *B2 = B + l * ldb;
	const int8_t *Zp2 = Zp ? Zp + (l / TRSM_INV_BLOCK) * TRSM_ZP_BYTES : nullptr;
	// the update C <- -Aoff X + a C, its A block packed ahead when Upk is
	// given; X and C start at rows xr and cr of the packed right-hand sides
	auto update = [&](size_t rows, size_t cols, const typename Elem<E>::T *Aoff, const typename Elem<E>::T *X,
			  typename Elem<E>::T *Cu, size_t xr, size_t cr) -> int {
		if (Upk)
			return zp_gemm_i8_batched_dev(ctx, E, E, p, I8Batch::single(), rows, nrhs, cols, p - 1, nullptr, lda,
						      X, ldb, a, Cu, ldb, Upk[(*ui)++],
						      pk ? pk->io(pk->PX, xr, pk->PB, cr) : I8PackIO());
		return zp_gemm_dev(ctx, E, p, rows, nrhs, cols, p - 1, Aoff, lda, X, ldb, a, Cu, ldb, algo);
	};
	if (!upper) {
		MLZ_TRY(trsm_rec<E>(ctx, p, upper, unit, l, nrhs, a, A11, lda, B1, ldb, W, algo, Zp, Upk, ui, pk, row0));
		MLZ_TRY(update(n - l, l, A + l * lda, B1, B2, row0, row0 + l));
		MLZ_TRY(trsm_rec<E>(ctx, p, upper, unit, n - l, nrhs, 1, A22, lda, B2, ldb, W + l * TRSM_LEAF, algo, Zp2,
				    Upk, ui, pk, row0 + l));
	} else {
		MLZ_TRY(trsm_rec<E>(ctx, p, upper, unit, n - l, nrhs, a, A22, lda, B2, ldb, W + l * TRSM_LEAF, algo, Zp2,
				    Upk, ui, pk, row0 + l));
		MLZ_TRY(update(l, n - l, A + l, B2, B1, row0 + l, row0));
		MLZ_TRY(trsm_rec<E>(ctx, p, upper, unit, l, nrhs, 1, A11, lda, B1, ldb, W, algo, Zp, Upk, ui, pk, row0));
	}
	return MLZ_OK;
}

// Diagonal checks and the inverses of all TRSM_LEAF-row diagonal blocks of A
// (n x n): on MLZ_OK *W points at workspace holding block b's inverse in rows
// b*TRSM_LEAF .., valid until the next call that uses WS_TRSM_0.
int zp_trsm_prepare(mlz_ctx *ctx, int elem, uint64_t p, int upper, int unit, size_t n, const void *A, size_t lda,
		    size_t *bad_diag, const uint32_t **W_out)
{
	const size_t nblk = (n + TRSM_LEAF - 1) / TRSM_LEAF;
	void *w = nullptr;
	const size_t dinv_bytes = round_up(n * sizeof(uint32_t), 256);
	const size_t w_bytes = nblk * TRSM_LEAF * TRSM_LEAF * sizeof(uint32_t);
	MLZ_TRY(ws_get(ctx, WS_TRSM_0, dinv_bytes + w_bytes + 256, &w));
	uint32_t *dinv = (uint32_t *) w;
	uint32_t *W = (uint32_t *) ((char *) w + dinv_bytes);
	unsigned long long *bad = (unsigned long long *) ((char *) W + w_bytes);
	if (!unit) {
		unsigned long long host[2];
		init_bad_kernel<<<1, 1, 0, ctx->stream>>>(bad);
		{
			ProfScope ps(ctx, "trsm_diag_inv");
			unsigned g = ceil_div(n, 256);
			if (elem == MLZ_U16)
				diag_inv_kernel<MLZ_U16><<<g, 256, 0, ctx->stream>>>((const uint16_t *) A, lda, (int) n, (uint32_t) p, dinv, bad);
			else if (elem == MLZ_U32)
				diag_inv_kernel<MLZ_U32><<<g, 256, 0, ctx->stream>>>((const uint32_t *) A, lda, (int) n, (uint32_t) p, dinv, bad);
			else
				diag_inv_kernel<MLZ_F64_SYM><<<g, 256, 0, ctx->stream>>>((const double *) A, lda, (int) n, (uint32_t) p, dinv, bad);
		}
		MLZ_TRY(check_launch(ctx, "diag_inv_kernel"));
		MLZ_HIP(ctx, hipMemcpyAsync(host, bad, sizeof host, hipMemcpyDeviceToHost, ctx->stream));
		MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
		if (host[0] != ~0ull) {
			size_t idx = (size_t) (upper ? host[1] : host[0]);
			if (bad_diag)
				*bad_diag = idx;
			return set_error(ctx, MLZ_E_NOT_INVERTIBLE, "diagonal entry %zu not invertible", idx);
		}
	}
	const uint32_t pp = (uint32_t) p;
	if (elem == MLZ_U16)
		MLZ_TRY(trsm_diag_inverses<MLZ_U16>(ctx, pp, upper, unit, n, A, lda, dinv, W));
	else if (elem == MLZ_U32)
		MLZ_TRY(trsm_diag_inverses<MLZ_U32>(ctx, pp, upper, unit, n, A, lda, dinv, W));
	else
		MLZ_TRY(trsm_diag_inverses<MLZ_F64_SYM>(ctx, pp, upper, unit, n, A, lda, dinv, W));
	*W_out = W;
	return MLZ_OK;
}

// Upper triangular banded A (canonical uint32, p <= 2^16): X = A^-1 B with
// the full TRSM_INV_BLOCK-row diagonal blocks inverted up front.  (Keeping
// B packed between the products here, as trsm_rec does, measured no faster
// in the 20000 x 30000 reduction, 21.6-21.7 against 21.55 ms: one packing of
// all of B up front costs about what the per-product packings did.)  The blocks
// are solved from the bottom: the partial block at the end (if any) by the
// 64-row chain, then block I as B_I <- B_I - A[I, i1 .. c1) X[i1 .. c1) (one
// gemm over its band, c1[I] one past its rows' last nonzero column) and
// X_I = Z_I B_I.  W: the 64-row inverses of zp_trsm_prepare for A;
// c1_64: band ends of the 64-row blocks (as zp_trsm_blocked_i8 takes them).
int zp_trsm_banded_inv_i8(mlz_ctx *ctx, uint64_t p, size_t n, size_t nrhs, const uint32_t *A, size_t lda,
			  uint32_t *B, size_t ldb, const uint32_t *W, const int *c1_64)
{
	constexpr size_t S = TRSM_INV_BLOCK;
	const size_t nfull = n / S, ntail = n - nfull * S;
	const uint32_t pp = (uint32_t) p;
	if (p > 65536 || nfull == 0)
		return set_error(ctx, MLZ_E_ARG, "banded inverse solve needs p <= 2^16 and n >= %zu", S);
	const size_t zbytes = round_up(nfull * S * S * sizeof(uint32_t), 256);
	const size_t tbytes = round_up(nfull * S * (S / 4) * sizeof(uint32_t), 256);
	void *w = nullptr;
	MLZ_TRY(ws_get(ctx, WS_TRSM_3, zbytes + tbytes + nfull * TRSM_ZP_BYTES, &w));
	uint32_t *Z = (uint32_t *) w;
	uint32_t *Tb = (uint32_t *) ((char *) w + zbytes);
	int8_t *Zp = (int8_t *) w + zbytes + tbytes;
	MLZ_TRY(trsm_block_inverses<MLZ_U32>(ctx, pp, 1, nfull, A, lda, W, Z, Tb));
	I8Batch bz = I8Batch::single();
	bz.count = (int) nfull;
	bz.outer[0] = (long long) (S * S);
	MLZ_TRY(zp_pack_i8_a_dev(ctx, MLZ_U32, pp, bz, S, S, Z, S, Zp, nullptr));
	if (ntail) {
		// the last rows: their band lies inside the tail block
		const size_t t0 = nfull * S, nb64 = (n + 63) / 64;
		std::vector<int> ct(nb64 - t0 / 64);
		for (size_t I = t0 / 64; I < nb64; ++I)
			ct[I - t0 / 64] = std::min((int) n, c1_64[I]) - (int) t0;
		MLZ_TRY(zp_trsm_blocked_i8(ctx, MLZ_U32, p, 1, ntail, nrhs, 1, A + t0 * lda + t0, lda, B + t0 * ldb, ldb,
					   W + t0 * TRSM_LEAF, ct.data()));
	}
	for (size_t I = nfull; I-- > 0;) {
		const size_t i0 = I * S, i1 = i0 + S;
		int c1 = (int) i1;
		for (size_t b = i0 / 64; b < i1 / 64; ++b)
			c1 = std::max(c1, c1_64[b]);
		c1 = std::min(c1, (int) n);
		if ((size_t) c1 > i1)
			MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, S, nrhs, c1 - i1, p - 1, A + i0 * lda + i1, lda, B + i1 * ldb, ldb,
					    1, B + i0 * ldb, ldb, MLZ_ALGO_INT8));
		MLZ_TRY(zp_gemm_i8_batched_dev(ctx, MLZ_U32, MLZ_U32, pp, I8Batch::single(), S, nrhs, S, 1, nullptr, S,
					       B + i0 * ldb, ldb, 0, B + i0 * ldb, ldb, Zp + I * TRSM_ZP_BYTES));
	}
	return MLZ_OK;
}

// The solve: for p <= 2^16 (INT8 gemms) and at least two full blocks of
// TRSM_INV_BLOCK rows, their inverses first, then the recursion with them as
// leaves; otherwise the recursion down to the 64-row chain.
template <int E>
static int trsm_apply_t(mlz_ctx *ctx, uint32_t p, int upper, int unit, size_t n, size_t nrhs, uint32_t a,
			const typename Elem<E>::T *A, size_t lda, typename Elem<E>::T *B, size_t ldb, const uint32_t *W,
			int algo)
{
	constexpr size_t S = TRSM_INV_BLOCK;
	const size_t nfull = n / S;
	if (p > 65536 || (algo != MLZ_ALGO_AUTO && algo != MLZ_ALGO_INT8) || nfull < 2)
		return trsm_rec<E>(ctx, p, upper, unit, n, nrhs, a, A, lda, B, ldb, W, algo);
	const size_t zbytes = round_up(nfull * S * S * sizeof(uint32_t), 256);
	const size_t tbytes = round_up(nfull * S * (S / 4) * sizeof(uint32_t), 256);   // largest doubling step
	// the updates' A blocks, packed on the second stream while the inverses
	// are computed (neither depends on B)
	std::vector<TrsmUpd> upd;
	trsm_updates(n, 0, upper, upd);
	size_t ubytes = 0;
	for (TrsmUpd &u : upd) {
		size_t item = 0;
		MLZ_TRY(zp_pack_i8_a_dev(ctx, E, p, I8Batch::single(), u.rows, u.cols, nullptr, lda, nullptr, &item));
		u.off = ubytes;
		ubytes += round_up(item, 256);
	}
	// the right-hand sides stay packed between the products when every block
	// is a full leaf and the columns fill whole 128-column tiles
	const bool packed_rhs = n % S == 0 && nrhs % 128 == 0 && !upd.empty();
	const size_t pkbytes = packed_rhs ? round_up(n * nrhs * 2, 256) : 0;
	void *w = nullptr;
	MLZ_TRY(ws_get(ctx, WS_TRSM_3, zbytes + tbytes + nfull * TRSM_ZP_BYTES + ubytes + 2 * pkbytes, &w));
	uint32_t *Z = (uint32_t *) w;
	uint32_t *Tb = (uint32_t *) ((char *) w + zbytes);
	int8_t *Zp = (int8_t *) w + zbytes + tbytes;
	int8_t *Up = Zp + nfull * TRSM_ZP_BYTES;
	TrsmPk pk{Up + ubytes, Up + ubytes + pkbytes, n * 64};
	std::vector<const int8_t *> upk(upd.size());
	struct Events {
		hipEvent_t e[2] = {nullptr, nullptr};
		~Events()
		{
			for (hipEvent_t x : e)
				if (x)
					(void) hipEventDestroy(x);
		}
	} ev;
	if (!upd.empty()) {
		for (int q = 0; q < 2; ++q)
			MLZ_HIP(ctx, hipEventCreateWithFlags(&ev.e[q], hipEventDisableTiming));
		// A may have been written by work queued before this call
		MLZ_HIP(ctx, hipEventRecord(ev.e[0], ctx->stream));
	}
	// The inverses are queued first, so that the context's stream has work
	// while the packing launches are issued.  (Packing from a point before the
	// 64-row diagonal inverses, beside them, measured no faster: 0.418-0.427
	// against 0.413-0.418 ms for 8192 x 1024; the packs slow the inverse
	// kernels as much as they gain.)
	MLZ_TRY(trsm_block_inverses<E>(ctx, p, upper, nfull, A, lda, W, Z, Tb));
	if (!upd.empty()) {
		hipStream_t s1;
		MLZ_TRY(aux_stream(ctx, &s1));
		MLZ_HIP(ctx, hipStreamWaitEvent(s1, ev.e[0], 0));
		for (size_t q = 0; q < upd.size(); ++q) {
			const TrsmUpd &u = upd[q];
			upk[q] = Up + u.off;
			MLZ_TRY(zp_pack_i8_a_dev(ctx, E, p, I8Batch::single(), u.rows, u.cols, A + u.r0 * lda + u.c0, lda,
						 Up + u.off, nullptr, s1));
		}
		if (packed_rhs)     // all of B as one packed B operand (k = n)
			MLZ_TRY(zp_pack_i8_b_dev(ctx, E, p, n, nrhs, B, ldb, pk.PB, s1));
		MLZ_HIP(ctx, hipEventRecord(ev.e[1], s1));
	}
	// the inverses as INT8 A operands, packed once for the leaves
	I8Batch bz = I8Batch::single();
	bz.count = (int) nfull;
	bz.outer[0] = (long long) (S * S);
	size_t item = 0;
	MLZ_TRY(zp_pack_i8_a_dev(ctx, MLZ_U32, p, bz, S, S, nullptr, S, nullptr, &item));
	if (item != TRSM_ZP_BYTES)
		return set_error(ctx, MLZ_E_ARG, "packed inverse size %zu, expected %zu", item, TRSM_ZP_BYTES);
	MLZ_TRY(zp_pack_i8_a_dev(ctx, MLZ_U32, p, bz, S, S, Z, S, Zp, nullptr));
	if (upd.empty())
		return trsm_rec<E>(ctx, p, upper, unit, n, nrhs, a, A, lda, B, ldb, W, algo, Zp);
	MLZ_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev.e[1], 0));
	size_t ui = 0;
	MLZ_TRY(trsm_rec<E>(ctx, p, upper, unit, n, nrhs, a, A, lda, B, ldb, W, algo, Zp, upk.data(), &ui,
			    packed_rhs ? &pk : nullptr, 0));
	if (ui != upd.size())
		return set_error(ctx, MLZ_E_ARG, "trsm used %zu of %zu packed update blocks", ui, upd.size());
	return MLZ_OK;
}

// The solve proper with the block inverses W of zp_trsm_prepare (for a
// diagonal sub-block starting at row i0 of the prepared matrix, pass
// W + i0 * TRSM_LEAF; i0 a multiple of TRSM_LEAF).
int zp_trsm_apply(mlz_ctx *ctx, int elem, uint64_t p, int upper, int unit, size_t n, size_t nrhs, uint64_t a,
		  const void *A, size_t lda, void *B, size_t ldb, const uint32_t *W, int algo)
{
	if (n == 0 || nrhs == 0)
		return MLZ_OK;
	a %= p;
	if (a == 0)                                  // level3-generic.tcc:434-435
		return zp_scal_dev(ctx, elem, p, n, nrhs, 0, B, ldb);
	const uint32_t pp = (uint32_t) p, aa = (uint32_t) a;
	if (elem == MLZ_U16)
		return trsm_apply_t<MLZ_U16>(ctx, pp, upper, unit, n, nrhs, aa, (const uint16_t *) A, lda, (uint16_t *) B,
					     ldb, W, algo);
	if (elem == MLZ_U32)
		return trsm_apply_t<MLZ_U32>(ctx, pp, upper, unit, n, nrhs, aa, (const uint32_t *) A, lda, (uint32_t *) B,
					     ldb, W, algo);
	return trsm_apply_t<MLZ_F64_SYM>(ctx, pp, upper, unit, n, nrhs, aa, (const double *) A, lda, (double *) B, ldb,
					 W, algo);
}

int zp_trsm_dev(mlz_ctx *ctx, int elem, uint64_t p, int upper, int unit, size_t n, size_t nrhs,
		uint64_t a, const void *A, size_t lda, void *B, size_t ldb, size_t *bad_diag, int algo)
{
	if (n == 0 || nrhs == 0)
		return MLZ_OK;
	if (a % p == 0)                              // level3-generic.tcc:434-435
		return zp_scal_dev(ctx, elem, p, n, nrhs, 0, B, ldb);
	const uint32_t *W = nullptr;
	MLZ_TRY(zp_trsm_prepare(ctx, elem, p, upper, unit, n, A, lda, bad_diag, &W));
	return zp_trsm_apply(ctx, elem, p, upper, unit, n, nrhs, a, A, lda, B, ldb, W, algo);
}

}  // namespace mlz

using namespace mlz;

extern "C" int mlz_zp_trsm(mlz_ctx *ctx, mlz_elem elem, uint64_t p, int upper, int unit_diag, size_t n,
			   size_t nrhs, uint64_t a, const void *A, size_t lda, void *B, size_t ldb,
			   size_t *bad_diag, int algo)
{
	if (!ctx)
		return MLZ_E_ARG;
	if (p < 2 || (elem == MLZ_U16 && p > 65535) || (elem == MLZ_U32 && p > 0xffffffffull) ||
	    (elem == MLZ_F64_SYM && p >= 67108864ull) || elem_size(elem) == 0)
		return set_error(ctx, MLZ_E_ARG, "bad modulus or element type");
	if ((n && !A) || (n && nrhs && !B))
		return set_error(ctx, MLZ_E_ARG, "null matrix pointer");
	if ((n && lda < n) || (n && nrhs && ldb < nrhs))
		return set_error(ctx, MLZ_E_ARG, "leading dimension smaller than row length");
	if (n > (size_t) INT32_MAX || nrhs > (size_t) INT32_MAX)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (ctx->ptr_mode == MLZ_PTR_DEVICE)
		return zp_trsm_dev(ctx, elem, p, upper, unit_diag, n, nrhs, a, A, lda, B, ldb, bad_diag, algo);
	if (n == 0 || nrhs == 0)
		return MLZ_OK;
	const size_t es = elem_size(elem);
	void *dA = nullptr, *dB = nullptr;
	MLZ_TRY(stage_in(ctx, WS_STAGE_A, A, n, n, lda, es, &dA));
	MLZ_TRY(stage_in(ctx, WS_STAGE_B, B, n, nrhs, ldb, es, &dB));
	int rc = zp_trsm_dev(ctx, elem, p, upper, unit_diag, n, nrhs, a, dA, n, dB, nrhs, bad_diag, algo);
	if (rc != MLZ_OK)
		return rc;
	return stage_out(ctx, B, dB, n, nrhs, ldb, es);
}
