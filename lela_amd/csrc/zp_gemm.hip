// lela_amd/csrc/zp_gemm.hip -- C <- a*A*B + b*C over Z/p on MI355X.
//
// Three exact kernel families (results are identical; only speed differs):
//
//  * INT8 (p <= 2^16).  Every residue v has a representative x in
//    [-32896, 32639] (v itself if v <= 32639, else v - p), and x + 32896 is
//    an unsigned 16-bit number u, so x = 256*(hi(u)-128) + (lo(u)-128) with
//    both limbs in [-128, 127].  A*B = 65536*HH + 256*(HL+LH) + LL where each
//    of HH, HL+LH, LL is an int8 x int8 -> int32 product on
//    v_mfma_i32_32x32x32_i8.  |HL+LH| <= 2*16384*k, so k <= 65535 keeps every
//    int32 partial sum exact; longer k is split into chunks that accumulate
//    through C.  Limbs are packed once per call into MFMA-fragment order so
//    that every global and LDS access is a contiguous 16-byte-per-lane load.
//  * PLAIN.  VALU uint64 multiply-accumulate (v_mad_u64_u32) over LDS tiles
//    with a single reduction per entry -- the delayed reduction of
//    ZpModule's dot (lela/blas/level1-modular.tcc:25-63); for p^2*k >= 2^64
//    the 2^64-wrap correction of lela/blas/level1-modular.tcc:135-158.
//  * FP64 (p < 2^26).  Symmetric A entries times B split into two centred
//    13-bit limbs, accumulated exactly on v_mfma_f64_16x16x4_f64 while the
//    partial sums stay below 2^53 (the role of block_size in
//    lela/ring/modular.h:661-669 and lela/blas/level3-modular.tcc:97-119).
//
// C update per entry follows lela/blas/level3-generic.tcc:309-311:
// c <- b*c + a*d mod p, with d = (A*B)_{ij} mod p.
#include "batch.hpp"

#include <mutex>

namespace mlz {

// ===========================================================================
// INT8 path
// ===========================================================================

// Fragment geometry for v_mfma_i32_32x32x32_i8: a fragment is 32 indices (rows
// of A or columns of B) x 32 k-values, 16 bytes per lane; lane l holds index
// l%32 and k-values 16*(l/32) .. 16*(l/32)+15.  A and B use the same
// (lane, byte) -> k assignment, so the hardware's k order is irrelevant.
// Packed operand layout: [iblock][kblock][limb(hi, lo)][64 lanes][16 bytes].
constexpr int I8_FRAG = 1024;
constexpr int I8_KMAX = 65504;   // multiple of 64, <= 65535

// Packs a 64 (index) x 64 (k) tile into 2 x 2 fragment pairs.
// TRANS = false: element(i, kk) = src[i*ld + kk]   (A, row-major m x k)
// TRANS = true : element(i, kk) = src[kk*ld + i]   (B, row-major k x n)
// Vector loads need a 16-byte aligned base and row stride.
inline bool aligned16(const void *p, size_t row_bytes)
{
	return ((uintptr_t) p % 16) == 0 && (row_bytes % 16) == 0;
}

// Eight consecutive elements of the contiguous dimension, one vector load
// when the whole run is inside the matrix and aligned (VEC), four 32-bit
// loads for a 4-byte aligned run of 16-bit entries, else per element.
template <int E, bool VEC>
__device__ inline void load8(const typename Elem<E>::T *ptr, int valid, uint32_t p, uint32_t (&out)[8])
{
	typedef typename Elem<E>::T T;
	if (VEC && valid >= 8) {
		if (sizeof(T) == 2) {
			const uint4 w = *(const uint4 *) ptr;
			const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
			for (int q = 0; q < 4; ++q) {
				out[2 * q] = ws[q] & 0xffffu;
				out[2 * q + 1] = ws[q] >> 16;
			}
			return;
		}
		if (sizeof(T) == 4) {
			const uint4 w0 = ((const uint4 *) ptr)[0], w1 = ((const uint4 *) ptr)[1];
			out[0] = w0.x; out[1] = w0.y; out[2] = w0.z; out[3] = w0.w;
			out[4] = w1.x; out[5] = w1.y; out[6] = w1.z; out[7] = w1.w;
			return;
		}
	}
	if (sizeof(T) == 2 && valid >= 8 && ((uintptr_t) ptr & 3) == 0) {
		// a 4-byte aligned run of 16-bit entries (rows of even length, say):
		// four 32-bit loads instead of eight 16-bit ones
		const uint32_t *w = (const uint32_t *) ptr;
#pragma unroll
		for (int q = 0; q < 4; ++q) {
			const uint32_t x = w[q];
			out[2 * q] = x & 0xffffu;
			out[2 * q + 1] = x >> 16;
		}
		return;
	}
#pragma unroll
	for (int q = 0; q < 8; ++q)
		out[q] = q < valid ? Elem<E>::load(ptr + q, p) : 0u;
}

// Packs the 64 x 64 tile at (i0, k0) with 256 threads; tile is 64 x 65 words
// of LDS.  Shared by the single-matrix and the batched pack kernels.
template <int E, bool TRANS, bool VEC>
__device__ __forceinline__ void i8_pack_tile(const typename Elem<E>::T *__restrict__ src, size_t ld, int idim,
					     int kdim, uint32_t p, int8_t *__restrict__ dst, int KB, int i0, int k0,
					     uint32_t (*tile)[65])
{
	const int tid = threadIdx.x;
	// 64 x 64 tile = 512 runs of 8 along the contiguous dimension, 2 per thread
#pragma unroll
	for (int e = 0; e < 2; ++e) {
		const int idx = e * 256 + tid;
		const int line = idx >> 3, off = (idx & 7) * 8;
		uint32_t v[8];
		if (!TRANS) {          // line = row i, runs along k
			const int gi = i0 + line, gk = k0 + off;
			const int valid = gi < idim ? min(8, kdim - gk) : 0;
			load8<E, VEC>(src + (size_t) gi * ld + gk, valid, p, v);
#pragma unroll
			for (int q = 0; q < 8; ++q)
				tile[line][off + q] = i8_limbs(q < valid ? v[q] : 0u, p);
		} else {               // line = k, runs along i
			const int gk = k0 + line, gi = i0 + off;
			const int valid = gk < kdim ? min(8, idim - gi) : 0;
			load8<E, VEC>(src + (size_t) gk * ld + gi, valid, p, v);
#pragma unroll
			for (int q = 0; q < 8; ++q)
				tile[off + q][line] = i8_limbs(q < valid ? v[q] : 0u, p);
		}
	}
	__syncthreads();
	const int f = tid >> 6, lane = tid & 63;
	const int isub = f >> 1, ksub = f & 1;
	const int i = isub * 32 + (lane & 31);
	const int kk0 = ksub * 32 + 16 * (lane >> 5);
	uint32_t hw[4], lw[4];
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		uint32_t h = 0, l = 0;
#pragma unroll
		for (int s = 0; s < 4; ++s) {
			uint32_t w = tile[i][kk0 + q * 4 + s];
			h |= ((w >> 8) & 0xffu) << (8 * s);
			l |= (w & 0xffu) << (8 * s);
		}
		hw[q] = h;
		lw[q] = l;
	}
	const size_t ib = (size_t) (i0 / 32 + isub), kb = (size_t) (k0 / 32 + ksub);
	int8_t *fr = dst + ((ib * KB + kb) * 2) * I8_FRAG + lane * 16;
	*(v4i *) fr = v4i{(int) hw[0], (int) hw[1], (int) hw[2], (int) hw[3]};
	*(v4i *) (fr + I8_FRAG) = v4i{(int) lw[0], (int) lw[1], (int) lw[2], (int) lw[3]};
}

template <int E, bool TRANS, bool VEC>
__global__ __launch_bounds__(256) void pack_i8_kernel(const typename Elem<E>::T *__restrict__ src,
						      size_t ld, int idim, int kdim, uint32_t p,
						      int8_t *__restrict__ dst, int KB, I8Batch bt, size_t dstride)
{
	__shared__ uint32_t tile[64][65];
	const int z = blockIdx.z;
	src += bt.offset(z, 0);
	dst += (size_t) z * dstride;
	i8_pack_tile<E, TRANS, VEC>(src, ld, idim, kdim, p, dst, KB, blockIdx.x * 64, blockIdx.y * 64, tile);
}

// Maps a linear workgroup id to an output tile.  Workgroups are dispatched
// round-robin over the 8 XCDs; the remap gives each XCD a contiguous range
// of tiles (its own L2 working set), then tiles are visited in groups of
// GM tile-rows so concurrently running workgroups share A and B panels.
__device__ inline void tile_of(int bid, int tiles_m, int tiles_n, int &tm, int &tn)
{
	const int total = tiles_m * tiles_n;
	if ((total & 7) == 0)
		bid = (bid & 7) * (total >> 3) + (bid >> 3);
	const int GM = 8;
	const int per_group = GM * tiles_n;
	const int group = bid / per_group;
	const int first_m = group * GM;
	const int gsz = min(tiles_m - first_m, GM);
	const int in_group = bid - group * per_group;
	tm = first_m + in_group % gsz;
	tn = in_group / gsz;
}

// Workgroup of WM x WN waves; each wave computes (32*WTM) x (32*WTN)
// outputs as WTM x WTN MFMA blocks of 32 x 32, with three int32 accumulators
// (HH, HL+LH, LL) each.  The k loop stages 32*KK k-values (KK k-blocks x 2
// limbs) of A and B per step through double-buffered LDS, the next step's
// global loads in flight during the current step's MFMAs.
template <int WTM, int WTN, int KK, int WM = 2, int WN = 2>
struct I8Tile {
	static constexpr int THREADS = 64 * WM * WN;
	static constexpr int BM = 32 * WTM * WM, BN = 32 * WTN * WN;  // workgroup tile
	static constexpr int RBA = WTM * WM, RBB = WTN * WN;          // 32-index blocks
	static constexpr int CHUNK = KK * 2048;                       // bytes per block per step
	static constexpr int STAGE = (RBA + RBB) * CHUNK;             // LDS bytes per step
	static constexpr int LOADS = STAGE / (THREADS * 16);          // 16-byte loads per thread
	static_assert(STAGE % (THREADS * 16) == 0, "stage must split into whole loads");
};

// The MFMAs of one staged k-step: A blocks at la, B blocks at lb, each
// KK k-blocks x 2 limbs x 1024 bytes (CHUNK bytes per 32-index block).
template <class T, int WTM, int WTN, int KK>
__device__ __forceinline__ void i8_stage_mfma(const int8_t *la, const int8_t *lb, int wr, int wc, int lane,
					      v16i (&acc)[3][WTM][WTN])
{
#pragma unroll
	for (int kk = 0; kk < KK; ++kk) {
		v4i fa[WTM][2];
#pragma unroll
		for (int i = 0; i < WTM; ++i)
#pragma unroll
			for (int l = 0; l < 2; ++l)
				fa[i][l] = *(const v4i *) (la + (wr * WTM + i) * T::CHUNK + kk * 2048 + l * 1024 + lane * 16);
#pragma unroll
		for (int j = 0; j < WTN; ++j) {
			const v4i fb0 = *(const v4i *) (lb + (wc * WTN + j) * T::CHUNK + kk * 2048 + lane * 16);
			const v4i fb1 = *(const v4i *) (lb + (wc * WTN + j) * T::CHUNK + kk * 2048 + 1024 + lane * 16);
#pragma unroll
			for (int i = 0; i < WTM; ++i)
				acc[0][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i][0], fb0, acc[0][i][j], 0, 0, 0);
#pragma unroll
			for (int i = 0; i < WTM; ++i)
				acc[1][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i][0], fb1, acc[1][i][j], 0, 0, 0);
#pragma unroll
			for (int i = 0; i < WTM; ++i)
				acc[2][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i][1], fb1, acc[2][i][j], 0, 0, 0);
#pragma unroll
			for (int i = 0; i < WTM; ++i)
				acc[1][i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i][1], fb0, acc[1][i][j], 0, 0, 0);
		}
	}
}

// The 32 x 32 block a wave holds (lane l, register r: row 8*(r/4) +
// 4*(l/32) + r%4, column l%32) written as the B operand fragments of a later
// product whose k runs along C's rows: fragment (column block, row block)
// lane L holds column L%32, rows 16*(L/32) .. +15.  Lanes l and l^32 swap
// half of their rows (one exchange of 8 values), then each lane writes its
// two limb words as 16 bytes each.
__device__ __forceinline__ void i8_pack_out(const uint32_t (&ov)[16], uint32_t p, int lane, int8_t *__restrict__ fr)
{
	const bool hi = lane >= 32;
	uint32_t rv[8];
#pragma unroll
	for (int r = 0; r < 8; ++r)
		rv[r] = (uint32_t) __shfl_xor((int) (hi ? ov[r] : ov[r + 8]), 32, 64);
	// rows 16*(l/32) + 4q .. +3 for q = 0..3
	uint32_t v[16];
#pragma unroll
	for (int t = 0; t < 4; ++t) {
		v[t] = hi ? rv[t] : ov[t];
		v[4 + t] = hi ? ov[8 + t] : rv[t];
		v[8 + t] = hi ? rv[4 + t] : ov[4 + t];
		v[12 + t] = hi ? ov[12 + t] : rv[4 + t];
	}
	uint32_t hw[4], lw[4];
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		uint32_t h = 0, l = 0;
#pragma unroll
		for (int t = 0; t < 4; ++t) {
			const uint32_t w = i8_limbs(v[4 * q + t], p);
			h |= ((w >> 8) & 0xffu) << (8 * t);
			l |= (w & 0xffu) << (8 * t);
		}
		hw[q] = h;
		lw[q] = l;
	}
	*(v4i *) (fr + lane * 16) = v4i{(int) hw[0], (int) hw[1], (int) hw[2], (int) hw[3]};
	*(v4i *) (fr + I8_FRAG + lane * 16) = v4i{(int) lw[0], (int) lw[1], (int) lw[2], (int) lw[3]};
}

// Epilogue of a workgroup tile.  32x32 accumulator layout: lane l,
// register r holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.  PK: the
// results are also written to Cpk as packed B fragments (i8_pack_out),
// 32-column blocks panelC bytes apart; the caller guarantees whole 32 x 32
// blocks.  (A template flag, not a null test: kept in every instantiation,
// the packing's registers slowed the 8192^3 product from 1.80 to 1.90 ms.)
template <int E, class T, int WTM, int WTN, bool PK = false>
__device__ __forceinline__ void i8_epilogue(const v16i (&acc)[3][WTM][WTN], typename Elem<E>::T *__restrict__ C,
					    size_t ldc, int m, int n, uint32_t a, uint32_t b, const ModP &mp, int tm,
					    int tn, int wr, int wc, int lane, int8_t *__restrict__ Cpk = nullptr,
					    size_t panelC = 0)
{
	const uint32_t p = (uint32_t) mp.p;
#pragma unroll
	for (int i = 0; i < WTM; ++i)
#pragma unroll
		for (int j = 0; j < WTN; ++j) {
			const int col = tn * T::BN + wc * 32 * WTN + j * 32 + (lane & 31);
			const int row0 = tm * T::BM + wr * 32 * WTM + i * 32;
			if (row0 + 32 <= m && col - (lane & 31) + 32 <= n) {
				// Whole 32 x 32 block inside C: the 16 loads of C are issued
				// together (one wait instead of one per element).
				typename Elem<E>::T *cb = C + (size_t) (row0 + 4 * (lane >> 5)) * ldc + col;
				uint32_t cv[16];
				if (b) {
#pragma unroll
					for (int r = 0; r < 16; ++r)
						cv[r] = Elem<E>::load(cb + (size_t) (8 * (r >> 2) + (r & 3)) * ldc, p);
				}
				uint32_t ov[16];
#pragma unroll
				for (int r = 0; r < 16; ++r) {
					const double v = fma((double) acc[0][i][j][r], 65536.0,
							     fma((double) acc[1][i][j][r], 256.0, (double) acc[2][i][j][r]));
					// a = 1: v + b c (|.| < 2^49) is exact, one reduction
					double t = a == 1 ? v : (double) a * (double) mp.mod_f64(v);
					if (b)
						t = fma((double) b, (double) cv[r], t);
					const uint32_t o = mp.mod_f64(t);
					Elem<E>::store(cb + (size_t) (8 * (r >> 2) + (r & 3)) * ldc, o, p);
					if constexpr (PK)
						ov[r] = o;
				}
				if constexpr (PK)
					i8_pack_out(ov, p, lane,
						    Cpk + (size_t) ((col - (lane & 31)) >> 5) * panelC + (size_t) (row0 >> 5) * 2 * I8_FRAG);
				continue;
			}
#pragma unroll
			for (int r = 0; r < 16; ++r) {
				const int row = row0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
				if (row < m && col < n) {
					// 65536*HH + 256*(HL+LH) + LL has |.| < 2^48, a*d + b*c < 2^33 and
					// (a = 1) v + b*c < 2^49 (p <= 2^16): exact doubles, reduced by mod_f64.
					const double v = fma((double) acc[0][i][j][r], 65536.0,
							     fma((double) acc[1][i][j][r], 256.0, (double) acc[2][i][j][r]));
					typename Elem<E>::T *cp = C + (size_t) row * ldc + col;
					double t = a == 1 ? v : (double) a * (double) mp.mod_f64(v);
					if (b)
						t = fma((double) b, (double) Elem<E>::load(cp, (uint32_t) mp.p), t);
					Elem<E>::store(cp, mp.mod_f64(t), (uint32_t) mp.p);
				}
			}
		}
}

// One output tile (tm, tn) of the INT8 product; lds holds 2 * STAGE bytes
// (double-buffered [A: RBA blocks | B: RBB blocks] x [KK k-blocks x 2 limbs x
// 1024 bytes]).  Shared by the single-matrix kernel and the batched one.
template <int E, int WTM, int WTN, int KK, int WM, int WN>
__device__ __forceinline__ void i8_tile(const int8_t *__restrict__ Ap, const int8_t *__restrict__ Bp, int KB,
					typename Elem<E>::T *__restrict__ C, size_t ldc, int m, int n, uint32_t a,
					uint32_t b, const ModP &mp, int tm, int tn, int8_t *lds_base)
{
	typedef I8Tile<WTM, WTN, KK, WM, WN> T;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wr = wave / WN, wc = wave % WN;

	const size_t panel = (size_t) KB * 2 * I8_FRAG;   // bytes per 32-index block
	const int8_t *Ag = Ap + (size_t) (tm * T::RBA) * panel;
	const int8_t *Bg = Bp + (size_t) (tn * T::RBB) * panel;
	// Load c of a step covers stage bytes [c*16*THREADS, (c+1)*16*THREADS);
	// thread tid moves bytes g = c*16*THREADS + tid*16, i.e. block g / CHUNK at
	// offset g % CHUNK.
	const int8_t *src[T::LOADS];
#pragma unroll
	for (int c = 0; c < T::LOADS; ++c) {
		const int g = c * 16 * T::THREADS + tid * 16;
		const int blk = g / T::CHUNK, o = g % T::CHUNK;
		src[c] = blk < T::RBA ? Ag + (size_t) blk * panel + o : Bg + (size_t) (blk - T::RBA) * panel + o;
	}

	v4i stage[T::LOADS];
	v16i acc[3][WTM][WTN];
#pragma unroll
	for (int t = 0; t < 3; ++t)
#pragma unroll
		for (int i = 0; i < WTM; ++i)
#pragma unroll
			for (int j = 0; j < WTN; ++j)
				acc[t][i][j] = v16i{0};

	const int KS = KB / KK;
#pragma unroll
	for (int c = 0; c < T::LOADS; ++c)
		stage[c] = *(const v4i *) src[c];
	for (int ks = 0; ks < KS; ++ks) {
		int8_t *la = lds_base + (ks & 1) * T::STAGE, *lb = la + T::RBA * T::CHUNK;
#pragma unroll
		for (int c = 0; c < T::LOADS; ++c)
			*(v4i *) (la + c * 16 * T::THREADS + tid * 16) = stage[c];
		__syncthreads();
		if (ks + 1 < KS) {
			const size_t o = (size_t) (ks + 1) * T::CHUNK;
#pragma unroll
			for (int c = 0; c < T::LOADS; ++c)
				stage[c] = *(const v4i *) (src[c] + o);
		}
		i8_stage_mfma<T, WTM, WTN, KK>(la, lb, wr, wc, lane, acc);
	}
	i8_epilogue<E, T, WTM, WTN>(acc, C, ldc, m, n, a, b, mp, tm, tn, wr, wc, lane);
}

template <int E, int WTM, int WTN, int KK, int OCC, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN, OCC) void gemm_i8_kernel(const int8_t *__restrict__ Ap,
							   const int8_t *__restrict__ Bp, int KB,
							   typename Elem<E>::T *__restrict__ C, size_t ldc,
							   int m, int n, uint32_t a, uint32_t b, ModP mp,
							   int tiles_m, int tiles_n, I8Batch bt, size_t sAp, size_t sBp)
{
	__shared__ __attribute__((aligned(16))) int8_t lds[2 * I8Tile<WTM, WTN, KK, WM, WN>::STAGE];
	const int z = blockIdx.y;
	Ap += (size_t) z * sAp;
	Bp += (size_t) z * sBp;
	C += bt.offset(z, 2);
	int tm, tn;
	tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
	i8_tile<E, WTM, WTN, KK, WM, WN>(Ap, Bp, KB, C, ldc, m, n, a, b, mp, tm, tn, lds);
}

// The same tile with the k-steps staged by direct global -> LDS loads
// (global_load_lds_dwordx4: lane l of a wave writes its 16 bytes at the
// wave's LDS address + 16*l, which is the register-staged layout above)
// through NS buffers, so that NS-1 steps are in flight while one is
// multiplied and no registers hold staged data.  The buffers are distinct
// LDS arrays, so the compiler can tell the loads into one from the reads of
// another and does not drain the loads in flight before each read.  Step j
// waits for its own loads with vmcnt (per wave, completed in issue order),
// then a barrier makes every wave's part visible and guarantees that the
// buffer refilled next (the one step j-1 read) is no longer read.
#define MLZ_AS_GLOBAL __attribute__((address_space(1)))
#define MLZ_AS_LDS __attribute__((address_space(3)))

// s_waitcnt vmcnt(N) with no expcnt / lgkmcnt condition (gfx9 encoding:
// vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
	static_assert(N >= 0 && N < 64, "vmcnt has 6 bits");
	__builtin_amdgcn_s_waitcnt((N & 0xf) | ((N >> 4) << 14) | (0x7 << 4) | (0xf << 8));
}

// This wave's LDS reads complete (lgkmcnt(0)), then the workgroup barrier.
// Global loads in flight are not waited for (the fence of __syncthreads()
// would drain them); the empty asm statements keep the compiler from moving
// memory accesses across the barrier.
__device__ __forceinline__ void lds_barrier()
{
	__builtin_amdgcn_s_waitcnt(0xf | (0x3 << 14) | (0x7 << 4));
	asm volatile("" ::: "memory");
	__builtin_amdgcn_s_barrier();
	asm volatile("" ::: "memory");
}

template <int E, int WTM, int WTN, int KK, int OCC, int WM, int WN, int NS, bool PK = false>
__global__ __launch_bounds__(64 * WM * WN, OCC) void gemm_i8_async_kernel(const int8_t *__restrict__ Ap,
								 const int8_t *__restrict__ Bp, int KB,
								 typename Elem<E>::T *__restrict__ C, size_t ldc,
								 int m, int n, uint32_t a, uint32_t b, ModP mp,
								 int tiles_m, int tiles_n, I8Batch bt, size_t sAp, size_t sBp,
								 size_t panelB, int8_t *__restrict__ Cpk, size_t panelC)
{
	typedef I8Tile<WTM, WTN, KK, WM, WN> T;
	{
		const int z = blockIdx.y;
		Ap += (size_t) z * sAp;
		Bp += (size_t) z * sBp;
		C += bt.offset(z, 2);
	}
	static_assert(NS == 3 || NS == 4, "three or four stage buffers");
	__shared__ __attribute__((aligned(16))) int8_t s0[T::STAGE], s1[T::STAGE], s2[T::STAGE],
		s3[NS > 3 ? T::STAGE : 16];
	int tm, tn;
	tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wr = wave / WN, wc = wave % WN;
	const int wave_off = __builtin_amdgcn_readfirstlane(wave) * 1024;

	// A's 32-row blocks are panel bytes apart; B's 32-column blocks panelB
	// (more than panel when B is a k-range of a larger packed operand)
	const size_t panel = (size_t) KB * 2 * I8_FRAG;
	const int8_t *Ag = Ap + (size_t) (tm * T::RBA) * panel;
	const int8_t *Bg = Bp + (size_t) (tn * T::RBB) * panelB;
	const int8_t *src[T::LOADS];
#pragma unroll
	for (int c = 0; c < T::LOADS; ++c) {
		const int g = c * 16 * T::THREADS + tid * 16;
		const int blk = g / T::CHUNK, o = g % T::CHUNK;
		src[c] = blk < T::RBA ? Ag + (size_t) blk * panel + o : Bg + (size_t) (blk - T::RBA) * panelB + o;
	}
	auto buf = [&](int s) -> int8_t * { return s == 0 ? s0 : s == 1 ? s1 : s == 2 ? s2 : s3; };
	auto issue = [&](int step, int8_t *dst) {
		const size_t o = (size_t) step * T::CHUNK;
#pragma unroll
		for (int c = 0; c < T::LOADS; ++c)
			__builtin_amdgcn_global_load_lds((const MLZ_AS_GLOBAL void *) (src[c] + o),
							 (MLZ_AS_LDS void *) (dst + c * 16 * T::THREADS + wave_off), 16, 0, 0);
	};

	v16i acc[3][WTM][WTN];
#pragma unroll
	for (int t = 0; t < 3; ++t)
#pragma unroll
		for (int i = 0; i < WTM; ++i)
#pragma unroll
			for (int j = 0; j < WTN; ++j)
				acc[t][i][j] = v16i{0};

	// Every step issues one stage of loads, so the wait count is the same
	// at every step: past the last step the loads repeat step KS - 1 into the
	// buffer just read, which nothing reads again.  (Conditional issues
	// would also make the compiler drain all loads before some reads.)
	// The main loop runs whole groups of NS steps with no branch inside, so
	// that the compiler sees the same loads in flight at every read; the
	// remaining steps follow it.
	const int KS = KB / KK;
	auto step = [&](int j, int s) {
		wait_vmcnt<T::LOADS * (NS - 2)>();   // step j's loads are done
		lds_barrier();
		issue(min(j + NS - 1, KS - 1), buf((s + NS - 1) % NS));
		i8_stage_mfma<T, WTM, WTN, KK>(buf(s), buf(s) + T::RBA * T::CHUNK, wr, wc, lane, acc);
	};
#pragma unroll
	for (int s = 0; s < NS - 1; ++s)
		issue(min(s, KS - 1), buf(s));
	int ks0 = 0;
	for (; ks0 + NS <= KS; ks0 += NS) {
#pragma unroll
		for (int s = 0; s < NS; ++s)
			step(ks0 + s, s);
	}
#pragma unroll
	for (int s = 0; s < NS - 1; ++s)
		if (ks0 + s < KS)
			step(ks0 + s, s);
	wait_vmcnt<0>();   // no load into LDS outlives the workgroup
	i8_epilogue<E, T, WTM, WTN, PK>(acc, C, ldc, m, n, a, b, mp, tm, tn, wr, wc, lane, Cpk, panelC);
}

// ===========================================================================
// No-allocation scratch.  mlz_zp_gemm_noalloc packs its operands into this
// file-scope device array instead of the context's workspace slots, so the
// call makes no hipMalloc.  Each device holds its own copy for the life of the
// process; its address on a device comes from hipGetSymbolAddress with that
// device current (the array's name in host code is only a host-side stand-in).
// Every call on a device shares the array: calls that overlap on one device
// (two contexts, two threads) race on it.
// ===========================================================================
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t) MLZ_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static bool g_scratch_done[NOALLOC_MAX_DEVICES];

int noalloc_scratch_init(mlz_ctx *ctx)
{
	if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
		return set_error(ctx, MLZ_E_UNSUPPORTED, "no-allocation scratch: device %d beyond %d", ctx->device,
				 NOALLOC_MAX_DEVICES);
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (!g_scratch_done[ctx->device]) {
		void *ptr = nullptr;
		// The caller (mlz_ctx_create) has made ctx->device current.
		hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess) {
			(void) hipGetLastError();
			return set_error(ctx, MLZ_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", ctx->device,
					 hipGetErrorString(e));
		}
		g_scratch_ptr[ctx->device] = ptr;
		g_scratch_done[ctx->device] = true;
	}
	ctx->fixed_ws = g_scratch_ptr[ctx->device];
	ctx->fixed_ws_bytes = NOALLOC_SCRATCH_BYTES;
	return MLZ_OK;
}

// Buffers for the packed A and B operands of one gemm launch: the workspace
// slots, or during a no-allocation call the scratch array (A at its start, B
// after it on a 256-byte boundary).  Both pack kernels write every byte they
// hand to the gemm kernel (padding included), so nothing needs clearing.
static int pack_buffers(mlz_ctx *ctx, size_t abytes, size_t bbytes, void **ap, void **bp)
{
	if (!ctx->noalloc) {
		MLZ_TRY(ws_get(ctx, WS_PACK_A, abytes, ap));
		return ws_get(ctx, WS_PACK_B, bbytes, bp);
	}
	const size_t boff = round_up(abytes, 256);
	if (!ctx->fixed_ws || boff + bbytes > ctx->fixed_ws_bytes)
		return set_error(ctx, MLZ_E_ARG,
				 "no-allocation gemm needs %zu bytes of packed operands, beyond the %zu-byte scratch",
				 boff + bbytes, ctx->fixed_ws_bytes);
	*ap = ctx->fixed_ws;
	*bp = (uint8_t *) ctx->fixed_ws + boff;
	return MLZ_OK;
}

// Tile variants of the INT8 kernel (index = ctx option "gemm.i8.variant"):
//   1: 128 x 128, 4 waves, 64 k per step
//   2: 128 x 256, 4 waves of 64 x 128, 32 k per step
//   3: 256 x 128, 4 waves of 128 x 64, 32 k per step
//   4: 128 x 128, 4 waves, 32 k per step, 2 waves/SIMD
//   5: 128 x 256, 8 waves of 64 x 64, 32 k per step, 2 waves/SIMD
//   6: 256 x 128, 8 waves of 64 x 64, 32 k per step, 2 waves/SIMD
//   7: as 5, staged by direct loads through 3 LDS buffers
//   8: as 5, staged by direct loads through 4 LDS buffers
//   9: as 5 with 64 k per step, direct loads through 3 LDS buffers
//  10: as 4, staged by direct loads through 4 LDS buffers
//  11: 64 x 64, 4 waves of 32 x 32, through 4 LDS buffers, 4 workgroups/CU
enum { I8V_AUTO = 0, I8V_COUNT = 12 };

// EA: storage of the A operand when it differs from B's and C's (E).
// Bpre: B already packed, its 32-column blocks bpanel bytes apart (a k-range
// of a larger packed operand); Cpk: C also written as packed B fragments,
// 32-column blocks cpanel bytes apart (i8_epilogue).  Both need one k chunk,
// one product and the multi-buffered kernels; Cpk needs m and n multiples of 32.
template <int E, int WTM, int WTN, int KK, int OCC, int WM = 2, int WN = 2, int NS = 0, int EA = E>
static int gemm_i8_launch_t(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, size_t k, uint32_t a,
			    const void *A, size_t lda, const void *B, size_t ldb, uint32_t b, void *C,
			    size_t ldc, const I8Batch &bt = I8Batch::single(), const int8_t *Apre = nullptr,
			    const int8_t *Bpre = nullptr, size_t bpanel = 0, int8_t *Cpk = nullptr, size_t cpanel = 0)
{
	typedef typename Elem<E>::T T;
	typedef typename Elem<EA>::T TA;
	if (Apre && k > (size_t) I8_KMAX)
		return set_error(ctx, MLZ_E_ARG, "a pre-packed A operand needs k <= %d", I8_KMAX);
	if ((Bpre || Cpk) && (k > (size_t) I8_KMAX || bt.count != 1 || NS == 0))
		return set_error(ctx, MLZ_E_ARG, "packed B input or packed output: one product with k <= %d", I8_KMAX);
	if (Cpk && (m % 32 || n % 32))
		return set_error(ctx, MLZ_E_ARG, "packed output needs m and n in multiples of 32");
	if (Bpre && bpanel < round_up(k, 64) * 64)
		return set_error(ctx, MLZ_E_ARG, "packed B: block stride %zu below its k", bpanel);
	if ((!Apre && !A) || (!Bpre && !B))
		return set_error(ctx, MLZ_E_ARG, "INT8 gemm: an operand is neither given nor pre-packed");
	typedef I8Tile<WTM, WTN, KK, WM, WN> Tile;
	const ModP mp = make_modp(p);
	const size_t mpad = round_up(m, Tile::BM), npad = round_up(n, Tile::BN);
	if (bt.count < 1 || bt.count > 65535)
		return set_error(ctx, MLZ_E_ARG, "batch of %d products (1 .. 65535)", bt.count);
	for (size_t k0 = 0; k0 < k; k0 += I8_KMAX) {
		const size_t kc = std::min((size_t) I8_KMAX, k - k0);
		const size_t kpad = round_up(kc, 64);
		const int KB = (int) (kpad / 32);
		const size_t sAp = mpad * kpad * 2, sBp = npad * kpad * 2;   // bytes per packed item
		void *ap = nullptr, *bp = nullptr;
		MLZ_TRY(pack_buffers(ctx, Apre ? 0 : sAp * bt.count, Bpre ? 0 : sBp * bt.count, &ap, &bp));
		if (Bpre)
			bp = (void *) Bpre;
		if (Apre)
			ap = (void *) Apre;
		else {
			ProfScope ps(ctx, "pack_i8_A");
			dim3 g(mpad / 64, kpad / 64, bt.count);
			const TA *src = (const TA *) A + k0;
			const bool vec = aligned16(src, lda * sizeof(TA)) &&
					 (bt.count == 1 || ((bt.outer[0] | bt.in[0]) * (long long) sizeof(TA)) % 16 == 0);
			if (vec)
				pack_i8_kernel<EA, false, true><<<g, 256, 0, ctx->stream>>>(src, lda, (int) m, (int) kc, p,
											    (int8_t *) ap, KB, bt, sAp);
			else
				pack_i8_kernel<EA, false, false><<<g, 256, 0, ctx->stream>>>(src, lda, (int) m, (int) kc, p,
											     (int8_t *) ap, KB, bt, sAp);
		}
		MLZ_TRY(check_launch(ctx, "pack_i8_kernel(A)"));
		if (!Bpre) {
			ProfScope ps(ctx, "pack_i8_B");
			dim3 g(npad / 64, kpad / 64, bt.count);
			const T *src = (const T *) B + k0 * ldb;
			// the pack kernel reads operand 0's offsets: hand it B's
			I8Batch bb = bt;
			bb.outer[0] = bt.outer[1];
			bb.in[0] = bt.in[1];
			const bool vec = aligned16(src, ldb * sizeof(T)) &&
					 (bt.count == 1 || ((bt.outer[1] | bt.in[1]) * (long long) sizeof(T)) % 16 == 0);
			if (vec)
				pack_i8_kernel<E, true, true><<<g, 256, 0, ctx->stream>>>(src, ldb, (int) n, (int) kc, p,
											  (int8_t *) bp, KB, bb, sBp);
			else
				pack_i8_kernel<E, true, false><<<g, 256, 0, ctx->stream>>>(src, ldb, (int) n, (int) kc, p,
											   (int8_t *) bp, KB, bb, sBp);
		}
		MLZ_TRY(check_launch(ctx, "pack_i8_kernel(B)"));
		const int tiles_m = (int) (mpad / Tile::BM), tiles_n = (int) (npad / Tile::BN);
		{
			ProfScope ps(ctx, "gemm_i8_mfma");
			dim3 g(tiles_m * tiles_n, bt.count);
			if constexpr (NS == 0)
				gemm_i8_kernel<E, WTM, WTN, KK, OCC, WM, WN><<<g, Tile::THREADS, 0, ctx->stream>>>(
					(const int8_t *) ap, (const int8_t *) bp, KB, (T *) C, ldc, (int) m, (int) n, a,
					k0 == 0 ? b : 1u, mp, tiles_m, tiles_n, bt, sAp, sBp);
			else if (Cpk) {
				// the batched entry's two tiles (64 x 64 and 128 x 128, four buffers)
				if constexpr (WM == 2 && WN == 2 && NS == 4 && WTM == WTN)
					gemm_i8_async_kernel<E, WTM, WTN, KK, OCC, WM, WN, NS, true>
						<<<g, Tile::THREADS, 0, ctx->stream>>>(
							(const int8_t *) ap, (const int8_t *) bp, KB, (T *) C, ldc, (int) m,
							(int) n, a, k0 == 0 ? b : 1u, mp, tiles_m, tiles_n, bt, sAp, sBp,
							Bpre ? bpanel : (size_t) KB * 2 * I8_FRAG, Cpk, cpanel);
				else
					return set_error(ctx, MLZ_E_ARG, "packed output: not for this tile variant");
			} else
				gemm_i8_async_kernel<E, WTM, WTN, KK, OCC, WM, WN, NS>
					<<<g, Tile::THREADS, 0, ctx->stream>>>(
						(const int8_t *) ap, (const int8_t *) bp, KB, (T *) C, ldc, (int) m, (int) n,
						a, k0 == 0 ? b : 1u, mp, tiles_m, tiles_n, bt, sAp, sBp,
						Bpre ? bpanel : (size_t) KB * 2 * I8_FRAG, nullptr, 0);
		}
		MLZ_TRY(check_launch(ctx, "gemm_i8_kernel"));
	}
	return MLZ_OK;
}

// Variant choice when none is forced.  The multi-buffered variants beat the
// register-staged ones at every measured shape (8192^3: 8 > 7 > 10 > 5 > 4;
// 2 and 3 spill registers).  The 8-wave 128 x 256 tile with four buffers wins
// once the output fills the 256 CUs with it and k is long (4096^3, 1024 x
// 8192 x 8192, 8192^3); with shorter k or fewer tiles the 128 x 128 tile at
// two workgroups per CU does (2048^3, 8192 x 8192 x 1024, 3000 x 5000 x 2000).
static int i8_variant_for(size_t m, size_t n, size_t k)
{
	const size_t tiles_wide = ceil_div(m, 128) * ceil_div(n, 256);
	if (tiles_wide >= 256 && k >= 4096)
		return 8;
	// fewer 128 x 128 tiles than CUs: the 64 x 64 tile fills the chip
	// (1024 x 1024 x 1024: 17.6 against 28.1 us; 2048 x 1024 x 2048: 34 against
	// 45 us; equal at 4096 x 1024 x 4096)
	if (ceil_div(m, 128) * ceil_div(n, 128) < 256)
		return 11;
	return 10;
}

template <int E>
static int gemm_i8_launch(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, size_t k, uint32_t a,
			  const void *A, size_t lda, const void *B, size_t ldb, uint32_t b, void *C,
			  size_t ldc)
{
	int v = ctx->i8_variant;
	if (v <= I8V_AUTO || v >= I8V_COUNT)
		v = i8_variant_for(m, n, k);
	switch (v) {
	case 2:
		return gemm_i8_launch_t<E, 2, 4, 1, 1>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 3:
		return gemm_i8_launch_t<E, 4, 2, 1, 1>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 4:
		return gemm_i8_launch_t<E, 2, 2, 1, 2>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 5:
		return gemm_i8_launch_t<E, 2, 2, 1, 2, 2, 4>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 6:
		return gemm_i8_launch_t<E, 2, 2, 1, 2, 4, 2>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 7:
		return gemm_i8_launch_t<E, 2, 2, 1, 2, 2, 4, 3>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 8:
		return gemm_i8_launch_t<E, 2, 2, 1, 2, 2, 4, 4>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 9:
		return gemm_i8_launch_t<E, 2, 2, 2, 2, 2, 4, 3>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 10:
		return gemm_i8_launch_t<E, 2, 2, 1, 2, 2, 2, 4>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case 11:
		return gemm_i8_launch_t<E, 1, 1, 1, 4, 2, 2, 4>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	default:
		return gemm_i8_launch_t<E, 2, 2, 2, 1>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	}
}

template <int E, int WTM, int WTN, int KK, int OCC, int WM, int WN, int NS>
static int gemm_i8_batched_ea(mlz_ctx *ctx, int elemA, uint32_t p, const I8Batch &bt, size_t m, size_t n, size_t k,
			      uint32_t a, const void *A, size_t lda, const void *B, size_t ldb, uint32_t b, void *C,
			      size_t ldc, const int8_t *Apre, const I8PackIO &io)
{
	if (elemA == MLZ_U16)
		return gemm_i8_launch_t<E, WTM, WTN, KK, OCC, WM, WN, NS, MLZ_U16>(ctx, p, m, n, k, a, A, lda, B, ldb, b,
										    C, ldc, bt, Apre, io.Bpre, io.bpanel,
										    io.Cpk, io.cpanel);
	if (elemA == MLZ_U32)
		return gemm_i8_launch_t<E, WTM, WTN, KK, OCC, WM, WN, NS, MLZ_U32>(ctx, p, m, n, k, a, A, lda, B, ldb, b,
										    C, ldc, bt, Apre, io.Bpre, io.bpanel,
										    io.Cpk, io.cpanel);
	return gemm_i8_launch_t<E, WTM, WTN, KK, OCC, WM, WN, NS, MLZ_F64_SYM>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C,
										ldc, bt, Apre, io.Bpre, io.bpanel, io.Cpk,
										io.cpanel);
}

template <int WTM, int WTN, int KK, int OCC, int WM, int WN, int NS>
static int gemm_i8_batched_e(mlz_ctx *ctx, int elem, int elemA, uint32_t p, const I8Batch &bt, size_t m, size_t n,
			     size_t k, uint32_t a, const void *A, size_t lda, const void *B, size_t ldb, uint32_t b,
			     void *C, size_t ldc, const int8_t *Apre, const I8PackIO &io)
{
	if (elem == MLZ_U16)
		return gemm_i8_batched_ea<MLZ_U16, WTM, WTN, KK, OCC, WM, WN, NS>(ctx, elemA, p, bt, m, n, k, a, A, lda, B,
										   ldb, b, C, ldc, Apre, io);
	if (elem == MLZ_U32)
		return gemm_i8_batched_ea<MLZ_U32, WTM, WTN, KK, OCC, WM, WN, NS>(ctx, elemA, p, bt, m, n, k, a, A, lda, B,
										   ldb, b, C, ldc, Apre, io);
	return gemm_i8_batched_ea<MLZ_F64_SYM, WTM, WTN, KK, OCC, WM, WN, NS>(ctx, elemA, p, bt, m, n, k, a, A, lda, B,
									       ldb, b, C, ldc, Apre, io);
}

// Batches use the 128 x 128 tile through four LDS buffers (variant 10), or
// the 64 x 64 one (variant 11) when the whole batch has fewer than 256
// tiles of 128 x 128: a batch is many small products, not one large one.
// elemA: storage of A.
int zp_gemm_i8_batched_dev(mlz_ctx *ctx, int elem, int elemA, uint32_t p, const I8Batch &bt, size_t m, size_t n,
			   size_t k, uint32_t a, const void *A, size_t lda, const void *B, size_t ldb, uint32_t b,
			   void *C, size_t ldc, const int8_t *Apre, const I8PackIO &io)
{
	if (elem_size(elem) == 0 || elem_size(elemA) == 0)
		return set_error(ctx, MLZ_E_ARG, "batched gemm: unknown element storage");
	if (p > 65536)
		return set_error(ctx, MLZ_E_UNSUPPORTED, "INT8 limb gemm needs p <= 2^16");
	if (m == 0 || n == 0 || bt.count == 0)
		return MLZ_OK;
	if (k == 0 || m > (size_t) INT32_MAX || n > (size_t) INT32_MAX)
		return set_error(ctx, MLZ_E_ARG, "batched gemm needs 0 < k and m, n < 2^31");
	a %= p;
	b %= p;
	if (Apre && bt.count > 1 && m % 128)
		return set_error(ctx, MLZ_E_ARG, "a batch of pre-packed A operands needs rows in multiples of 128");
	// the 64 x 64 tile unless the batch has 256 or more 128 x 128 tiles
	if ((size_t) bt.count * ceil_div(m, 128) * ceil_div(n, 128) < 256)
		return gemm_i8_batched_e<1, 1, 1, 4, 2, 2, 4>(ctx, elem, elemA, p, bt, m, n, k, a, A, lda, B, ldb, b, C, ldc,
							      Apre, io);
	return gemm_i8_batched_e<2, 2, 1, 2, 2, 2, 4>(ctx, elem, elemA, p, bt, m, n, k, a, A, lda, B, ldb, b, C, ldc, Apre,
						      io);
}

// Packs A (rows x k per item, the A operand layout of the INT8 kernels) for a
// batch of bt.count items into dst, item z at z * (bytes per item); returns
// the bytes per item through *item_bytes.  For A operands used by several
// products (zp_gemm_i8_batched_dev's Apre).
int zp_pack_i8_a_dev(mlz_ctx *ctx, int elem, uint32_t p, const I8Batch &bt, size_t rows, size_t k, const void *A,
		     size_t lda, int8_t *dst, size_t *item_bytes, hipStream_t stream)
{
	const size_t mpad = round_up(rows, 128), kpad = round_up(k, 64);
	const size_t bytes = mpad * kpad * 2;
	if (item_bytes)
		*item_bytes = bytes;
	if (!dst)
		return MLZ_OK;
	if (k > (size_t) I8_KMAX || p > 65536)
		return set_error(ctx, MLZ_E_ARG, "INT8 packing needs k <= %d and p <= 2^16", I8_KMAX);
	if (!stream)
		stream = ctx->stream;
	const int KB = (int) (kpad / 32);
	dim3 g(mpad / 64, kpad / 64, bt.count);
	const size_t es = elem_size(elem);
	// 16-byte loads when every row of every item starts aligned
	const bool vec = aligned16(A, lda * es) && (bt.count == 1 || ((bt.outer[0] | bt.in[0]) * (long long) es) % 16 == 0);
	ProfScope ps(ctx, "pack_i8_A");
#define PACK(EE, TT, V) pack_i8_kernel<EE, false, V><<<g, 256, 0, stream>>>((const TT *) A, lda, (int) rows, \
									   (int) k, p, dst, KB, bt, bytes)
	if (elem == MLZ_U16) {
		if (vec) PACK(MLZ_U16, uint16_t, true); else PACK(MLZ_U16, uint16_t, false);
	} else if (elem == MLZ_U32) {
		if (vec) PACK(MLZ_U32, uint32_t, true); else PACK(MLZ_U32, uint32_t, false);
	} else {
		if (vec) PACK(MLZ_F64_SYM, double, true); else PACK(MLZ_F64_SYM, double, false);
	}
#undef PACK
	return check_launch(ctx, "pack_i8_kernel(A, batch)");
}

// Packs B (k x cols, the B operand layout of the INT8 kernels: 32-column
// blocks of round_up(k, 64) * 64 bytes) into dst.
int zp_pack_i8_b_dev(mlz_ctx *ctx, int elem, uint32_t p, size_t k, size_t cols, const void *B, size_t ldb,
		     int8_t *dst, hipStream_t stream)
{
	if (k > (size_t) I8_KMAX || p > 65536)
		return set_error(ctx, MLZ_E_ARG, "INT8 packing needs k <= %d and p <= 2^16", I8_KMAX);
	if (!stream)
		stream = ctx->stream;
	const size_t npad = round_up(cols, 64), kpad = round_up(k, 64);
	const int KB = (int) (kpad / 32);
	dim3 g(npad / 64, kpad / 64, 1);
	const bool vec = aligned16(B, ldb * elem_size(elem));
	ProfScope ps(ctx, "pack_i8_B");
#define PACK(EE, TT, V) pack_i8_kernel<EE, true, V><<<g, 256, 0, stream>>>((const TT *) B, ldb, (int) cols, (int) k, \
									  p, dst, KB, I8Batch::single(), 0)
	if (elem == MLZ_U16) {
		if (vec) PACK(MLZ_U16, uint16_t, true); else PACK(MLZ_U16, uint16_t, false);
	} else if (elem == MLZ_U32) {
		if (vec) PACK(MLZ_U32, uint32_t, true); else PACK(MLZ_U32, uint32_t, false);
	} else {
		if (vec) PACK(MLZ_F64_SYM, double, true); else PACK(MLZ_F64_SYM, double, false);
	}
#undef PACK
	return check_launch(ctx, "pack_i8_kernel(B)");
}

// ===========================================================================
// Rank-k update C <- C - G*W for small k (the trailing update of a panel of
// the echelon form): the output is read and written once per call while k
// is tiny, so the tiled kernel's LDS staging and barriers are overhead.
// Each wave takes one 32-row block and a strip of NBLK 32-column blocks;
// fragments come straight from the packed operands (16 bytes per lane, no
// LDS), the limb products go through the same three accumulators, and the
// epilogue subtracts in place.  No barriers: a wave past the last row block
// just returns.
// ===========================================================================
template <int UPD_NBLK, typename TC>
__global__ __launch_bounds__(256) void rank_update_i8_kernel(const int8_t *__restrict__ Gp,
							     const int8_t *__restrict__ Wp, int KB,
							     TC *__restrict__ C, size_t ldc, int m, int n,
							     ModP mp, int strips_n)
{
	const int lane = threadIdx.x & 63;
	const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int rb = wid / strips_n, sn = wid - rb * strips_n;
	if (rb * 32 >= m)
		return;
	const size_t panel = (size_t) KB * 2 * I8_FRAG;
	const int8_t *ga = Gp + (size_t) rb * panel + lane * 16;
	const int cb0 = sn * UPD_NBLK;
	const int nblk = min(UPD_NBLK, (n + 31) / 32 - cb0);
	// the strip's outputs are loaded first so that their latency overlaps the MFMAs
	uint32_t cv[UPD_NBLK][16];
#pragma unroll
	for (int j = 0; j < UPD_NBLK; ++j) {
		const int col = (cb0 + j) * 32 + (lane & 31);
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			const int row = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
			cv[j][r] = (j < nblk && row < m && col < n) ? (uint32_t) C[(size_t) row * ldc + col] : 0u;
		}
	}
	v16i acc[3][UPD_NBLK];
#pragma unroll
	for (int t = 0; t < 3; ++t)
#pragma unroll
		for (int j = 0; j < UPD_NBLK; ++j)
			acc[t][j] = v16i{0};
	for (int kb = 0; kb < KB; ++kb) {
		const v4i ah = *(const v4i *) (ga + (size_t) kb * 2 * I8_FRAG);
		const v4i al = *(const v4i *) (ga + (size_t) kb * 2 * I8_FRAG + I8_FRAG);
#pragma unroll
		for (int j = 0; j < UPD_NBLK; ++j) {
			if (j < nblk) {
				const int8_t *gb = Wp + (size_t) (cb0 + j) * panel + (size_t) kb * 2 * I8_FRAG + lane * 16;
				const v4i bh = *(const v4i *) gb, bl = *(const v4i *) (gb + I8_FRAG);
				acc[0][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, acc[0][j], 0, 0, 0);
				acc[1][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bl, acc[1][j], 0, 0, 0);
				acc[2][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bl, acc[2][j], 0, 0, 0);
				acc[1][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh, acc[1][j], 0, 0, 0);
			}
		}
	}
	const uint32_t p = (uint32_t) mp.p;
#pragma unroll
	for (int j = 0; j < UPD_NBLK; ++j) {
		const int col = (cb0 + j) * 32 + (lane & 31);
		if (j >= nblk || col >= n)
			continue;
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			const int row = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
			if (row < m) {
				const double v = fma((double) acc[0][j][r], 65536.0,
						     fma((double) acc[1][j][r], 256.0, (double) acc[2][j][r]));
				const uint32_t d = mp.mod_f64(v);
				const uint32_t c = cv[j][r];
				C[(size_t) row * ldc + col] = (TC) (c >= d ? c - d : c + p - d);
			}
		}
	}
}

// The same update for full 32 x 32 tiles of a 16-byte aligned C: each wave
// takes TPW tiles side by side (one 32-row strip), loads them straight into
// LDS first (global_load_lds_dwordx4: a 16-byte piece per lane, NQ loads of
// RPL rows per tile; no registers hold them, so the loads of all TPW tiles are
// in flight while the MFMAs run), then per tile reads C in the accumulator
// layout, writes the results back to LDS and stores them as 16-byte pieces.
// The G fragments are loaded once per strip.  A wave touches only its own LDS
// tiles and its LDS operations complete in order: no barriers.  C holds
// uint32 or (p <= 2^16) uint16 entries.
template <int TPW, typename TC>
__global__ __launch_bounds__(256) void rank_update_i8_lds_kernel(const int8_t *__restrict__ Gp,
								 const int8_t *__restrict__ Wp, int KB,
								 TC *__restrict__ C, size_t ldc, int m, int sns,
								 ModP mp)
{
	constexpr int EPL = 16 / (int) sizeof(TC);   // entries per 16-byte piece
	constexpr int LPR = 32 / EPL;                // lanes per tile row
	constexpr int RPL = 64 / LPR;                // tile rows per load
	constexpr int NQ = 32 / RPL;                 // loads per tile
	__shared__ __attribute__((aligned(16))) TC tile[4 * TPW * 32 * 32];
	const int lane = threadIdx.x & 63;
	const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int wid = blockIdx.x * 4 + wv;
	const int rb = wid / sns, sn = wid - rb * sns;
	if (rb * 32 >= m)
		return;
	TC *T = tile + wv * (TPW * 1024);
	TC *Cw = C + (size_t) (rb * 32) * ldc + sn * (32 * TPW);
	const int rows = min(32, m - rb * 32);     // a last partial row block: rows past m are
						   // loaded from the last row and not stored
	// tile j, load q: rows RPL q .. RPL q + RPL-1; lane l: row RPL q + l/LPR,
	// columns 32j + EPL(l%LPR) .. +EPL-1, landing at T + 1024j + 32 RPL q + EPL l
	// (row-major tiles)
#pragma unroll
	for (int j = 0; j < TPW; ++j)
#pragma unroll
		for (int q = 0; q < NQ; ++q)
			__builtin_amdgcn_global_load_lds((const MLZ_AS_GLOBAL void *) (Cw + (size_t) min(RPL * q + lane / LPR, rows - 1) *
										  ldc + 32 * j + EPL * (lane % LPR)),
							 (MLZ_AS_LDS void *) (T + 1024 * j + 32 * RPL * q), 16, 0, 0);
	const size_t panel = (size_t) KB * 2 * I8_FRAG;
	const int8_t *ga = Gp + (size_t) rb * panel + lane * 16;
	const uint32_t p = (uint32_t) mp.p;
	const int col = lane & 31;
	for (int j = 0; j < TPW; ++j) {
		const int8_t *gb = Wp + (size_t) (sn * TPW + j) * panel + lane * 16;
		v16i acc0 = v16i{0}, acc1 = v16i{0}, acc2 = v16i{0};
		for (int kb = 0; kb < KB; ++kb) {
			const v4i ah = *(const v4i *) (ga + (size_t) kb * 2 * I8_FRAG);
			const v4i al = *(const v4i *) (ga + (size_t) kb * 2 * I8_FRAG + I8_FRAG);
			const v4i bh = *(const v4i *) (gb + (size_t) kb * 2 * I8_FRAG);
			const v4i bl = *(const v4i *) (gb + (size_t) kb * 2 * I8_FRAG + I8_FRAG);
			acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, acc0, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bl, acc1, 0, 0, 0);
			acc2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bl, acc2, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh, acc1, 0, 0, 0);
		}
		// all of this wave's loads are done, the tiles among them
		wait_vmcnt<0>();
		asm volatile("" ::: "memory");
		TC *Tj = T + 1024 * j;
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
			const double v = fma((double) acc0[r], 65536.0, fma((double) acc1[r], 256.0, (double) acc2[r]));
			const uint32_t d = mp.mod_f64(v);
			const uint32_t c = Tj[row * 32 + col];
			Tj[row * 32 + col] = (TC) (c >= d ? c - d : c + p - d);
		}
#pragma unroll
		for (int q = 0; q < NQ; ++q)
			if (RPL * q + lane / LPR < rows)
				*(uint4 *) (Cw + (size_t) (RPL * q + lane / LPR) * ldc + 32 * j + EPL * (lane % LPR)) =
					*(const uint4 *) (Tj + 32 * RPL * q + EPL * lane);
	}
}

// C (m x n, canonical uint32, leading dimension ldc) <- C - G*W mod p with G
// m x k (ldg) and W k x n (ldw), p <= 2^16, k <= I8_KMAX.
int zp_rank_update_i8_dev(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, size_t k, const uint32_t *G, size_t ldg,
			  const uint32_t *W, size_t ldw, uint32_t *C, size_t ldc)
{
	if (m == 0 || n == 0 || k == 0)
		return MLZ_OK;
	if (p > 65536 || k > (size_t) I8_KMAX)
		return set_error(ctx, MLZ_E_UNSUPPORTED, "rank update needs p <= 2^16 and k <= %d", I8_KMAX);
	const size_t mpad = round_up(m, 64), npad = round_up(n, 64), kpad = round_up(k, 64);
	const int KB = (int) (kpad / 32);
	void *gp = nullptr, *wp = nullptr;
	MLZ_TRY(ws_get(ctx, WS_PACK_A, mpad * kpad * 2, &gp));
	MLZ_TRY(ws_get(ctx, WS_PACK_B, npad * kpad * 2, &wp));
	{
		ProfScope ps(ctx, "pack_i8_A");
		dim3 g(mpad / 64, kpad / 64);
		if (aligned16(G, ldg * 4))
			pack_i8_kernel<MLZ_U32, false, true><<<g, 256, 0, ctx->stream>>>(G, ldg, (int) m, (int) k, p,
											 (int8_t *) gp, KB, I8Batch::single(), 0);
		else
			pack_i8_kernel<MLZ_U32, false, false><<<g, 256, 0, ctx->stream>>>(G, ldg, (int) m, (int) k, p,
											  (int8_t *) gp, KB, I8Batch::single(), 0);
	}
	{
		ProfScope ps(ctx, "pack_i8_B");
		dim3 g(npad / 64, kpad / 64);
		if (aligned16(W, ldw * 4))
			pack_i8_kernel<MLZ_U32, true, true><<<g, 256, 0, ctx->stream>>>(W, ldw, (int) n, (int) k, p,
											(int8_t *) wp, KB, I8Batch::single(), 0);
		else
			pack_i8_kernel<MLZ_U32, true, false><<<g, 256, 0, ctx->stream>>>(W, ldw, (int) n, (int) k, p,
											 (int8_t *) wp, KB, I8Batch::single(), 0);
	}
	MLZ_TRY(check_launch(ctx, "pack_i8_kernel"));
	return zp_rank_update_i8_packed(ctx, p, m, n, KB, (const int8_t *) gp, (const int8_t *) wp, C, ldc);
}

template <typename TC>
static int rank_update_i8_reg(mlz_ctx *ctx, const ModP &mp, size_t m, size_t n, int KB, const int8_t *gp,
			      const int8_t *wp, TC *C, size_t ldc, hipStream_t stream);

template <typename TC>
static int rank_update_i8_packed(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, int KB, const int8_t *gp,
				 const int8_t *wp, TC *C, size_t ldc, hipStream_t stream)
{
	if (!stream)
		stream = ctx->stream;
	if (m == 0 || n == 0)
		return MLZ_OK;
	const ModP mp = make_modp(p);
	ProfScope ps(ctx, stream == ctx->stream ? "rank_update_i8" : nullptr);
	// full tiles of an aligned C on the LDS-staged kernel, the edge strips below
	// tiles per wave: two when the rows start on 128-byte boundaries, else one
	// (measured); the context option fixes it
	const int tpw = ctx->upd_lds >= 0 ? ctx->upd_lds : ((ldc * sizeof(TC)) % 128 == 0 ? 2 : 1);
	const size_t rbs = ceil_div(m, 32), sns = tpw > 0 ? n / (32 * (size_t) tpw) : 0, cbs = sns * (size_t) tpw;
	if (tpw > 0 && sns && (ldc * sizeof(TC)) % 16 == 0 && ((uintptr_t) C & 15) == 0) {
		const unsigned grid = ceil_div(rbs * sns, 4);
		if (tpw == 1)
			rank_update_i8_lds_kernel<1, TC><<<grid, 256, 0, stream>>>(gp, wp, KB, C, ldc, (int) m, (int) sns, mp);
		else if (tpw == 2)
			rank_update_i8_lds_kernel<2, TC><<<grid, 256, 0, stream>>>(gp, wp, KB, C, ldc, (int) m, (int) sns, mp);
		else
			rank_update_i8_lds_kernel<4, TC><<<grid, 256, 0, stream>>>(gp, wp, KB, C, ldc, (int) m, (int) sns, mp);
		MLZ_TRY(check_launch(ctx, "rank_update_i8_lds_kernel"));
		const size_t panel = (size_t) KB * 2 * I8_FRAG;
		if (n > cbs * 32)        // the last columns
			MLZ_TRY(rank_update_i8_reg(ctx, mp, m, n - cbs * 32, KB, gp, wp + cbs * panel, C + cbs * 32, ldc, stream));
		return MLZ_OK;
	}
	return rank_update_i8_reg(ctx, mp, m, n, KB, gp, wp, C, ldc, stream);
}

int zp_rank_update_i8_packed(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, int KB, const int8_t *gp,
			     const int8_t *wp, uint32_t *C, size_t ldc, hipStream_t stream)
{
	return rank_update_i8_packed(ctx, p, m, n, KB, gp, wp, C, ldc, stream);
}

int zp_rank_update_i8_packed(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, int KB, const int8_t *gp,
			     const int8_t *wp, uint16_t *C, size_t ldc, hipStream_t stream)
{
	return rank_update_i8_packed(ctx, p, m, n, KB, gp, wp, C, ldc, stream);
}

// The register-staged kernel over all of C.
template <typename TC>
static int rank_update_i8_reg(mlz_ctx *ctx, const ModP &mp, size_t m, size_t n, int KB, const int8_t *gp,
			      const int8_t *wp, TC *C, size_t ldc, hipStream_t stream)
{
	const size_t mpad = round_up(m, 64);
	const int nblk = ctx->upd_nblk == 2 || ctx->upd_nblk == 4 ? ctx->upd_nblk : 1;
	const int strips_n = (int) ceil_div(n, 32 * nblk);
	const size_t waves = (mpad / 32) * (size_t) strips_n;
	const unsigned grid = ceil_div(waves, 4);
	if (nblk == 1)
		rank_update_i8_kernel<1, TC><<<grid, 256, 0, stream>>>((const int8_t *) gp, (const int8_t *) wp, KB, C, ldc,
								       (int) m, (int) n, mp, strips_n);
	else if (nblk == 4)
		rank_update_i8_kernel<4, TC><<<grid, 256, 0, stream>>>((const int8_t *) gp, (const int8_t *) wp, KB, C, ldc,
								       (int) m, (int) n, mp, strips_n);
	else
		rank_update_i8_kernel<2, TC><<<grid, 256, 0, stream>>>((const int8_t *) gp, (const int8_t *) wp, KB, C, ldc,
								       (int) m, (int) n, mp, strips_n);
	return check_launch(ctx, "rank_update_i8_kernel");
}

// ===========================================================================
// PLAIN path
// ===========================================================================

// The output tile at (row0, col0) of the plain kernel; As and Bs are 16 x 64
// words of LDS each.  Shared by the single-matrix and the batched kernels.
template <int E, bool WIDE>
__device__ __forceinline__ void plain_tile(const typename Elem<E>::T *__restrict__ A, size_t lda,
					   const typename Elem<E>::T *__restrict__ B, size_t ldb,
					   typename Elem<E>::T *__restrict__ C, size_t ldc, int m, int n, int k,
					   uint32_t a, uint32_t b, const ModP &mp, uint64_t two_64, int row0, int col0,
					   uint32_t (*As)[64], uint32_t (*Bs)[64])
{
	const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
	const uint32_t p = (uint32_t) mp.p;
	uint64_t acc[4][4];
#pragma unroll
	for (int r = 0; r < 4; ++r)
#pragma unroll
		for (int c = 0; c < 4; ++c)
			acc[r][c] = 0;
	for (int k0 = 0; k0 < k; k0 += 16) {
		// A tile 64 x 16: thread loads 4 entries, k fastest for coalescing.
#pragma unroll
		for (int e = 0; e < 4; ++e) {
			int idx = e * 256 + threadIdx.x;
			int r = idx >> 4, kk = idx & 15;
			int gr = row0 + r, gk = k0 + kk;
			As[kk][r] = (gr < m && gk < k) ? Elem<E>::load(A + (size_t) gr * lda + gk, p) : 0u;
		}
#pragma unroll
		for (int e = 0; e < 4; ++e) {
			int idx = e * 256 + threadIdx.x;
			int kk = idx >> 6, c = idx & 63;
			int gk = k0 + kk, gc = col0 + c;
			Bs[kk][c] = (gk < k && gc < n) ? Elem<E>::load(B + (size_t) gk * ldb + gc, p) : 0u;
		}
		__syncthreads();
#pragma unroll
		for (int kk = 0; kk < 16; ++kk) {
			uint4 av = *(const uint4 *) &As[kk][ty * 4];
			uint4 bv = *(const uint4 *) &Bs[kk][tx * 4];
			uint32_t ar[4] = {av.x, av.y, av.z, av.w};
			uint32_t br[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
			for (int r = 0; r < 4; ++r)
#pragma unroll
				for (int c = 0; c < 4; ++c) {
					uint64_t t = (uint64_t) ar[r] * br[c];
					if (WIDE) {
						acc[r][c] += t;
						if (acc[r][c] < t)
							acc[r][c] += two_64;
					} else {
						acc[r][c] += t;
					}
				}
		}
		__syncthreads();
	}
#pragma unroll
	for (int r = 0; r < 4; ++r) {
		const int row = row0 + ty * 4 + r;
		if (row >= m)
			continue;
#pragma unroll
		for (int c = 0; c < 4; ++c) {
			const int col = col0 + tx * 4 + c;
			if (col >= n)
				continue;
			uint32_t d = (uint32_t) mp.reduce(acc[r][c]);
			typename Elem<E>::T *cp = C + (size_t) row * ldc + col;
			uint32_t cij = b ? mp.mul(Elem<E>::load(cp, p), b) : 0u;
			Elem<E>::store(cp, mp.axpy(a, d, cij), p);
		}
	}
}

// 64 x 64 output tile, 16 x 16 threads, 4 x 4 entries per thread, k staged
// 16 at a time through LDS as canonical uint32.
template <int E, bool WIDE>
__global__ __launch_bounds__(256) void gemm_plain_kernel(const typename Elem<E>::T *__restrict__ A, size_t lda,
							 const typename Elem<E>::T *__restrict__ B, size_t ldb,
							 typename Elem<E>::T *__restrict__ C, size_t ldc,
							 int m, int n, int k, uint32_t a, uint32_t b, ModP mp,
							 uint64_t two_64)
{
	__shared__ __attribute__((aligned(16))) uint32_t As[16][64];
	__shared__ __attribute__((aligned(16))) uint32_t Bs[16][64];
	plain_tile<E, WIDE>(A, lda, B, ldb, C, ldc, m, n, k, a, b, mp, two_64, blockIdx.y * 64, blockIdx.x * 64, As,
			    Bs);
}

template <int E>
static int gemm_plain_launch(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, size_t k, uint32_t a,
			     const void *A, size_t lda, const void *B, size_t ldb, uint32_t b, void *C,
			     size_t ldc)
{
	typedef typename Elem<E>::T T;
	const ModP mp = make_modp(p);
	uint64_t t = 2;
	for (int i = 0; i < 6; ++i)
		t = (t * t) % p;  // 2^64 mod p (lela/ring/modular.h:616-625)
	const unsigned __int128 bound = (unsigned __int128) (p - 1) * (p - 1) * (unsigned __int128) k;
	const bool wide = bound >= ((unsigned __int128) 1 << 64);
	dim3 g(ceil_div(n, 64), ceil_div(m, 64));
	ProfScope ps(ctx, "gemm_plain");
	if (wide)
		gemm_plain_kernel<E, true><<<g, 256, 0, ctx->stream>>>((const T *) A, lda, (const T *) B, ldb,
								       (T *) C, ldc, (int) m, (int) n, (int) k,
								       a, b, mp, t);
	else
		gemm_plain_kernel<E, false><<<g, 256, 0, ctx->stream>>>((const T *) A, lda, (const T *) B, ldb,
									(T *) C, ldc, (int) m, (int) n, (int) k,
									a, b, mp, t);
	return check_launch(ctx, "gemm_plain_kernel");
}

// ===========================================================================
// FP64 path
// ===========================================================================

// v_mfma_f64_16x16x4_f64 fragments: A fragment 16 rows x 4 k, lane l holds
// row l%16, k l/16; B fragment 4 k x 16 columns, lane l holds column l%16,
// k l/16.  The 16 x 16 result: lane l, register r holds row l/16 + 4*r,
// column l%16 (measured on gfx950).  Packed A: [rblock16][kblock4][64]
// doubles (symmetric residues); packed B: [cblock16][kblock4][limb(hi, lo)][64].
constexpr int F64_LIMB_BITS = 13;

__device__ inline int32_t sym_of(uint32_t v, uint32_t p)
{
	return (2ull * v > (unsigned long long) p) ? (int32_t) ((int64_t) v - p) : (int32_t) v;
}

template <int E>
__global__ __launch_bounds__(256) void pack_f64_a_kernel(const typename Elem<E>::T *__restrict__ src, size_t ld,
							 int rows, int kdim, uint32_t p, double *__restrict__ dst,
							 int KB4, int total)
{
	int t = blockIdx.x * 256 + threadIdx.x;
	if (t >= total)
		return;
	// t enumerates (rb, kb, lane) in destination order.
	int lane = t & 63, rest = t >> 6;
	int kb = rest % KB4, rb = rest / KB4;
	int row = rb * 16 + (lane & 15), kk = kb * 4 + (lane >> 4);
	double v = 0.0;
	if (row < rows && kk < kdim)
		v = (double) sym_of(Elem<E>::load(src + (size_t) row * ld + kk, p), p);
	dst[t] = v;
}

template <int E>
__global__ __launch_bounds__(256) void pack_f64_b_kernel(const typename Elem<E>::T *__restrict__ src, size_t ld,
							 int cols, int kdim, uint32_t p, double *__restrict__ dst,
							 int KB4, int total)
{
	int t = blockIdx.x * 256 + threadIdx.x;
	if (t >= total)
		return;
	int lane = t & 63, rest = t >> 6;
	int kb = rest % KB4, cb = rest / KB4;
	int col = cb * 16 + (lane & 15), kk = kb * 4 + (lane >> 4);
	int32_t s = 0;
	if (col < cols && kk < kdim)
		s = sym_of(Elem<E>::load(src + (size_t) kk * ld + col, p), p);
	// Centred split s = hi*2^13 + lo, lo in [-4096, 4095].
	int32_t lo = ((s & 8191) ^ 4096) - 4096;
	int32_t hi = (s - lo) >> F64_LIMB_BITS;
	size_t base = ((size_t) cb * KB4 + kb) * 128 + lane;
	dst[base] = (double) hi;
	dst[base + 64] = (double) lo;
}

// Workgroup: 4 waves, each computing a 32 x 64 tile (2 x 4 blocks of 16 x 16,
// two limbs).  Waves read fragments straight from the packed arrays (L1/L2
// serve the reuse; the FP64 MFMA rate, not bandwidth, bounds this kernel).
template <int E>
__global__ __launch_bounds__(256) void gemm_f64_kernel(const double *__restrict__ Ap, const double *__restrict__ Bp,
						       int KB4, typename Elem<E>::T *__restrict__ C, size_t ldc,
						       int m, int n, uint32_t a, uint32_t b, ModP mp, int tiles_m,
						       int tiles_n)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	int tm, tn;
	tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
	// Workgroup tile 64 x 128: waves 2 x 2 of 32 x 64.
	const int wr = wave >> 1, wc = wave & 1;
	const int rb0 = tm * 4 + wr * 2;      // 16-row blocks
	const int cb0 = tn * 8 + wc * 4;      // 16-column blocks
	const double *Ag = Ap + (size_t) rb0 * KB4 * 64 + lane;
	const double *Bg = Bp + (size_t) cb0 * KB4 * 128 + lane;
	v4d acc[2][4][2];
#pragma unroll
	for (int i = 0; i < 2; ++i)
#pragma unroll
		for (int j = 0; j < 4; ++j)
#pragma unroll
			for (int l = 0; l < 2; ++l)
				acc[i][j][l] = v4d{0.0, 0.0, 0.0, 0.0};
	for (int kb = 0; kb < KB4; ++kb) {
		double fa[2], fb[4][2];
#pragma unroll
		for (int i = 0; i < 2; ++i)
			fa[i] = Ag[((size_t) i * KB4 + kb) * 64];
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			fb[j][0] = Bg[((size_t) j * KB4 + kb) * 128];
			fb[j][1] = Bg[((size_t) j * KB4 + kb) * 128 + 64];
		}
#pragma unroll
		for (int i = 0; i < 2; ++i)
#pragma unroll
			for (int j = 0; j < 4; ++j)
#pragma unroll
				for (int l = 0; l < 2; ++l)
					acc[i][j][l] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j][l], acc[i][j][l], 0, 0, 0);
	}
	const uint32_t p = (uint32_t) mp.p;
	// |partial sums| < 2^53; shift by a multiple of p before reducing.
	const uint64_t off = ((((uint64_t) 1 << 53) + p - 1) / p) * p;
#pragma unroll
	for (int i = 0; i < 2; ++i)
#pragma unroll
		for (int j = 0; j < 4; ++j) {
			const int col = (cb0 + j) * 16 + (lane & 15);
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const int row = (rb0 + i) * 16 + (lane >> 4) + 4 * r;
				if (row < m && col < n) {
					int64_t hi = (int64_t) acc[i][j][0][r], lo = (int64_t) acc[i][j][1][r];
					uint64_t h = mp.reduce((uint64_t) (hi + (int64_t) off));
					uint64_t l = mp.reduce((uint64_t) (lo + (int64_t) off));
					uint32_t d = (uint32_t) mp.reduce((h << F64_LIMB_BITS) + l);
					typename Elem<E>::T *cp = C + (size_t) row * ldc + col;
					uint32_t cij = b ? mp.mul(Elem<E>::load(cp, p), b) : 0u;
					Elem<E>::store(cp, mp.axpy(a, d, cij), p);
				}
			}
		}
}

template <int E>
static int gemm_f64_launch(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, size_t k, uint32_t a,
			   const void *A, size_t lda, const void *B, size_t ldb, uint32_t b, void *C,
			   size_t ldc)
{
	typedef typename Elem<E>::T T;
	const ModP mp = make_modp(p);
	// Exactness: |a| <= (p-1)/2, |limb| <= 4097, so each product is below
	// amax*4097 and kchunk products sum below 2^53.
	const uint64_t amax = (p - 1) / 2 + 1;
	size_t kchunk = (size_t) ((((uint64_t) 1 << 53) - 1) / (amax * 4097ull));
	kchunk = kchunk / 4 * 4;
	if (kchunk < 4)
		return set_error(ctx, MLZ_E_UNSUPPORTED, "FP64 carrier needs p < 2^26 (p=%u)", p);
	const size_t mpad = round_up(m, 64), npad = round_up(n, 128);
	for (size_t k0 = 0; k0 < k; k0 += kchunk) {
		const size_t kc = std::min(kchunk, k - k0);
		const int KB4 = (int) ceil_div(kc, 4);
		void *ap = nullptr, *bp = nullptr;
		MLZ_TRY(pack_buffers(ctx, mpad * KB4 * 4 * sizeof(double), npad * KB4 * 4 * 2 * sizeof(double),
				     &ap, &bp));
		{
			ProfScope ps(ctx, "pack_f64_A");
			int total = (int) (mpad / 16 * KB4 * 64);
			pack_f64_a_kernel<E><<<ceil_div(total, 256), 256, 0, ctx->stream>>>(
				(const T *) A + k0, lda, (int) m, (int) kc, p, (double *) ap, KB4, total);
		}
		MLZ_TRY(check_launch(ctx, "pack_f64_a_kernel"));
		{
			ProfScope ps(ctx, "pack_f64_B");
			int total = (int) (npad / 16 * KB4 * 64);
			pack_f64_b_kernel<E><<<ceil_div(total, 256), 256, 0, ctx->stream>>>(
				(const T *) B + k0 * ldb, ldb, (int) n, (int) kc, p, (double *) bp, KB4, total);
		}
		MLZ_TRY(check_launch(ctx, "pack_f64_b_kernel"));
		const int tiles_m = (int) (mpad / 64), tiles_n = (int) (npad / 128);
		{
			ProfScope ps(ctx, "gemm_f64_mfma");
			gemm_f64_kernel<E><<<tiles_m * tiles_n, 256, 0, ctx->stream>>>(
				(const double *) ap, (const double *) bp, KB4, (T *) C, ldc, (int) m, (int) n, a,
				k0 == 0 ? b : 1u, mp, tiles_m, tiles_n);
		}
		MLZ_TRY(check_launch(ctx, "gemm_f64_kernel"));
	}
	return MLZ_OK;
}

// ===========================================================================
// scal and dispatch
// ===========================================================================

template <int E>
__global__ void scal_kernel(typename Elem<E>::T *C, size_t ldc, int m, int n, uint32_t a, ModP mp)
{
	size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= (size_t) m * n)
		return;
	int r = (int) (idx / n), c = (int) (idx % n);
	typename Elem<E>::T *cp = C + (size_t) r * ldc + c;
	Elem<E>::store(cp, a ? mp.mul(Elem<E>::load(cp, (uint32_t) mp.p), a) : 0u, (uint32_t) mp.p);
}

int zp_scal_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, uint64_t a, void *C, size_t ldc)
{
	if (m == 0 || n == 0)
		return MLZ_OK;
	const ModP mp = make_modp(p);
	unsigned g = ceil_div(m * n, 256);
	if (elem == MLZ_U16)
		scal_kernel<MLZ_U16><<<g, 256, 0, ctx->stream>>>((uint16_t *) C, ldc, (int) m, (int) n, (uint32_t) a, mp);
	else if (elem == MLZ_U32)
		scal_kernel<MLZ_U32><<<g, 256, 0, ctx->stream>>>((uint32_t *) C, ldc, (int) m, (int) n, (uint32_t) a, mp);
	else
		scal_kernel<MLZ_F64_SYM><<<g, 256, 0, ctx->stream>>>((double *) C, ldc, (int) m, (int) n, (uint32_t) a, mp);
	return check_launch(ctx, "scal_kernel");
}

int pick_gemm_algo(uint64_t p, int elem, size_t m, size_t n, size_t k, int algo)
{
	(void) elem;
	if (algo != MLZ_ALGO_AUTO)
		return algo;
	const size_t work = m * n * k;
	if (p <= 65536 && work >= ((size_t) 1 << 18))
		return MLZ_ALGO_INT8;
	return MLZ_ALGO_PLAIN;
}

template <int E>
static int gemm_dispatch(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, size_t k, uint32_t a,
			 const void *A, size_t lda, const void *B, size_t ldb, uint32_t b, void *C,
			 size_t ldc, int algo)
{
	switch (algo) {
	case MLZ_ALGO_INT8:
		return gemm_i8_launch<E>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case MLZ_ALGO_FP64:
		return gemm_f64_launch<E>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	case MLZ_ALGO_PLAIN:
		return gemm_plain_launch<E>(ctx, p, m, n, k, a, A, lda, B, ldb, b, C, ldc);
	default:
		return set_error(ctx, MLZ_E_ARG, "unknown gemm algo %d", algo);
	}
}

int check_modulus(mlz_ctx *ctx, int elem, uint64_t p)
{
	if (p < 2)
		return set_error(ctx, MLZ_E_ARG, "modulus %llu < 2", (unsigned long long) p);
	if (elem == MLZ_U16 && p > 65535)
		return set_error(ctx, MLZ_E_ARG, "Modular<uint16> needs p < 2^16 (ModularTraits<uint16>::valid_modulus)");
	if (elem == MLZ_U32 && p > 0xffffffffull)
		return set_error(ctx, MLZ_E_ARG, "Modular<uint32> needs p < 2^32");
	if (elem == MLZ_F64_SYM && p >= 67108864ull)  // ModularTraits<double>::valid_modulus
		return set_error(ctx, MLZ_E_ARG, "Modular<double> needs p < 2^26");
	if (elem != MLZ_U16 && elem != MLZ_U32 && elem != MLZ_F64_SYM)
		return set_error(ctx, MLZ_E_ARG, "unknown element type %d", elem);
	return MLZ_OK;
}

int zp_gemm_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, size_t k, uint64_t a,
		const void *A, size_t lda, const void *B, size_t ldb, uint64_t b, void *C, size_t ldc,
		int algo)
{
	MLZ_TRY(check_modulus(ctx, elem, p));
	if (m == 0 || n == 0)
		return MLZ_OK;
	a %= p;
	b %= p;
	if (k == 0 || a == 0)
		return zp_scal_dev(ctx, elem, p, m, n, b, C, ldc);
	if (m > (size_t) INT32_MAX || n > (size_t) INT32_MAX || k > (size_t) INT32_MAX)
		return set_error(ctx, MLZ_E_ARG, "dimension too large");
	algo = pick_gemm_algo(p, elem, m, n, k, algo);
	if (algo == MLZ_ALGO_INT8 && p > 65536)
		return set_error(ctx, MLZ_E_UNSUPPORTED, "INT8 limb gemm needs p <= 2^16");
	if (algo == MLZ_ALGO_FP64 && p >= 67108864ull)
		return set_error(ctx, MLZ_E_UNSUPPORTED, "FP64 carrier gemm needs p < 2^26");
	const uint32_t pp = (uint32_t) p, aa = (uint32_t) a, bb = (uint32_t) b;
	if (elem == MLZ_U16)
		return gemm_dispatch<MLZ_U16>(ctx, pp, m, n, k, aa, A, lda, B, ldb, bb, C, ldc, algo);
	if (elem == MLZ_U32)
		return gemm_dispatch<MLZ_U32>(ctx, pp, m, n, k, aa, A, lda, B, ldb, bb, C, ldc, algo);
	return gemm_dispatch<MLZ_F64_SYM>(ctx, pp, m, n, k, aa, A, lda, B, ldb, bb, C, ldc, algo);
}

// ===========================================================================
// Batched kernels: persistent workgroups take work items (one output tile of
// one instance, or one 64 x 64 pack tile) from an atomic counter until the
// list is exhausted.  The lists are ordered largest instance first, so the
// tail of the launch holds the small tiles.  With L.tcost set, every item adds
// its workgroup's duration (s_memrealtime, 100 MHz) to its instance's total.
// ===========================================================================

// Items claimed per atomic by the pack and INT8 tile kernels (their items
// are small and numerous).
constexpr int BATCH_PACK_CHUNK = 8;
constexpr int BATCH_I8_CHUNK = 2;

// Pack items: x = 0 (A) / 1 (B) | 2 (16-byte aligned rows), y = 64-index block,
// z = 64-k block.
template <int E>
__global__ __launch_bounds__(256) void batch_pack_i8_kernel(const BatchGemm *__restrict__ tasks, BatchLaunch L,
							    uint32_t p)
{
	typedef typename Elem<E>::T T;
	__shared__ uint32_t tile[64][65];
	__shared__ int slot;
	for (;;) {
		const int c0 = batch_next(L, &slot, BATCH_PACK_CHUNK);
		if (c0 < 0)
			break;
		const int c1 = min(c0 + BATCH_PACK_CHUNK, L.nitems);
		for (int it = c0; it < c1; ++it) {
			const BatchItem w = L.items[it];
			const BatchGemm &t = tasks[w.task];
			const uint64_t t0 = wall_clock64();
			const bool isB = w.x & 1, vec = w.x & 2;
			if (!isB) {
				if (vec)
					i8_pack_tile<E, false, true>((const T *) t.A, t.lda, t.m, t.k, p, t.Ap, t.KB, w.y * 64,
								     w.z * 64, tile);
				else
					i8_pack_tile<E, false, false>((const T *) t.A, t.lda, t.m, t.k, p, t.Ap, t.KB, w.y * 64,
								      w.z * 64, tile);
			} else {
				if (vec)
					i8_pack_tile<E, true, true>((const T *) t.B, t.ldb, t.n, t.k, p, t.Bp, t.KB, w.y * 64,
								    w.z * 64, tile);
				else
					i8_pack_tile<E, true, false>((const T *) t.B, t.ldb, t.n, t.k, p, t.Bp, t.KB, w.y * 64,
								     w.z * 64, tile);
			}
			__syncthreads();
			batch_charge(L, w.task, t0);
		}
	}
}

// INT8 tile items: x = tile row, y = tile column.
template <int E, int WTM, int WTN, int KK, int OCC>
__global__ __launch_bounds__(256, OCC) void batch_i8_kernel(const BatchGemm *__restrict__ tasks, BatchLaunch L,
							    ModP mp)
{
	typedef typename Elem<E>::T T;
	__shared__ __attribute__((aligned(16))) int8_t lds[2 * I8Tile<WTM, WTN, KK>::STAGE];
	__shared__ int slot;
	for (;;) {
		const int c0 = batch_next(L, &slot, BATCH_I8_CHUNK);
		if (c0 < 0)
			break;
		const int c1 = min(c0 + BATCH_I8_CHUNK, L.nitems);
		for (int it = c0; it < c1; ++it) {
			const BatchItem w = L.items[it];
			const BatchGemm &t = tasks[w.task];
			const uint64_t t0 = wall_clock64();
			i8_tile<E, WTM, WTN, KK, 2, 2>(t.Ap, t.Bp, t.KB, (T *) t.C, t.ldc, t.m, t.n, t.a, t.b, mp, w.x, w.y,
						       lds);
			__syncthreads();
			batch_charge(L, w.task, t0);
		}
	}
}

// Plain tile items: x = 64-row block, y = 64-column block.
template <int E, bool WIDE>
__global__ __launch_bounds__(256) void batch_plain_kernel(const BatchGemm *__restrict__ tasks, BatchLaunch L, ModP mp,
							  uint64_t two_64)
{
	typedef typename Elem<E>::T T;
	__shared__ __attribute__((aligned(16))) uint32_t As[16][64];
	__shared__ __attribute__((aligned(16))) uint32_t Bs[16][64];
	__shared__ int slot;
	for (;;) {
		const int it = batch_next(L, &slot);
		if (it < 0)
			break;
		const BatchItem w = L.items[it];
		const BatchGemm &t = tasks[w.task];
		const uint64_t t0 = wall_clock64();
		plain_tile<E, WIDE>((const T *) t.A, t.lda, (const T *) t.B, t.ldb, (T *) t.C, t.ldc, t.m, t.n, t.k, t.a,
				    t.b, mp, two_64, w.x * 64, w.y * 64, As, Bs);
		__syncthreads();
		batch_charge(L, w.task, t0);
	}
}

// Sparse-A tile items: x = 16-row block, y = 256-column block.  Wave w takes
// rows 4w..4w+3 of the block; for each row it reads A 64 entries at a time,
// and for every nonzero a_il (found by ballot, the loop is wave-uniform) adds
// a_il * B[l][cols] to the lane's four columns (lane, lane+64, +128, +192).
// Work per row is proportional to its nonzeros, not to k.
template <int E, bool WIDE>
__global__ __launch_bounds__(256) void batch_sparse_kernel(const BatchGemm *__restrict__ tasks, BatchLaunch L,
							   ModP mp, uint64_t two_64)
{
	typedef typename Elem<E>::T T;
	__shared__ int slot;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint32_t p = (uint32_t) mp.p;
	for (;;) {
		const int it = batch_next(L, &slot);
		if (it < 0)
			break;
		const BatchItem w = L.items[it];
		const BatchGemm &t = tasks[w.task];
		const uint64_t t0 = wall_clock64();
		const T *A = (const T *) t.A, *B = (const T *) t.B;
		T *C = (T *) t.C;
		const int col0 = w.y * 256 + lane;
		for (int rr = 0; rr < 4; ++rr) {
			const int row = w.x * 16 + wave * 4 + rr;
			if (row >= t.m)
				break;
			uint64_t acc[4] = {0, 0, 0, 0};
			const T *arow = A + (size_t) row * t.lda;
			for (int kc = 0; kc < t.k; kc += 64) {
				const int l = kc + lane;
				const uint32_t x = l < t.k ? Elem<E>::load(arow + l, p) : 0u;
				uint64_t mask = __ballot(x != 0);
				while (mask) {
					const int s = __ffsll((unsigned long long) mask) - 1;
					mask &= mask - 1;
					const uint32_t xv = (uint32_t) __builtin_amdgcn_readlane((int) x, s);
					const T *brow = B + (size_t) (kc + s) * t.ldb;
#pragma unroll
					for (int q = 0; q < 4; ++q) {
						const int col = col0 + 64 * q;
						if (col < t.n) {
							const uint64_t pr = (uint64_t) xv * Elem<E>::load(brow + col, p);
							acc[q] += pr;
							if (WIDE && acc[q] < pr)
								acc[q] += two_64;
						}
					}
				}
			}
#pragma unroll
			for (int q = 0; q < 4; ++q) {
				const int col = col0 + 64 * q;
				if (col < t.n) {
					const uint32_t d = (uint32_t) mp.reduce(acc[q]);
					T *cp = C + (size_t) row * t.ldc + col;
					const uint32_t cij = t.b ? mp.mul(Elem<E>::load(cp, p), t.b) : 0u;
					Elem<E>::store(cp, mp.axpy(t.a, d, cij), p);
				}
			}
		}
		__syncthreads();
		batch_charge(L, w.task, t0);
	}
}

static uint64_t two64_mod(uint32_t p)
{
	uint64_t t = 2;
	for (int i = 0; i < 6; ++i)
		t = (t * t) % p;
	return t;
}

static bool plain_wide(uint32_t p, uint64_t kmax)
{
	return (unsigned __int128) (p - 1) * (p - 1) * (unsigned __int128) kmax >= ((unsigned __int128) 1 << 64);
}

template <int E>
static int batch_gemm_launch_t(hipStream_t stream, int variant, uint32_t p, uint64_t kmax, const BatchGemm *tasks,
			       BatchLaunch L, int grid)
{
	const ModP mp = make_modp(p);
	const uint64_t t64 = two64_mod(p);
	const bool wide = plain_wide(p, kmax);
	switch (variant) {
	case BG_PLAIN:
		if (wide)
			batch_plain_kernel<E, true><<<grid, 256, 0, stream>>>(tasks, L, mp, t64);
		else
			batch_plain_kernel<E, false><<<grid, 256, 0, stream>>>(tasks, L, mp, t64);
		break;
	case BG_SPARSE:
		if (wide)
			batch_sparse_kernel<E, true><<<grid, 256, 0, stream>>>(tasks, L, mp, t64);
		else
			batch_sparse_kernel<E, false><<<grid, 256, 0, stream>>>(tasks, L, mp, t64);
		break;
	case BG_I8S:
		batch_i8_kernel<E, 1, 1, 2, 4><<<grid, 256, 0, stream>>>(tasks, L, mp);
		break;
	case BG_I8L:
		batch_i8_kernel<E, 2, 2, 1, 2><<<grid, 256, 0, stream>>>(tasks, L, mp);
		break;
	default:
		return MLZ_E_ARG;
	}
	return hipGetLastError() == hipSuccess ? MLZ_OK : MLZ_E_HIP;
}

int batch_gemm_launch(hipStream_t stream, int elem, int variant, uint32_t p, uint64_t kmax, const BatchGemm *tasks,
		      BatchLaunch L, int grid)
{
	if (elem == MLZ_U16)
		return batch_gemm_launch_t<MLZ_U16>(stream, variant, p, kmax, tasks, L, grid);
	if (elem == MLZ_U32)
		return batch_gemm_launch_t<MLZ_U32>(stream, variant, p, kmax, tasks, L, grid);
	return batch_gemm_launch_t<MLZ_F64_SYM>(stream, variant, p, kmax, tasks, L, grid);
}

int batch_pack_i8_launch(hipStream_t stream, int elem, uint32_t p, const BatchGemm *tasks, BatchLaunch L, int grid)
{
	if (elem == MLZ_U16)
		batch_pack_i8_kernel<MLZ_U16><<<grid, 256, 0, stream>>>(tasks, L, p);
	else if (elem == MLZ_U32)
		batch_pack_i8_kernel<MLZ_U32><<<grid, 256, 0, stream>>>(tasks, L, p);
	else
		batch_pack_i8_kernel<MLZ_F64_SYM><<<grid, 256, 0, stream>>>(tasks, L, p);
	return hipGetLastError() == hipSuccess ? MLZ_OK : MLZ_E_HIP;
}

template <typename K>
static int occ_of(K kernel, int threads)
{
	int n = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1)
		n = 1;
	return n;
}

template <int E>
static int batch_gemm_occ(int variant, uint32_t p, uint64_t kmax)
{
	const bool wide = plain_wide(p, kmax);
	switch (variant) {
	case BG_PLAIN:
		return wide ? occ_of(batch_plain_kernel<E, true>, 256) : occ_of(batch_plain_kernel<E, false>, 256);
	case BG_SPARSE:
		return wide ? occ_of(batch_sparse_kernel<E, true>, 256) : occ_of(batch_sparse_kernel<E, false>, 256);
	case BG_I8S:
		return occ_of(batch_i8_kernel<E, 1, 1, 2, 4>, 256);
	case BG_I8L:
		return occ_of(batch_i8_kernel<E, 2, 2, 1, 2>, 256);
	default:
		return occ_of(batch_pack_i8_kernel<E>, 256);
	}
}

int batch_gemm_wgs_per_cu(int elem, int variant, uint32_t p, uint64_t kmax)
{
	if (elem == MLZ_U16)
		return batch_gemm_occ<MLZ_U16>(variant, p, kmax);
	if (elem == MLZ_U32)
		return batch_gemm_occ<MLZ_U32>(variant, p, kmax);
	return batch_gemm_occ<MLZ_F64_SYM>(variant, p, kmax);
}

}  // namespace mlz

using namespace mlz;

extern "C" int mlz_zp_gemm(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, size_t k,
			   uint64_t a, const void *A, size_t lda, const void *B, size_t ldb, uint64_t b,
			   void *C, size_t ldc, int algo)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_TRY(check_modulus(ctx, elem, p));
	if ((m && k && !A) || (k && n && !B) || (m && n && !C))
		return set_error(ctx, MLZ_E_ARG, "null matrix pointer");
	if ((m && k && lda < k) || (k && n && ldb < n) || (m && n && ldc < n))
		return set_error(ctx, MLZ_E_ARG, "leading dimension smaller than row length");
	if (algo < MLZ_ALGO_AUTO || algo > MLZ_ALGO_FP64)
		return set_error(ctx, MLZ_E_ARG, "unknown gemm algo %d", algo);
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (ctx->ptr_mode == MLZ_PTR_DEVICE)
		return zp_gemm_dev(ctx, elem, p, m, n, k, a, A, lda, B, ldb, b, C, ldc, algo);
	if (m == 0 || n == 0)
		return MLZ_OK;
	const size_t es = elem_size(elem);
	void *dA = nullptr, *dB = nullptr, *dC = nullptr;
	if (k) {
		MLZ_TRY(stage_in(ctx, WS_STAGE_A, A, m, k, lda, es, &dA));
		MLZ_TRY(stage_in(ctx, WS_STAGE_B, B, k, n, ldb, es, &dB));
	}
	MLZ_TRY(stage_in(ctx, WS_STAGE_C, C, m, n, ldc, es, &dC));
	MLZ_TRY(zp_gemm_dev(ctx, elem, p, m, n, k, a, dA, k, dB, n, b, dC, n, algo));
	return stage_out(ctx, C, dC, m, n, ldc, es);
}

extern "C" int mlz_zp_gemm_noalloc(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, size_t k,
				   uint64_t a, const void *A, size_t lda, const void *B, size_t ldb, uint64_t b,
				   void *C, size_t ldc, int algo)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_TRY(check_modulus(ctx, elem, p));
	if ((m && k && !A) || (k && n && !B) || (m && n && !C))
		return set_error(ctx, MLZ_E_ARG, "null matrix pointer");
	if ((m && k && lda < k) || (k && n && ldb < n) || (m && n && ldc < n))
		return set_error(ctx, MLZ_E_ARG, "leading dimension smaller than row length");
	if (algo < MLZ_ALGO_AUTO || algo > MLZ_ALGO_FP64)
		return set_error(ctx, MLZ_E_ARG, "unknown gemm algo %d", algo);
	if (!ctx->fixed_ws)
		return set_error(ctx, MLZ_E_HIP, "no-allocation scratch has no device address on device %d",
				 ctx->device);
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	ctx->noalloc = true;
	const int rc = zp_gemm_dev(ctx, elem, p, m, n, k, a, A, lda, B, ldb, b, C, ldc, algo);
	ctx->noalloc = false;
	return rc;
}
