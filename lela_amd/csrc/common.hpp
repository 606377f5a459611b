// lela_amd/csrc/common.hpp -- context, workspace, error handling and modular
// arithmetic shared by every kernel file of liblela_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mlz.h"

namespace mlz {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Barrett reduction modulo p < 2^32.
// mu = floor((2^64-1)/p); for x < 2^64, q = hi64(x*mu) satisfies
// x/p - 2.5 < q <= x/p, so r = x - q*p lies in [0, 3p) and two conditional
// subtractions bring it to [0, p).
// ---------------------------------------------------------------------------
struct ModP {
	uint64_t p;
	uint64_t mu;
	double pinv;     // 1/p
	int small;       // p < 2^26: products are exact in a double
	int prime;       // p is prime: inverses by Fermat
	float pinvf;     // 1/p in single precision (mul16)

	__host__ __device__ inline uint64_t reduce(uint64_t x) const
	{
#if defined(__HIP_DEVICE_COMPILE__)
		uint64_t q = __umul64hi(x, mu);
#else
		uint64_t q = (uint64_t) (((unsigned __int128) x * mu) >> 64);
#endif
		uint64_t r = x - q * p;
		r = r >= p ? r - p : r;
		r = r >= p ? r - p : r;
		return r;
	}
	// (x*y) mod p for x, y < p < 2^32.  For p < 2^26 the product is exact in a
	// double; q = floor(x*y/p) is then off by at most one and the remainder
	// x*y - q*p is exact (fma), so one correction each way suffices.
	__host__ __device__ inline uint32_t mul_small(uint32_t x, uint32_t y) const
	{
		const double prod = (double) x * (double) y;
		const double q = floor(prod * pinv);
		double r = fma(-q, (double) p, prod);
		r = r < 0.0 ? r + (double) p : r;
		r = r >= (double) p ? r - (double) p : r;
		return (uint32_t) r;
	}
	// (x*y) mod p for x, y < p <= 2^16: the product is below 2^32 (exact in
	// 24-bit integer multiplies), and the quotient estimate from single
	// precision has relative error below 2^-22 of a quotient below 2^16, so
	// it is off by at most one; the remainder, below 2^17 in magnitude, is
	// exact in 32-bit two's complement.
	__host__ __device__ inline uint32_t mul16(uint32_t x, uint32_t y) const
	{
		const uint32_t pp = (uint32_t) p;
#if defined(__HIP_DEVICE_COMPILE__)
		const uint32_t prod = __umul24(x, y);
		const uint32_t q = (uint32_t) ((float) prod * pinvf);
		int32_t r = (int32_t) (prod - __umul24(q, pp));
#else
		const uint32_t prod = x * y;
		const uint32_t q = (uint32_t) ((float) prod * pinvf);
		int32_t r = (int32_t) (prod - q * pp);
#endif
		r += r < 0 ? (int32_t) pp : 0;
		r -= r >= (int32_t) pp ? (int32_t) pp : 0;
		return (uint32_t) r;
	}
	// (f*g + x) mod p for f, x < p and g <= p <= 2^16: the argument is below
	// p^2, so the same single-precision quotient applies.
	__host__ __device__ inline uint32_t mad16(uint32_t x, uint32_t f, uint32_t g) const
	{
		const uint32_t pp = (uint32_t) p;
#if defined(__HIP_DEVICE_COMPILE__)
		const uint32_t v = __umul24(f, g) + x;
		const uint32_t q = (uint32_t) ((float) v * pinvf);
		int32_t r = (int32_t) (v - __umul24(q, pp));
#else
		const uint32_t v = f * g + x;
		const uint32_t q = (uint32_t) ((float) v * pinvf);
		int32_t r = (int32_t) (v - q * pp);
#endif
		r += r < 0 ? (int32_t) pp : 0;
		r -= r >= (int32_t) pp ? (int32_t) pp : 0;
		return (uint32_t) r;
	}
	__host__ __device__ inline uint32_t mul(uint32_t x, uint32_t y) const
	{
		if (small)
			return mul_small(x, y);
		return (uint32_t) reduce((uint64_t) x * y);
	}
	// Multiplication with the path fixed at compile time (kernels templated
	// on SMALL = (p < 2^26) avoid evaluating both paths).
	template <bool SMALL>
	__host__ __device__ inline uint32_t mulT(uint32_t x, uint32_t y) const
	{
		return SMALL ? mul_small(x, y) : (uint32_t) reduce((uint64_t) x * y);
	}
	// x mod p for an integer-valued double |x| < 2^53 and p < 2^26: the
	// quotient estimate floor(x/p) is off by at most one, the remainder
	// x - q*p is exact (fma), one correction each way.
	__host__ __device__ inline uint32_t mod_f64(double x) const
	{
		const double pd = (double) p;
		const double q = floor(x * pinv);
		double r = fma(-q, pd, x);
		r = r < 0.0 ? r + pd : r;
		r = r >= pd ? r - pd : r;
		return (uint32_t) r;
	}
	// (x*y - f*g) mod p for x, y, f, g < p.  SMALL (p < 2^26): both products
	// are exact doubles and so is their difference (|.| < 2^52), one
	// reduction; otherwise two Barrett products.
	template <bool SMALL>
	__host__ __device__ inline uint32_t mul_subT(uint32_t x, uint32_t y, uint32_t f, uint32_t g) const
	{
		if (SMALL)
			return mod_f64(fma((double) x, (double) y, -((double) f * (double) g)));
		return sub(mulT<false>(x, y), mulT<false>(f, g));
	}
	// x^e mod p
	__host__ __device__ inline uint32_t pow(uint32_t x, uint64_t e) const
	{
		uint32_t r = (uint32_t) (1 % p);
		while (e) {
			if (e & 1)
				r = mul(r, x);
			x = mul(x, x);
			e >>= 1;
		}
		return r;
	}
	// Inverse of x mod p, 0 when none exists: Fermat for prime p, else the
	// extended Euclidean algorithm.
	__host__ __device__ inline uint32_t inv(uint32_t x) const
	{
		x = (uint32_t) (x % p);
		if (x == 0)
			return 0;
		if (prime)
			return pow(x, p - 2);
		int64_t r0 = x, r1 = (int64_t) p, s0 = 1, s1 = 0;
		while (r1 != 0) {
			int64_t q = r0 / r1, t;
			t = r0 - q * r1; r0 = r1; r1 = t;
			t = s0 - q * s1; s0 = s1; s1 = t;
		}
		if (r0 != 1)
			return 0;
		s0 %= (int64_t) p;
		if (s0 < 0)
			s0 += (int64_t) p;
		return (uint32_t) s0;
	}
	// (a*x + y) mod p for a, x, y < p < 2^32 (a*x + y < 2^64).
	__host__ __device__ inline uint32_t axpy(uint32_t a, uint32_t x, uint32_t y) const
	{
		return (uint32_t) reduce((uint64_t) a * x + y);
	}
	__host__ __device__ inline uint32_t add(uint32_t x, uint32_t y) const
	{
		uint64_t s = (uint64_t) x + y;
		return (uint32_t) (s >= p ? s - p : s);
	}
	__host__ __device__ inline uint32_t sub(uint32_t x, uint32_t y) const
	{
		return x >= y ? x - y : (uint32_t) (x + p - y);
	}
	__host__ __device__ inline uint32_t neg(uint32_t x) const
	{
		return x == 0 ? 0u : (uint32_t) (p - x);
	}
};

// Deterministic Miller-Rabin for 32-bit n (bases 2, 7, 61 suffice below 2^32).
inline bool is_prime_u32(uint64_t n)
{
	if (n < 2)
		return false;
	for (uint64_t q : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 61ull})
		if (n % q == 0)
			return n == q;
	uint64_t d = n - 1;
	int s = 0;
	while ((d & 1) == 0) {
		d >>= 1;
		++s;
	}
	for (uint64_t a : {2ull, 7ull, 61ull}) {
		uint64_t x = 1, b = a % n, e = d;
		while (e) {
			if (e & 1)
				x = (uint64_t) ((unsigned __int128) x * b % n);
			b = (uint64_t) ((unsigned __int128) b * b % n);
			e >>= 1;
		}
		if (x == 1 || x == n - 1)
			continue;
		bool comp = true;
		for (int r = 1; r < s && comp; ++r) {
			x = (uint64_t) ((unsigned __int128) x * x % n);
			if (x == n - 1)
				comp = false;
		}
		if (comp)
			return false;
	}
	return true;
}

inline ModP make_modp(uint64_t p)
{
	ModP m;
	m.p = p;
	m.mu = ~(uint64_t) 0 / p;
	m.pinv = 1.0 / (double) p;
	m.pinvf = (float) (1.0 / (double) p);
	m.small = p < (1ull << 26);
	m.prime = is_prime_u32(p);
	return m;
}

// Signed 8-bit limbs of a residue v < p <= 2^16 for the INT8 MFMA kernels:
// u = x + 32896 with x the representative in [-32896, 32639] (v itself if
// v <= 32639, else v - p); the bytes of u ^ 0x8080 are the limbs (hi, lo), so
// x = 256*hi + lo with both in [-128, 127].  Zero maps to (0, 0).
__device__ inline uint32_t i8_limbs(uint32_t v, uint32_t p)
{
	int32_t x = v > 32639u ? (int32_t) v - (int32_t) p : (int32_t) v;
	return (uint32_t) (x + 32896) ^ 0x8080u;
}

// Element storage <-> canonical uint32 in [0, p).
template <int E> struct Elem;
template <> struct Elem<MLZ_U16> {
	typedef uint16_t T;
	__device__ static inline uint32_t load(const T *ptr, uint32_t) { return *ptr; }
	__device__ static inline void store(T *ptr, uint32_t v, uint32_t) { *ptr = (T) v; }
};
template <> struct Elem<MLZ_U32> {
	typedef uint32_t T;
	__device__ static inline uint32_t load(const T *ptr, uint32_t) { return *ptr; }
	__device__ static inline void store(T *ptr, uint32_t v, uint32_t) { *ptr = v; }
};
template <> struct Elem<MLZ_F64_SYM> {
	typedef double T;
	// Symmetric representative s in [-(p-1)/2, (p-1)/2] -> canonical.
	__device__ static inline uint32_t load(const T *ptr, uint32_t p)
	{
		long long s = (long long) *ptr;
		return (uint32_t) (s < 0 ? s + (long long) p : s);
	}
	// Canonical -> symmetric: ModularTraits<double>::valid_rep keeps
	// v <= p/2 (lela/ring/modular.h:310-326).
	__device__ static inline void store(T *ptr, uint32_t v, uint32_t p)
	{
		*ptr = (2ull * v > (unsigned long long) p) ? (double) ((long long) v - (long long) p) : (double) v;
	}
};

inline size_t elem_size(int e)
{
	return e == MLZ_U16 ? 2 : e == MLZ_U32 ? 4 : e == MLZ_F64_SYM ? 8 : 0;
}

// ---------------------------------------------------------------------------
// Context.
// ---------------------------------------------------------------------------
enum WsSlot {
	WS_STAGE_A, WS_STAGE_B, WS_STAGE_C,      // host-pointer staging
	WS_PACK_A, WS_PACK_B,                    // gemm operand packing
	WS_TRSM_0, WS_TRSM_1, WS_TRSM_2, WS_TRSM_3,   // trsm temporaries
	WS_RREF_0, WS_RREF_1, WS_RREF_2, WS_RREF_3, WS_RREF_4, WS_RREF_5,
	WS_FL_0, WS_FL_1, WS_FL_2, WS_FL_3, WS_FL_4, WS_FL_5, WS_FL_6, WS_FL_7,
	WS_GF2_0, WS_GF2_1, WS_GF2_2, WS_GF2_3,
	WS_BATCH_0, WS_BATCH_1, WS_BATCH_2, WS_BATCH_3,
	WS_SMALL,
	WS_INVTAB,                               // inverses mod p (p <= 2^16), see invtab_p
	WS_NUM_SLOTS
};

struct ProfRec {
	std::string name;
	uint64_t count;
	double ms;
};

struct PendingEv {
	const char *name;
	hipEvent_t e0, e1;
};

}  // namespace mlz

struct mlz_ctx {
	int device;
	hipStream_t stream;
	hipStream_t aux_stream;  // second stream (default priority) for overlapped work; created on first use
	hipStream_t copy_stream; // third stream (default priority) for small result copies; created on first use
	int ptr_mode;
	std::string last_error;
	void *ws[mlz::WS_NUM_SLOTS];
	size_t ws_bytes[mlz::WS_NUM_SLOTS];
	bool prof_on;
	std::vector<mlz::PendingEv> pending;
	std::vector<hipEvent_t> free_events;
	std::vector<mlz::ProfRec> prof;
	hipEvent_t timer0, timer1;
	int num_cus;
	void *host_pinned;
	size_t host_pinned_bytes;
	int i8_variant;      // INT8 gemm tile variant, 0 = chosen by shape
	int upd_nblk;        // 32-column blocks per wave of the rank-k update (1, 2, 4)
	int upd_lds;         // INT8 rank-k update: full tiles staged in LDS, this many per wave (1, 2, 4); 0: registers
	                     // only; -1: chosen by the row alignment
	int rref_fast_threads;   // threads of the echelon form's fused pivot-search kernel (256, 512, 1024)
	uint32_t invtab_p;       // modulus whose inverse table WS_INVTAB holds (0 = none)
	int gf2_kernel;          // GF(2) gemm: 0 Four Russians (default), 1 the AND/XOR-popcount dot form
	struct mlz_tuner *tuner;   // autotuner table and cost-model weights (batch.hip), lazily created
	// No-allocation calls (mlz_zp_gemm_noalloc): while `noalloc` is set, gemm
	// operand packing takes its buffers from `fixed_ws` (this device's copy of
	// the library's static scratch array) instead of the workspace slots, and
	// no profiling events are created.
	bool noalloc;
	void *fixed_ws;
	size_t fixed_ws_bytes;
};

namespace mlz {

int set_error(mlz_ctx *ctx, int code, const char *fmt, ...);

#define MLZ_HIP(ctx, expr)                                                              \
	do {                                                                            \
		hipError_t _e = (expr);                                                 \
		if (_e != hipSuccess)                                                   \
			return ::mlz::set_error((ctx), MLZ_E_HIP, "%s:%d %s: %s", __FILE__,     \
						__LINE__, #expr, hipGetErrorString(_e)); \
	} while (0)

#define MLZ_TRY(expr)                           \
	do {                                    \
		int _rc = (expr);               \
		if (_rc != MLZ_OK)              \
			return _rc;             \
	} while (0)

// Device buffer of at least `bytes` in a workspace slot (grows, never shrinks
// until mlz_ctx_trim).
int ws_get(mlz_ctx *ctx, int slot, size_t bytes, void **out);

// Profiling brackets around a kernel launch on ctx->stream.
void prof_begin(mlz_ctx *ctx, const char *name, PendingEv *ev);
void prof_end(mlz_ctx *ctx, PendingEv *ev);

struct ProfScope {
	mlz_ctx *ctx;
	PendingEv ev;
	ProfScope(mlz_ctx *c, const char *name) : ctx(c) { prof_begin(ctx, name, &ev); }
	~ProfScope() { prof_end(ctx, &ev); }
};

int check_launch(mlz_ctx *ctx, const char *what);

// A batch of equally shaped products C_z <- a A_z B_z + b C_z (z < count):
// operand w (0 = A, 1 = B, 2 = C) of item z sits at element offset
// (z / inner) * outer[w] + (z % inner) * in[w] from the base pointer, so
// blocks along a diagonal (two strides) are one batch.  count = 1 is the
// ordinary single product.
struct I8Batch {
	int count, inner;
	long long outer[3], in[3];
	__host__ __device__ long long offset(int z, int w) const
		{ return (long long) (z / inner) * outer[w] + (long long) (z % inner) * in[w]; }
	static I8Batch single()
	{
		I8Batch b;
		b.count = b.inner = 1;
		for (int w = 0; w < 3; ++w)
			b.outer[w] = b.in[w] = 0;
		return b;
	}
};

// Internal device-pointer entry points shared between the kernel files.
// All pointers are device pointers; work is queued on ctx->stream.
int zp_gemm_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, size_t k,
		uint64_t a, const void *A, size_t lda, const void *B, size_t ldb,
		uint64_t b, void *C, size_t ldc, int algo);
int zp_scal_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, uint64_t a,
		void *C, size_t ldc);
int zp_trsm_dev(mlz_ctx *ctx, int elem, uint64_t p, int upper, int unit, size_t n, size_t nrhs,
		uint64_t a, const void *A, size_t lda, void *B, size_t ldb, size_t *bad_diag, int algo);
int zp_trsm_prepare(mlz_ctx *ctx, int elem, uint64_t p, int upper, int unit, size_t n, const void *A, size_t lda,
		    size_t *bad_diag, const uint32_t **W_out);
int zp_trsm_apply(mlz_ctx *ctx, int elem, uint64_t p, int upper, int unit, size_t n, size_t nrhs, uint64_t a,
		  const void *A, size_t lda, void *B, size_t ldb, const uint32_t *W, int algo);
constexpr size_t TRSM_LEAF_ROWS = 64;   // block size of zp_trsm_prepare's inverses
// X = a A^-1 B by 64-row blocks on INT8 MFMA (p <= 2^16), W from
// zp_trsm_prepare for the same A; bend (upper only, or null = n): one past
// the last column block I's band needs (a banded A skips the zero columns).
int zp_trsm_blocked_i8(mlz_ctx *ctx, int elem, uint64_t p, int upper, size_t n, size_t nrhs, uint64_t a,
		       const void *A, size_t lda, void *B, size_t ldb, const uint32_t *W, const int *bend);
// Upper banded A (canonical uint32, p <= 2^16, n >= 1024): X = A^-1 B with
// the 1024-row diagonal blocks inverted up front; c1_64 as for
// zp_trsm_blocked_i8.
int zp_trsm_banded_inv_i8(mlz_ctx *ctx, uint64_t p, size_t n, size_t nrhs, const uint32_t *A, size_t lda,
			  uint32_t *B, size_t ldb, const uint32_t *W, const int *c1_64);
int zp_rref_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, void *A, size_t lda,
		size_t *rank_out, uint32_t *pivots_out, int algo);
// Non-reduced echelon forms with LELA's transposition pivoting (zp_lu.hip).
// store = false: A is only read (rank, det and the lists are still returned).
// pivcols (host, may be null): the pivot column of each step.
int zp_echelon_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, void *A, size_t lda, int mode,
		   size_t *rank_out, uint64_t *det_out, uint32_t *perm_rows, uint32_t *perm_cols, int algo,
		   bool store = true, uint32_t *pivcols = nullptr);
// Table of inverses mod p <= 2^16 in device memory (0 where none exists).
int zp_inverse_table(mlz_ctx *ctx, const ModP &mp, const uint16_t **out);
int gf2_gemm_dev(mlz_ctx *ctx, size_t m, size_t n, size_t k, int a, const uint64_t *A, size_t lda,
		 const uint64_t *B, size_t ldb, int b, uint64_t *C, size_t ldc);
int gf2_trsm_dev(mlz_ctx *ctx, int upper, int unit, size_t n, size_t nrhs, int a, const uint64_t *A, size_t lda,
		 uint64_t *B, size_t ldb, size_t *bad_diag);
int gf2_rref_dev(mlz_ctx *ctx, size_t m, size_t n, uint64_t *A, size_t lda, size_t *rank_out,
		 uint32_t *pivots_out);
int pick_gemm_algo(uint64_t p, int elem, size_t m, size_t n, size_t k, int algo);
// bt.count products of one shape on INT8 MFMA (p <= 2^16), operands at the
// batch's offsets from A, B and C (element storage E, device pointers).
// elemA: A's storage (any of the three; elem is B's and C's).
// Apre (may be null): A already packed by zp_pack_i8_a_dev (then A is unused).
// Packed operands across products (one product, k <= I8_KMAX): Bpre, B
// already packed with its 32-column blocks bpanel bytes apart (a k-range of a
// larger packed operand: the pointer is offset by 2048 bytes per 32 rows of
// k); Cpk, C also written in that layout with blocks cpanel bytes apart (m, n
// multiples of 32).  B's columns up to the tile padding must exist in Bpre.
struct I8PackIO {
	const int8_t *Bpre = nullptr;
	size_t bpanel = 0;
	int8_t *Cpk = nullptr;
	size_t cpanel = 0;
};
int zp_gemm_i8_batched_dev(mlz_ctx *ctx, int elem, int elemA, uint32_t p, const I8Batch &bt, size_t m, size_t n,
			   size_t k, uint32_t a, const void *A, size_t lda, const void *B, size_t ldb, uint32_t b,
			   void *C, size_t ldc, const int8_t *Apre = nullptr, const I8PackIO &io = I8PackIO());
int zp_pack_i8_a_dev(mlz_ctx *ctx, int elem, uint32_t p, const I8Batch &bt, size_t rows, size_t k, const void *A,
		     size_t lda, int8_t *dst, size_t *item_bytes, hipStream_t stream = nullptr);
int zp_pack_i8_b_dev(mlz_ctx *ctx, int elem, uint32_t p, size_t k, size_t cols, const void *B, size_t ldb,
		     int8_t *dst, hipStream_t stream = nullptr);
int zp_rank_update_i8_dev(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, size_t k, const uint32_t *G, size_t ldg,
			  const uint32_t *W, size_t ldw, uint32_t *C, size_t ldc);
// The same update with G (m x 32*KB) and W (32*KB x n) already packed as INT8
// limb fragments ([32-index block][k block][limb][64 lanes][16 bytes]).
int zp_rank_update_i8_packed(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, int KB, const int8_t *Gp,
			     const int8_t *Wp, uint32_t *C, size_t ldc, hipStream_t stream = nullptr);
int zp_rank_update_i8_packed(mlz_ctx *ctx, uint32_t p, size_t m, size_t n, int KB, const int8_t *Gp,
			     const int8_t *Wp, uint16_t *C, size_t ldc, hipStream_t stream = nullptr);
// The context's second stream, created on first use (default priority).
int aux_stream(mlz_ctx *ctx, hipStream_t *out);
int copy_stream(mlz_ctx *ctx, hipStream_t *out);
int check_modulus(mlz_ctx *ctx, int elem, uint64_t p);
void tuner_free(mlz_ctx *ctx);
// Looks up (once per device, with the device current) the address of the
// static scratch array that mlz_zp_gemm_noalloc uses; called by mlz_ctx_create.
int noalloc_scratch_init(mlz_ctx *ctx);

// Host <-> device staging of a strided matrix (rows x cols elements of esize
// bytes, leading dimension ld elements) into a dense device buffer with
// leading dimension cols.
int stage_in(mlz_ctx *ctx, int slot, const void *host, size_t rows, size_t cols, size_t ld,
	     size_t esize, void **dev);
int stage_out(mlz_ctx *ctx, void *host, const void *dev, size_t rows, size_t cols, size_t ld,
	      size_t esize);

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned) ((a + b - 1) / b); }
inline size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

}  // namespace mlz
