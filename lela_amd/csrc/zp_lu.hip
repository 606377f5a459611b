// lela_amd/csrc/zp_lu.hip -- non-reduced row-echelon forms over Z/p with
// LELA's transposition pivoting, on MI355X.
//
// Replaces Elimination::echelonize (lela/algorithms/elimination.tcc:47-139,
// with and without compute_L) and Elimination::pluq (:266-370) for dense
// matrices with DensePivotStrategy (lela/algorithms/pivot-strategy.tcc:19-41):
// step i takes the first column (from the previous pivot's on) in which some
// row at a position >= i is nonzero, the first such row as pivot, and swaps
// it with position i.  Unlike the reduced form, the result depends on this
// order, so it is reproduced exactly; the arithmetic is blocked.
//
// The working copy is kept in position order (a transposition swaps two rows
// of it).  Per panel of up to LU_W columns from column c at position i:
//  1. one workgroup loads the panel part of the LU_K rows at positions
//     i .. i+LU_K-1 (the window) and runs the elimination above on them,
//     column by column: the first nonzero row of the window is the first
//     nonzero row overall, because the window holds the first positions in
//     order.  It returns the pivots, the window's new order, L11 (the
//     multipliers among the panel's pivot rows) and U11 (the pivot rows on
//     the pivot columns after elimination).
//  2. The window rows are put in that order.  With P the panel's pivot rows
//     as they stood (now at positions i ..), one thread per column computes
//     U = L11^-1 P (forward substitution: the new pivot rows) and
//     W = U11^-1 U = P_J^-1 P (back substitution).
//  3. Every row below the new pivots becomes X_j - X_j[J] W: one rank-LU_W
//     update of the contiguous rows below.  This is the unique result of the
//     column-by-column elimination's row operations: each row below changes by
//     the combination of pivot rows that clears the pivot columns, X_j[J]
//     P_J^-1.
//  4. If a column of the panel has no nonzero in the window (and rows below
//     it exist), the panel ends there; after its update a search over every
//     remaining row finds the next pivot (first column, then first row), and
//     its transposition is made before the next panel.
// compute_L: the working copy carries [A | I_m]; with orig[t] the original
// index of the row at position t, the augmented part of row t at column
// orig[s] is T[t][s] of T P A = U, the value LELA leaves in A[t][s] left of
// the pivots.  pluq: the multipliers of the rows below are X_j[J] U11^-1 (one
// thread per row), the pivot rows' own are L11; they are kept in a buffer
// whose rows move with the working copy's; the column order comes from the
// pivot columns.
#include "common.hpp"

#include <climits>

namespace mlz {

constexpr int LU_W = 64;     // panel width (columns)
constexpr int LU_K = 64;     // window rows (one per lane)
constexpr int LU_NT = 512;   // threads of the window kernel: 8 waves of 8 panel columns
constexpr int LU_SEARCH_COLS = 2048;   // columns per launch of the pivot search
constexpr int LU_NARROW = 128;         // columns of a panel's update done before the next panel (multiple of 32)
constexpr uint32_t LU_EV_NONE = ~0u;    // a column without a nonzero in the window
constexpr int LU_SPIN_MAX = 1 << 22;   // polls of a column's flag before the kernel reports an error

// Device outputs of one window kernel.
struct LuPanel {
	int rp;            // pivots found
	int stop;          // 1: a column of the panel needs the rows beyond the window
	int stop_col;      // that column (absolute)
	int bad;           // 1: a pivot had no inverse (composite modulus); 2: a wave stopped waiting
	int steps[LU_W];   // window position of step s's pivot row (before its swap)
	int J[LU_W];       // pivot column of step s (absolute)
	uint32_t x[LU_W];  // pivot value of step s
	uint32_t dinv[LU_W];   // its inverse
	int wperm[LU_K];   // window order on exit: original window index at each position
};

// One panel: window rows i0 .. i0+kw-1 (kw = min(LU_K, m - i0)), columns
// c .. c+wc-1 (pivots searched in them only).  Outputs, with rp pivots and P
// the panel's pivot rows as they stood (in step order), LU_W x LU_W each and
// zero outside the first rp rows:
//   LiM rows 0 .. 63:   L11^-1, the combination of P giving each new pivot row;
//   LiM rows 64 .. 127: P_J^-1 (P on the pivot columns J, inverted);
//   L11: unit lower triangular, the multipliers among the pivot rows (L11[t][s],
//        s < t: the multiple of new pivot row s subtracted from pivot row t),
//        so that P_J = L11 U11 with U11 the new pivot rows on J.
//
// Lane l holds the row at window position l, wave w the panel columns 8w ..
// 8w+7.  Column cc: the wave holding it takes the first lane >= cur with a
// nonzero there (one ballot), the pivot's inverse (from an LDS copy of the
// table when p <= 2^16), stores the multipliers and the column's event, then
// sets the column's flag.  Every other wave waits for that flag only (not for
// each other), exchanges the two rows if the pivot is not at cur (lanes cur
// and first: the transposition), reads the pivot row from lane cur and
// updates its columns from cc on.  Nothing in LDS is rewritten within a
// panel while a wave may still read it: each column has its own event, each
// step its own multipliers, and the transpositions' effect on the stored
// multipliers is replayed afterwards.  Then L11^-1 and P_J^-1 = U11^-1 L11^-1
// come from a blocked substitution over 8 x 8 blocks with all threads: the
// diagonal blocks are inverted first, then each block row is one product.
// (Doubling both triangular inverses from the 8 x 8 blocks, then one product
// U11^-1 L11^-1, has 8 barriers instead of 32 but measured slower, 21.95
// against 21.45 ms for the 20000 x 30000 reduction: the sums bound this part.)
template <bool SMALL, bool TAB, typename TX>
__global__ __launch_bounds__(LU_NT) void lu_window_kernel(const TX *__restrict__ X, size_t ldx, int m, int c,
							  int wc, int i0, ModP mp, const uint16_t *__restrict__ invtab,
							  LuPanel *__restrict__ out, uint32_t *__restrict__ L11,
							  uint32_t *__restrict__ LiM)
{
	using FT = typename std::conditional<TAB, uint16_t, uint32_t>::type;
	constexpr int FP = TAB ? LU_W + 2 : LU_W + 1;       // padded row of Lm (odd word stride)
	constexpr int XP = LU_W + 4;                        // padded rows of the post-pass matrices
	constexpr int MAT = LU_W * XP;                      // words per matrix
	constexpr size_t POST = (size_t) (5 * MAT + 8 * XP + 2 * 512) * 4;
	constexpr size_t BIG = TAB ? (65536 * 2 > POST ? 65536 * 2 : POST) : POST;
	__shared__ __align__(16) unsigned char big[BIG];    // the inverse table, then the post-pass
	__shared__ FT Lm[LU_K * FP];                        // multipliers by position and step
	__shared__ uint32_t evc[LU_W][2];                   // per column: pivot lane (or none), inverse
	__shared__ int ready[LU_W];                         // per column: its event is published
	__shared__ uint32_t sdinv[LU_W], sx[LU_W];
	__shared__ int ord[LU_K], jl[LU_W], sfirst[LU_W];
	const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
	const int kw = min(LU_K, m - i0);
	const bool covers = i0 + LU_K >= m;
	const uint32_t pp = (uint32_t) mp.p;
	auto mul = [&](uint32_t a, uint32_t b) -> uint32_t { return TAB ? mp.mul16(a, b) : mp.mulT<SMALL>(a, b); };
	auto upd = [&](uint32_t a, uint32_t f, uint32_t pv) -> uint32_t {     // a - f pv
		return TAB ? mp.mad16(a, f, pp - pv) : mp.sub(a, mp.mulT<SMALL>(f, pv));
	};
	uint32_t x[8];
#pragma unroll
	for (int jj = 0; jj < 8; ++jj) {
		const int col = 8 * w + jj;
		x[jj] = (l < kw && col < wc) ? (uint32_t) X[(size_t) (i0 + l) * ldx + c + col] : 0u;
	}
	if (tid < LU_K) {
		ord[tid] = tid;
		ready[tid] = 0;
	}
	if (TAB) {   // the table's first p entries, 16 bytes per load, all loads issued at once
		constexpr int VPT = 65536 * 2 / 16 / LU_NT;
		const int nv = ((int) mp.p + 7) / 8;
		const uint4 *src = reinterpret_cast<const uint4 *>(invtab);
		uint4 *dst = reinterpret_cast<uint4 *>(big);
		uint4 tv[VPT];
#pragma unroll
		for (int q = 0; q < VPT; ++q)
			tv[q] = src[min(q * LU_NT + tid, nv - 1)];
#pragma unroll
		for (int q = 0; q < VPT; ++q) {
			const int e = q * LU_NT + tid;
			if (e < nv)
				dst[e] = tv[q];
		}
	}
	const uint16_t *itab = reinterpret_cast<const uint16_t *>(big);
	__shared__ int sfail;
	if (tid == 0)
		sfail = 0;
	__syncthreads();
	int cur = 0, stop = 0, stop_col = 0, bad = 0;
	bool done = false, fast = false;
	// Fast path for a full window (LU_K rows, LU_W columns): no transposition
	// happens while every step's diagonal entry is nonzero -- row s is then
	// the first nonzero row of column s among positions >= s, and column s the
	// first column with a nonzero -- so the panel is eliminated in 8-column
	// blocks: the wave owning block k eliminates its columns in registers
	// (cross-lane reads, no barrier between its steps), and after one barrier
	// the waves to its right apply the block's multipliers.  A zero diagonal
	// entry (or a pivot without an inverse) sends the panel to the
	// column-by-column search below, from the original rows.
	if (kw == LU_K && wc == LU_W) {
		for (int k = 0; k < 8; ++k) {
			if (__builtin_amdgcn_readfirstlane(w) == k) {
#pragma unroll
				for (int j = 0; j < 8; ++j) {
					const int s = 8 * k + j;
					const uint32_t xs = __builtin_amdgcn_readlane(x[j], s);
					const uint32_t inv = xs == 0u ? 0u : (TAB ? (uint32_t) itab[xs] : mp.inv(xs));
					if (inv == 0u) {
						if (l == 0)
							sfail = 1;
						break;
					}
					const uint32_t f = l > s ? mul(x[j], inv) : 0u;
					Lm[l * FP + s] = (FT) f;
#pragma unroll
					for (int jj = 0; jj < 8; ++jj)
						if (jj > j)
							x[jj] = upd(x[jj], f, __builtin_amdgcn_readlane(x[jj], s));
					x[j] = l > s ? 0u : x[j];
					if (l == 0) {
						jl[s] = s;
						sdinv[s] = inv;
						sfirst[s] = s;
						sx[s] = xs;
					}
				}
			}
			__syncthreads();
			if (sfail)
				break;
			if (w > k) {
#pragma unroll
				for (int j = 0; j < 8; ++j) {
					const int s = 8 * k + j;
					const uint32_t f = (uint32_t) Lm[l * FP + s];
#pragma unroll
					for (int jj = 0; jj < 8; ++jj)
						x[jj] = upd(x[jj], f, __builtin_amdgcn_readlane(x[jj], s));
				}
			}
		}
		if (sfail) {             // start over from the window as loaded
#pragma unroll
			for (int jj = 0; jj < 8; ++jj)
				x[jj] = (uint32_t) X[(size_t) (i0 + l) * ldx + c + 8 * w + jj];
			__syncthreads();     // every wave is past its reads of the fast path's multipliers
		} else {
			cur = LU_K;
			done = fast = true;
		}
	}
	// The owner's search of column cc_ at step s: the event, the multipliers
	// by position (after the exchange of positions s and first) and the step
	// records, then the column's flag (a wave's LDS operations complete in order).
	// The owner keeps its own copy (o_e0, o_inv, o_f) and reads nothing back.
	uint32_t o_e0 = LU_EV_NONE, o_inv = 0u, o_f = 0u;
	auto publish = [&](int cc_, int s, uint32_t v) {
		const uint64_t bal = __ballot(v != 0u && l >= s);
		if (bal == 0) {
			o_e0 = LU_EV_NONE;
			if (l == 0)
				evc[cc_][0] = LU_EV_NONE;
		} else {
			const int first = __builtin_ctzll(bal);
			const uint32_t xp = __builtin_amdgcn_readlane(v, first);
			const uint32_t vc = __builtin_amdgcn_readlane(v, s);
			const uint32_t inv = TAB ? (uint32_t) itab[xp] : mp.inv(xp);
			const uint32_t a = l == first ? vc : v;
			o_e0 = (uint32_t) first;
			o_inv = inv;
			o_f = l > s ? mul(a, inv) : 0u;
			Lm[l * FP + s] = (FT) o_f;
			if (l == 0) {
				evc[cc_][0] = (uint32_t) first;
				evc[cc_][1] = inv;
				jl[s] = cc_;
				sdinv[s] = inv;
				sfirst[s] = first;
				sx[s] = xp;
			}
		}
		__builtin_amdgcn_wave_barrier();
		if (l == 0)
			__hip_atomic_store(&ready[cc_], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
	};
	for (int cb = 0; cb < wc && !done; cb += 8) {
		const int ow = cb >> 3;
		const bool own = __builtin_amdgcn_readfirstlane(w) == ow;   // a scalar test
		// The owner of columns cb .. cb+7 runs at a raised wave priority and
		// one column ahead: after each event it updates the next column
		// first, publishes it, then updates its other columns.
		if (own) {
			__builtin_amdgcn_s_setprio(2);
			if (cur < kw)
				publish(cb, cur, x[0]);
		}
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			const int cc = cb + j;
			if (cc >= wc) {
				done = true;
				break;
			}
			if (cur == kw) {     // every window row is a pivot row
				if (!covers) {
					stop = 1;
					stop_col = c + cc;
				}
				done = true;
				break;
			}
			if (!own) {
				int spins = 0;
				while (__hip_atomic_load(&ready[cc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0 &&
				       spins < LU_SPIN_MAX) {
					__builtin_amdgcn_s_sleep(1);
					++spins;
				}
				__builtin_amdgcn_wave_barrier();
				if (spins >= LU_SPIN_MAX) {   // never expected: reported, not waited for
					bad = 2;
					done = true;
					break;
				}
			}
			uint32_t e0, inv, f;
			if (own) {
				e0 = o_e0;
				inv = o_inv;
				f = o_f;
			} else {
				e0 = __builtin_amdgcn_readfirstlane(evc[cc][0]);
				inv = __builtin_amdgcn_readfirstlane(evc[cc][1]);
				f = (uint32_t) Lm[l * FP + cur];
			}
			if (e0 == LU_EV_NONE) {
				if (!covers) {       // rows beyond the window may have one
					stop = 1;
					stop_col = c + cc;
					done = true;
					break;
				}
				// a zero column: no pivot (pivot-strategy.tcc:31-39)
				if (own && j < 7 && cc + 1 < wc)
					publish(cc + 1, cur, x[j + 1]);
				continue;
			}
			const int first = (int) e0;
			if (inv == 0u) {             // not invertible: the modulus is composite
				bad = 1;
				done = true;
				break;
			}
			if (first != cur) {          // the transposition (cur, first)
				const int src = l == cur ? first : (l == first ? cur : l);
#pragma unroll
				for (int jj = 0; jj < 8; ++jj)
					x[jj] = __shfl(x[jj], src, 64);
				if (tid == 0) {
					const int t0 = ord[cur];
					ord[cur] = ord[first];
					ord[first] = t0;
				}
			}
			const bool ahead = own && j < 7 && cc + 1 < wc && cur + 1 < kw;
			if (ahead) {
				x[j + 1] = upd(x[j + 1], f, __builtin_amdgcn_readlane(x[j + 1], cur));
				publish(cc + 1, cur + 1, x[j + 1]);
			}
#pragma unroll
			for (int jj = 0; jj < 8; ++jj)
				if (8 * w + jj >= cc && !(ahead && jj == j + 1))   // the pivot row is zero left of cc
					x[jj] = upd(x[jj], f, __builtin_amdgcn_readlane(x[jj], cur));
			++cur;
		}
		if (own)
			__builtin_amdgcn_s_setprio(0);
	}
	const int rp = cur;
	if (bad == 2)
		ready[0] = -1;           // a wave gave up waiting
	__syncthreads();             // the table is no longer read
	if (ready[0] == -1)
		bad = 2;
	// The multipliers of step s were stored by position at step s; the
	// transpositions of later steps move two of them each (the fast path makes none).
	if (!fast && tid < rp)
		for (int t = tid + 1; t < rp; ++t) {
			const int ft = sfirst[t];
			if (ft != t) {
				const FT t0 = Lm[t * FP + tid];
				Lm[t * FP + tid] = Lm[ft * FP + tid];
				Lm[ft * FP + tid] = t0;
			}
		}
	uint32_t *Xf = reinterpret_cast<uint32_t *>(big);   // the panel after elimination
	uint32_t *Ls = Xf + MAT, *Us = Ls + MAT, *Li = Us + MAT, *Ms = Li + MAT, *Rt = Ms + MAT;
	uint32_t *IL = Rt + 8 * XP, *IU = IL + 512;          // inverses of the 8 x 8 diagonal blocks
#pragma unroll
	for (int jj = 0; jj < 8; ++jj)
		Xf[l * XP + 8 * w + jj] = x[jj];
	if (tid >= rp && tid < LU_W)
		sdinv[tid] = 1u;
	__syncthreads();
	if (tid < LU_K)
		out->wperm[tid] = tid < kw ? ord[tid] : tid;
	if (tid < rp) {
		out->steps[tid] = sfirst[tid];
		out->J[tid] = c + jl[tid];
		out->x[tid] = sx[tid];
		out->dinv[tid] = sdinv[tid];
	}
	if (tid == 0) {
		out->rp = rp;
		out->stop = stop;
		out->stop_col = stop_col;
		out->bad = bad;
	}
	if (bad)
		return;
	// L11 and U11, completed by the identity past rp
	for (int idx = tid; idx < LU_W * LU_W; idx += LU_NT) {
		const int t = idx / LU_W, s_ = idx % LU_W;
		const uint32_t lv = t == s_ ? 1u : ((s_ < t && t < rp) ? (uint32_t) Lm[t * FP + s_] : 0u);
		Ls[t * XP + s_] = lv;
		Us[t * XP + s_] = t == s_ ? (t < rp ? Xf[t * XP + jl[t]] : 1u)
					  : ((t < s_ && s_ < rp) ? Xf[t * XP + jl[s_]] : 0u);
		L11[idx] = t < rp ? lv : 0u;
	}
	__syncthreads();
	// sums of products: raw products below 2^52 for p < 2^26 (64 terms below
	// 2^58), reduced products otherwise
	auto prod = [&](uint32_t a, uint32_t b) -> uint64_t {
		const uint64_t v = (uint64_t) a * b;
		return SMALL ? v : mp.reduce(v);
	};
	if (tid < 128) {             // column u of the inverse of diagonal block B
		const int B = (tid & 63) >> 3, u = tid & 7, o = 8 * B;
		uint32_t y[8];
		if (tid < 64) {          // unit lower: y_r = -sum_{u <= q < r} L_rq y_q
#pragma unroll
			for (int r = 0; r < 8; ++r) {
				uint64_t acc = 0;
#pragma unroll
				for (int q = 0; q < r; ++q)
					if (q >= u)
						acc += prod(Ls[(o + r) * XP + o + q], y[q]);
				y[r] = r < u ? 0u : (r == u ? 1u : mp.neg((uint32_t) mp.reduce(acc)));
			}
#pragma unroll
			for (int r = 0; r < 8; ++r)
				IL[B * 64 + r * 8 + u] = y[r];
		} else {                 // upper: y_r = -(sum_{r < q <= u} U_rq y_q) / U_rr, y_u = 1 / U_uu
#pragma unroll
			for (int r = 7; r >= 0; --r) {
				uint64_t acc = 0;
#pragma unroll
				for (int q = r + 1; q < 8; ++q)
					if (q <= u)
						acc += prod(Us[(o + r) * XP + o + q], y[q]);
				y[r] = r > u ? 0u
					     : (r == u ? sdinv[o + r] : mul(mp.neg((uint32_t) mp.reduce(acc)), sdinv[o + r]));
			}
#pragma unroll
			for (int r = 0; r < 8; ++r)
				IU[B * 64 + r * 8 + u] = y[r];
		}
	}
	__syncthreads();
	const int r8 = tid >> 6, jc = tid & 63;
	// L11 Li = I by block rows, top down: Li_B = IL_B (E_B - sum_{C<B} L_BC Li_C)
	for (int B = 0; B < 8; ++B) {
		const int t = 8 * B + r8;
		uint64_t acc = 0;
		for (int qb = 0; qb < B; ++qb)
#pragma unroll
			for (int u = 0; u < 8; ++u)
				acc += prod(Ls[t * XP + 8 * qb + u], Li[(8 * qb + u) * XP + jc]);
		Rt[r8 * XP + jc] = mp.sub(t == jc ? 1u : 0u, (uint32_t) mp.reduce(acc));
		__syncthreads();
		acc = 0;
#pragma unroll
		for (int u = 0; u < 8; ++u)
			acc += prod(IL[B * 64 + r8 * 8 + u], Rt[u * XP + jc]);
		Li[t * XP + jc] = (uint32_t) mp.reduce(acc);
		__syncthreads();
	}
	// U11 M = Li by block rows, bottom up: M_B = IU_B (Li_B - sum_{C>B} U_BC M_C)
	for (int B = 7; B >= 0; --B) {
		const int t = 8 * B + r8;
		uint64_t acc = 0;
		for (int qb = B + 1; qb < 8; ++qb)
#pragma unroll
			for (int u = 0; u < 8; ++u)
				acc += prod(Us[t * XP + 8 * qb + u], Ms[(8 * qb + u) * XP + jc]);
		Rt[r8 * XP + jc] = mp.sub(Li[t * XP + jc], (uint32_t) mp.reduce(acc));
		__syncthreads();
		acc = 0;
#pragma unroll
		for (int u = 0; u < 8; ++u)
			acc += prod(IU[B * 64 + r8 * 8 + u], Rt[u * XP + jc]);
		Ms[t * XP + jc] = (uint32_t) mp.reduce(acc);
		__syncthreads();
	}
	for (int idx = tid; idx < LU_W * LU_W; idx += LU_NT) {
		const int t = idx / LU_W, s_ = idx % LU_W;
		LiM[idx] = t < rp ? Li[t * XP + s_] : 0u;
		LiM[LU_W * LU_W + idx] = t < rp ? Ms[t * XP + s_] : 0u;
	}
}

// Rows i0 .. i0+kw-1 of Y (ld ncols) put in the window's new order, through
// the temporary T (kw x ncols): first gather, then copy back.
template <typename TY>
__global__ void lu_window_gather_kernel(const TY *__restrict__ Y, size_t ld, int ncols, int i0,
					const LuPanel *__restrict__ st, TY *__restrict__ T)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
	if (col < ncols)
		T[(size_t) q * ncols + col] = Y[(size_t) (i0 + st->wperm[q]) * ld + col];
}

template <typename TY>
__global__ void lu_window_put_kernel(TY *__restrict__ Y, size_t ld, int ncols, int i0, const TY *__restrict__ T)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
	if (col < ncols)
		Y[(size_t) (i0 + q) * ld + col] = T[(size_t) q * ncols + col];
}

// orig[i0 + q] <- orig[i0 + wperm[q]] (one workgroup of LU_K threads).
__global__ void lu_window_orig_kernel(int *__restrict__ orig, int i0, int kw, const LuPanel *__restrict__ st)
{
	__shared__ int old[LU_K];
	const int q = threadIdx.x;
	if (q < kw)
		old[q] = orig[i0 + q];
	__syncthreads();
	if (q < kw)
		orig[i0 + q] = old[st->wperm[q]];
}

// pluq: the pivot rows' own multipliers, Lb[i0 + t][i0 + s] = L11[t][s] (s < t < rp).
__global__ void lu_pivot_mult_kernel(const uint32_t *__restrict__ L11, int i0, const LuPanel *__restrict__ st,
				     uint32_t *__restrict__ Lb, size_t ldl)
{
	const int t = blockIdx.x, s_ = threadIdx.x;
	if (t < st->rp && s_ < t)
		Lb[(size_t) (i0 + t) * ldl + i0 + s_] = L11[t * LU_W + s_];
}

// The new pivot rows: X[i0 + t][c + col] = U[t][col] for t < rp.
__global__ void lu_scatter_pivots_kernel(uint32_t *__restrict__ X, size_t ldx, int c, int nc, int i0,
					 const LuPanel *__restrict__ st, const uint32_t *__restrict__ U)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
	if (col >= nc || t >= st->rp)
		return;
	X[(size_t) (i0 + t) * ldx + c + col] = U[(size_t) t * nc + col];
}

// G[q][t] = X[i1 + q][J_t] for t < rp, zero for t >= rp (G is mb x LU_W).
// Rows i1 .. i1+mb-1 with i1 >= i0; those at positions below i0 + rp (the new
// pivot rows) get zeros, so an update from i0 on leaves them unchanged.
__global__ void lu_gather_mult_kernel(const uint32_t *__restrict__ X, size_t ldx, int mb, int i1, int i0,
				      const LuPanel *__restrict__ st, uint32_t *__restrict__ G)
{
	const int q = blockIdx.x * (blockDim.x / LU_W) + threadIdx.x / LU_W, t = threadIdx.x % LU_W;
	if (q >= mb)
		return;
	const int rp = st->rp;
	G[(size_t) q * LU_W + t] = (t < rp && i1 + q >= i0 + rp) ? X[(size_t) (i1 + q) * ldx + st->J[t]] : 0u;
}

// The panel's update for p <= 2^16, fused (the INT8 rank update follows).
// Workgroup j owns the columns [c + TC b, c + TC b + TC), b = cb0 + j, of the
// window rows:
// it reads them in the window's new order (position q: window row wperm[q]),
// forms [L11^-1; P_J^-1] P on v_mfma_f64_16x16x4f64 (64 terms below 2^32:
// exact in a double), writes the new pivot rows and the remaining window rows
// back in that order, and packs W = P_J^-1 P as the update's INT8 B fragments
// (KB = 2).  Every read of a slice precedes its writes, and slices are
// disjoint, so the reordering is done in place.
template <int TC, typename TX>
__global__ __launch_bounds__(256) void lu_wb_i8_kernel(TX *__restrict__ X, size_t ldx, int nc, int c, int i0,
						       int kw, const LuPanel *__restrict__ st,
						       const uint32_t *__restrict__ LiM, int8_t *__restrict__ Wp, ModP mp,
						       int cb0)
{
	static_assert(TC == 32, "column slice of 32");
	__shared__ uint32_t Ps[LU_K][TC + 1], Pm[2 * LU_W][LU_W + 1], Os[2 * LU_W][TC + 1];
	__shared__ int wq[LU_K];
	const uint32_t p = (uint32_t) mp.p;
	const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
	const int cbk = cb0 + (int) blockIdx.x;     // the slice's block index (W fragments too)
	const int j0 = cbk * TC;
	const int rp = st->rp;
	// LiM (32 KB) as 16-byte loads, issued before the window order is waited
	// for, so that its latency overlaps the dependent loads of the rows
	{
		uint4 lv[2 * LU_W * LU_W / 4 / 256];
#pragma unroll
		for (int u = 0; u < 2 * LU_W * LU_W / 4 / 256; ++u)
			lv[u] = reinterpret_cast<const uint4 *>(LiM)[tid + 256 * u];
		if (tid < LU_K)
			wq[tid] = st->wperm[tid];
#pragma unroll
		for (int u = 0; u < 2 * LU_W * LU_W / 4 / 256; ++u) {
			const int e = 4 * (tid + 256 * u), r = e / LU_W, cc = e % LU_W;
			Pm[r][cc] = lv[u].x;
			Pm[r][cc + 1] = lv[u].y;
			Pm[r][cc + 2] = lv[u].z;
			Pm[r][cc + 3] = lv[u].w;
		}
	}
	__syncthreads();
	for (int e = tid; e < LU_K * TC; e += 256) {
		const int q = e / TC, col = e % TC, gc = j0 + col;
		Ps[q][col] = (q < kw && gc < nc) ? (uint32_t) X[(size_t) (i0 + wq[q]) * ldx + c + gc] : 0u;
	}
	__syncthreads();
	// wave w: output rows 32w .. 32w+31 as two 16-row blocks, TC/16 column
	// blocks each.  A fragment: lane l holds (row l%16, k l/16); B fragment:
	// (k l/16, column l%16); result register r: (row l/16 + 4r, column l%16).
	{
		constexpr int NJ = TC / 16;
		typedef double v4d __attribute__((ext_vector_type(4)));
		v4d acc[2][NJ];
#pragma unroll
		for (int b = 0; b < 2; ++b)
#pragma unroll
			for (int j = 0; j < NJ; ++j)
				acc[b][j] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
		for (int s_ = 0; s_ < LU_W; s_ += 4) {
			const int kk = s_ + (lane >> 4);
			const double a0 = (double) Pm[32 * w + (lane & 15)][kk];
			const double a1 = (double) Pm[32 * w + 16 + (lane & 15)][kk];
#pragma unroll
			for (int j = 0; j < NJ; ++j) {
				const double bv = (double) Ps[kk][16 * j + (lane & 15)];
				acc[0][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv, acc[0][j], 0, 0, 0);
				acc[1][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv, acc[1][j], 0, 0, 0);
			}
		}
#pragma unroll
		for (int b = 0; b < 2; ++b)
#pragma unroll
			for (int j = 0; j < NJ; ++j)
#pragma unroll
				for (int r = 0; r < 4; ++r)
					Os[32 * w + 16 * b + (lane >> 4) + 4 * r][16 * j + (lane & 15)] = mp.mod_f64(acc[b][j][r]);
	}
	__syncthreads();
	for (int e = tid; e < LU_K * TC; e += 256) {
		const int q = e / TC, col = e % TC, gc = j0 + col;
		if (q < kw && gc < nc)
			X[(size_t) (i0 + q) * ldx + c + gc] = (TX) (q < rp ? Os[q][col] : Ps[q][col]);
	}
	// B fragments: piece (column block, k block q, lane) holds column
	// 32 cb + lane%32 and k = 32q + 16(lane/32) .. +15
	if (tid < 128) {
		const int q = tid >> 6, cc = lane & 31, t = 32 * q + 16 * (lane >> 5);
		uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
		for (int b = 0; b < 16; ++b) {
			const uint32_t lb = i8_limbs(Os[LU_W + t + b][cc], p);
			hw[b >> 2] |= ((lb >> 8) & 0xffu) << (8 * (b & 3));
			lw[b >> 2] |= (lb & 0xffu) << (8 * (b & 3));
		}
		int8_t *fr = Wp + ((size_t) cbk * 2 + q) * 2 * 1024 + lane * 16;
		*(uint4 *) fr = make_uint4(hw[0], hw[1], hw[2], hw[3]);
		*(uint4 *) (fr + 1024) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
	}
}

// G for the update of rows i0 .. i0+mb-1: G[r][t] = X[i0 + r][J_t] for t < rp,
// zero on the new pivot rows (r < rp), packed as the update's INT8 A fragments
// (KB = 2).  Thread: the 16-byte piece (row block, k block, lane) of one
// workgroup's 64 rows.
template <typename TX>
__global__ __launch_bounds__(256) void lu_gpack_i8_kernel(const TX *__restrict__ X, size_t ldx, int mb, int i0,
							  int c, const LuPanel *__restrict__ st, int8_t *__restrict__ Gp,
							  uint32_t p)
{
	const int tid = threadIdx.x, lane = tid & 63, q = (tid >> 6) & 1, rb = 2 * blockIdx.x + (tid >> 7);
	const int rp = st->rp;
	const int r = 32 * rb + (lane & 31), s0 = 32 * q + 16 * (lane >> 5);
	const bool live = r < mb && r >= rp;
	const TX *row = X + (size_t) (i0 + r) * ldx;
	uint32_t xv[16];
	// a full-rank panel has J_t = c + t: sixteen consecutive entries, 16-byte
	// loads when the row stride and c keep them aligned
	constexpr int EPL = 16 / (int) sizeof(TX);
	if (rp == LU_W && st->J[LU_W - 1] == c + LU_W - 1 && (ldx % EPL) == 0 && (c % EPL) == 0) {
		const uint4 *src = (const uint4 *) (row + c + s0);
#pragma unroll
		for (int qq = 0; qq < 16 / EPL; ++qq) {
			const uint4 u = live ? src[qq] : make_uint4(0u, 0u, 0u, 0u);
			const uint32_t wv[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
			for (int e = 0; e < EPL; ++e)
				xv[EPL * qq + e] = EPL == 4 ? wv[e] : (wv[e >> 1] >> (16 * (e & 1))) & 0xffffu;
		}
	} else {
#pragma unroll
		for (int b = 0; b < 16; ++b)
			xv[b] = (live && s0 + b < rp) ? (uint32_t) row[st->J[s0 + b]] : 0u;
	}
	uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
	for (int b = 0; b < 16; ++b) {
		const uint32_t lb = i8_limbs(xv[b], p);
		hw[b >> 2] |= ((lb >> 8) & 0xffu) << (8 * (b & 3));
		lw[b >> 2] |= (lb & 0xffu) << (8 * (b & 3));
	}
	int8_t *fr = Gp + ((size_t) rb * 2 + q) * 2 * 1024 + lane * 16;
	*(uint4 *) fr = make_uint4(hw[0], hw[1], hw[2], hw[3]);
	*(uint4 *) (fr + 1024) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
}

// First nonzero at positions >= i0, columns [c0, c1): key = col << 32 | pos.
template <typename TX>
__global__ void lu_search_kernel(const TX *__restrict__ X, size_t ldx, int m, int i0, int c0, int c1,
				 unsigned long long *__restrict__ key)
{
	const int col = c0 + blockIdx.x * blockDim.x + threadIdx.x;
	if (col >= c1)
		return;
	for (int pos = i0 + blockIdx.y; pos < m; pos += gridDim.y) {
		if (X[(size_t) pos * ldx + col] != 0) {
			atomicMin(key, ((unsigned long long) col << 32) | (unsigned) pos);
			return;
		}
	}
}

__global__ void lu_key_init_kernel(unsigned long long *key) { *key = ~0ull; }

// Swap rows a and b of Y (ncols entries) -- one transposition.
template <typename TY>
__global__ void lu_swap_rows_kernel(TY *__restrict__ Y, size_t ld, int ncols, int a, int b)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x;
	if (col >= ncols)
		return;
	const TY t = Y[(size_t) a * ld + col];
	Y[(size_t) a * ld + col] = Y[(size_t) b * ld + col];
	Y[(size_t) b * ld + col] = t;
}

__global__ void lu_iota_kernel(int *orig, int m)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < m)
		orig[i] = i;
}

// Working copy: canonical residues of A in columns 0 .. n-1, for compute_L the
// identity in columns n .. n+m-1.
// Columns ntot .. ldx-1 (the padding of a row) are zero.
template <int E, typename TX>
__global__ void lu_load_kernel(const typename Elem<E>::T *__restrict__ A, size_t lda, int m, int n, int ntot, uint32_t p,
			       TX *__restrict__ X, size_t ldx)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, row = blockIdx.y;
	if ((size_t) col >= ldx)
		return;
	X[(size_t) row * ldx + col] = (TX) (col < n ? Elem<E>::load(A + (size_t) row * lda + col, p)
						    : (col - n == row && col < ntot ? 1u : 0u));
}

// Output row t.  mode U: the working copy.  compute_L: U from the diagonal
// on, T[t][s] = X[t][n + orig[s]] left of it.  pluq: L left of the diagonal
// (first `rank` columns), U in the column order corder.
template <int E, typename TX>
__global__ void lu_store_kernel(const TX *__restrict__ X, size_t ldx, int m, int n, int rank, int mode,
				const int *__restrict__ orig, const int *__restrict__ corder,
				const uint32_t *__restrict__ L, size_t ldl, uint32_t p, typename Elem<E>::T *__restrict__ A,
				size_t lda)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
	if (col >= n)
		return;
	const size_t row = (size_t) t;
	uint32_t v;
	if (mode == MLZ_ECH_U)
		v = X[row * ldx + col];
	else if (mode == MLZ_ECH_TRANSFORM)
		v = col < t ? X[row * ldx + n + orig[col]] : X[row * ldx + col];
	else
		v = col < t ? (col < rank ? L[row * ldl + col] : 0u) : X[row * ldx + corder[col]];
	Elem<E>::store(A + (size_t) t * lda + col, v, p);
}

// TX: the working copy's entry type.  uint16 serves the fused path (p <= 2^16,
// INT8 update, no pluq) and halves the bytes each panel's update moves; every
// other path keeps uint32.
template <typename TX>
static int echelon_impl(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, void *A, size_t lda, int mode,
			size_t *rank_out, uint64_t *det_out, uint32_t *perm_rows, uint32_t *perm_cols, int algo,
			bool store, uint32_t *pivcols)
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
	// rows of the working copy start on whole 64-entry boundaries; the columns
	// past ntot are zero, so the packed update can run over whole 64-column strips
	const size_t ldx = round_up(ntot, 64);
	const size_t rmax = std::min(m, n);
	const bool pluq = mode == MLZ_ECH_PLUQ;
	// workspace: X (m x ntot); the window temporary (LU_K x max(ntot, rmax)) and
	// UW (2 LU_W x ntot: the new pivot rows, then W = P_J^-1 P); G (m x LU_W)
	// and for pluq Lb (m x rmax); L11, LiM (2 LU_W x LU_W), Ui, orig, corder, state
	void *w0, *w1, *w2, *w3;
	MLZ_TRY(ws_get(ctx, WS_RREF_0, m * ldx * sizeof(TX), &w0));
	const size_t sXw = round_up(LU_K * std::max(ntot, rmax) * 4, 256), sUW = round_up(2 * LU_W * ntot * 4, 256);
	MLZ_TRY(ws_get(ctx, WS_RREF_1, sXw + sUW, &w1));
	const size_t sG = round_up(m * LU_W * 4, 256);
	MLZ_TRY(ws_get(ctx, WS_RREF_2, sG + (pluq ? m * rmax * 4 : 0), &w2));
	const size_t sq = round_up(LU_W * LU_W * 4, 256);
	const size_t small = 6 * sq + round_up(m * 4, 256) + round_up(n * 4, 256) + 2 * round_up(sizeof(LuPanel), 256) + 256;
	MLZ_TRY(ws_get(ctx, WS_RREF_3, small, &w3));
	TX *X = (TX *) w0;
	uint32_t *Xw = (uint32_t *) w1, *UW = (uint32_t *) ((char *) w1 + sXw);
	uint32_t *G = (uint32_t *) w2, *Lb = (uint32_t *) ((char *) w2 + sG);
	char *q = (char *) w3;
	// the panel state and LiM alternate between two copies (a panel's update
	// may still read them while the next panel's kernel runs)
	uint32_t *L11 = (uint32_t *) q; q += sq;
	uint32_t *LiMb[2];
	LiMb[0] = (uint32_t *) q; q += 2 * sq;
	LiMb[1] = (uint32_t *) q; q += 2 * sq;
	uint32_t *Ui = (uint32_t *) q; q += sq;
	int *orig = (int *) q; q += round_up(m * 4, 256);
	int *corder = (int *) q; q += round_up(n * 4, 256);
	LuPanel *stb[2];
	stb[0] = (LuPanel *) q; q += round_up(sizeof(LuPanel), 256);
	stb[1] = (LuPanel *) q; q += round_up(sizeof(LuPanel), 256);
	unsigned long long *key = (unsigned long long *) q;
	LuPanel *st = stb[0];
	uint32_t *LiM = LiMb[0];

	const uint16_t *invtab = nullptr;
	if (p <= 65536)
		MLZ_TRY(zp_inverse_table(ctx, mp, &invtab));
	{
		ProfScope ps(ctx, "lu_load");
		lu_iota_kernel<<<ceil_div(m, 256), 256, 0, ctx->stream>>>(orig, (int) m);
		dim3 g(ceil_div(ldx, 256), (unsigned) m);
		if (elem == MLZ_U16)
			lu_load_kernel<MLZ_U16, TX><<<g, 256, 0, ctx->stream>>>((const uint16_t *) A, lda, (int) m, (int) n,
									     (int) ntot, pp, X, ldx);
		else if (elem == MLZ_U32)
			lu_load_kernel<MLZ_U32, TX><<<g, 256, 0, ctx->stream>>>((const uint32_t *) A, lda, (int) m, (int) n,
									     (int) ntot, pp, X, ldx);
		else
			lu_load_kernel<MLZ_F64_SYM, TX><<<g, 256, 0, ctx->stream>>>((const double *) A, lda, (int) m, (int) n,
										 (int) ntot, pp, X, ldx);
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
	// The panel state comes back through pinned memory and an event, so the
	// host can queue the panel's update behind the window kernel first.
	struct Res {
		LuPanel *h = nullptr;
		hipEvent_t ev = nullptr, wev = nullptr;
		hipStream_t cs = nullptr;
		~Res()
		{
			if (cs)
				(void) hipStreamSynchronize(cs);
			if (wev)
				(void) hipEventDestroy(wev);
			if (ev)
				(void) hipEventDestroy(ev);
			if (h)
				(void) hipHostFree(h);
		}
	} res;
	MLZ_HIP(ctx, hipHostMalloc((void **) &res.h, sizeof(LuPanel), hipHostMallocDefault));
	MLZ_HIP(ctx, hipEventCreateWithFlags(&res.ev, hipEventDisableTiming));
	MLZ_HIP(ctx, hipEventCreateWithFlags(&res.wev, hipEventDisableTiming));
	// the state goes back on a stream of its own, so that the copy does not
	// hold up the panel's update queued behind the window kernel
	hipStream_t cs;
	MLZ_TRY(copy_stream(ctx, &cs));
	res.cs = cs;
	auto swap_rows = [&](int a, int b) -> int {
		lu_swap_rows_kernel<TX><<<ceil_div(ntot, 256), 256, 0, ctx->stream>>>(X, ldx, (int) ntot, a, b);
		lu_swap_rows_kernel<int><<<1, 1, 0, ctx->stream>>>(orig, 1, 1, a, b);
		if (pluq)
			lu_swap_rows_kernel<uint32_t><<<ceil_div(rmax, 256), 256, 0, ctx->stream>>>(Lb, rmax, (int) rmax, a, b);
		return check_launch(ctx, "lu_swap_rows_kernel");
	};
	// The update of the panel at (i, c).  With rp < 0 (not known yet) every
	// size is taken at its maximum: the rows from i on (the pivot rows get
	// zero multipliers), all LU_W rows of P (LiM is zero past rp).
	// p <= 2^16 without pluq: the fused kernels and the packed INT8 update,
	// split in two.  The narrow part (the first LU_NARROW columns from c, which
	// hold the next panel) runs on the context's stream; the wide part (the
	// rest) on the second stream, so that it overlaps the next panel's kernel.
	// The next narrow part waits for this wide part (their columns meet); the
	// packed operands alternate between two copies.
	const bool fused = upd_i8 && !pluq;
	int8_t *Gpk[2] = {nullptr, nullptr}, *Wpk[2] = {nullptr, nullptr};
	struct Overlap {
		hipStream_t s1 = nullptr;
		hipEvent_t nev[2] = {nullptr, nullptr}, uev[2] = {nullptr, nullptr};
		~Overlap()
		{
			if (s1)
				(void) hipStreamSynchronize(s1);
			for (int b = 0; b < 2; ++b)
				for (hipEvent_t e : {nev[b], uev[b]})
					if (e)
						(void) hipEventDestroy(e);
		}
	} ov;
	bool wide_live = false;      // a wide part is queued on the second stream
	int wide_b = 0;              // ... with this copy
	if (fused) {
		void *gq, *wq;
		const size_t gb = round_up(round_up(m, 64) * LU_W * 2, 256), wb = round_up(round_up(ntot, 64) * LU_W * 2, 256);
		MLZ_TRY(ws_get(ctx, WS_PACK_A, 2 * gb, &gq));
		MLZ_TRY(ws_get(ctx, WS_PACK_B, 2 * wb, &wq));
		Gpk[0] = (int8_t *) gq;
		Gpk[1] = (int8_t *) gq + gb;
		Wpk[0] = (int8_t *) wq;
		Wpk[1] = (int8_t *) wq + wb;
		hipStream_t s1;
		MLZ_TRY(aux_stream(ctx, &s1));
		ov.s1 = s1;
		for (int b = 0; b < 2; ++b) {
			MLZ_HIP(ctx, hipEventCreateWithFlags(&ov.nev[b], hipEventDisableTiming));
			MLZ_HIP(ctx, hipEventCreateWithFlags(&ov.uev[b], hipEventDisableTiming));
		}
	}
	// the context's stream waits for the queued wide part
	auto join_wide = [&]() -> int {
		if (wide_live)
			MLZ_HIP(ctx, hipStreamWaitEvent(ctx->stream, ov.uev[wide_b], 0));
		wide_live = false;
		return MLZ_OK;
	};
	// Columns c .. ldx-1, the zero padding included, so that the wide update
	// runs over whole 64-column strips (a separate launch for the last columns
	// was on the critical path: 24.3 -> 24.1 ms for the 5000 x 15000 D of the
	// 20000 x 30000 reduction).  Starting the wide part's W, or all of the wide
	// part, beside the narrow chain measured slower (25.8 and 24.5 ms): the
	// narrow chain is latency-bound and loses more to the shared CUs than the
	// overlap gains.
	auto update_fused = [&](int kw, int b) -> int {
		const size_t nc = ldx - c, mb = m - i;
		const size_t nn = std::min(nc, (size_t) LU_NARROW);
		const int nbn = (int) (nn / 32);
		const bool wide = nc > nn;
		MLZ_TRY(join_wide());
		{
			ProfScope ps(ctx, "lu_update");
			if (mode == MLZ_ECH_TRANSFORM)   // the row origins matter for T only
				lu_window_orig_kernel<<<1, LU_K, 0, ctx->stream>>>(orig, (int) i, kw, st);
			lu_wb_i8_kernel<32, TX><<<ceil_div(nn, 32), 256, 0, ctx->stream>>>(X, ldx, (int) nc, (int) c, (int) i, kw,
											    st, LiM, Wpk[b], mp, 0);
			lu_gpack_i8_kernel<<<ceil_div(mb, 64), 256, 0, ctx->stream>>>(X, ldx, (int) mb, (int) i, (int) c, st,
											Gpk[b], pp);
			MLZ_TRY(check_launch(ctx, "lu_gpack_i8_kernel"));
			MLZ_TRY(zp_rank_update_i8_packed(ctx, pp, mb, nn, 2, Gpk[b], Wpk[b], X + i * ldx + c, ldx));
			if (wide) {
				MLZ_HIP(ctx, hipEventRecord(ov.nev[b], ctx->stream));
				MLZ_HIP(ctx, hipStreamWaitEvent(ov.s1, ov.nev[b], 0));
				lu_wb_i8_kernel<32, TX><<<ceil_div(nc, 32) - nbn, 256, 0, ov.s1>>>(X, ldx, (int) nc, (int) c, (int) i, kw,
												    st, LiM, Wpk[b], mp, nbn);
				MLZ_TRY(zp_rank_update_i8_packed(ctx, pp, mb, nc - nn, 2, Gpk[b], Wpk[b] + (size_t) nbn * 4096,
								 X + i * ldx + c + nn, ldx, ov.s1));
				MLZ_HIP(ctx, hipEventRecord(ov.uev[b], ov.s1));
				wide_live = true;
				wide_b = b;
			}
		}
		return MLZ_OK;
	};
	auto update = [&](int kw, int rp) -> int {
		if constexpr (sizeof(TX) != sizeof(uint32_t)) {   // not reached: narrow copies are fused
			(void) kw;
			(void) rp;
			return set_error(ctx, MLZ_E_ARG, "lu: the general update needs a uint32 working copy");
		} else {
			const size_t nc = ntot - c;
			const size_t rb = rp < 0 ? i : i + (size_t) rp, mb = m - rb;
			const size_t kk = rp < 0 ? (size_t) kw : (size_t) rp;
			ProfScope ps(ctx, "lu_update");
			// the window rows in their new order (the working copy and, for pluq, the multipliers)
			lu_window_gather_kernel<uint32_t><<<dim3(ceil_div(ntot, 256), kw), 256, 0, ctx->stream>>>(
				X, ldx, (int) ntot, (int) i, st, Xw);
			lu_window_put_kernel<uint32_t><<<dim3(ceil_div(ntot, 256), kw), 256, 0, ctx->stream>>>(
				X, ldx, (int) ntot, (int) i, Xw);
			if (mode == MLZ_ECH_TRANSFORM)
				lu_window_orig_kernel<<<1, LU_K, 0, ctx->stream>>>(orig, (int) i, kw, st);
			if (pluq && i > 0) {
				lu_window_gather_kernel<uint32_t><<<dim3(ceil_div(rmax, 256), kw), 256, 0, ctx->stream>>>(
					Lb, rmax, (int) rmax, (int) i, st, Xw);
				lu_window_put_kernel<uint32_t><<<dim3(ceil_div(rmax, 256), kw), 256, 0, ctx->stream>>>(
					Lb, rmax, (int) rmax, (int) i, Xw);
			}
			MLZ_TRY(check_launch(ctx, "lu_window_put_kernel"));
			// UW = [L11^-1; P_J^-1] P: the new pivot rows and W
			MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, 2 * LU_W, nc, kk, 1, LiM, LU_W, X + i * ldx + c, ldx, 0, UW, nc,
					    algo));
			if (mb) {
				lu_gather_mult_kernel<<<ceil_div(mb, 256 / LU_W), 256, 0, ctx->stream>>>(X, ldx, (int) mb, (int) rb,
													   (int) i, st, G);
				MLZ_TRY(check_launch(ctx, "lu_gather_mult_kernel"));
			}
			if (pluq) {
				if (mb) {
					// the rows below: x U11 = X_j[J], U11^-1 = P_J^-1 L11
					MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, LU_W, LU_W, LU_W, 1, LiM + LU_W * LU_W, LU_W, L11, LU_W, 0,
							    Ui, LU_W, algo));
					MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, mb, kk, kk, 1, G, LU_W, Ui, LU_W, 0, Lb + rb * rmax + i,
							    rmax, algo));
				}
				lu_pivot_mult_kernel<<<LU_W, LU_W, 0, ctx->stream>>>(L11, (int) i, st, Lb, rmax);
				MLZ_TRY(check_launch(ctx, "lu_pivot_mult_kernel"));
			}
			if (mb) {
				uint32_t *Xb = X + rb * ldx + c;
				const uint32_t *Wp = UW + (size_t) LU_W * nc;
				if (upd_i8)
					MLZ_TRY(zp_rank_update_i8_dev(ctx, pp, mb, nc, LU_W, G, LU_W, Wp, nc, Xb, ldx));
				else
					MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, mb, nc, LU_W, p - 1, G, LU_W, Wp, nc, 1, Xb, ldx, algo));
			}
			lu_scatter_pivots_kernel<<<dim3(ceil_div(nc, 256), LU_W), 256, 0, ctx->stream>>>(X, ldx, (int) c, (int) nc,
													(int) i, st, UW);
			return check_launch(ctx, "lu_scatter_pivots_kernel");
		}
	};
	int panel = 0;
	while (i < m && c < n) {
		const int wc = (int) std::min((size_t) LU_W, n - c);
		const int kw = (int) std::min((size_t) LU_K, m - i);
		const int pb = fused ? (panel & 1) : 0;
		++panel;
		st = stb[pb];
		LiM = LiMb[pb];
		{
			ProfScope ps(ctx, "lu_window");
			if constexpr (sizeof(TX) != sizeof(uint32_t))   // p <= 2^16: the table is there
				lu_window_kernel<true, true, TX><<<1, LU_NT, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, wc,
												(int) i, mp, invtab, st, L11, LiM);
			else if (invtab)
				lu_window_kernel<true, true, TX><<<1, LU_NT, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, wc,
												(int) i, mp, invtab, st, L11, LiM);
			else if (mp.small)
				lu_window_kernel<true, false, TX><<<1, LU_NT, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, wc,
												 (int) i, mp, invtab, st, L11, LiM);
			else
				lu_window_kernel<false, false, TX><<<1, LU_NT, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, wc,
												  (int) i, mp, invtab, st, L11, LiM);
		}
		MLZ_TRY(check_launch(ctx, "lu_window_kernel"));
		MLZ_HIP(ctx, hipEventRecord(res.wev, ctx->stream));
		MLZ_HIP(ctx, hipStreamWaitEvent(cs, res.wev, 0));
		MLZ_HIP(ctx, hipMemcpyAsync(res.h, st, sizeof(LuPanel), hipMemcpyDeviceToHost, cs));
		MLZ_HIP(ctx, hipEventRecord(res.ev, cs));
		// pluq's sizes depend on rp; otherwise the update is queued at once
		const bool early = !pluq;
		if (early)
			MLZ_TRY(fused ? update_fused(kw, pb) : update(kw, -1));
		MLZ_HIP(ctx, hipEventSynchronize(res.ev));
		const LuPanel &h = *res.h;
		if (h.bad) {
			MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
			if (h.bad == 2)
				return set_error(ctx, MLZ_E_HIP, "lu_window_kernel: a wave stopped waiting for a column");
			return set_error(ctx, MLZ_E_NOT_INVERTIBLE, "Could not invert pivot-element in the ring");
		}
		const int rp = h.rp;
		for (int s = 0; s < rp; ++s) {
			steps.push_back(s == 0 && pending >= 0 ? (int) pending : (int) i + h.steps[s]);
			pcols.push_back(h.J[s]);
			det = mp.mul((uint32_t) det, h.x[s]);
		}
		if (rp > 0) {
			pending = -1;
			if (!early)
				MLZ_TRY(update(kw, rp));
		}
		i += rp;
		if (!h.stop) {
			c += wc;
			continue;
		}
		// the window ran out: the next pivot (first column, then first
		// position) among all remaining rows, searched in column chunks
		c = h.stop_col;
		MLZ_TRY(join_wide());
		bool found = false;
		while (i < m && c < n) {
			const size_t c1 = std::min(n, c + LU_SEARCH_COLS);
			unsigned long long k = ~0ull;
			lu_key_init_kernel<<<1, 1, 0, ctx->stream>>>(key);
			lu_search_kernel<<<dim3(ceil_div(c1 - c, 256), (unsigned) std::min<size_t>(m - i, 64)), 256, 0,
					   ctx->stream>>>(X, ldx, (int) m, (int) i, (int) c, (int) c1, key);
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
				MLZ_TRY(swap_rows((int) i, (int) pos));
				pending = (long) pos;
			}
			break;
		}
		if (!found)
			break;
	}
	MLZ_TRY(join_wide());
	const size_t rank = steps.size();
	if (perm_rows)
		for (size_t s = 0; s < rank; ++s)
			perm_rows[s] = (uint32_t) steps[s];
	if (pivcols)
		for (size_t s = 0; s < rank; ++s)
			pivcols[s] = (uint32_t) pcols[s];
	// pluq's column order: the pivot columns swapped into place step by step
	std::vector<int> co(n), where(n);
	for (size_t j = 0; j < n; ++j)
		co[j] = where[j] = (int) j;
	for (size_t s = 0; s < rank; ++s) {
		const int cpos = where[pcols[s]];
		if (perm_cols)
			perm_cols[s] = (uint32_t) cpos;
		const int a = co[s], b = co[cpos];
		co[s] = b;
		co[cpos] = a;
		where[b] = (int) s;
		where[a] = cpos;
	}
	if (pluq && store)
		MLZ_HIP(ctx, hipMemcpyAsync(corder, co.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	if (store) {
		ProfScope ps(ctx, "lu_store");
		dim3 g(ceil_div(n, 256), (unsigned) m);
#define STORE(EE, TT) lu_store_kernel<EE, TX><<<g, 256, 0, ctx->stream>>>(X, ldx, (int) m, (int) n, (int) rank, mode, \
								     orig, corder, Lb, rmax, pp, (TT *) A, lda)
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

int zp_echelon_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, void *A, size_t lda, int mode,
		   size_t *rank_out, uint64_t *det_out, uint32_t *perm_rows, uint32_t *perm_cols, int algo, bool store,
		   uint32_t *pivcols)
{
	const bool fused = p <= 65536 && (algo == MLZ_ALGO_AUTO || algo == MLZ_ALGO_INT8) && mode != MLZ_ECH_PLUQ;
	return fused ? echelon_impl<uint16_t>(ctx, elem, p, m, n, A, lda, mode, rank_out, det_out, perm_rows, perm_cols,
					      algo, store, pivcols)
		     : echelon_impl<uint32_t>(ctx, elem, p, m, n, A, lda, mode, rank_out, det_out, perm_rows, perm_cols,
					      algo, store, pivcols);
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

