// lela_amd/csrc/zp_rref.hip -- reduced row-echelon form over Z/p on MI355X.
//
// Replaces the pivot-row reduction of Elimination::echelonize_reduced
// (lela/algorithms/elimination.tcc:141-264) and GaussJordan::echelonize_reduced
// (lela/algorithms/gauss-jordan.tcc:413-466) behind EchelonForm::echelonize
// (lela/solutions/echelon-form.h:103-154).  The reduced row-echelon form, the
// rank and the pivot columns are unique, so any pivot order gives LELA's
// result; this one is built for the GPU:
//
// Blocked Gauss-Jordan over column panels of RREF_W columns.  Per panel
//  1. pivot search among the rows that are not pivots yet, restricted to the
//     panel: one workgroup reduces the first RREF_W such rows together with
//     the identity; if that reaches full panel rank its rows are the pivot
//     rows and step 2's inverse comes out of the same elimination (if it
//     covers every remaining row, its pivots are the answer).  Otherwise a
//     tournament in one kernel: each
//     workgroup reduces a 128-row chunk and keeps its pivot rows, and
//     workgroups merge pairwise up a tree (the second to arrive at a node
//     merges, via an atomic arrival counter).  Either way the chosen rows span
//     the panel's row space, so their pivot columns J are its rank profile.
//  2. W = P_J^-1 * P (P = the k pivot rows from column c on, P_J its columns
//     J): one workgroup inverts P_J (unless step 1 did), a gemm forms W
//     (identity on J).
//  3. every other row i: A_i <- A_i - A_i[J] * W, one gemm over all rows (the
//     new pivot rows get a zero multiplier row; for p <= 2^16 a rank-64
//     update on INT8 limbs), then W is written into the pivot rows.
// Eliminations inside a workgroup keep the matrix in registers: thread
// (g, c) owns column c of the rows of group g, so each thread's updates are
// independent.  Steps are fraction free (no inverse on the critical path),
// except in step 1's fast path for p <= 2^16: there the pivot row is
// normalised with an inverse from a table in LDS and the other rows are
// updated without reduction (one fma per entry).  Pivots are checked for
// invertibility (LELA's check, lela/algorithms/elimination.tcc:217-220).
// For n >= 2 * RREF_NB the panels run over RREF_NB-column blocks: a panel
// updates only its block, and the columns right of the block get one gemm
// per block (512 columns measured best of 256, 512 and 768 at 5000 x 15000).
// Rank, pivot rows and pivot columns stay on the device; the host only checks
// the rank every few panels to stop early.  A final kernel writes the pivot
// rows in pivot-column order followed by zero rows (standard layout).
#include "batch.hpp"

#include <climits>
#include <type_traits>

namespace mlz {

constexpr int RREF_W = 64;        // panel width (columns)
constexpr size_t RREF_NB = 512;   // column block of the blocked trailing update
constexpr int RREF_CHUNK = 128;   // rows per tournament leaf (= two merged lists)
constexpr int RREF_MAXLVL = 24;

// Device-side state of one RREF call.
struct RrefState {
	int *is_pivot;       // [m]: 0 = not a pivot row, else 1 + panel index
	int *rank;           // [1]
	int *piv_rows;       // [min(m,n)] in pivot-column order
	int *piv_cols;       // [min(m,n)]
	int *k;              // [1] pivots found in the current panel
	int *J;              // [RREF_W] panel-relative pivot columns
	int *prow;           // [RREF_W] their rows
	int *bad;            // [1] set when a pivot is not invertible (composite p)
	int *need_tree;      // [1] the fast path did not settle the panel
	int *inv_done;       // [1] the fast path also produced P_J^-1
	int *rank0;          // [1] rank when the current column block began
	int *lists_rows;     // [levels][nodes0][RREF_W]
	int *lists_cnt;      // [levels][nodes0]
	int *counters;       // [levels][nodes0]
};

// Elimination with the matrix in registers.  The block has NG = blockDim/NC
// row groups of RPG rows; thread (g = tid/NC, c = tid%NC) holds v[i] = entry
// (row g*RPG + i, column c).  Columns 0..nsearch-1 are processed in order; the
// pivot of a column is the live row of lowest index with a nonzero there.
// reduced = false: live rows are cleared below/above (echelon form, values not
// normalised).  reduced = true: Gauss-Jordan, every other row is cleared and
// the pivot row scaled to 1.  Returns the pivot count; out_rows/out_cols list
// the pivots in order.  alive[] must be set for rows < nrows on entry.
template <int NC, int RPG, bool SMALL>
__device__ int reg_eliminate(uint32_t (&v)[RPG], int nrows, int nsearch, bool reduced, const ModP &mp,
			     int *alive, uint32_t *f, uint32_t *pivrow, int *s_min, int *out_rows, int *out_cols,
			     int *bad)
{
	// Fraction-free steps: a row r that has x_r != 0 in the pivot column
	// becomes x * row_r - x_r * pivot_row, x the pivot value (x is invertible
	// for prime p, so row spaces and ranks are unchanged).  No inverse sits
	// on the critical path; reduced mode normalises the pivot rows once at
	// the end, all inverses in parallel.  s_min is two words used in turn.
	const int tid = threadIdx.x, c = tid % NC, g = tid / NC, rbase = g * RPG;
	if (tid == 0) {
		s_min[0] = INT_MAX;
		s_min[1] = INT_MAX;
	}
	__syncthreads();
	int cnt = 0;
	for (int col = 0; col < nsearch; ++col) {
		if (c == col) {
			int found = INT_MAX;
#pragma unroll
			for (int i = RPG - 1; i >= 0; --i) {
				const int r = rbase + i;
				if (r < nrows && v[i] != 0 && alive[r])
					found = r;
			}
			if (found != INT_MAX)
				atomicMin(&s_min[col & 1], found);
		}
		__syncthreads();
		const int pr = s_min[col & 1];
		if (tid == 0)
			s_min[(col + 1) & 1] = INT_MAX;
		if (pr != INT_MAX) {
			if (pr / RPG == g) {
				uint32_t val = 0;
#pragma unroll
				for (int i = 0; i < RPG; ++i)
					if (rbase + i == pr)
						val = v[i];
				pivrow[c] = val;
			}
			if (c == col) {
#pragma unroll
				for (int i = 0; i < RPG; ++i) {
					const int r = rbase + i;
					f[r] = (r != pr && r < nrows && (reduced || alive[r])) ? v[i] : 0u;
				}
			}
			if (tid == 0) {
				alive[pr] = 0;
				out_rows[cnt] = pr;
				out_cols[cnt] = col;
			}
		}
		__syncthreads();
		if (pr == INT_MAX)
			continue;
		// In echelon mode columns up to col are never read again (the pivot
		// test of a later column only reads that column), so they are not
		// updated.
		if (reduced || c > col) {
			const uint32_t pc = pivrow[c], x = pivrow[col];
#pragma unroll
			for (int i = 0; i < RPG; ++i) {
				const uint32_t fr = f[rbase + i];
				if (fr)
					v[i] = mp.mul_subT<SMALL>(x, v[i], fr, pc);
			}
		}
		++cnt;
	}
	__syncthreads();
	if (reduced) {
		// pivrow -> column-to-pivot map, f -> inverse of each pivot row's lead
		int *colpos = (int *) pivrow;
		for (int t = tid; t < NC; t += blockDim.x)
			colpos[t] = -1;
		__syncthreads();
		for (int s = tid; s < cnt; s += blockDim.x)
			colpos[out_cols[s]] = s;
		__syncthreads();
		const int s = colpos[c];
		if (s >= 0 && out_rows[s] / RPG == g) {
			uint32_t d = 0;
#pragma unroll
			for (int i = 0; i < RPG; ++i)
				if (rbase + i == out_rows[s])
					d = v[i];
			const uint32_t di = mp.inv(d);
			if (di == 0)
				*bad = 1;
			f[out_rows[s]] = di;
		}
		__syncthreads();
#pragma unroll
		for (int i = 0; i < RPG; ++i) {
			const int r = rbase + i;
			if (r < nrows && !alive[r])
				v[i] = mp.mulT<SMALL>(v[i], f[r]);
		}
		__syncthreads();
	}
	return cnt;
}

// Gauss-Jordan with lazy reduction, for p <= 2^16 (reduced mode only).  The
// entries are integer-valued doubles.  A chosen pivot row is normalised at
// once (times the inverse of its lead) and every other row r becomes
// v_r - f_r * g with f_r = v_r[col] mod p and g the normalised pivot row,
// both below p, so after s steps |v| < p + s * (p-1)^2 < 2^53 for s <= 64:
// one fma per entry and step and no reduction until the end (where v is
// reduced to [0, p)).
//
// The pivot search runs on wave 0 (lane r = row r) one step ahead: the
// owners of column col+2 publish it (Q) after their update of step col, and
// after step col's pivot row is known wave 0 forms column col+1's values
// itself, Q - f * g (the same fma the owners apply), reduces them, picks the
// lowest live nonzero row by a ballot, and its lane looks up the lead's
// inverse in inv_tab (LDS).  Two barriers per column: A (pivot, multipliers,
// inverse published) and B (normalised pivot row published).  Q, the
// multipliers (as doubles), the pivot row index and the inverse are double
// buffered by column parity.  A lead without inverse (composite p) sets
// *bad.  Requires nrows <= 64.
template <int NC, int RPG>
__device__ int reg_gj_wave(double (&v)[RPG], int nrows, int nsearch, const ModP &mp, double (*Q)[64],
			   double (*fd)[64], int *s_pr, uint32_t *s_y, uint32_t *pivrow, int *out_rows, int *out_cols,
			   int *bad, const uint16_t *inv_tab)
{
	const int tid = threadIdx.x, c = tid % NC, g = tid / NC, rbase = g * RPG;
	// wave 0's search over the column values x (lane = row); cnt_w and
	// alive_w are wave 0's own view
	bool alive_w = tid < nrows;
	int cnt_w = 0;
	auto search = [&](int col, double x) {
		const uint32_t xr = mp.mod_f64(x);
		const bool cand = alive_w && xr != 0;
		const unsigned long long bal = __ballot(cand);
		const int pr = bal ? __ffsll((long long) bal) - 1 : -1;
		const int b = col & 1;
		fd[b][tid] = tid != pr ? (double) xr : 0.0;
		if (tid == pr) {
			const uint32_t y = inv_tab[xr];
			s_y[b] = y;
			if (y == 0)
				*bad = 1;
			out_rows[cnt_w] = pr;
			out_cols[cnt_w] = col;
		}
		if (tid == 0)
			s_pr[b] = pr;
		if (pr >= 0) {
			alive_w = alive_w && tid != pr;
			++cnt_w;
		}
	};
	// prologue: columns 0 and 1 published, column 0 searched
	for (int q = 0; q < 2 && q < nsearch; ++q)
		if (c == q) {
#pragma unroll
			for (int i = 0; i < RPG; ++i)
				Q[q][rbase + i] = v[i];
		}
	__syncthreads();
	if (tid < 64)
		search(0, Q[0][tid]);
	int cnt = 0;
	for (int col = 0; col < nsearch; ++col) {
		const int b = col & 1;
		__syncthreads();                                   // A: step col's pivot is known
		const int pr = s_pr[b];
		if (pr >= 0 && pr / RPG == g) {
			const uint32_t y = s_y[b];
			double xv = 0.0;
#pragma unroll
			for (int i = 0; i < RPG; ++i)
				if (rbase + i == pr)
					xv = v[i];
			const uint32_t gv = mp.mul_small(mp.mod_f64(xv), y);
			pivrow[c] = gv;
#pragma unroll
			for (int i = 0; i < RPG; ++i)
				if (rbase + i == pr)
					v[i] = (double) gv;
		}
		__syncthreads();                                   // B: the pivot row is published
		if (pr >= 0) {
			const double gc = (double) pivrow[c];
#pragma unroll
			for (int i = 0; i < RPG; ++i)
				v[i] = fma(-fd[b][rbase + i], gc, v[i]);
			++cnt;
		}
		if (col + 2 < nsearch && c == col + 2) {
#pragma unroll
			for (int i = 0; i < RPG; ++i)
				Q[b][rbase + i] = v[i];
		}
		if (col + 1 < nsearch && tid < 64) {
			// the pivot row's own entry is its normalised value
			double x = Q[(col + 1) & 1][tid];
			if (pr >= 0)
				x = tid == pr ? (double) pivrow[col + 1] : fma(-fd[b][tid], (double) pivrow[col + 1], x);
			search(col + 1, x);
		}
	}
#pragma unroll
	for (int i = 0; i < RPG; ++i)
		v[i] = (double) mp.mod_f64(v[i]);
	__syncthreads();
	return cnt;
}

// Inverse of an 8 x 8 block B (LDS, row stride ldb, residues below p) on one
// wave, by fraction-free Gauss-Jordan on [B | I] with row pivoting inside the
// block: lane 8r + c holds entries (r, c) and (r, 8 + c), rows exchange values
// by cross-lane reads, so no step needs a workgroup barrier.  On success
// T[8 j + s] = (B^-1)[j][s]; returns false (on every lane) when B is
// singular or a lead has no inverse (composite p).
__device__ bool wave_inv8(const uint32_t *B, int ldb, uint32_t *T, const ModP &mp, const uint16_t *inv_tab)
{
	const int lane = threadIdx.x & 63, r = lane >> 3, c = lane & 7;
	uint32_t m0 = B[r * ldb + c], m1 = r == c ? 1u : 0u;
	bool used = false;
	int pj[8];
#pragma unroll
	for (int j = 0; j < 8; ++j) {
		const unsigned long long bal = __ballot(c == j && !used && m0 != 0);
		if (bal == 0)
			return false;
		const int lp = __ffsll((long long) bal) - 1, pr = lp >> 3;
		pj[j] = pr;
		const uint32_t x = __shfl(m0, lp);
		const uint32_t fr = __shfl(m0, 8 * r + j);
		const uint32_t p0 = __shfl(m0, 8 * pr + c), p1 = __shfl(m1, 8 * pr + c);
		if (r == pr) {
			used = true;
		} else {
			m0 = mp.mul_subT<true>(x, m0, fr, p0);
			m1 = mp.mul_subT<true>(x, m1, fr, p1);
		}
	}
	// row pj[j] is the pivot of column j, with its lead in column j
	uint32_t lead = 0;
	int myj = 0;
#pragma unroll
	for (int j = 0; j < 8; ++j) {
		const uint32_t lj = __shfl(m0, 8 * pj[j] + j);
		if (r == pj[j]) {
			lead = lj;
			myj = j;
		}
	}
	const uint32_t y = inv_tab[lead];
	if (__ballot(y == 0))
		return false;
	T[8 * myj + c] = mp.mul_small(m1, y);
	return true;
}

// [S | I] -> [I | S^-1] for a 64 x 64 S by Gauss-Jordan over eight 8-column
// blocks (1024 threads; thread (g, c) holds column c of rows 8g .. 8g+7 as
// integer-valued doubles).  Block q: the block's rows and columns go to LDS
// reduced, one wave inverts the 8 x 8 diagonal block (wave_inv8), the block's
// rows become U = T * X[block rows] and every other row r gets
// X[r] - X[r, block] * U (8 products below p^2 each: after eight blocks
// |X| < p + 64 p^2 < 2^53, exact).  Row 8q + j ends as the pivot row of
// column 8q + j.  Returns false (uniformly, with v partly transformed) when
// a diagonal block is singular.
__device__ bool block_gj64(double (&v)[8], const ModP &mp, const uint16_t *tab, uint32_t (*Ls)[8],
			   uint32_t (*XR)[128], uint32_t (*U)[128], uint32_t *T, int *s_ok)
{
	const int tid = threadIdx.x, c = tid & 127, g = tid >> 7;
	for (int q = 0; q < 8; ++q) {
		if (c >= 8 * q && c < 8 * q + 8) {
#pragma unroll
			for (int i = 0; i < 8; ++i)
				Ls[8 * g + i][c - 8 * q] = mp.mod_f64(v[i]);
		}
		if (g == q) {
#pragma unroll
			for (int i = 0; i < 8; ++i)
				XR[i][c] = mp.mod_f64(v[i]);
		}
		__syncthreads();
		if (tid < 64) {
			const bool ok = wave_inv8(&Ls[8 * q][0], 8, T, mp, tab);
			if (tid == 0)
				*s_ok = ok;
		}
		__syncthreads();
		if (!*s_ok)
			return false;
		{
			const int j = g;
			const uint4 t0 = *(const uint4 *) &T[8 * j], t1 = *(const uint4 *) &T[8 * j + 4];
			const uint32_t tj[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
			double a = 0.0;
#pragma unroll
			for (int s = 0; s < 8; ++s)
				a = fma((double) tj[s], (double) XR[s][c], a);
			U[j][c] = mp.mod_f64(a);
		}
		__syncthreads();
		double uc[8];
#pragma unroll
		for (int s = 0; s < 8; ++s)
			uc[s] = (double) U[s][c];
#pragma unroll
		for (int i = 0; i < 8; ++i) {
			const int r = 8 * g + i;
			if (g == q) {
				v[i] = uc[i];
			} else {
				const uint4 l0 = *(const uint4 *) &Ls[r][0], l1 = *(const uint4 *) &Ls[r][4];
				const uint32_t lr[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
				double acc = 0.0;
#pragma unroll
				for (int s = 0; s < 8; ++s)
					acc = fma((double) lr[s], uc[s], acc);
				v[i] -= acc;
			}
		}
		__syncthreads();
	}
	return true;
}

// Inverses mod p for p <= 2^16 (0 where none exists).
__global__ void invtab_kernel(uint16_t *__restrict__ tab, ModP mp)
{
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x < mp.p)
		tab[x] = (uint16_t) mp.inv(x);
}

// The context's inverse table for p (built on first use per modulus).
static int get_invtab(mlz_ctx *ctx, const ModP &mp, const uint16_t **out)
{
	void *w = nullptr;
	MLZ_TRY(ws_get(ctx, WS_INVTAB, 65536 * sizeof(uint16_t), &w));
	if (ctx->invtab_p != (uint32_t) mp.p) {
		invtab_kernel<<<ceil_div(mp.p, 256), 256, 0, ctx->stream>>>((uint16_t *) w, mp);
		MLZ_TRY(check_launch(ctx, "invtab_kernel"));
		ctx->invtab_p = (uint32_t) mp.p;
	}
	*out = (const uint16_t *) w;
	return MLZ_OK;
}

int zp_inverse_table(mlz_ctx *ctx, const ModP &mp, const uint16_t **out)
{
	return get_invtab(ctx, mp, out);
}

__device__ inline int nodes_at(int nodes0, int level)
{
	int n = nodes0;
	for (int l = 0; l < level; ++l)
		n = (n + 1) / 2;
	return n;
}

// Fast path, one workgroup: S = the first wc rows that are not pivots yet (in
// row order), on the panel's wc columns, is reduced together with the
// identity, [S | I] -> [Pi | R].  If S has full panel rank its rows are the
// panel's pivot rows and P_J^-1 comes out directly: with Q the order of the
// rows by pivot column, P_J = Q S, so P_J^-1 = S^-1 Q^T, whose entry (c, t)
// is R[row of pivot c][row of pivot t].  If S is rank deficient but holds
// every live row, the pivots are still final and the inverse kernel builds
// P_J^-1; otherwise the tournament decides.
//
// LAZY (p <= 2^16): reg_gj_wave with the inverse table copied into LDS;
// otherwise fraction-free steps with a reduction per step (any p).
template <bool SMALL, int TPB, bool LAZY>
__global__ __launch_bounds__(TPB) void rref_fast_kernel(const uint32_t *__restrict__ A, size_t lda, int m, int c,
						       int wc, ModP mp, RrefState st, uint32_t *__restrict__ Pinv,
						       const uint16_t *__restrict__ invtab)
{
	constexpr int W = RREF_W, NC = 2 * RREF_W, RPG = W / (TPB / NC);
	__shared__ int rowid[W], alive[W], outr[W], outc[W], pos[W];
	__shared__ uint32_t f[W], pivrow[NC], s_y[2];
	__shared__ int s_min[2], s_cnt, s_wave[16], s_pr[2];
	__shared__ __attribute__((aligned(16))) uint16_t tab[LAZY ? 65536 : 8];
	__shared__ double s_q[LAZY ? 2 : 1][64], s_fd[LAZY ? 2 : 1][64];
	constexpr bool BLK = LAZY && TPB == 1024;   // block Gauss-Jordan first (full 64 x 64 S)
	__shared__ __attribute__((aligned(16))) uint32_t b_l[BLK ? 64 : 1][8], b_xr[BLK ? 8 : 1][128],
		b_u[BLK ? 8 : 1][128], b_t[BLK ? 64 : 4];
	__shared__ int b_ok;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	if (LAZY)
		for (int e = tid; e < (int) ((mp.p + 7) / 8); e += TPB)
			((uint4 *) tab)[e] = ((const uint4 *) invtab)[e];
	if (tid == 0)
		s_cnt = 0;
	__syncthreads();
	int scanned_all = 1;
	const int nwaves = blockDim.x >> 6;
	for (int base = 0; base < m; base += blockDim.x) {
		const int row = base + tid;
		const bool live = row < m && st.is_pivot[row] == 0;
		const unsigned long long bal = __ballot(live);
		if (lane == 0)
			s_wave[wave] = __popcll(bal);
		__syncthreads();
		int before = s_cnt;
		for (int w = 0; w < wave; ++w)
			before += s_wave[w];
		const int slot = before + __popcll(bal & ((1ull << lane) - 1));
		if (live && slot < wc)
			rowid[slot] = row;
		__syncthreads();
		if (tid == 0)
			for (int w = 0; w < nwaves; ++w)
				s_cnt += s_wave[w];
		__syncthreads();
		if (s_cnt >= wc) {
			// rows beyond the collected ones remain unseen unless this was the end
			scanned_all = (s_cnt == wc && base + (int) blockDim.x >= m);
			break;
		}
	}
	const int nr = min(s_cnt, wc);
	for (int r = tid; r < W; r += blockDim.x)
		alive[r] = r < nr;
	const int ci = tid % NC, g = tid / NC;
	uint32_t v[RPG];
#pragma unroll
	for (int i = 0; i < RPG; ++i) {
		const int r = g * RPG + i;
		if (ci < W)
			v[i] = (r < nr && ci < wc) ? A[(size_t) rowid[r] * lda + c + ci] : 0u;
		else
			v[i] = (ci - W == r) ? 1u : 0u;
	}
	int cnt;
	if (!LAZY) {
		cnt = reg_eliminate<NC, RPG, SMALL>(v, W, wc, true, mp, alive, f, pivrow, s_min, outr, outc, st.bad);
	} else {
		double vd[RPG];
#pragma unroll
		for (int i = 0; i < RPG; ++i)
			vd[i] = (double) v[i];
		bool blk = false;
		if constexpr (BLK) {
			if (nr == W && wc == W) {
				blk = block_gj64(vd, mp, tab, b_l, b_xr, b_u, b_t, &b_ok);
				if (blk) {
					// row t is the pivot row of column t
					for (int s = tid; s < W; s += TPB) {
						outr[s] = s;
						outc[s] = s;
					}
#pragma unroll
					for (int i = 0; i < RPG; ++i)
						vd[i] = (double) mp.mod_f64(vd[i]);
					cnt = W;
				} else {
#pragma unroll
					for (int i = 0; i < RPG; ++i)
						vd[i] = (double) v[i];
				}
			}
		}
		if (!blk)
			cnt = reg_gj_wave<NC, RPG>(vd, nr, wc, mp, s_q, s_fd, s_pr, s_y, pivrow, outr, outc, st.bad, tab);
#pragma unroll
		for (int i = 0; i < RPG; ++i)
			v[i] = (uint32_t) vd[i];
	}
	const bool full = cnt == wc;
	const bool done = full || scanned_all;
	if (full) {
		for (int s = tid; s < cnt; s += blockDim.x)
			pos[outr[s]] = s;
		__syncthreads();
		if (ci >= W) {
#pragma unroll
			for (int i = 0; i < RPG; ++i) {
				const int r = g * RPG + i;
				if (r < wc && ci - W < wc)
					Pinv[(size_t) pos[r] * W + pos[ci - W]] = v[i];
			}
		}
		for (int e = tid; e < W * W; e += blockDim.x) {
			const int r = e / W, q = e % W;
			if (r >= wc || q >= wc)
				Pinv[e] = r == q ? 1u : 0u;
		}
	}
	for (int t = tid; t < RREF_W; t += blockDim.x) {
		if (done) {
			st.J[t] = t < cnt ? outc[t] : 0;
			st.prow[t] = t < cnt ? rowid[outr[t]] : 0;
		}
	}
	if (tid == 0) {
		if (done)
			*st.k = cnt;
		*st.need_tree = done ? 0 : 1;
		*st.inv_done = full ? 1 : 0;
	}
}

// Inverse of P_J (k x k, padded with the identity to RREF_W; entry (r, s) =
// A[prow[r]][c + J[s]]) by Gauss-Jordan on [P_J | I] held in registers (128
// columns x 2 groups of 32 rows), by one 1024-thread workgroup.  Run by the
// tournament's root (or its first workgroup when the fast path found the
// pivots without P_J^-1), so the panel needs no launch of its own for it.
template <bool SMALL>
__device__ void rref_inv_body(const uint32_t *__restrict__ A, size_t lda, int c, const ModP &mp, RrefState st,
			      uint32_t *__restrict__ Pinv)
{
	constexpr int W = RREF_W, NC = 2 * RREF_W, RPG = RREF_W / (1024 / NC);
	__shared__ int alive[W], outr[W], outc[W], pos[W];
	__shared__ uint32_t f[W], pivrow[NC];
	__shared__ int s_min[2];
	const int tid = threadIdx.x, cc = tid % NC, g = tid / NC;
	const int k = *st.k;
	uint32_t v[RPG];
#pragma unroll
	for (int i = 0; i < RPG; ++i) {
		const int r = g * RPG + i;
		uint32_t x;
		if (cc < W)
			x = (r < k && cc < k) ? A[(size_t) st.prow[r] * lda + c + st.J[cc]] : (r == cc ? 1u : 0u);
		else
			x = (cc - W == r) ? 1u : 0u;
		v[i] = x;
	}
	for (int r = tid; r < W; r += blockDim.x) {
		alive[r] = 1;
		pos[r] = r;
	}
	__syncthreads();
	const int cnt = reg_eliminate<NC, RPG, SMALL>(v, W, k, true, mp, alive, f, pivrow, s_min, outr, outc, st.bad);
	if (tid == 0 && cnt < k)
		*st.bad = 1;
	for (int s = tid; s < cnt; s += blockDim.x)
		pos[outr[s]] = outc[s];
	__syncthreads();
	if (cc >= W) {
#pragma unroll
		for (int i = 0; i < RPG; ++i) {
			const int r = g * RPG + i;
			Pinv[(size_t) pos[r] * W + (cc - W)] = v[i];
		}
	}
}

template <bool SMALL>
__global__ __launch_bounds__(1024) void rref_tourney_kernel(const uint32_t *__restrict__ A, size_t lda, int m,
							   int c, int wc, ModP mp, RrefState st, int nodes0,
							   uint32_t *__restrict__ Pinv)
{
	constexpr int RPG = RREF_CHUNK / (1024 / RREF_W);
	__shared__ int rowid[RREF_CHUNK], alive[RREF_CHUNK];
	__shared__ uint32_t f[RREF_CHUNK], pivrow[RREF_W];
	__shared__ int outr[RREF_W], outc[RREF_W];
	__shared__ int s_min[2], s_go, s_nr, s_n0;
	if (*st.need_tree == 0) {
		// the fast path settled the pivots; P_J^-1 may still be missing
		if (blockIdx.x == 0 && *st.inv_done == 0)
			rref_inv_body<SMALL>(A, lda, c, mp, st, Pinv);
		return;
	}
	const int tid = threadIdx.x, cc = tid % RREF_W, g = tid / RREF_W;
	int node = blockIdx.x, level = 0;
	uint32_t v[RPG];

	// Leaf: rows [node*CHUNK, ...) that are not pivots yet.
	const int r0 = node * RREF_CHUNK;
	const int nr0 = min(RREF_CHUNK, m - r0);
	for (int r = tid; r < RREF_CHUNK; r += blockDim.x) {
		rowid[r] = r0 + r;
		alive[r] = r < nr0 && st.is_pivot[r0 + r] == 0;
	}
#pragma unroll
	for (int i = 0; i < RPG; ++i) {
		const int r = g * RPG + i;
		v[i] = (r < nr0 && cc < wc) ? A[(size_t) (r0 + r) * lda + c + cc] : 0u;
	}
	__syncthreads();
	int cnt = reg_eliminate<RREF_W, RPG, SMALL>(v, nr0, wc, false, mp, alive, f, pivrow, s_min, outr, outc, st.bad);
	for (int t = tid; t < cnt; t += blockDim.x)
		outr[t] = rowid[outr[t]];
	__syncthreads();

	for (;;) {
		const int here = nodes_at(nodes0, level);
		if (here == 1) {
			// Root: publish the panel's pivots and reset the arrival counters.
			for (int t = tid; t < RREF_W; t += blockDim.x) {
				st.J[t] = t < cnt ? outc[t] : 0;
				st.prow[t] = t < cnt ? outr[t] : 0;
			}
			if (tid == 0)
				*st.k = cnt;
			for (int t = tid; t < RREF_MAXLVL * nodes0; t += blockDim.x)
				st.counters[t] = 0;
			__syncthreads();   // the published list is read back by the inversion
			rref_inv_body<SMALL>(A, lda, c, mp, st, Pinv);
			return;
		}
		int *lr = st.lists_rows + ((size_t) level * nodes0 + node) * RREF_W;
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
		// Merge both children's lists (re-reading the rows' panel values).
		const int c0 = parent * 2, c1 = parent * 2 + 1;
		if (tid == 0) {
			const int n0 = ((volatile int *) st.lists_cnt)[(size_t) level * nodes0 + c0];
			const int n1 = c1 < here ? ((volatile int *) st.lists_cnt)[(size_t) level * nodes0 + c1] : 0;
			s_nr = n0 + n1;
			s_n0 = n0;
		}
		__syncthreads();
		const int nr = s_nr, n0 = s_n0;
		for (int r = tid; r < RREF_CHUNK; r += blockDim.x) {
			if (r < nr) {
				const int child = r < n0 ? c0 : c1;
				const int pos = r < n0 ? r : r - n0;
				rowid[r] = ((volatile int *) st.lists_rows)[((size_t) level * nodes0 + child) * RREF_W + pos];
			}
			alive[r] = r < nr;
		}
		__syncthreads();
#pragma unroll
		for (int i = 0; i < RPG; ++i) {
			const int r = g * RPG + i;
			v[i] = (r < nr && cc < wc) ? A[(size_t) rowid[r] * lda + c + cc] : 0u;
		}
		cnt = reg_eliminate<RREF_W, RPG, SMALL>(v, nr, wc, false, mp, alive, f, pivrow, s_min, outr, outc, st.bad);
		for (int t = tid; t < cnt; t += blockDim.x)
			outr[t] = rowid[outr[t]];
		__syncthreads();
		++level;
		node = parent;
	}
}

// P[t][col] = A[prow[t]][c + col] (t < k, else 0); marks the new pivot rows and
// appends them to the pivot lists (pivot column coff + c + J[t]).  With a
// tracker block at column inj of A (inj >= 0), pivot t also gets 1 added in
// tracker column inj + (pivots found since the block began) + t.
__global__ void rref_gather_kernel(const uint32_t *__restrict__ A, size_t lda, int nc, int c, int panel,
				   RrefState st, uint32_t *__restrict__ P, int coff, int inj, uint32_t p)
{
	const int t = blockIdx.y;
	const int k = *st.k;
	const int col = blockIdx.x * blockDim.x + threadIdx.x;
	if (col < nc) {
		uint32_t x = t < k ? A[(size_t) st.prow[t] * lda + c + col] : 0u;
		if (inj >= 0 && t < k && c + col == inj + (*st.rank - *st.rank0) + t)
			x = x + 1 == p ? 0u : x + 1;
		P[(size_t) t * nc + col] = x;
	}
	if (blockIdx.x == 0 && threadIdx.x == 0 && t < k) {
		const int r = *st.rank;
		st.is_pivot[st.prow[t]] = panel + 1;
		st.piv_rows[r + t] = st.prow[t];
		st.piv_cols[r + t] = coff + c + st.J[t];
	}
}

// The panel step after the pivot search, fused for p <= 2^16 (INT8 trailing
// update).  Workgroup j owns P's columns [TC j, TC j + TC) (TC = 32 or 64;
// narrower column slices give the launch more workgroups): it gathers them from
// the pivot rows (plus the tracker's 1, as rref_gather_kernel), forms
// W = P_J^-1 * P for them (64 terms below 2^32 each: exact in a double),
// writes W into the pivot rows (the update leaves those rows unchanged: their
// multipliers are zero) and packs W as the update's INT8 B fragments
// (KB = 2).  Workgroup 0 also marks the pivot rows and appends the pivot
// lists.
template <int TC>
__global__ __launch_bounds__(256) void rref_wb_i8_kernel(uint32_t *__restrict__ A, size_t lda, int nc, int c,
							 int panel, RrefState st, const uint32_t *__restrict__ Pinv,
							 int8_t *__restrict__ Wp, int coff, int inj, ModP mp)
{
	constexpr int W = RREF_W;
	static_assert(TC == 32 || TC == 64, "column slice of 32 or 64");
	__shared__ uint32_t Ps[W][TC + 1], Pi[W][W + 1];
	__shared__ int prow[W];
	const uint32_t p = (uint32_t) mp.p;
	const int tid = threadIdx.x;
	const int j0 = blockIdx.x * TC;
	// P_J^-1 loaded first, all at once, so that its latency overlaps the
	// dependent loads of the pivot rows
	uint32_t pv[W * W / 256];
#pragma unroll
	for (int u = 0; u < W * W / 256; ++u)
		pv[u] = Pinv[tid + 256 * u];
	const int k = *st.k;
	const int kb = *st.rank - *st.rank0;
	if (tid < W)
		prow[tid] = tid < k ? st.prow[tid] : -1;
#pragma unroll
	for (int u = 0; u < W * W / 256; ++u) {
		const int e = tid + 256 * u;
		Pi[e / W][e % W] = pv[u];
	}
	__syncthreads();
	for (int e = tid; e < W * TC; e += 256) {
		const int t = e / TC, col = e % TC, gc = j0 + col;
		uint32_t x = 0;
		if (t < k && gc < nc) {
			x = A[(size_t) prow[t] * lda + c + gc];
			if (inj >= 0 && c + gc == inj + kb + t)
				x = x + 1 == p ? 0u : x + 1;
		}
		Ps[t][col] = x;
	}
	__syncthreads();
	// W on v_mfma_f64_16x16x4f64: wave w computes rows 16w .. 16w+15 as
	// TC/16 16 x 16 blocks.  A fragment: lane l holds (row l%16, k l/16); B
	// fragment: (k l/16, column l%16); result register r of lane l: (row
	// l/16 + 4r, column l%16).
	{
		constexpr int NJ = TC / 16;
		const int lane = tid & 63, w = tid >> 6;
		typedef double v4d __attribute__((ext_vector_type(4)));
		v4d acc[NJ];
#pragma unroll
		for (int j = 0; j < NJ; ++j)
			acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
		for (int s = 0; s < W; s += 4) {
			const double a = (double) Pi[16 * w + (lane & 15)][s + (lane >> 4)];
#pragma unroll
			for (int j = 0; j < NJ; ++j)
				acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, (double) Ps[s + (lane >> 4)][16 * j + (lane & 15)],
									     acc[j], 0, 0, 0);
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < NJ; ++j)
#pragma unroll
			for (int r = 0; r < 4; ++r)
				Ps[16 * w + (lane >> 4) + 4 * r][16 * j + (lane & 15)] = mp.mod_f64(acc[j][r]);
	}
	__syncthreads();
	for (int e = tid; e < W * TC; e += 256) {
		const int t = e / TC, cc = e % TC, gc = j0 + cc;
		if (t < k && gc < nc)
			A[(size_t) prow[t] * lda + c + gc] = Ps[t][cc];
	}
	// B fragments: piece (column block cb, k block q, lane) holds column
	// 32cb + lane%32 and k = 32q + 16(lane/32) .. +15
	if (tid < 128 * (TC / 32)) {
		const int lane = tid & 63, q = (tid >> 6) & 1, cb = tid >> 7;
		const int cc = 32 * cb + (lane & 31), t = 32 * q + 16 * (lane >> 5);
		uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
		for (int b = 0; b < 16; ++b) {
			const uint32_t l = i8_limbs(Ps[t + b][cc], p);
			hw[b >> 2] |= ((l >> 8) & 0xffu) << (8 * (b & 3));
			lw[b >> 2] |= (l & 0xffu) << (8 * (b & 3));
		}
		int8_t *fr = Wp + ((size_t) ((TC / 32) * blockIdx.x + cb) * 2 + q) * 2 * 1024 + lane * 16;
		*(uint4 *) fr = make_uint4(hw[0], hw[1], hw[2], hw[3]);
		*(uint4 *) (fr + 1024) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
	}
	if (blockIdx.x == 0 && tid < k) {
		const int r = *st.rank;
		st.is_pivot[prow[tid]] = panel + 1;
		st.piv_rows[r + tid] = prow[tid];
		st.piv_cols[r + tid] = coff + c + st.J[tid];
	}
}

// G[i][s] = A[i][c + J[s]] (s < k; zero on this panel's pivot rows), packed
// as the update's INT8 A fragments (KB = 2), then rank += k.  Thread: the
// 16-byte piece (row block, k block, lane) of one workgroup's 64 rows.
__global__ __launch_bounds__(256) void rref_gpack_i8_kernel(const uint32_t *__restrict__ A, size_t lda, int m,
							    int c, int panel, RrefState st, int8_t *__restrict__ Gp,
							    uint32_t p)
{
	const int tid = threadIdx.x, lane = tid & 63, q = (tid >> 6) & 1, rb = 2 * blockIdx.x + (tid >> 7);
	const int k = *st.k;
	const int i = 32 * rb + (lane & 31), s0 = 32 * q + 16 * (lane >> 5);
	const bool live = i < m && st.is_pivot[i] != panel + 1;
	// A full 64-column panel of full rank has J[s] = s (J increases, so
	// J[63] = 63 says it): the 16 multipliers are then consecutive columns
	// inside the panel, four 16-byte loads when the row stride keeps them
	// aligned (c and s0 are multiples of 16)
	uint32_t xv[16];
	if (k == RREF_W && st.J[RREF_W - 1] == RREF_W - 1 && (lda & 3) == 0) {
		const uint4 *src = (const uint4 *) (A + (size_t) i * lda + c + s0);
#pragma unroll
		for (int qq = 0; qq < 4; ++qq) {
			const uint4 u = live ? src[qq] : make_uint4(0u, 0u, 0u, 0u);
			xv[4 * qq] = u.x;
			xv[4 * qq + 1] = u.y;
			xv[4 * qq + 2] = u.z;
			xv[4 * qq + 3] = u.w;
		}
	} else {
#pragma unroll
		for (int b = 0; b < 16; ++b)
			xv[b] = (live && s0 + b < k) ? A[(size_t) i * lda + c + st.J[s0 + b]] : 0u;
	}
	uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
	for (int b = 0; b < 16; ++b) {
		const uint32_t l = i8_limbs(xv[b], p);
		hw[b >> 2] |= ((l >> 8) & 0xffu) << (8 * (b & 3));
		lw[b >> 2] |= (l & 0xffu) << (8 * (b & 3));
	}
	int8_t *fr = Gp + ((size_t) rb * 2 + q) * 2 * 1024 + lane * 16;
	*(uint4 *) fr = make_uint4(hw[0], hw[1], hw[2], hw[3]);
	*(uint4 *) (fr + 1024) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
	if (blockIdx.x == 0 && tid == 0)
		*st.rank += k;
}

// G[i][s] = A[i][c + J[s]] for s < k, except 0 on this panel's pivot rows.
__global__ void rref_gatherG_kernel(const uint32_t *__restrict__ A, size_t lda, int m, int c, int panel,
				    RrefState st, uint32_t *__restrict__ G)
{
	const int i = blockIdx.x * (blockDim.x / RREF_W) + threadIdx.x / RREF_W;
	const int s = threadIdx.x % RREF_W;
	if (i >= m)
		return;
	const int k = *st.k;
	uint32_t v = 0;
	if (s < k && st.is_pivot[i] != panel + 1)
		v = A[(size_t) i * lda + c + st.J[s]];
	G[(size_t) i * RREF_W + s] = v;
}

// Writes W into the pivot rows.
__global__ void rref_scatter_kernel(uint32_t *__restrict__ A, size_t lda, int nc, int c, RrefState st,
				    const uint32_t *__restrict__ W)
{
	const int t = blockIdx.y;
	const int k = *st.k;
	const int col = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < k && col < nc)
		A[(size_t) st.prow[t] * lda + c + col] = W[(size_t) t * nc + col];
}

__global__ void rref_advance_kernel(RrefState st)
{
	*st.rank += *st.k;
}

// Column blocks: X (m x xw) <- [A[:, c0 .. c0+nbw) | 0], and rank0 <- rank.
__global__ void rref_block_load_kernel(const uint32_t *__restrict__ A, size_t lda, int m, int c0, int nbw,
				       uint32_t *__restrict__ X, int xw, RrefState st)
{
	// x covers the columns j, grid.y strides over the rows
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
		*st.rank0 = *st.rank;
	if (j >= xw)
		return;
	for (int i = blockIdx.y; i < m; i += gridDim.y)
		X[(size_t) i * xw + j] = j < nbw ? A[(size_t) i * lda + c0 + j] : 0u;
}

// The block's columns back into A.
__global__ void rref_block_store_kernel(const uint32_t *__restrict__ X, int xw, int m, int nbw,
					uint32_t *__restrict__ A, size_t lda, int c0)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nbw)
		return;
	for (int i = blockIdx.y; i < m; i += gridDim.y)
		A[(size_t) i * lda + c0 + j] = X[(size_t) i * xw + j];
}

// After a column block: TR[t] = the trailing columns (from c0 + nbw on) of
// the block's t-th pivot row (zero for t beyond the block's pivots), and the
// tracker entry of that row loses the 1 its pivot step added, so that the
// trailing columns of every row become A + E * TR.
__global__ void rref_block_trail_kernel(uint32_t *__restrict__ X, int xw, int nbw, const uint32_t *__restrict__ A,
					size_t lda, int c0, int nt, RrefState st, uint32_t *__restrict__ TR, uint32_t p)
{
	const int t = blockIdx.y;
	const int kb = *st.rank - *st.rank0;
	const int col = blockIdx.x * blockDim.x + threadIdx.x;
	const int row = t < kb ? st.piv_rows[*st.rank0 + t] : -1;
	if (col < nt)
		TR[(size_t) t * nt + col] = row >= 0 ? A[(size_t) row * lda + c0 + nbw + col] : 0u;
	if (blockIdx.x == 0 && threadIdx.x == 0 && row >= 0) {
		uint32_t *e = X + (size_t) row * xw + nbw + t;
		*e = *e == 0 ? p - 1 : *e - 1;
	}
}

// Working copy W (m x n, row stride n, canonical residues) of A (stride ld)
// and back.  Four columns per thread, grid.y strides over the rows; VEC (n,
// ld and both pointers aligned to four entries) moves them with one vector
// load and one vector store.
constexpr size_t RREF_COPY_ROWS = 1024;   // grid.y of the copy launches

template <int E, bool VEC>
__global__ void rref_load_kernel(const typename Elem<E>::T *__restrict__ src, size_t ld, int m, int n,
				 uint32_t p, uint32_t *__restrict__ dst)
{
	typedef typename Elem<E>::T T;
	typedef T T4 __attribute__((ext_vector_type(4)));
	typedef uint32_t U4 __attribute__((ext_vector_type(4)));
	const int c0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
	if (c0 >= n)
		return;
	for (int r = blockIdx.y; r < m; r += gridDim.y) {
		const T *s = src + (size_t) r * ld + c0;
		uint32_t *d = dst + (size_t) r * n + c0;
		if (VEC) {   // n % 4 == 0: the four columns exist
			const T4 x = *(const T4 *) s;
			const T t[4] = {x.x, x.y, x.z, x.w};
			U4 y;
			y.x = Elem<E>::load(&t[0], p);
			y.y = Elem<E>::load(&t[1], p);
			y.z = Elem<E>::load(&t[2], p);
			y.w = Elem<E>::load(&t[3], p);
			*(U4 *) d = y;
		} else {
#pragma unroll
			for (int q = 0; q < 4; ++q)
				if (c0 + q < n)
					d[q] = Elem<E>::load(s + q, p);
		}
	}
}

// Output row t: pivot row t for t < rank, zero row after.
template <int E, bool VEC>
__global__ void rref_store_kernel(const uint32_t *__restrict__ W, int m, int n, uint32_t p, RrefState st,
				  typename Elem<E>::T *__restrict__ dst, size_t ld)
{
	typedef typename Elem<E>::T T;
	typedef T T4 __attribute__((ext_vector_type(4)));
	typedef uint32_t U4 __attribute__((ext_vector_type(4)));
	const int c0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
	if (c0 >= n)
		return;
	const int rank = *st.rank;
	for (int t = blockIdx.y; t < m; t += gridDim.y) {
		T *o = dst + (size_t) t * ld + c0;
		const uint32_t *w = t < rank ? W + (size_t) st.piv_rows[t] * n + c0 : nullptr;
		if (VEC) {
			const U4 v = w ? *(const U4 *) w : U4{0u, 0u, 0u, 0u};
			T x[4];
			Elem<E>::store(&x[0], v.x, p);
			Elem<E>::store(&x[1], v.y, p);
			Elem<E>::store(&x[2], v.z, p);
			Elem<E>::store(&x[3], v.w, p);
			*(T4 *) o = T4{x[0], x[1], x[2], x[3]};
		} else {
#pragma unroll
			for (int q = 0; q < 4; ++q)
				if (c0 + q < n)
					Elem<E>::store(o + q, w ? w[q] : 0u, p);
		}
	}
}

__global__ void rref_init_kernel(RrefState st, int m, int cnt_counters)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < m)
		st.is_pivot[i] = 0;
	if (i < cnt_counters)
		st.counters[i] = 0;
	if (i == 0) {
		*st.rank = 0;
		*st.rank0 = 0;
		*st.k = 0;
		*st.bad = 0;
		*st.need_tree = 0;
		*st.inv_done = 0;
	}
}

// A -> W (LOAD) or the pivot rows of W -> A, vectorised when n, lda and A
// are aligned to four entries (W is workspace, aligned).
template <int E, bool LOAD>
static void rref_copy_launch(mlz_ctx *ctx, typename Elem<E>::T *A, size_t lda, size_t m, size_t n, uint32_t p,
			     uint32_t *W, const RrefState &st)
{
	typedef typename Elem<E>::T T;
	const dim3 grid(ceil_div(ceil_div(n, 4), 256), (unsigned) std::min(m, RREF_COPY_ROWS));
	const bool vec = n % 4 == 0 && lda % 4 == 0 && (uintptr_t) A % (4 * sizeof(T)) == 0;
	if (LOAD) {
		if (vec)
			rref_load_kernel<E, true><<<grid, 256, 0, ctx->stream>>>(A, lda, (int) m, (int) n, p, W);
		else
			rref_load_kernel<E, false><<<grid, 256, 0, ctx->stream>>>(A, lda, (int) m, (int) n, p, W);
	} else {
		if (vec)
			rref_store_kernel<E, true><<<grid, 256, 0, ctx->stream>>>(W, (int) m, (int) n, p, st, A, lda);
		else
			rref_store_kernel<E, false><<<grid, 256, 0, ctx->stream>>>(W, (int) m, (int) n, p, st, A, lda);
	}
}

int zp_rref_dev(mlz_ctx *ctx, int elem, uint64_t p, size_t m, size_t n, void *A, size_t lda, size_t *rank_out,
		uint32_t *pivots_out, int algo)
{
	if (m == 0 || n == 0) {
		if (rank_out)
			*rank_out = 0;
		return MLZ_OK;
	}
	const ModP mp = make_modp(p);
	const uint32_t pp = (uint32_t) p;
	const int nodes0 = (int) ceil_div(m, RREF_CHUNK);
	const size_t mn = m * n;
	const size_t rmax = std::min(m, n);
	if (nodes0 > (1 << (RREF_MAXLVL - 2)))
		return set_error(ctx, MLZ_E_ARG, "too many rows for the pivot tournament");
	// workspace
	void *w0, *w1, *w2, *w3;
	MLZ_TRY(ws_get(ctx, WS_RREF_0, mn * sizeof(uint32_t), &w0));                 // working matrix
	// P (or W's INT8 fragments), W, Pinv; G (or its INT8 fragments)
	const size_t pbytes = std::max((size_t) RREF_W * n * sizeof(uint32_t), round_up(n, 64) * 2 * RREF_W);
	MLZ_TRY(ws_get(ctx, WS_RREF_1, pbytes + (size_t) RREF_W * n * sizeof(uint32_t) +
				      (size_t) RREF_W * RREF_W * sizeof(uint32_t), &w1));
	MLZ_TRY(ws_get(ctx, WS_RREF_2, std::max(m * RREF_W * sizeof(uint32_t), round_up(m, 64) * 2 * RREF_W), &w2));
	const size_t ints = m + 1 + 2 * rmax + 1 + 2 * RREF_W + 4 + (size_t) RREF_MAXLVL * nodes0 * RREF_W +
			    2 * (size_t) RREF_MAXLVL * nodes0;
	MLZ_TRY(ws_get(ctx, WS_RREF_3, ints * sizeof(int), &w3));
	uint32_t *Wm = (uint32_t *) w0;
	uint32_t *P = (uint32_t *) w1, *Wb = (uint32_t *) ((char *) w1 + pbytes), *Pinv = Wb + (size_t) RREF_W * n;
	uint32_t *G = (uint32_t *) w2;
	int8_t *Wp = (int8_t *) P, *Gp = (int8_t *) G;
	RrefState st;
	int *q = (int *) w3;
	st.is_pivot = q; q += m;
	st.rank = q; q += 1;
	st.piv_rows = q; q += rmax;
	st.piv_cols = q; q += rmax;
	st.k = q; q += 1;
	st.J = q; q += RREF_W;
	st.prow = q; q += RREF_W;
	st.bad = q; q += 1;
	st.need_tree = q; q += 1;
	st.inv_done = q; q += 1;
	st.rank0 = q; q += 1;
	st.lists_rows = q; q += (size_t) RREF_MAXLVL * nodes0 * RREF_W;
	st.lists_cnt = q; q += (size_t) RREF_MAXLVL * nodes0;
	st.counters = q;

	{
		const int cc = RREF_MAXLVL * nodes0;
		rref_init_kernel<<<ceil_div(std::max(m, (size_t) cc), 256), 256, 0, ctx->stream>>>(st, (int) m, cc);
		ProfScope ps(ctx, "rref_load");
		if (elem == MLZ_U16)
			rref_copy_launch<MLZ_U16, true>(ctx, (uint16_t *) A, lda, m, n, pp, Wm, st);
		else if (elem == MLZ_U32)
			rref_copy_launch<MLZ_U32, true>(ctx, (uint32_t *) A, lda, m, n, pp, Wm, st);
		else
			rref_copy_launch<MLZ_F64_SYM, true>(ctx, (double *) A, lda, m, n, pp, Wm, st);
		MLZ_TRY(check_launch(ctx, "rref_load_kernel"));
	}
	// trailing updates on INT8 limbs whenever the gemm family allows them
	const bool upd_i8 = p <= 65536 && (algo == MLZ_ALGO_AUTO || algo == MLZ_ALGO_INT8);
	// pivot search with lazy reduction and the inverse table for p <= 2^16
	const uint16_t *invtab = nullptr;
	if (p <= 65536)
		MLZ_TRY(get_invtab(ctx, mp, &invtab));
	int panel = 0;
	// One panel of X (row stride ldx, xn columns): pivots of columns
	// [c, c + RREF_W), then every row's columns from c on are updated.
	// coff = X's column offset in A; inj = tracker column of X or -1.
	auto panel_step = [&](uint32_t *X, size_t ldx, size_t xn, size_t c, int coff, int inj) -> int {
		const int wc = (int) std::min((size_t) RREF_W, xn - c);
		const int nc = (int) (xn - c);
		{
			ProfScope ps(ctx, "rref_pivots");
#define FAST(TPB, LAZY) rref_fast_kernel<true, TPB, LAZY><<<1, TPB, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, wc, \
										       mp, st, Pinv, invtab)
			if (mp.small) {
				const int threads = ctx->rref_fast_threads;
				if (invtab)
					threads == 256 ? FAST(256, true) : threads == 512 ? FAST(512, true) : FAST(1024, true);
				else
					threads == 256 ? FAST(256, false) : threads == 512 ? FAST(512, false) : FAST(1024, false);
				rref_tourney_kernel<true><<<nodes0, 1024, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, wc, mp, st, nodes0,
											   Pinv);
			} else {
				rref_fast_kernel<false, 1024, false><<<1, 1024, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, wc, mp,
												   st, Pinv, nullptr);
				rref_tourney_kernel<false><<<nodes0, 1024, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, wc, mp, st,
											    nodes0, Pinv);
			}
#undef FAST
		}
		MLZ_TRY(check_launch(ctx, "rref_tourney_kernel"));   // it also forms P_J^-1 when the fast path did not
		if (upd_i8) {
			// W formed, written and packed in one kernel, G gathered and
			// packed in another, then the INT8 update
			ProfScope ps(ctx, "rref_update");
			rref_wb_i8_kernel<32><<<ceil_div(nc, 32), 256, 0, ctx->stream>>>(X, ldx, nc, (int) c, panel, st, Pinv,
											 Wp, coff, inj, mp);
			rref_gpack_i8_kernel<<<ceil_div(m, 64), 256, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, panel, st,
											 Gp, pp);
			MLZ_TRY(check_launch(ctx, "rref_gpack_i8_kernel"));
			MLZ_TRY(zp_rank_update_i8_packed(ctx, pp, m, nc, 2, Gp, Wp, X + c, ldx));
			++panel;
			return MLZ_OK;
		}
		rref_gather_kernel<<<dim3(ceil_div(nc, 256), RREF_W), 256, 0, ctx->stream>>>(X, ldx, nc, (int) c, panel, st, P,
											coff, inj, pp);
		MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, RREF_W, nc, RREF_W, 1, Pinv, RREF_W, P, nc, 0, Wb, nc, algo));
		rref_gatherG_kernel<<<ceil_div(m, 256 / RREF_W), 256, 0, ctx->stream>>>(X, ldx, (int) m, (int) c, panel, st, G);
		MLZ_TRY(check_launch(ctx, "rref_gatherG_kernel"));
		MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, m, nc, RREF_W, p - 1, G, RREF_W, Wb, nc, 1, X + c, ldx, algo));
		rref_scatter_kernel<<<dim3(ceil_div(nc, 256), RREF_W), 256, 0, ctx->stream>>>(X, ldx, nc, (int) c, st, Wb);
		rref_advance_kernel<<<1, 1, 0, ctx->stream>>>(st);
		MLZ_TRY(check_launch(ctx, "rref_scatter_kernel"));
		++panel;
		return MLZ_OK;
	};
	auto rank_now = [&](int *r) -> int {
		MLZ_HIP(ctx, hipMemcpyAsync(r, st.rank, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
		MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
		return MLZ_OK;
	};
	if (n < 2 * RREF_NB) {
		// narrow matrices: panels over the whole width
		for (size_t c = 0; c < n; c += RREF_W) {
			MLZ_TRY(panel_step(Wm, n, n, c, 0, -1));
			if ((panel & 7) == 0 || c + RREF_W >= n) {
				int r = 0;
				MLZ_TRY(rank_now(&r));
				if ((size_t) r >= m)
					break;
			}
		}
	} else {
		// Column blocks of RREF_NB: the block's panels update only the block
		// (X = its columns, plus a tracker block E that starts at zero and
		// receives a 1 in column j when the block's j-th pivot row is chosen).
		// Every row's trailing columns then depend on the original ones only
		// through the block's pivot rows TR: they become A + E * TR (with the
		// pivot rows' own 1 taken out), one gemm of inner size RREF_NB instead
		// of RREF_NB / RREF_W rank-RREF_W passes over the trailing matrix.
		const size_t xmax = 2 * RREF_NB;
		void *w4, *w5;
		MLZ_TRY(ws_get(ctx, WS_RREF_4, m * xmax * sizeof(uint32_t), &w4));
		MLZ_TRY(ws_get(ctx, WS_RREF_5, RREF_NB * (n - RREF_NB) * sizeof(uint32_t), &w5));
		uint32_t *X = (uint32_t *) w4, *TR = (uint32_t *) w5;
		for (size_t c0 = 0; c0 < n; c0 += RREF_NB) {
			const size_t nbw = std::min(RREF_NB, n - c0), nt = n - c0 - nbw;
			const size_t xw = nt ? 2 * nbw : nbw;
			{
				ProfScope ps(ctx, "rref_block");
				rref_block_load_kernel<<<dim3(ceil_div(xw, 256), (unsigned) std::min(m, RREF_COPY_ROWS)), 256, 0,
							 ctx->stream>>>(Wm, n, (int) m, (int) c0, (int) nbw, X, (int) xw, st);
			}
			MLZ_TRY(check_launch(ctx, "rref_block_load_kernel"));
			for (size_t c = 0; c < nbw; c += RREF_W)
				MLZ_TRY(panel_step(X, xw, nt ? xw : nbw, c, (int) c0, nt ? (int) nbw : -1));
			{
				ProfScope ps(ctx, "rref_block");
				rref_block_store_kernel<<<dim3(ceil_div(nbw, 256), (unsigned) std::min(m, RREF_COPY_ROWS)), 256, 0,
							  ctx->stream>>>(X, (int) xw, (int) m, (int) nbw, Wm, n, (int) c0);
				if (nt)
					rref_block_trail_kernel<<<dim3(ceil_div(nt, 256), nbw), 256, 0, ctx->stream>>>(
						X, (int) xw, (int) nbw, Wm, n, (int) c0, (int) nt, st, TR, pp);
			}
			MLZ_TRY(check_launch(ctx, "rref_block_trail_kernel"));
			if (nt)
				MLZ_TRY(zp_gemm_dev(ctx, MLZ_U32, p, m, nt, nbw, 1, X + nbw, xw, TR, nt, 1, Wm + c0 + nbw, n, algo));
			int r = 0;
			MLZ_TRY(rank_now(&r));
			if ((size_t) r >= m)
				break;
		}
	}
	int r = 0, bad = 0;
	MLZ_HIP(ctx, hipMemcpyAsync(&r, st.rank, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(&bad, st.bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	if (bad)
		return set_error(ctx, MLZ_E_NOT_INVERTIBLE, "Could not invert pivot-element in the ring");
	{
		ProfScope ps(ctx, "rref_store");
		if (elem == MLZ_U16)
			rref_copy_launch<MLZ_U16, false>(ctx, (uint16_t *) A, lda, m, n, pp, Wm, st);
		else if (elem == MLZ_U32)
			rref_copy_launch<MLZ_U32, false>(ctx, (uint32_t *) A, lda, m, n, pp, Wm, st);
		else
			rref_copy_launch<MLZ_F64_SYM, false>(ctx, (double *) A, lda, m, n, pp, Wm, st);
		MLZ_TRY(check_launch(ctx, "rref_store_kernel"));
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

// ===========================================================================
// Batched echelon forms: one workgroup per instance (NT threads: 1024 for the
// VALU-update variants, 512 for the MFMA ones, whose accumulators need the
// larger register budget).
//
// Instances of a batch are small (up to BR_MAXM rows) and many, so each one
// runs start to finish inside one workgroup taken from the work list, with
// all synchronisation by barriers.  Per panel of W columns:
//  1. pivot search: the rows that are not pivots are visited in order; each
//     round eliminates the current basis of the panel (at most W rows, kept
//     first so they stay pivots) together with the next live rows, up to
//     BR_CH rows, in registers (reg_eliminate), and keeps the pivot rows.  It
//     stops at full panel rank or when every row has been seen; the basis then
//     spans the panel's live rows, so its pivot columns J are their rank
//     profile.
//  2. P_J^-1 from [P_J | I] in registers, then Wb = P_J^-1 * P over columns
//     c..n-1 (identity on J);
//  3. every other row i, 64-row block at a time: its multipliers A_i[J] go to
//     LDS, then A_i <- A_i - A_i[J] * Wb; finally Wb is written into the pivot
//     rows.
// Products are summed exactly before one reduction per entry: in double
// precision when W*(p-1)^2 < 2^53 (FMA, full-rate FP64), else in uint64 with
// the 2^64-wrap correction.
// ===========================================================================
constexpr int BR_CH = 128;   // rows per pivot-search elimination
constexpr int BR_RB = 64;    // rows per update block (multipliers in LDS), VALU update
constexpr int BR_RB8 = 128;  // rows per update block, MFMA update (multipliers as limbs)
constexpr int BR_KC = 8;     // multipliers per register chunk
constexpr int BR_RQ = 8;     // rows per update unit of one wave

template <int W, bool FMA, bool I8>
struct BrLds {
	typedef typename std::conditional<FMA, double, uint32_t>::type GT;
	// MFMA update: the 64-row block's multipliers as INT8 A fragments,
	// [32-row block][k block][limb][64 lanes][16 bytes]
	__attribute__((aligned(16))) int8_t Ga[I8 ? BR_RB8 / 32 : 1][I8 ? W / 32 : 1][2][64][16];
	int rowid[BR_CH], alive[BR_CH], tmp[BR_CH];
	uint32_t f[BR_CH];
	uint32_t pivrow[2 * W];
	int outr[W], outc[W], J[W], pos[W];
	long long srcoff[W];
	int s_min[2];
	int slot, nr, cursor;
	uint8_t is_piv[BR_MAXM];     // 0 live, 1 pivot of an earlier panel, 2 of this panel
	uint8_t rownz[I8 ? BR_RB8 : BR_RB];
	uint32_t Pinv[W][W];
	GT G[I8 ? 1 : BR_RB][W];   // multipliers of the VALU update
};

// x (an exact integer below 2^53) mod p for p < 2^26.
__device__ __forceinline__ uint32_t br_mod_f64(double x, const ModP &mp)
{
	const double pd = (double) mp.p;
	const double q = floor(x * mp.pinv);
	double r = fma(-q, pd, x);
	r = r < 0.0 ? r + pd : r;
	r = r >= pd ? r - pd : r;
	return (uint32_t) r;
}

// (x - (65536 hi + 256 mid + lo)) mod p for x in [0, p), p <= 2^16 and INT8
// MFMA limb sums of up to 64-term dot products of signed bytes (|hi|, |lo| <=
// 2^20, |mid| <= 2^21).  mid and lo fold into x in int32 (|t| < 2^30); the
// rest is one exact double (|v| < 2^37) and a round-to-nearest quotient: the
// product v * (1/p) is within 2^-17 of v / p, so |v - q p| <= p/2 + 1, r + p
// lies in [0, 2p) and one unsigned min step finishes it.
__device__ __forceinline__ uint32_t br_sub_mod(uint32_t x, int hi, int mid, int lo, const ModP &mp)
{
	const int t = (int) x - (mid * 256 + lo);
	const double v = fma((double) hi, -65536.0, (double) t);
	const double q = rint(v * mp.pinv);
	const uint32_t p = (uint32_t) mp.p;
	const uint32_t u = (uint32_t) ((int) fma(-q, (double) p, v) + (int) p);
	return min(u, u - p);
}

// out[r * ld + j] (SUB: -)= sum_{s < k} G[r][s] * src[srcoff[s] + j] mod p for
// rows r < nrows with rownz[r] set and columns j < ncols.  Units of BR_RQ rows
// x 64 columns are spread over the waves; a lane owns one column and keeps
// BR_KC source values in registers while the multipliers (uniform across the
// wave) are read from LDS.
template <int NT, int W, bool FMA, bool I8, bool SUB, typename ST, typename OT>
__device__ __forceinline__ void br_apply(BrLds<W, FMA, I8> &S, int k, int nrows, int ncols, const ST *src, OT *out,
					 size_t ld, const ModP &mp, uint64_t two_64)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int nchunks = (ncols + 63) >> 6, nq = (nrows + BR_RQ - 1) / BR_RQ;
	const uint32_t p = (uint32_t) mp.p;
	for (int u = wave; u < nchunks * nq; u += NT / 64) {
		const int rq = u / nchunks, jc = u - rq * nchunks;
		const int j = jc * 64 + lane, r0 = rq * BR_RQ;
		const bool colok = j < ncols;
		if (FMA) {
			double acc[BR_RQ];
#pragma unroll
			for (int r = 0; r < BR_RQ; ++r)
				acc[r] = 0.0;
			for (int s0 = 0; s0 < k; s0 += BR_KC) {
				double wv[BR_KC];
#pragma unroll
				for (int q = 0; q < BR_KC; ++q)
					wv[q] = (s0 + q < k && colok) ? (double) src[S.srcoff[s0 + q] + j] : 0.0;
#pragma unroll
				for (int r = 0; r < BR_RQ; ++r) {
					if (r0 + r < nrows && S.rownz[r0 + r]) {
						const double *g = (const double *) &S.G[r0 + r][s0];
#pragma unroll
						for (int q = 0; q < BR_KC; ++q)
							acc[r] = fma(g[q], wv[q], acc[r]);
					}
				}
			}
#pragma unroll
			for (int r = 0; r < BR_RQ; ++r) {
				if (r0 + r < nrows && S.rownz[r0 + r] && colok) {
					OT *o = out + (size_t) (r0 + r) * ld + j;
					const uint32_t x = br_mod_f64(acc[r], mp);
					*o = (OT) (SUB ? mp.sub(*o, x) : x);
				}
			}
		} else {
			uint64_t acc[BR_RQ];
#pragma unroll
			for (int r = 0; r < BR_RQ; ++r)
				acc[r] = 0;
			for (int s0 = 0; s0 < k; s0 += BR_KC) {
				uint32_t wv[BR_KC];
#pragma unroll
				for (int q = 0; q < BR_KC; ++q)
					wv[q] = (s0 + q < k && colok) ? src[S.srcoff[s0 + q] + j] : 0u;
#pragma unroll
				for (int r = 0; r < BR_RQ; ++r) {
					if (r0 + r < nrows && S.rownz[r0 + r]) {
						const uint32_t *g = (const uint32_t *) &S.G[r0 + r][s0];
#pragma unroll
						for (int q = 0; q < BR_KC; ++q) {
							const uint64_t t = (uint64_t) g[q] * wv[q];
							acc[r] += t;
							if (acc[r] < t)
								acc[r] += two_64;
						}
					}
				}
			}
#pragma unroll
			for (int r = 0; r < BR_RQ; ++r) {
				if (r0 + r < nrows && S.rownz[r0 + r] && colok) {
					OT *o = out + (size_t) (r0 + r) * ld + j;
					const uint32_t x = (uint32_t) mp.reduce(acc[r]);
					*o = (OT) (SUB ? mp.sub(*o, x) : x);
				}
			}
		}
		(void) p;
	}
}

// Wb = P_J^-1 * P on INT8 MFMA, P the k pivot rows of the working copy
// (columns c .. c+nc-1), P_J^-1 as A fragments in S.Ga (entries outside k x k
// zero).  A wave takes a 32-column block: it loads the block's P fragments
// straight from the pivot rows, multiplies, writes the reduced pivot rows back
// into the working copy and the block's Wb fragments (the update's B operand,
// [32-column block][k block][limb][64 lanes][16 bytes], lane l holding column
// l%32 and k-values 16*(l/32) .. +15) into Wbp.  Rows k .. W-1 of Wb are zero.
template <int NT, int W, bool FMA>
__device__ __forceinline__ void br_wb_i8(BrLds<W, FMA, true> &S, int k, int nc, uint16_t *__restrict__ Wm, size_t n,
					 int c, const ModP &mp, int8_t *__restrict__ Wbp)
{
	constexpr int KB = W / 32;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int cl = lane & 31, hi = lane >> 5;
	const int ncb = (nc + 31) / 32;
	const uint32_t p = (uint32_t) mp.p;
	for (int cb = wave; cb < ncb; cb += NT / 64) {
		const int col = cb * 32 + cl;
		const bool colok = col < nc;
		uint16_t *const base = Wm + c + col;
		v4i bh[KB], bl[KB];
#pragma unroll
		for (int kb = 0; kb < KB; ++kb) {
			uint32_t x[16];
#pragma unroll
			for (int q = 0; q < 16; ++q) {
				const int s = kb * 32 + 16 * hi + q;
				x[q] = (s < k && colok) ? base[(size_t) S.rowid[s] * n] : 0u;
			}
			uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
			for (int q = 0; q < 16; ++q) {
				const uint32_t l = i8_limbs(x[q], p);
				hw[q >> 2] |= ((l >> 8) & 0xffu) << (8 * (q & 3));
				lw[q >> 2] |= (l & 0xffu) << (8 * (q & 3));
			}
			bh[kb] = v4i{(int) hw[0], (int) hw[1], (int) hw[2], (int) hw[3]};
			bl[kb] = v4i{(int) lw[0], (int) lw[1], (int) lw[2], (int) lw[3]};
		}
#pragma unroll
		for (int rb = 0; rb < KB; ++rb) {
			v16i acc0 = v16i{0}, acc1 = v16i{0}, acc2 = v16i{0};
#pragma unroll
			for (int kb = 0; kb < KB; ++kb) {
				const v4i ah = *(const v4i *) &S.Ga[rb][kb][0][lane][0];
				const v4i al = *(const v4i *) &S.Ga[rb][kb][1][lane][0];
				acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh[kb], acc0, 0, 0, 0);
				acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bl[kb], acc1, 0, 0, 0);
				acc2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bl[kb], acc2, 0, 0, 0);
				acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh[kb], acc1, 0, 0, 0);
			}
			// lane l holds column cl, rows 8(r/4) + 4 hi + r%4 of the block
			uint32_t d[16];
#pragma unroll
			for (int r = 0; r < 16; ++r) {
				const double v = fma((double) acc0[r], 65536.0, fma((double) acc1[r], 256.0, (double) acc2[r]));
				d[r] = mp.mod_f64(v);
				const int s = rb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
				if (s < k && colok)
					base[(size_t) S.rowid[s] * n] = (uint16_t) d[r];
			}
			// fragment order wants rows 16 hi + q in lane l: the halves swap
			// rows 4-7 / 12-15 (lower half) against 16-19 / 24-27 (upper)
			uint32_t f[16];
#pragma unroll
			for (int i = 0; i < 8; ++i) {
				const uint32_t y = (uint32_t) __shfl_xor((int) (hi ? d[i] : d[8 + i]), 32);
				const int q = (i & 3) + 8 * (i >> 2);   // y's place: rows 4-7 / 12-15 or 16-19 / 24-27
				if (hi) {
					f[q] = y;
					f[q + 4] = d[8 + i];
				} else {
					f[q] = d[(i & 3) + 4 * (i >> 2)];
					f[q + 4] = y;
				}
			}
			uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
			for (int q = 0; q < 16; ++q) {
				const uint32_t l = i8_limbs(f[q], p);
				hw[q >> 2] |= ((l >> 8) & 0xffu) << (8 * (q & 3));
				lw[q >> 2] |= (l & 0xffu) << (8 * (q & 3));
			}
			int8_t *fr = Wbp + (size_t) (cb * KB + rb) * 2 * 1024 + lane * 16;
			*(v4i *) fr = v4i{(int) hw[0], (int) hw[1], (int) hw[2], (int) hw[3]};
			*(v4i *) (fr + 1024) = v4i{(int) lw[0], (int) lw[1], (int) lw[2], (int) lw[3]};
		}
	}
}

// C (nb <= BR_RB8 rows x nc columns; row stride ldc a multiple of 4 and rows
// 8-byte aligned) -= G * Wb on INT8 MFMA, G from S.Ga, Wb from Wbp.  It is
// computed transposed, C^T -= Wb^T G^T: Wbp's fragments are the A operand and
// S.Ga's the B operand as they stand, and lane l of a 32 x 32 unit then owns
// row l%32 of C and four runs of four consecutive columns, 8 g + 4 (l/32) +
// 0..3, loaded and stored as 8-byte vectors.  Rows whose multipliers are all
// zero are not written; the row padding (columns nc .. ldc-1) may be.
// Software pipelined: a unit's outputs and Wb fragments are loaded while the
// previous unit's MFMAs and epilogue run (addresses clamped into the block,
// so the loads need no branch; clamped values are never written back).
template <int NT, int W, bool FMA>
__device__ __forceinline__ void br_update_i8(BrLds<W, FMA, true> &S, int nb, int nc, const int8_t *__restrict__ Wbp,
					     uint16_t *__restrict__ C, size_t ldc, const ModP &mp)
{
	constexpr int KB = W / 32;
	const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int ncb = (nc + 31) / 32, nrb = (nb + 31) / 32, nu = nrb * ncb;
	const int rl = lane & 31, c4 = 4 * (lane >> 5), clast = (nc - 1) & ~3;
	uint2 cv[4];
	v4i bh[KB], bl[KB];
	auto fetch = [&](int rb, int cb, uint2(&cx)[4], v4i(&hx)[KB], v4i(&lx)[KB]) {
		const uint16_t *cc = C + (size_t) min(rb * 32 + rl, nb - 1) * ldc;
#pragma unroll
		for (int g = 0; g < 4; ++g)
			cx[g] = *(const uint2 *) (cc + min(cb * 32 + 8 * g + c4, clast));
#pragma unroll
		for (int kb = 0; kb < KB; ++kb) {
			const int8_t *gb = Wbp + (size_t) (cb * KB + kb) * 2 * 1024 + lane * 16;
			hx[kb] = *(const v4i *) gb;
			lx[kb] = *(const v4i *) (gb + 1024);
		}
	};
	// unit u = wave + (NT/64) i as (row block, column block), stepped forward
	// without a division (wave-uniform)
	auto advance = [&](int &rb, int &cb) {
		cb += NT / 64;
		while (cb >= ncb) {
			cb -= ncb;
			++rb;
		}
	};
	int rb = wave / ncb, cb = wave - rb * ncb, rbn = rb, cbn = cb;
	advance(rbn, cbn);
	if (wave < nu)
		fetch(rb, cb, cv, bh, bl);
	for (int u = wave; u < nu; u += NT / 64) {
		uint2 cn[4];
		v4i nh[KB], nl[KB];
		const bool more = u + NT / 64 < nu;   // else the current unit again (loads stay in bounds)
		fetch(more ? rbn : rb, more ? cbn : cb, cn, nh, nl);
		v16i acc0 = v16i{0}, acc1 = v16i{0}, acc2 = v16i{0};
#pragma unroll
		for (int kb = 0; kb < KB; ++kb) {
			const v4i ah = *(const v4i *) &S.Ga[rb][kb][0][lane][0];
			const v4i al = *(const v4i *) &S.Ga[rb][kb][1][lane][0];
			acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(bh[kb], ah, acc0, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(bl[kb], ah, acc1, 0, 0, 0);
			acc2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(bl[kb], al, acc2, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(bh[kb], al, acc1, 0, 0, 0);
		}
		const int row = rb * 32 + rl;
		if (row < nb && S.rownz[row]) {
			uint16_t *const out = C + (size_t) row * ldc;
#pragma unroll
			for (int g = 0; g < 4; ++g) {
				const int col0 = cb * 32 + 8 * g + c4;
				if (col0 < nc) {
					const uint32_t xw[2] = {cv[g].x, cv[g].y};
					uint32_t yw[2] = {0u, 0u};
#pragma unroll
					for (int j = 0; j < 4; ++j) {
						const int r = 4 * g + j;
						const uint32_t x = (xw[j >> 1] >> (16 * (j & 1))) & 0xffffu;
						yw[j >> 1] |= br_sub_mod(x, acc0[r], acc1[r], acc2[r], mp) << (16 * (j & 1));
					}
					*(uint2 *) (out + col0) = make_uint2(yw[0], yw[1]);
				}
			}
		}
#pragma unroll
		for (int g = 0; g < 4; ++g)
			cv[g] = cn[g];
#pragma unroll
		for (int kb = 0; kb < KB; ++kb) {
			bh[kb] = nh[kb];
			bl[kb] = nl[kb];
		}
		rb = rbn;
		cb = cbn;
		advance(rbn, cbn);
	}
}

// Waves per SIMD the batched echelon kernel is compiled for: the INT8 variant
// with 32-column panels fits two 512-thread workgroups per CU (2 x 20 KB of
// LDS, 128 VGPRs), so one instance's pivot search overlaps the other's update.
constexpr int br_waves_per_simd(int W, bool I8, int NT)
{
	return I8 && W == 32 ? 4 : NT / 256;
}

template <int E, int W, bool SMALL, bool FMA, bool I8, int NT>
__global__ __launch_bounds__(NT, br_waves_per_simd(W, I8, NT)) void batch_rref_kernel(const BatchRref *__restrict__ tasks, BatchLaunch L,
								ModP mp, uint64_t two_64, int *__restrict__ ranks,
								int *__restrict__ bad)
{
	typedef typename Elem<E>::T T;
	typedef typename BrLds<W, FMA, I8>::GT GT;
	// working copy: residues fit 16 bits for the MFMA variants (p <= 2^16)
	typedef typename std::conditional<I8, uint16_t, uint32_t>::type WT;
	constexpr int RPG = BR_CH / (NT / W);             // pivot search: rows per thread
	constexpr int NCI = 2 * W, RPGI = W / (NT / NCI);  // inverse: rows per thread
	__shared__ BrLds<W, FMA, I8> S;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const uint32_t p = (uint32_t) mp.p;
	for (;;) {
		const int it = batch_next(L, &S.slot);
		if (it < 0)
			break;
		const BatchItem w = L.items[it];
		const BatchRref &t = tasks[w.task];
		const uint64_t t0 = wall_clock64();
		const int m = t.m, n = t.n;
		// working copy, rows padded to a multiple of 4 entries for the MFMA
		// variants (8-byte aligned rows for the update's vector accesses)
		const int ldw = br_ldw(n, I8);
		WT *Wm = (WT *) t.Wm;
		uint32_t *Wb = t.Wb;
		int *badp = bad + w.task;
		// 16-bit storage into the 16-bit working copy: 4-byte aligned rows move
		// as pairs (rows of the copy are 8-byte aligned)
		constexpr bool PAIRS = I8 && sizeof(T) == 2;
		for (int r = wave; r < m; r += NT / 64) {
			const T *src = (const T *) t.A + (size_t) r * t.lda;
			WT *dst = Wm + (size_t) r * ldw;
			if (PAIRS && ((uintptr_t) src & 3) == 0) {
				for (int c = 2 * lane; c < n; c += 128) {
					if (c + 1 < n)
						*(uint32_t *) (dst + c) = *(const uint32_t *) (src + c);
					else
						dst[c] = (WT) Elem<E>::load(src + c, p);
				}
			} else {
				for (int c = lane; c < n; c += 64)
					dst[c] = (WT) Elem<E>::load(src + c, p);
			}
		}
		for (int r = tid; r < m; r += NT)
			S.is_piv[r] = 0;
		__syncthreads();
		int rank = 0;
		for (int c = 0; c < n && rank < m; c += W) {
			const int wc = min(W, n - c), nc = n - c;
			// ---- 1. pivot search.  The first pass reduces [S | I] for S = the
			// first wc live rows; full panel rank gives the pivot rows and P_J^-1
			// at once.  Otherwise passes of up to BR_CH rows from the top find the
			// rank profile, and step 2 inverts P_J.
			int cnt = 0;
			bool have_inv = false;
			for (int round = 0;; ++round) {
				const bool fast = round == 0;
				const int lim = fast ? wc : BR_CH;
				if (wave == 0) {
					int have = fast || round == 1 ? 0 : cnt, cur = fast || round == 1 ? 0 : S.cursor;
					while (have < lim && cur < m) {
						const int row = cur + lane;
						const bool live = row < m && S.is_piv[row] == 0;
						const uint64_t bal = __ballot(live);
						const int before = __popcll(bal & ((1ull << lane) - 1ull));
						const int room = lim - have, tot = __popcll(bal);
						if (live && before < room)
							S.rowid[have + before] = row;
						if (tot <= room) {
							have += tot;
							cur += 64;
						} else {
							uint64_t bb = bal;
							for (int q = 0; q < room - 1; ++q)
								bb &= bb - 1;
							have = lim;
							cur += __ffsll((unsigned long long) bb);
						}
					}
					if (lane == 0) {
						S.nr = have;
						S.cursor = cur;
					}
				}
				__syncthreads();
				const int nr = S.nr;
				const bool seen_all = S.cursor >= m;
				if (fast) {
					const int ci = tid % NCI, gi = tid / NCI;
					uint32_t v[RPGI];
#pragma unroll
					for (int i = 0; i < RPGI; ++i) {
						const int r = gi * RPGI + i;
						if (ci < W)
							v[i] = (r < nr && ci < wc) ? Wm[(size_t) S.rowid[r] * ldw + c + ci] : 0u;
						else
							v[i] = (ci - W == r) ? 1u : 0u;
					}
					for (int r = tid; r < W; r += NT)
						S.alive[r] = r < nr;
					const int kf = reg_eliminate<NCI, RPGI, SMALL>(v, W, wc, true, mp, S.alive, S.f, S.pivrow,
										       S.s_min, S.outr, S.outc, badp);
					if (kf == wc) {
						// P_J = Q S, P_J^-1 = S^-1 Q^T: entry (c, t) = R[row of pivot c][row of pivot t]
						for (int q = tid; q < kf; q += NT) {
							S.pos[S.outr[q]] = q;
							S.tmp[q] = S.rowid[S.outr[q]];
						}
						__syncthreads();
						if (ci >= W) {
#pragma unroll
							for (int i = 0; i < RPGI; ++i) {
								const int r = gi * RPGI + i;
								if (r < wc && ci - W < wc)
									S.Pinv[S.pos[r]][S.pos[ci - W]] = v[i];
							}
						}
						for (int q = tid; q < kf; q += NT) {
							S.rowid[q] = S.tmp[q];
							S.J[q] = S.outc[q];
						}
						__syncthreads();
						cnt = kf;
						have_inv = true;
						break;
					}
					__syncthreads();
					continue;   // rank deficient: search from the top
				}
				for (int r = tid; r < BR_CH; r += NT)
					S.alive[r] = r < nr;
				{
					const int cc = tid % W, g = tid / W;
					uint32_t v[RPG];
#pragma unroll
					for (int i = 0; i < RPG; ++i) {
						const int r = g * RPG + i;
						v[i] = (r < nr && cc < wc) ? Wm[(size_t) S.rowid[r] * ldw + c + cc] : 0u;
					}
					cnt = reg_eliminate<W, RPG, SMALL>(v, nr, wc, false, mp, S.alive, S.f, S.pivrow, S.s_min,
									   S.outr, S.outc, badp);
				}
				for (int q = tid; q < cnt; q += NT)
					S.tmp[q] = S.rowid[S.outr[q]];
				__syncthreads();
				for (int q = tid; q < cnt; q += NT) {
					S.rowid[q] = S.tmp[q];
					S.J[q] = S.outc[q];
				}
				__syncthreads();
				if (cnt == wc || seen_all)
					break;
			}
			const int k = cnt;
			if (k == 0)
				continue;
			// ---- 2. P_J^-1 (unless the first pass produced it), then Wb = P_J^-1 * P
			if (!have_inv) {
				const int ci = tid % NCI, gi = tid / NCI;
				uint32_t v[RPGI];
#pragma unroll
				for (int i = 0; i < RPGI; ++i) {
					const int r = gi * RPGI + i;
					uint32_t x;
					if (ci < W)
						x = (r < k && ci < k) ? Wm[(size_t) S.rowid[r] * ldw + c + S.J[ci]] : (r == ci ? 1u : 0u);
					else
						x = (ci - W == r) ? 1u : 0u;
					v[i] = x;
				}
				for (int r = tid; r < W; r += NT) {
					S.alive[r] = 1;
					S.pos[r] = r;
				}
				const int ki = reg_eliminate<NCI, RPGI, SMALL>(v, W, k, true, mp, S.alive, S.f, S.pivrow, S.s_min,
									       S.outr, S.outc, badp);
				if (tid == 0 && ki < k)
					*badp = 1;
				for (int q = tid; q < ki; q += NT)
					S.pos[S.outr[q]] = S.outc[q];
				__syncthreads();
				if (ci >= W) {
#pragma unroll
					for (int i = 0; i < RPGI; ++i)
						S.Pinv[S.pos[gi * RPGI + i]][ci - W] = v[i];
				}
			}
			for (int s = tid; s < k; s += NT) {
				S.srcoff[s] = (long long) S.rowid[s] * ldw + c;
				S.is_piv[S.rowid[s]] = 2;
				t.piv_rows[rank + s] = S.rowid[s];
				t.piv_cols[rank + s] = c + S.J[s];
			}
			for (int r = tid; r < BR_RB; r += NT)
				S.rownz[r] = 1;
			__syncthreads();
			if constexpr (I8) {
				// P_J^-1 as A fragments, then Wb on MFMA straight into the pivot
				// rows and the update's fragments
				for (int e = tid; e < W * W; e += NT) {
					const int r = e / W, s = e % W;
					const uint32_t l = i8_limbs(r < k && s < k ? S.Pinv[r][s] : 0u, p);
					const int ln = (r & 31) + 32 * ((s & 31) >> 4), by = s & 15;
					S.Ga[r >> 5][s >> 5][0][ln][by] = (int8_t) (l >> 8);
					S.Ga[r >> 5][s >> 5][1][ln][by] = (int8_t) l;
				}
				__syncthreads();
				br_wb_i8<NT, W, FMA>(S, k, nc, Wm, ldw, c, mp, t.Wbp);
				__syncthreads();
			} else {
				for (int e = tid; e < k * W; e += NT) {
					const int r = e / W, s = e % W;
					S.G[r][s] = (GT) (s < k ? S.Pinv[r][s] : 0u);
				}
				__syncthreads();
				br_apply<NT, W, FMA, I8, false>(S, k, k, nc, Wm, Wb, n, mp, two_64);
				__syncthreads();
			}
			// ---- 3. eliminate column set J from every other row
			for (int s = tid; s < k; s += NT)
				S.srcoff[s] = (long long) s * n;
			constexpr int RB = I8 ? BR_RB8 : BR_RB;
			for (int i0 = 0; i0 < m; i0 += RB) {
				const int nb = min(RB, m - i0);
				if constexpr (I8) {
					// multipliers straight into A fragments: element (r, s) is byte
					// s%16 of lane r%32 + 32*((s%32)/16) of fragment (r/32, s/32)
					for (int r = tid; r < RB; r += NT)
						S.rownz[r] = 0;
					__syncthreads();
					// four consecutive multipliers of a row per thread, one 32-bit LDS
					// word per limb.  A full-rank panel has J[s] = s (the search takes
					// the columns in order), and the four are one 8-byte load of the
					// row (rows are 8-byte aligned, the padding keeps it in bounds).
					const bool ident = k == wc;
					for (int e = tid; e < ((nb + 31) & ~31) * (W / 4); e += NT) {
						const int r = e / (W / 4), s0 = 4 * (e % (W / 4));
						const int i = i0 + r;
						const bool live = r < nb && S.is_piv[i] != 2;
						const WT *row = Wm + (size_t) i * ldw + c;
						uint32_t g[4];
						if (ident) {
							const uint2 v = live && s0 < k ? *(const uint2 *) (row + s0) : make_uint2(0u, 0u);
#pragma unroll
							for (int j = 0; j < 4; ++j)
								g[j] = s0 + j < k ? ((j < 2 ? v.x : v.y) >> (16 * (j & 1))) & 0xffffu : 0u;
						} else {
#pragma unroll
							for (int j = 0; j < 4; ++j)
								g[j] = live && s0 + j < k ? row[S.J[s0 + j]] : 0u;
						}
						if (g[0] | g[1] | g[2] | g[3])
							S.rownz[r] = 1;
						uint32_t hw = 0, lw = 0;
#pragma unroll
						for (int j = 0; j < 4; ++j) {
							const uint32_t l = i8_limbs(g[j], p);
							hw |= ((l >> 8) & 0xffu) << (8 * j);
							lw |= (l & 0xffu) << (8 * j);
						}
						const int ln = (r & 31) + 32 * ((s0 & 31) >> 4), by = s0 & 15;
						*(uint32_t *) &S.Ga[r >> 5][s0 >> 5][0][ln][by] = hw;
						*(uint32_t *) &S.Ga[r >> 5][s0 >> 5][1][ln][by] = lw;
					}
					__syncthreads();
					br_update_i8<NT, W, FMA>(S, nb, nc, t.Wbp, (uint16_t *) Wm + (size_t) i0 * ldw + c, ldw, mp);
					__syncthreads();
				} else {
					for (int e = tid; e < nb * W; e += NT) {
						const int r = e / W, s = e % W;
						const int i = i0 + r;
						uint32_t g = 0;
						if (s < k && S.is_piv[i] != 2)
							g = Wm[(size_t) i * ldw + c + S.J[s]];
						S.G[r][s] = (GT) g;
					}
					__syncthreads();
					// rownz[r]: some multiplier of row i0 + r is nonzero (a wave per row)
					for (int r = wave; r < nb; r += NT / 64) {
						const bool nz = lane < k && S.G[r][lane] != (GT) 0;
						const uint64_t bal = __ballot(nz);
						if (lane == 0)
							S.rownz[r] = bal != 0;
					}
					__syncthreads();
					br_apply<NT, W, FMA, I8, true>(S, k, nb, nc, Wb, Wm + (size_t) i0 * ldw + c, ldw, mp, two_64);
					__syncthreads();
				}
			}
			if constexpr (!I8)
				for (int e = tid; e < k * nc; e += NT) {
					const int s = e / nc, j = e % nc;
					Wm[(size_t) S.rowid[s] * ldw + c + j] = (WT) Wb[(size_t) s * n + j];
				}
			for (int s = tid; s < k; s += NT)
				S.is_piv[S.rowid[s]] = 1;
			for (int r = tid; r < (I8 ? BR_RB8 : BR_RB); r += NT)
				S.rownz[r] = 1;
			rank += k;
			__syncthreads();
		}
		// ---- standard layout: pivot rows in pivot-column order, then zero rows
		for (int r = wave; r < m; r += NT / 64) {
			const int src = r < rank ? t.piv_rows[r] : -1;
			T *dst = (T *) t.A + (size_t) r * t.lda;
			if (PAIRS && ((uintptr_t) dst & 3) == 0) {
				const WT *srow = Wm + (size_t) (src >= 0 ? src : 0) * ldw;
				for (int c = 2 * lane; c < n; c += 128) {
					if (c + 1 < n)
						*(uint32_t *) (dst + c) = src >= 0 ? *(const uint32_t *) (srow + c) : 0u;
					else
						Elem<E>::store(dst + c, src >= 0 ? (uint32_t) srow[c] : 0u, p);
				}
			} else {
				for (int c = lane; c < n; c += 64) {
					const uint32_t x = src >= 0 ? Wm[(size_t) src * ldw + c] : 0u;
					Elem<E>::store(dst + c, x, p);
				}
			}
		}
		if (t.pivots)
			for (int s = tid; s < rank; s += NT)
				t.pivots[s] = (uint32_t) t.piv_cols[s];
		if (tid == 0)
			ranks[w.task] = rank;
		__syncthreads();
		batch_charge(L, w.task, t0);
	}
}

template <int E, int W>
static void batch_rref_launch_w(hipStream_t stream, bool mfma, const ModP &mp, uint64_t two_64,
				const BatchRref *tasks, int *ranks, int *bad, BatchLaunch L, int grid)
{
	const bool small = mp.p < (1ull << 26);
	const bool fma = (unsigned __int128) W * (mp.p - 1) * (mp.p - 1) < ((unsigned __int128) 1 << 53);
	if (mfma)   // p <= 2^16: the fused and FP64 paths apply
		batch_rref_kernel<E, W, true, true, true, 512><<<grid, 512, 0, stream>>>(tasks, L, mp, two_64, ranks, bad);
	else if (fma)
		batch_rref_kernel<E, W, true, true, false, 1024><<<grid, 1024, 0, stream>>>(tasks, L, mp, two_64, ranks, bad);
	else if (small)
		batch_rref_kernel<E, W, true, false, false, 1024><<<grid, 1024, 0, stream>>>(tasks, L, mp, two_64, ranks,
												  bad);
	else
		batch_rref_kernel<E, W, false, false, false, 1024><<<grid, 1024, 0, stream>>>(tasks, L, mp, two_64, ranks,
												   bad);
}

template <int E>
static void batch_rref_launch_e(hipStream_t stream, int variant, const ModP &mp, uint64_t two_64,
				const BatchRref *tasks, int *ranks, int *bad, BatchLaunch L, int grid)
{
	if (br_width(variant) == 32)
		batch_rref_launch_w<E, 32>(stream, br_mfma(variant), mp, two_64, tasks, ranks, bad, L, grid);
	else
		batch_rref_launch_w<E, 64>(stream, br_mfma(variant), mp, two_64, tasks, ranks, bad, L, grid);
}

int batch_rref_launch(hipStream_t stream, int elem, int variant, uint32_t p, const BatchRref *tasks, int *ranks,
		      int *bad, BatchLaunch L, int grid)
{
	const ModP mp = make_modp(p);
	uint64_t two_64 = 2;
	for (int i = 0; i < 6; ++i)
		two_64 = (two_64 * two_64) % p;
	if (elem == MLZ_U16)
		batch_rref_launch_e<MLZ_U16>(stream, variant, mp, two_64, tasks, ranks, bad, L, grid);
	else if (elem == MLZ_U32)
		batch_rref_launch_e<MLZ_U32>(stream, variant, mp, two_64, tasks, ranks, bad, L, grid);
	else
		batch_rref_launch_e<MLZ_F64_SYM>(stream, variant, mp, two_64, tasks, ranks, bad, L, grid);
	return hipGetLastError() == hipSuccess ? MLZ_OK : MLZ_E_HIP;
}

int batch_rref_wgs_per_cu(int elem, int variant, uint32_t p)
{
	(void) elem;
	(void) p;
	// INT8 with 32-column panels: two 512-thread workgroups per CU; the others
	// take a CU each (1024 threads, or 256 VGPRs for INT8 with 64 columns)
	return variant == BR_I8_W32 ? 2 : 1;
}

}  // namespace mlz

using namespace mlz;

extern "C" int mlz_zp_rref(mlz_ctx *ctx, mlz_elem elem, uint64_t p, size_t m, size_t n, void *A, size_t lda,
			   size_t *rank, uint32_t *pivots, int algo)
{
	if (!ctx)
		return MLZ_E_ARG;
	if (p < 2 || (elem == MLZ_U16 && p > 65535) || (elem == MLZ_U32 && p > 0xffffffffull) ||
	    (elem == MLZ_F64_SYM && p >= 67108864ull) || elem_size(elem) == 0)
		return set_error(ctx, MLZ_E_ARG, "bad modulus or element type");
	if (m && n && (!A || lda < n))
		return set_error(ctx, MLZ_E_ARG, "bad matrix pointer or leading dimension");
	if (m > (size_t) INT32_MAX / 2 || n > (size_t) INT32_MAX / 2 || m * n > ((size_t) 1 << 34))
		return set_error(ctx, MLZ_E_ARG, "matrix too large");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (ctx->ptr_mode == MLZ_PTR_DEVICE)
		return zp_rref_dev(ctx, elem, p, m, n, A, lda, rank, pivots, algo);
	if (m == 0 || n == 0) {
		if (rank)
			*rank = 0;
		return MLZ_OK;
	}
	const size_t es = elem_size(elem);
	void *dA = nullptr;
	MLZ_TRY(stage_in(ctx, WS_STAGE_A, A, m, n, lda, es, &dA));
	MLZ_TRY(zp_rref_dev(ctx, elem, p, m, n, dA, n, rank, pivots, algo));
	return stage_out(ctx, A, dA, m, n, lda, es);
}
