// lela_amd/csrc/batch.hip -- batches of independent Z/p gemm and echelon-form
// instances: variant choice (fixed, autotuner table, cost model), work lists,
// device workspace, hipGraph capture, per-task timing, and the autotuner.
//
// LELA runs such instances one call at a time (BLAS3::gemm,
// lela/blas/level3.h:81-83; EchelonForm::echelonize,
// lela/solutions/echelon-form.h:103-154); the results here are those calls'
// results instance by instance.  What a batch adds is throughput: every
// kernel variant is a persistent launch over a work list that mixes the
// tiles of all instances that chose it, so 4096 instances of 32..1024 rows
// fill the 256 CUs instead of queueing 4096 launches each too small to.
//
// Execution of a plan:
//   zero counters -> fork -> branch 0: INT8 pack, INT8 64x64, INT8 128x128
//                            branch 1: plain      branch 2: sparse-A
//                            branch 3: rref 64-column panels (VALU, MFMA)
//                            branch 4: rref 32-column panels (VALU, MFMA)
//                 -> join
// The branches are independent streams joined by events, so in a captured
// graph they are parallel nodes; MLZ_BATCH_SERIAL puts everything on the
// context's stream (used for timing, so variants do not share CUs).
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

// Context-owned state of the autotuner and the cost model.
struct mlz_tuner {
	std::vector<int8_t> table[3];       // by op (1, 2): 16 x 16 x 16 buckets, -1 empty
	bool has_table[3] = {false, false, false};
	std::vector<float> model[3];        // by op: MLZ_MODEL_PARAMS weights
	bool has_model[3] = {false, false, false};
};

namespace mlz {

// Cost-model kernels (mlp.hip).
int mlp_predict_dev(mlz_ctx *ctx, const float *params, const float *x, size_t n, float *y);

constexpr int TB = 16;   // buckets per dimension of the table

void tuner_free(mlz_ctx *ctx)
{
	delete ctx->tuner;
	ctx->tuner = nullptr;
}

static mlz_tuner *tuner_of(mlz_ctx *ctx)
{
	if (!ctx->tuner) {
		ctx->tuner = new mlz_tuner();
		for (int op = 1; op <= 2; ++op) {
			ctx->tuner->table[op].assign(TB * TB * TB, -1);
			ctx->tuner->model[op].assign(MLZ_MODEL_PARAMS, 0.0f);
		}
	}
	return ctx->tuner;
}

static int log2_floor(uint64_t x)
{
	int l = 0;
	while (x > 1) {
		x >>= 1;
		++l;
	}
	return l;
}

static uint64_t task_k(const mlz_task &t)
{
	return t.op == MLZ_TASK_GEMM ? t.k : std::min(t.m, t.n);
}

static int bucket_of(const mlz_task &t)
{
	const int lm = std::min(log2_floor(t.m), TB - 1), ln = std::min(log2_floor(t.n), TB - 1);
	const int lk = std::min(log2_floor(task_k(t)), TB - 1);
	return (lm * TB + ln) * TB + lk;
}

static int nvariants(int op)
{
	return op == MLZ_TASK_GEMM ? BG_NVAR : BR_NVAR;
}

static bool variant_ok(const mlz_task &t, int v, uint64_t p)
{
	if (t.op == MLZ_TASK_GEMM) {
		if (v < 0 || v >= BG_NVAR)
			return false;
		if (v == BG_I8S || v == BG_I8L)
			return p <= 65536 && t.k <= 65504;
		return true;
	}
	if (v == BR_I8_W64 || v == BR_I8_W32)
		return p <= 65536;
	return v >= 0 && v < BR_NVAR;
}

// Shape rule used when neither a table entry nor a model applies.
static int rule_variant(const mlz_task &t, uint64_t p)
{
	if (t.op == MLZ_TASK_RREF)
		return p <= 65536 ? BR_I8_W64 : BR_W32;
	if (p <= 65536 && t.k <= 65504)
		return (uint64_t) t.m * t.n * t.k >= ((uint64_t) 1 << 24) ? BG_I8L : BG_I8S;
	return BG_PLAIN;
}

static int table_variant(mlz_tuner *tu, const mlz_task &t, uint64_t p)
{
	if (!tu || !tu->has_table[t.op])
		return -1;
	const std::vector<int8_t> &tab = tu->table[t.op];
	const int b = bucket_of(t);
	if (tab[b] >= 0 && variant_ok(t, tab[b], p))
		return tab[b];
	// nearest filled bucket (L1 distance in log2 space)
	const int bm = b / (TB * TB), bn = (b / TB) % TB, bk = b % TB;
	int best = -1, bestd = std::numeric_limits<int>::max();
	for (int e = 0; e < TB * TB * TB; ++e) {
		if (tab[e] < 0 || !variant_ok(t, tab[e], p))
			continue;
		const int d = std::abs(e / (TB * TB) - bm) + std::abs((e / TB) % TB - bn) + std::abs(e % TB - bk);
		if (d < bestd) {
			bestd = d;
			best = tab[e];
		}
	}
	return best;
}

void task_features(const mlz_task &t, uint64_t p, float *x)
{
	const double kk = (double) task_k(t);
	x[0] = (float) ((std::log2((double) t.m) - 7.5) / 1.5);
	x[1] = (float) ((std::log2((double) t.n) - 7.5) / 1.5);
	x[2] = (float) ((std::log2(kk) - 7.5) / 1.5);
	// log2 of the density: the sparse-A gemm wins below about 2^-8 of nonzeros
	x[3] = (float) ((std::log2(std::max((double) t.density, 1.0 / 65536.0)) + 8.0) / 8.0);
	x[4] = (float) ((std::log2((double) p) - 16.0) / 8.0);
}

}  // namespace mlz

using namespace mlz;

struct mlz_batch {
	int elem = 0;
	uint64_t p = 0;
	int flags = 0;
	size_t ntasks = 0;
	std::vector<int> variant;
	std::vector<int> rref_slot;         // task -> index in the rank array (-1: gemm)
	char *arena = nullptr;
	size_t zero_words = 0;              // 32-bit words cleared before each execution
	int *d_bad = nullptr;
	int *d_ranks = nullptr;
	unsigned long long *d_tcost = nullptr;   // [ntasks] main kernels, [ntasks] INT8 pack
	struct Launch {
		int kind;      // 0 pack, 1 gemm, 2 rref
		int variant;
		int branch;
		int grid;
		BatchLaunch L;
	};
	std::vector<Launch> launches;
	BatchGemm *d_gemm = nullptr;
	BatchRref *d_rref = nullptr;
	float occ_gemm[BG_NVAR + 1] = {1, 1, 1, 1, 1};
	float occ_rref[BR_NVAR] = {1, 1};
	uint64_t kmax = 1;
	hipStream_t side[4] = {nullptr, nullptr, nullptr, nullptr};
	hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
};

namespace mlz {

__global__ void batch_zero_kernel(uint32_t *__restrict__ p, size_t words)
{
	for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t) gridDim.x * blockDim.x)
		p[i] = 0u;
}

static void free_plan(mlz_batch *b)
{
	if (b->exec)
		(void) hipGraphExecDestroy(b->exec);
	if (b->graph)
		(void) hipGraphDestroy(b->graph);
	for (auto &s : b->side)
		if (s)
			(void) hipStreamDestroy(s);
	for (auto &e : b->ev)
		if (e)
			(void) hipEventDestroy(e);
	if (b->arena)
		(void) hipFree(b->arena);
	delete b;
}

static int enqueue(mlz_ctx *ctx, mlz_batch *b)
{
	hipStream_t s0 = ctx->stream;
	if (b->zero_words) {
		const unsigned g = (unsigned) std::min<size_t>(1024, ceil_div(b->zero_words, 256));
		batch_zero_kernel<<<g, 256, 0, s0>>>((uint32_t *) b->arena, b->zero_words);
		MLZ_TRY(check_launch(ctx, "batch_zero_kernel"));
	}
	const bool serial = (b->flags & MLZ_BATCH_SERIAL) != 0;
	bool used[5] = {false, false, false, false, false};
	for (auto &l : b->launches)
		used[l.branch] = true;
	if (!serial) {
		MLZ_HIP(ctx, hipEventRecord(b->ev[0], s0));
		for (int br = 1; br < 5; ++br)
			if (used[br])
				MLZ_HIP(ctx, hipStreamWaitEvent(b->side[br - 1], b->ev[0], 0));
	}
	for (auto &l : b->launches) {
		hipStream_t s = (serial || l.branch == 0) ? s0 : b->side[l.branch - 1];
		int rc;
		if (l.kind == 0)
			rc = batch_pack_i8_launch(s, b->elem, (uint32_t) b->p, b->d_gemm, l.L, l.grid);
		else if (l.kind == 1)
			rc = batch_gemm_launch(s, b->elem, l.variant, (uint32_t) b->p, b->kmax, b->d_gemm, l.L, l.grid);
		else
			rc = batch_rref_launch(s, b->elem, l.variant, (uint32_t) b->p, b->d_rref, b->d_ranks, b->d_bad, l.L,
					       l.grid);
		if (rc != MLZ_OK)
			return set_error(ctx, rc, "batched kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
	}
	if (!serial) {
		for (int br = 1; br < 5; ++br)
			if (used[br]) {
				MLZ_HIP(ctx, hipEventRecord(b->ev[br], b->side[br - 1]));
				MLZ_HIP(ctx, hipStreamWaitEvent(s0, b->ev[br], 0));
			}
	}
	return MLZ_OK;
}

static int choose_variants(mlz_ctx *ctx, uint64_t p, mlz_task *tasks, size_t nt, int policy)
{
	mlz_tuner *tu = ctx->tuner;
	// cost model: one inference launch per op over its tasks
	std::vector<float> pred[3];
	std::vector<size_t> idx[3];
	if (policy == MLZ_POLICY_MODEL) {
		for (int op = 1; op <= 2; ++op) {
			if (!tu || !tu->has_model[op])
				continue;
			std::vector<float> x;
			for (size_t i = 0; i < nt; ++i)
				if (tasks[i].op == op && tasks[i].variant < 0) {
					idx[op].push_back(i);
					float f[5];
					task_features(tasks[i], p, f);
					x.insert(x.end(), f, f + 5);
				}
			if (idx[op].empty())
				continue;
			pred[op].resize(idx[op].size() * MLZ_MAX_VARIANTS);
			MLZ_TRY(mlp_predict_dev(ctx, tu->model[op].data(), x.data(), idx[op].size(), pred[op].data()));
		}
	}
	std::vector<long> pos(nt, -1);
	for (int op = 1; op <= 2; ++op)
		for (size_t j = 0; j < idx[op].size(); ++j)
			pos[idx[op][j]] = (long) j;
	for (size_t i = 0; i < nt; ++i) {
		mlz_task &t = tasks[i];
		if (t.variant >= 0) {
			if (!variant_ok(t, t.variant, p))
				return set_error(ctx, MLZ_E_UNSUPPORTED, "task %zu: variant %d does not apply", i, t.variant);
			continue;
		}
		if (policy == MLZ_POLICY_FIXED)
			return set_error(ctx, MLZ_E_ARG, "task %zu: policy FIXED needs a variant", i);
		int v = -1;
		if (policy == MLZ_POLICY_MODEL && pos[i] >= 0) {
			const float *y = &pred[t.op][(size_t) pos[i] * MLZ_MAX_VARIANTS];
			float best = std::numeric_limits<float>::infinity();
			for (int c = 0; c < nvariants(t.op); ++c)
				if (variant_ok(t, c, p) && y[c] < best) {
					best = y[c];
					v = c;
				}
		}
		if (v < 0)
			v = table_variant(tu, t, p);
		if (v < 0)
			v = rule_variant(t, p);
		t.variant = v;
	}
	return MLZ_OK;
}

static int validate(mlz_ctx *ctx, int elem, uint64_t p, const mlz_task *tasks, size_t nt)
{
	MLZ_TRY(check_modulus(ctx, elem, p));
	if (nt && !tasks)
		return set_error(ctx, MLZ_E_ARG, "null task array");
	if (nt > (size_t) INT32_MAX / 4)
		return set_error(ctx, MLZ_E_ARG, "too many tasks");
	for (size_t i = 0; i < nt; ++i) {
		const mlz_task &t = tasks[i];
		if (t.m == 0 || t.n == 0 || t.m > (1u << 24) || t.n > (1u << 24))
			return set_error(ctx, MLZ_E_ARG, "task %zu: bad dimensions", i);
		if (t.variant < -1 || t.variant >= MLZ_MAX_VARIANTS)
			return set_error(ctx, MLZ_E_ARG, "task %zu: bad variant %d", i, t.variant);
		if (!(t.density >= 0.0f && t.density <= 1.0f))
			return set_error(ctx, MLZ_E_ARG, "task %zu: density outside [0, 1]", i);
		if (t.op == MLZ_TASK_GEMM) {
			if (t.k == 0 || t.k > (1u << 24))
				return set_error(ctx, MLZ_E_ARG, "task %zu: bad inner dimension", i);
			if (!t.A || !t.B || !t.C)
				return set_error(ctx, MLZ_E_ARG, "task %zu: null matrix pointer", i);
			if (t.lda < t.k || t.ldb < t.n || t.ldc < t.n)
				return set_error(ctx, MLZ_E_ARG, "task %zu: leading dimension smaller than row length", i);
		} else if (t.op == MLZ_TASK_RREF) {
			if (t.m > (uint32_t) BR_MAXM)
				return set_error(ctx, MLZ_E_ARG, "task %zu: batched echelon forms take at most %d rows", i,
						 BR_MAXM);
			if (!t.A || t.lda < t.n)
				return set_error(ctx, MLZ_E_ARG, "task %zu: bad matrix pointer or leading dimension", i);
		} else {
			return set_error(ctx, MLZ_E_ARG, "task %zu: unknown op %d", i, t.op);
		}
	}
	return MLZ_OK;
}

// Byte layout helper: reserves aligned ranges of the arena.
struct Layout {
	size_t off = 0;
	size_t take(size_t bytes)
	{
		const size_t o = off;
		off = round_up(off + std::max<size_t>(bytes, 1), 256);
		return o;
	}
};

static int plan_build(mlz_ctx *ctx, int elem, uint64_t p, mlz_task *tasks, size_t nt, int policy, int flags,
		      mlz_batch **out)
{
	MLZ_TRY(validate(ctx, elem, p, tasks, nt));
	MLZ_TRY(choose_variants(ctx, p, tasks, nt, policy));
	mlz_batch *b = new mlz_batch();
	b->elem = elem;
	b->p = p;
	b->flags = flags;
	b->ntasks = nt;
	b->variant.resize(nt);
	b->rref_slot.assign(nt, -1);
	const size_t es = elem_size(elem);

	// descriptor indices, ordered largest first within each kernel's list
	std::vector<size_t> order(nt);
	std::iota(order.begin(), order.end(), 0);
	auto work = [&](size_t i) {
		const mlz_task &t = tasks[i];
		return t.op == MLZ_TASK_GEMM ? (double) t.m * t.n * t.k : (double) t.m * t.n * std::min(t.m, t.n);
	};
	std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work(x) > work(y); });
	std::vector<size_t> gemm_of(nt, 0), rref_of(nt, 0);
	size_t ng = 0, nr = 0;
	for (size_t i = 0; i < nt; ++i) {
		b->variant[i] = tasks[i].variant;
		if (tasks[i].op == MLZ_TASK_GEMM) {
			gemm_of[i] = ng++;
			b->kmax = std::max<uint64_t>(b->kmax, tasks[i].k);
		} else {
			b->rref_slot[i] = (int) i;
			rref_of[i] = nr++;
		}
	}

	// work lists
	std::vector<BatchItem> items[BG_NVAR + 1 + BR_NVAR];   // gemm variants, pack, rref variants
	const int PACK = BG_NVAR, RR0 = BG_NVAR + 1;
	Layout lay;
	// zero region: counters (one per launch, at most 7), bad [nt], tcost [2 nt]
	const size_t counters_off = 0;
	const size_t bad_off = 64;
	const size_t tcost_off = round_up(bad_off + nt * sizeof(int), 8);
	const size_t zero_bytes = round_up(tcost_off + 2 * nt * sizeof(unsigned long long), 256);
	lay.off = zero_bytes;
	const size_t ranks_off = lay.take(nt * sizeof(int));
	const size_t gdesc_off = lay.take(ng * sizeof(BatchGemm));
	const size_t rdesc_off = lay.take(nr * sizeof(BatchRref));
	std::vector<BatchGemm> gd(ng);
	std::vector<BatchRref> rd(nr);
	std::vector<size_t> ap_off(nt, 0), bp_off(nt, 0), wm_off(nt, 0), wb_off(nt, 0), wbp_off(nt, 0), pr_off(nt, 0);
	for (size_t oi = 0; oi < nt; ++oi) {
		const size_t i = order[oi];
		const mlz_task &t = tasks[i];
		const int v = t.variant;
		if (t.op == MLZ_TASK_GEMM) {
			BatchGemm &g = gd[gemm_of[i]];
			g.A = t.A;
			g.B = t.B;
			g.C = t.C;
			g.lda = t.lda;
			g.ldb = t.ldb;
			g.ldc = t.ldc;
			g.m = (int) t.m;
			g.n = (int) t.n;
			g.k = (int) t.k;
			g.a = (uint32_t) (t.a % p);
			g.b = (uint32_t) (t.b % p);
			g.Ap = g.Bp = nullptr;
			g.KB = 0;
			const int task = (int) gemm_of[i];
			const int bm = bg_tile_m(v), bn = bg_tile_n(v);
			const int tm = (int) ceil_div(t.m, bm), tn = (int) ceil_div(t.n, bn);
			if (v == BG_I8S || v == BG_I8L) {
				const size_t mpad = (size_t) tm * bm, npad = (size_t) tn * bn, kpad = round_up(t.k, 64);
				g.KB = (int) (kpad / 32);
				ap_off[i] = lay.take(mpad * kpad * 2);
				bp_off[i] = lay.take(npad * kpad * 2);
				const int va = ((uintptr_t) t.A % 16 == 0 && (t.lda * es) % 16 == 0) ? 2 : 0;
				const int vb = ((uintptr_t) t.B % 16 == 0 && (t.ldb * es) % 16 == 0) ? 2 : 0;
				for (int y = 0; y < (int) (mpad / 64); ++y)
					for (int z = 0; z < (int) (kpad / 64); ++z)
						items[PACK].push_back({task, 0 | va, y, z});
				for (int y = 0; y < (int) (npad / 64); ++y)
					for (int z = 0; z < (int) (kpad / 64); ++z)
						items[PACK].push_back({task, 1 | vb, y, z});
			}
			for (int x = 0; x < tm; ++x)
				for (int y = 0; y < tn; ++y)
					items[v].push_back({task, x, y, 0});
		} else {
			BatchRref &r = rd[rref_of[i]];
			r.A = t.A;
			r.lda = t.lda;
			r.pivots = t.pivots;
			r.m = (int) t.m;
			r.n = (int) t.n;
			wm_off[i] = lay.take((size_t) t.m * br_ldw((int) t.n, br_mfma(v)) * (br_mfma(v) ? 2 : 4));
			wb_off[i] = br_mfma(v) ? 0 : lay.take((size_t) br_width(v) * t.n * 4);   // VALU variants only
			wbp_off[i] = br_mfma(v) ? lay.take(br_wbp_bytes(v, t.n)) : 0;
			pr_off[i] = lay.take(2 * (size_t) std::min(t.m, t.n) * 4);
			items[RR0 + v].push_back({(int) rref_of[i], 0, 0, 0});
		}
	}
	size_t items_off[BG_NVAR + 1 + BR_NVAR];
	for (int l = 0; l < BG_NVAR + 1 + BR_NVAR; ++l)
		items_off[l] = lay.take(items[l].size() * sizeof(BatchItem));
	const size_t total = lay.off;

	hipError_t e = hipMalloc((void **) &b->arena, total);
	if (e != hipSuccess) {
		free_plan(b);
		return set_error(ctx, MLZ_E_NOMEM, "batch workspace of %zu bytes: %s", total, hipGetErrorString(e));
	}
	char *A0 = b->arena;
	for (size_t i = 0; i < nt; ++i) {
		if (tasks[i].op == MLZ_TASK_GEMM) {
			BatchGemm &g = gd[gemm_of[i]];
			if (g.KB) {
				g.Ap = (int8_t *) (A0 + ap_off[i]);
				g.Bp = (int8_t *) (A0 + bp_off[i]);
			}
		} else {
			BatchRref &r = rd[rref_of[i]];
			r.Wm = A0 + wm_off[i];
			r.Wb = br_mfma(tasks[i].variant) ? nullptr : (uint32_t *) (A0 + wb_off[i]);
			r.Wbp = br_mfma(tasks[i].variant) ? (int8_t *) (A0 + wbp_off[i]) : nullptr;
			r.piv_rows = (int *) (A0 + pr_off[i]);
			r.piv_cols = r.piv_rows + std::min(tasks[i].m, tasks[i].n);
		}
	}
	b->zero_words = zero_bytes / 4;
	b->d_bad = (int *) (A0 + bad_off);
	b->d_tcost = (unsigned long long *) (A0 + tcost_off);
	b->d_ranks = (int *) (A0 + ranks_off);
	b->d_gemm = (BatchGemm *) (A0 + gdesc_off);
	b->d_rref = (BatchRref *) (A0 + rdesc_off);
	// ranks and bad flags are indexed by rref descriptor; map back to tasks
	for (size_t i = 0; i < nt; ++i)
		b->rref_slot[i] = tasks[i].op == MLZ_TASK_RREF ? (int) rref_of[i] : -1;
	// tcost is indexed by descriptor too: gemm descriptors [0, ng), rref [ng, ng + nr)
	unsigned long long *tc_gemm = b->d_tcost, *tc_rref = b->d_tcost + ng, *tc_pack = b->d_tcost + nt;

	auto fail = [&](int rc) {
		free_plan(b);
		return rc;
	};
	if (ng && hipMemcpy(b->d_gemm, gd.data(), ng * sizeof(BatchGemm), hipMemcpyHostToDevice) != hipSuccess)
		return fail(set_error(ctx, MLZ_E_HIP, "upload of gemm descriptors failed"));
	if (nr && hipMemcpy(b->d_rref, rd.data(), nr * sizeof(BatchRref), hipMemcpyHostToDevice) != hipSuccess)
		return fail(set_error(ctx, MLZ_E_HIP, "upload of echelon descriptors failed"));
	for (int l = 0; l < BG_NVAR + 1 + BR_NVAR; ++l)
		if (!items[l].empty() &&
		    hipMemcpy(A0 + items_off[l], items[l].data(), items[l].size() * sizeof(BatchItem),
			      hipMemcpyHostToDevice) != hipSuccess)
			return fail(set_error(ctx, MLZ_E_HIP, "upload of work lists failed"));

	// launches: branch 0 = pack + INT8, 1 = plain, 2 = sparse, 3/4 = rref
	const bool timing = (flags & MLZ_BATCH_TIMING) != 0;
	int counter = 0;
	auto add = [&](int kind, int variant, int branch, int list, unsigned long long *tc, int wgs) {
		mlz_batch::Launch L;
		L.kind = kind;
		L.variant = variant;
		L.branch = branch;
		L.L.items = (const BatchItem *) (A0 + items_off[list]);
		L.L.nitems = (int) items[list].size();
		L.L.queue = (int *) (A0 + counters_off) + counter++;
		L.L.tcost = timing ? tc : nullptr;
		L.grid = (int) std::min<size_t>(items[list].size(), (size_t) ctx->num_cus * wgs);
		b->launches.push_back(L);
	};
	const uint32_t p32 = (uint32_t) p;
	for (int v = 0; v <= BG_NVAR; ++v)
		b->occ_gemm[v] = (float) batch_gemm_wgs_per_cu(elem, v, p32, b->kmax);
	for (int v = 0; v < BR_NVAR; ++v)
		b->occ_rref[v] = (float) batch_rref_wgs_per_cu(elem, v, p32);
	if (!items[PACK].empty())
		add(0, 0, 0, PACK, tc_pack, (int) b->occ_gemm[BG_NVAR]);
	const int gemm_branch[BG_NVAR] = {1, 2, 0, 0};
	for (int v : {BG_I8L, BG_I8S, BG_PLAIN, BG_SPARSE})
		if (!items[v].empty())
			add(1, v, gemm_branch[v], v, tc_gemm, (int) b->occ_gemm[v]);
	for (int v = 0; v < BR_NVAR; ++v)
		if (!items[RR0 + v].empty())
			add(2, v, 3 + (v & 1), RR0 + v, tc_rref, (int) b->occ_rref[v]);

	if (!(flags & MLZ_BATCH_SERIAL)) {
		for (auto &s : b->side)
			if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
				return fail(set_error(ctx, MLZ_E_HIP, "stream creation failed"));
		for (auto &ev : b->ev)
			if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
				return fail(set_error(ctx, MLZ_E_HIP, "event creation failed"));
	}
	if (flags & MLZ_BATCH_GRAPH) {
		if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
			return fail(set_error(ctx, MLZ_E_HIP, "hipStreamBeginCapture failed"));
		const int rc = enqueue(ctx, b);
		hipGraph_t g = nullptr;
		const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
		b->graph = g;
		if (rc != MLZ_OK)
			return fail(rc);
		if (ce != hipSuccess || !g)
			return fail(set_error(ctx, MLZ_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ce)));
		if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
			return fail(set_error(ctx, MLZ_E_HIP, "hipGraphInstantiate failed"));
	}
	*out = b;
	return MLZ_OK;
}

// Per-descriptor cost of a task in microseconds of one CU.
static float task_cost_us(const mlz_batch *b, const mlz_task &t, unsigned long long main_ticks,
			  unsigned long long pack_ticks)
{
	const double us = 0.01;   // s_memrealtime runs at 100 MHz
	if (t.op == MLZ_TASK_GEMM)
		return (float) (main_ticks * us / b->occ_gemm[t.variant] + pack_ticks * us / b->occ_gemm[BG_NVAR]);
	return (float) (main_ticks * us / b->occ_rref[t.variant]);
}

static int plan_finish(mlz_ctx *ctx, mlz_batch *b, mlz_task *tasks, float *cost_us, size_t *bad_task)
{
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	const size_t nt = b->ntasks;
	if (nt == 0)
		return MLZ_OK;
	size_t ng = 0;
	for (size_t i = 0; i < nt; ++i)
		ng += tasks[i].op == MLZ_TASK_GEMM;
	const size_t nr = nt - ng;
	std::vector<int> ranks(nr), bad(nr);
	if (nr) {
		MLZ_HIP(ctx, hipMemcpy(ranks.data(), b->d_ranks, nr * sizeof(int), hipMemcpyDeviceToHost));
		MLZ_HIP(ctx, hipMemcpy(bad.data(), b->d_bad, nr * sizeof(int), hipMemcpyDeviceToHost));
	}
	std::vector<unsigned long long> tc;
	if (cost_us && (b->flags & MLZ_BATCH_TIMING)) {
		tc.resize(2 * nt);
		MLZ_HIP(ctx, hipMemcpy(tc.data(), b->d_tcost, 2 * nt * sizeof(unsigned long long), hipMemcpyDeviceToHost));
	}
	int rc = MLZ_OK;
	size_t gi = 0;
	for (size_t i = 0; i < nt; ++i) {
		mlz_task &t = tasks[i];
		t.variant = b->variant[i];
		size_t d;
		if (t.op == MLZ_TASK_RREF) {
			d = (size_t) b->rref_slot[i];
			t.rank = (size_t) ranks[d];
			if (bad[d] && rc == MLZ_OK) {
				rc = set_error(ctx, MLZ_E_NOT_INVERTIBLE, "task %zu: a pivot has no inverse mod p", i);
				if (bad_task)
					*bad_task = i;
			}
			if (!tc.empty())
				cost_us[i] = task_cost_us(b, t, tc[ng + d], 0);
		} else {
			d = gi++;
			if (!tc.empty())
				cost_us[i] = task_cost_us(b, t, tc[d], tc[nt + d]);
		}
	}
	return rc;
}

}  // namespace mlz

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int mlz_batch_plan(mlz_ctx *ctx, mlz_elem elem, uint64_t p, mlz_task *tasks, size_t ntasks, int policy,
			      int flags, mlz_batch **plan)
{
	if (!ctx || !plan)
		return MLZ_E_ARG;
	*plan = nullptr;
	if (policy < MLZ_POLICY_FIXED || policy > MLZ_POLICY_MODEL)
		return set_error(ctx, MLZ_E_ARG, "unknown policy %d", policy);
	if (flags & ~(MLZ_BATCH_GRAPH | MLZ_BATCH_TIMING | MLZ_BATCH_SERIAL))
		return set_error(ctx, MLZ_E_ARG, "unknown batch flags %d", flags);
	if (ctx->ptr_mode != MLZ_PTR_DEVICE)
		return set_error(ctx, MLZ_E_ARG, "batches run on device pointers (mlz_ctx_set_ptr_mode MLZ_PTR_DEVICE)");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	return plan_build(ctx, elem, p, tasks, ntasks, policy, flags, plan);
}

extern "C" int mlz_batch_launch(mlz_ctx *ctx, mlz_batch *plan)
{
	if (!ctx || !plan)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	if (plan->ntasks == 0)
		return MLZ_OK;
	if (plan->exec) {
		MLZ_HIP(ctx, hipGraphLaunch(plan->exec, ctx->stream));
		return MLZ_OK;
	}
	return enqueue(ctx, plan);
}

extern "C" int mlz_batch_finish(mlz_ctx *ctx, mlz_batch *plan, mlz_task *tasks, float *cost_us, size_t *bad_task)
{
	if (!ctx || !plan || (plan->ntasks && !tasks))
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	return plan_finish(ctx, plan, tasks, cost_us, bad_task);
}

extern "C" int mlz_batch_destroy(mlz_ctx *ctx, mlz_batch *plan)
{
	if (!ctx || !plan)
		return MLZ_E_ARG;
	(void) hipSetDevice(ctx->device);
	(void) hipStreamSynchronize(ctx->stream);
	free_plan(plan);
	return MLZ_OK;
}

extern "C" int mlz_batch_run(mlz_ctx *ctx, mlz_elem elem, uint64_t p, mlz_task *tasks, size_t ntasks, int policy)
{
	mlz_batch *b = nullptr;
	MLZ_TRY(mlz_batch_plan(ctx, elem, p, tasks, ntasks, policy, 0, &b));
	int rc = mlz_batch_launch(ctx, b);
	if (rc == MLZ_OK)
		rc = mlz_batch_finish(ctx, b, tasks, nullptr, nullptr);
	mlz_batch_destroy(ctx, b);
	return rc;
}

extern "C" int mlz_task_features(const mlz_task *task, uint64_t p, float *x5)
{
	if (!task || !x5 || task->m == 0 || task->n == 0 || p < 2 ||
	    (task->op == MLZ_TASK_GEMM && task->k == 0))
		return MLZ_E_ARG;
	task_features(*task, p, x5);
	return MLZ_OK;
}

extern "C" int mlz_tune(mlz_ctx *ctx, mlz_elem elem, uint64_t p, mlz_task *tasks, size_t ntasks, int reps,
			float *cost_us)
{
	if (!ctx || !cost_us || reps < 1)
		return MLZ_E_ARG;
	if (ctx->ptr_mode != MLZ_PTR_DEVICE)
		return set_error(ctx, MLZ_E_ARG, "batches run on device pointers (mlz_ctx_set_ptr_mode MLZ_PTR_DEVICE)");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	MLZ_TRY(validate(ctx, elem, p, tasks, ntasks));
	const size_t es = elem_size(elem);
	for (size_t i = 0; i < ntasks * MLZ_MAX_VARIANTS; ++i)
		cost_us[i] = std::numeric_limits<float>::quiet_NaN();
	// echelon-form inputs are reduced in place: keep a copy to restore before every run
	size_t backup_bytes = 0;
	for (size_t i = 0; i < ntasks; ++i)
		if (tasks[i].op == MLZ_TASK_RREF)
			backup_bytes += round_up((size_t) tasks[i].m * tasks[i].n * es, 256);
	char *backup = nullptr;
	if (backup_bytes)
		MLZ_HIP(ctx, hipMalloc((void **) &backup, backup_bytes));
	auto copy_all = [&](bool save) -> int {
		size_t off = 0;
		for (size_t i = 0; i < ntasks; ++i) {
			const mlz_task &t = tasks[i];
			if (t.op != MLZ_TASK_RREF)
				continue;
			const size_t row = (size_t) t.n * es;
			hipError_t e = save ? hipMemcpy2DAsync(backup + off, row, t.A, t.lda * es, row, t.m,
							       hipMemcpyDeviceToDevice, ctx->stream)
					    : hipMemcpy2DAsync(t.A, t.lda * es, backup + off, row, row, t.m,
							       hipMemcpyDeviceToDevice, ctx->stream);
			if (e != hipSuccess)
				return set_error(ctx, MLZ_E_HIP, "copy of an echelon-form input failed");
			off += round_up((size_t) t.m * t.n * es, 256);
		}
		return MLZ_OK;
	};
	int rc = copy_all(true);
	std::vector<mlz_task> sub;
	std::vector<size_t> who;
	std::vector<float> c, samples;
	for (int v = 0; v < MLZ_MAX_VARIANTS && rc == MLZ_OK; ++v) {
		sub.clear();
		who.clear();
		for (size_t i = 0; i < ntasks; ++i)
			if (variant_ok(tasks[i], v, p)) {
				sub.push_back(tasks[i]);
				sub.back().variant = v;
				who.push_back(i);
			}
		if (sub.empty())
			continue;
		mlz_batch *b = nullptr;
		rc = plan_build(ctx, elem, p, sub.data(), sub.size(), MLZ_POLICY_FIXED, MLZ_BATCH_TIMING | MLZ_BATCH_SERIAL,
				&b);
		if (rc != MLZ_OK)
			break;
		c.assign(sub.size(), 0.0f);
		samples.assign(sub.size() * reps, 0.0f);
		for (int r = -1; r < reps && rc == MLZ_OK; ++r) {   // r = -1: warm-up
			rc = copy_all(false);
			if (rc == MLZ_OK)
				rc = enqueue(ctx, b);
			if (rc == MLZ_OK)
				rc = plan_finish(ctx, b, sub.data(), c.data(), nullptr);
			if (rc == MLZ_E_NOT_INVERTIBLE)
				rc = MLZ_OK;   // timing is still meaningful; results are not used
			if (r >= 0)
				for (size_t j = 0; j < sub.size(); ++j)
					samples[j * reps + r] = c[j];
		}
		free_plan(b);
		for (size_t j = 0; j < sub.size() && rc == MLZ_OK; ++j) {
			float *s = &samples[j * reps];
			std::nth_element(s, s + reps / 2, s + reps);
			cost_us[who[j] * MLZ_MAX_VARIANTS + v] = s[reps / 2];
		}
	}
	if (rc == MLZ_OK)
		rc = copy_all(false);
	if (backup) {
		(void) hipStreamSynchronize(ctx->stream);
		(void) hipFree(backup);
	}
	if (rc != MLZ_OK)
		return rc;
	// table: per op and bucket, the variant of least total cost over its tasks
	mlz_tuner *tu = tuner_of(ctx);
	for (int op = 1; op <= 2; ++op) {
		std::vector<double> sum(TB * TB * TB * MLZ_MAX_VARIANTS, 0.0);
		std::vector<int> cnt(TB * TB * TB, 0), ok(TB * TB * TB * MLZ_MAX_VARIANTS, 0);
		bool any = false;
		for (size_t i = 0; i < ntasks; ++i) {
			if (tasks[i].op != op)
				continue;
			const int bk = bucket_of(tasks[i]);
			++cnt[bk];
			for (int v = 0; v < nvariants(op); ++v) {
				const float x = cost_us[i * MLZ_MAX_VARIANTS + v];
				if (!std::isnan(x)) {
					sum[bk * MLZ_MAX_VARIANTS + v] += x;
					++ok[bk * MLZ_MAX_VARIANTS + v];
				}
			}
		}
		std::vector<int8_t> &tab = tu->table[op];
		tab.assign(TB * TB * TB, -1);
		for (int bk = 0; bk < TB * TB * TB; ++bk) {
			if (!cnt[bk])
				continue;
			double best = std::numeric_limits<double>::infinity();
			for (int v = 0; v < nvariants(op); ++v)
				if (ok[bk * MLZ_MAX_VARIANTS + v] == cnt[bk] && sum[bk * MLZ_MAX_VARIANTS + v] < best) {
					best = sum[bk * MLZ_MAX_VARIANTS + v];
					tab[bk] = (int8_t) v;
				}
			any = any || tab[bk] >= 0;
		}
		tu->has_table[op] = any;
	}
	return MLZ_OK;
}

extern "C" int mlz_tune_table_get(mlz_ctx *ctx, int op, int8_t *table)
{
	if (!ctx || !table || (op != MLZ_TASK_GEMM && op != MLZ_TASK_RREF))
		return MLZ_E_ARG;
	mlz_tuner *tu = tuner_of(ctx);
	std::copy(tu->table[op].begin(), tu->table[op].end(), table);
	return MLZ_OK;
}

extern "C" int mlz_tune_table_set(mlz_ctx *ctx, int op, const int8_t *table)
{
	if (!ctx || !table || (op != MLZ_TASK_GEMM && op != MLZ_TASK_RREF))
		return MLZ_E_ARG;
	mlz_tuner *tu = tuner_of(ctx);
	bool any = false;
	for (int e = 0; e < TB * TB * TB; ++e) {
		if (table[e] < -1 || table[e] >= nvariants(op))
			return set_error(ctx, MLZ_E_ARG, "table entry %d: variant %d out of range", e, table[e]);
		any = any || table[e] >= 0;
	}
	tu->table[op].assign(table, table + TB * TB * TB);
	tu->has_table[op] = any;
	return MLZ_OK;
}

// Model weights live with the table; mlp.hip trains and evaluates them.
namespace mlz {
std::vector<float> *model_slot(mlz_ctx *ctx, int op, bool **trained)
{
	mlz_tuner *tu = tuner_of(ctx);
	*trained = &tu->has_model[op];
	return &tu->model[op];
}
}  // namespace mlz
