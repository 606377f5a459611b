// lela_amd/csrc/mlp.hip -- the batch cost model: an MLP 5 -> 32 -> 32 -> 4
// (ReLU, FP32) that predicts log2 of each kernel variant's cost of a task from
// its features (mlz_task_features), trained on measured costs.
//
// Training is full-batch Adam on the masked mean squared error
//   L = (1/M) sum_{s,v} mask[s][v] (y[s][v] - t[s][v])^2,  M = sum mask,
// with two kernels per step:
//  * mlp_grad_kernel: a workgroup takes GT samples, one per thread: forward
//    pass, loss, and backward pass (dy = 2 mask (y - t) / M; dz2 = W3 dy
//    masked by h2 > 0; dz1 = W2 dz2 masked by h1 > 0), leaves each sample's
//    x, h1, dz1, h2, dz2, dy in LDS, and then thread t sums the
//    outer-product terms of parameters t, t + GT, ... over the workgroup's
//    samples (dW1 = x dz1^T, dW2 = h1 dz2^T, dW3 = h2 dy^T, biases = sums),
//    writing one partial gradient per workgroup;
//  * mlp_update_kernel: one thread per parameter adds the partials in a
//    fixed order (the gradient is reproducible bit for bit) and, when
//    training, applies the Adam step.
// Inference (mlp_predict_kernel) is the forward pass, one thread per task.
#include "common.hpp"

#include <cmath>

namespace mlz {

constexpr int NI = 5, NH = 32, NO = MLZ_MAX_VARIANTS;
constexpr int OW1 = 0, OB1 = OW1 + NI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
	      OB3 = OW3 + NH * NO, NP = OB3 + NO;
static_assert(NP == MLZ_MODEL_PARAMS, "parameter layout");
constexpr int GT = 128;   // samples per workgroup of the gradient kernel

std::vector<float> *model_slot(mlz_ctx *ctx, int op, bool **trained);

__device__ __forceinline__ void mlp_forward(const float *w, const float (&x)[NI], float (&h1)[NH], float (&h2)[NH],
					    float (&y)[NO])
{
#pragma unroll
	for (int j = 0; j < NH; ++j) {
		float a = w[OB1 + j];
#pragma unroll
		for (int f = 0; f < NI; ++f)
			a = fmaf(x[f], w[OW1 + f * NH + j], a);
		h1[j] = fmaxf(a, 0.0f);
	}
#pragma unroll
	for (int j = 0; j < NH; ++j) {
		float a = w[OB2 + j];
#pragma unroll
		for (int i = 0; i < NH; ++i)
			a = fmaf(h1[i], w[OW2 + i * NH + j], a);
		h2[j] = fmaxf(a, 0.0f);
	}
#pragma unroll
	for (int v = 0; v < NO; ++v) {
		float a = w[OB3 + v];
#pragma unroll
		for (int i = 0; i < NH; ++i)
			a = fmaf(h2[i], w[OW3 + i * NO + v], a);
		y[v] = a;
	}
}

// partial: [gridDim.x][NP + 1], the last entry of a row is the loss part.
__global__ __launch_bounds__(GT) void mlp_grad_kernel(const float *__restrict__ params, const float *__restrict__ X,
						      const float *__restrict__ Y, const uint8_t *__restrict__ M,
						      int n, float inv_cnt, float *__restrict__ partial)
{
	__shared__ float w[NP];
	__shared__ float sx[GT][NI], sdy[GT][NO];
	__shared__ float sh1[GT][NH + 1], sdz1[GT][NH + 1], sh2[GT][NH + 1], sdz2[GT][NH + 1];
	__shared__ float sloss[GT];
	const int tid = threadIdx.x;
	for (int q = tid; q < NP; q += GT)
		w[q] = params[q];
	__syncthreads();
	const int s = blockIdx.x * GT + tid;
	const bool valid = s < n;
	float x[NI], h1[NH], h2[NH], y[NO], dy[NO], dz2[NH], dz1[NH];
#pragma unroll
	for (int f = 0; f < NI; ++f)
		x[f] = valid ? X[(size_t) s * NI + f] : 0.0f;
	mlp_forward(w, x, h1, h2, y);
	float loss = 0.0f;
#pragma unroll
	for (int v = 0; v < NO; ++v) {
		const bool on = valid && M[(size_t) s * NO + v];
		const float d = on ? y[v] - Y[(size_t) s * NO + v] : 0.0f;
		loss = fmaf(d * d, inv_cnt, loss);
		dy[v] = 2.0f * d * inv_cnt;
	}
#pragma unroll
	for (int i = 0; i < NH; ++i) {
		float g = 0.0f;
#pragma unroll
		for (int v = 0; v < NO; ++v)
			g = fmaf(w[OW3 + i * NO + v], dy[v], g);
		dz2[i] = h2[i] > 0.0f ? g : 0.0f;
	}
#pragma unroll
	for (int i = 0; i < NH; ++i) {
		float g = 0.0f;
#pragma unroll
		for (int j = 0; j < NH; ++j)
			g = fmaf(w[OW2 + i * NH + j], dz2[j], g);
		dz1[i] = h1[i] > 0.0f ? g : 0.0f;
	}
	if (!valid) {
#pragma unroll
		for (int i = 0; i < NH; ++i)
			h1[i] = h2[i] = 0.0f;
	}
#pragma unroll
	for (int f = 0; f < NI; ++f)
		sx[tid][f] = x[f];
#pragma unroll
	for (int v = 0; v < NO; ++v)
		sdy[tid][v] = dy[v];
#pragma unroll
	for (int i = 0; i < NH; ++i) {
		sh1[tid][i] = h1[i];
		sdz1[tid][i] = dz1[i];
		sh2[tid][i] = h2[i];
		sdz2[tid][i] = dz2[i];
	}
	sloss[tid] = loss;
	__syncthreads();
	float *out = partial + (size_t) blockIdx.x * (NP + 1);
	for (int q = tid; q < NP; q += GT) {
		float acc = 0.0f;
		if (q < OB1) {
			const int f = q / NH, j = q % NH;
			for (int t = 0; t < GT; ++t)
				acc = fmaf(sx[t][f], sdz1[t][j], acc);
		} else if (q < OW2) {
			const int j = q - OB1;
			for (int t = 0; t < GT; ++t)
				acc += sdz1[t][j];
		} else if (q < OB2) {
			const int i = (q - OW2) / NH, j = (q - OW2) % NH;
			for (int t = 0; t < GT; ++t)
				acc = fmaf(sh1[t][i], sdz2[t][j], acc);
		} else if (q < OW3) {
			const int j = q - OB2;
			for (int t = 0; t < GT; ++t)
				acc += sdz2[t][j];
		} else if (q < OB3) {
			const int i = (q - OW3) / NO, v = (q - OW3) % NO;
			for (int t = 0; t < GT; ++t)
				acc = fmaf(sh2[t][i], sdy[t][v], acc);
		} else {
			const int v = q - OB3;
			for (int t = 0; t < GT; ++t)
				acc += sdy[t][v];
		}
		out[q] = acc;
	}
	if (tid == 0) {
		float l = 0.0f;
		for (int t = 0; t < GT; ++t)
			l += sloss[t];
		out[NP] = l;
	}
}

// grad[q] = sum of the partials (fixed order); with adam != 0 also the Adam
// step on params (beta1 0.9, beta2 0.999, eps 1e-8; c1, c2 the bias
// corrections 1/(1 - beta^t)).
__global__ void mlp_update_kernel(const float *__restrict__ partial, int nblocks, float *__restrict__ grad,
				  float *__restrict__ params, float *__restrict__ m1, float *__restrict__ m2, int adam,
				  float lr, float c1, float c2)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q > NP)
		return;
	float g = 0.0f;
	for (int b = 0; b < nblocks; ++b)
		g += partial[(size_t) b * (NP + 1) + q];
	grad[q] = g;
	if (adam && q < NP) {
		const float a = 0.9f * m1[q] + 0.1f * g;
		const float v = 0.999f * m2[q] + 0.001f * g * g;
		m1[q] = a;
		m2[q] = v;
		params[q] -= lr * (a * c1) / (sqrtf(v * c2) + 1e-8f);
	}
}

__global__ __launch_bounds__(256) void mlp_predict_kernel(const float *__restrict__ params,
							  const float *__restrict__ X, int n, float *__restrict__ Y)
{
	__shared__ float w[NP];
	for (int q = threadIdx.x; q < NP; q += blockDim.x)
		w[q] = params[q];
	__syncthreads();
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	float x[NI], h1[NH], h2[NH], y[NO];
#pragma unroll
	for (int f = 0; f < NI; ++f)
		x[f] = X[(size_t) s * NI + f];
	mlp_forward(w, x, h1, h2, y);
#pragma unroll
	for (int v = 0; v < NO; ++v)
		Y[(size_t) s * NO + v] = y[v];
}

__global__ void mlp_fill_kernel(float *__restrict__ p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0.0f;
}

// Device buffers of one model call, carved from one workspace slot.
struct MlpBufs {
	float *params, *m1, *m2, *grad, *partial, *X, *Y;
	uint8_t *M;
	int nblocks;
};

static int mlp_bufs(mlz_ctx *ctx, size_t n, MlpBufs &b)
{
	b.nblocks = (int) ceil_div(std::max<size_t>(n, 1), GT);
	const size_t fl = 4 * round_up(NP + 1, 64) + (size_t) b.nblocks * (NP + 1) + round_up(n * NI, 64) +
			  round_up(n * NO, 64);
	void *p = nullptr;
	MLZ_TRY(ws_get(ctx, WS_BATCH_0, fl * sizeof(float) + round_up(n * NO, 256) + 1024, &p));
	float *f = (float *) p;
	const size_t stride = round_up(NP + 1, 64);
	b.params = f;
	b.m1 = f + stride;
	b.m2 = f + 2 * stride;
	b.grad = f + 3 * stride;
	b.partial = f + 4 * stride;
	b.X = b.partial + (size_t) b.nblocks * (NP + 1);
	b.X = (float *) round_up((uintptr_t) b.X, 256);
	b.Y = b.X + round_up(n * NI, 64);
	b.M = (uint8_t *) (b.Y + round_up(n * NO, 64));
	return MLZ_OK;
}

int mlp_predict_dev(mlz_ctx *ctx, const float *params, const float *x, size_t n, float *y)
{
	if (n == 0)
		return MLZ_OK;
	if (n > (size_t) INT32_MAX / NI)
		return set_error(ctx, MLZ_E_ARG, "too many samples");
	MlpBufs b;
	MLZ_TRY(mlp_bufs(ctx, n, b));
	MLZ_HIP(ctx, hipMemcpyAsync(b.params, params, NP * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(b.X, x, n * NI * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
	mlp_predict_kernel<<<ceil_div(n, 256), 256, 0, ctx->stream>>>(b.params, b.X, (int) n, b.Y);
	MLZ_TRY(check_launch(ctx, "mlp_predict_kernel"));
	MLZ_HIP(ctx, hipMemcpyAsync(y, b.Y, n * NO * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MLZ_OK;
}

static int mlp_upload(mlz_ctx *ctx, const MlpBufs &b, const float *x, const float *y, const uint8_t *mask, size_t n,
		      size_t *cnt)
{
	MLZ_HIP(ctx, hipMemcpyAsync(b.X, x, n * NI * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(b.Y, y, n * NO * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(b.M, mask, n * NO, hipMemcpyHostToDevice, ctx->stream));
	size_t c = 0;
	for (size_t i = 0; i < n * NO; ++i)
		c += mask[i] != 0;
	*cnt = c;
	return MLZ_OK;
}

static uint64_t splitmix(uint64_t &s)
{
	uint64_t z = (s += 0x9e3779b97f4a7c15ull);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

}  // namespace mlz

using namespace mlz;

static bool model_op_ok(int op)
{
	return op == MLZ_TASK_GEMM || op == MLZ_TASK_RREF;
}

extern "C" int mlz_model_get(mlz_ctx *ctx, int op, float *params)
{
	if (!ctx || !params || !model_op_ok(op))
		return MLZ_E_ARG;
	bool *trained;
	std::vector<float> *w = model_slot(ctx, op, &trained);
	std::copy(w->begin(), w->end(), params);
	return MLZ_OK;
}

extern "C" int mlz_model_set(mlz_ctx *ctx, int op, const float *params)
{
	if (!ctx || !params || !model_op_ok(op))
		return MLZ_E_ARG;
	for (int q = 0; q < NP; ++q)
		if (!std::isfinite(params[q]))
			return set_error(ctx, MLZ_E_ARG, "model parameter %d is not finite", q);
	bool *trained;
	std::vector<float> *w = model_slot(ctx, op, &trained);
	w->assign(params, params + NP);
	*trained = true;
	return MLZ_OK;
}

extern "C" int mlz_model_predict(mlz_ctx *ctx, int op, const float *x, size_t n, float *y)
{
	if (!ctx || !model_op_ok(op) || (n && (!x || !y)))
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	bool *trained;
	std::vector<float> *w = model_slot(ctx, op, &trained);
	return mlp_predict_dev(ctx, w->data(), x, n, y);
}

extern "C" int mlz_model_grad(mlz_ctx *ctx, const float *params, const float *x, const float *y,
			      const uint8_t *mask, size_t n, float *grad, float *loss)
{
	if (!ctx || !params || !grad || n == 0 || !x || !y || !mask)
		return MLZ_E_ARG;
	if (n > (size_t) INT32_MAX / NI)
		return set_error(ctx, MLZ_E_ARG, "too many samples");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	MlpBufs b;
	MLZ_TRY(mlp_bufs(ctx, n, b));
	size_t cnt = 0;
	MLZ_TRY(mlp_upload(ctx, b, x, y, mask, n, &cnt));
	MLZ_HIP(ctx, hipMemcpyAsync(b.params, params, NP * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
	const float inv = cnt ? 1.0f / (float) cnt : 0.0f;
	mlp_grad_kernel<<<b.nblocks, GT, 0, ctx->stream>>>(b.params, b.X, b.Y, b.M, (int) n, inv, b.partial);
	mlp_update_kernel<<<ceil_div(NP + 1, 256), 256, 0, ctx->stream>>>(b.partial, b.nblocks, b.grad, b.params, b.m1,
									  b.m2, 0, 0.0f, 1.0f, 1.0f);
	MLZ_TRY(check_launch(ctx, "mlp_grad_kernel"));
	std::vector<float> g(NP + 1);
	MLZ_HIP(ctx, hipMemcpyAsync(g.data(), b.grad, (NP + 1) * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	std::copy(g.begin(), g.begin() + NP, grad);
	if (loss)
		*loss = g[NP];
	return MLZ_OK;
}

extern "C" int mlz_model_train(mlz_ctx *ctx, int op, const float *x, const float *y, const uint8_t *mask, size_t n,
			       int steps, float lr, uint64_t seed, float *loss)
{
	if (!ctx || !model_op_ok(op) || n == 0 || !x || !y || !mask || steps < 0 || !(lr > 0.0f))
		return MLZ_E_ARG;
	if (n > (size_t) INT32_MAX / NI)
		return set_error(ctx, MLZ_E_ARG, "too many samples");
	MLZ_HIP(ctx, hipSetDevice(ctx->device));
	// initial weights: uniform(-r, r), r = sqrt(6 / (fan_in + fan_out));
	// output biases = mean target of each variant
	std::vector<float> w0(NP, 0.0f);
	uint64_t st = seed;
	auto fill = [&](int off, int fin, int fout) {
		const double r = std::sqrt(6.0 / (fin + fout));
		for (int q = 0; q < fin * fout; ++q)
			w0[off + q] = (float) ((double) (splitmix(st) >> 11) * 0x1.0p-53 * 2.0 * r - r);
	};
	fill(OW1, NI, NH);
	fill(OW2, NH, NH);
	fill(OW3, NH, NO);
	for (int v = 0; v < NO; ++v) {
		double s = 0.0;
		size_t c = 0;
		for (size_t i = 0; i < n; ++i)
			if (mask[i * NO + v]) {
				s += y[i * NO + v];
				++c;
			}
		w0[OB3 + v] = c ? (float) (s / c) : 0.0f;
	}
	MlpBufs b;
	MLZ_TRY(mlp_bufs(ctx, n, b));
	size_t cnt = 0;
	MLZ_TRY(mlp_upload(ctx, b, x, y, mask, n, &cnt));
	if (cnt == 0)
		return set_error(ctx, MLZ_E_ARG, "no measured targets");
	MLZ_HIP(ctx, hipMemcpyAsync(b.params, w0.data(), NP * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
	mlp_fill_kernel<<<ceil_div(NP, 256), 256, 0, ctx->stream>>>(b.m1, NP);
	mlp_fill_kernel<<<ceil_div(NP, 256), 256, 0, ctx->stream>>>(b.m2, NP);
	const float inv = 1.0f / (float) cnt;
	double p1 = 1.0, p2 = 1.0;
	for (int t = 1; t <= steps; ++t) {
		p1 *= 0.9;
		p2 *= 0.999;
		mlp_grad_kernel<<<b.nblocks, GT, 0, ctx->stream>>>(b.params, b.X, b.Y, b.M, (int) n, inv, b.partial);
		mlp_update_kernel<<<ceil_div(NP + 1, 256), 256, 0, ctx->stream>>>(
			b.partial, b.nblocks, b.grad, b.params, b.m1, b.m2, 1, lr, (float) (1.0 / (1.0 - p1)),
			(float) (1.0 / (1.0 - p2)));
	}
	MLZ_TRY(check_launch(ctx, "mlp training kernels"));
	// loss at the final weights
	mlp_grad_kernel<<<b.nblocks, GT, 0, ctx->stream>>>(b.params, b.X, b.Y, b.M, (int) n, inv, b.partial);
	mlp_update_kernel<<<ceil_div(NP + 1, 256), 256, 0, ctx->stream>>>(b.partial, b.nblocks, b.grad, b.params, b.m1,
									  b.m2, 0, 0.0f, 1.0f, 1.0f);
	MLZ_TRY(check_launch(ctx, "mlp_grad_kernel"));
	std::vector<float> w(NP);
	float l = 0.0f;
	MLZ_HIP(ctx, hipMemcpyAsync(w.data(), b.params, NP * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipMemcpyAsync(&l, b.grad + NP, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	for (int q = 0; q < NP; ++q)
		if (!std::isfinite(w[q]))
			return set_error(ctx, MLZ_E_ARG, "training diverged (parameter %d not finite)", q);
	bool *trained;
	std::vector<float> *slot = model_slot(ctx, op, &trained);
	*slot = w;
	*trained = true;
	if (loss)
		*loss = l;
	return MLZ_OK;
}
