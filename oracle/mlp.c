/* oracle/mlp.c -- CPU restatement of the batch cost model (test infrastructure
 * only; see oracle.h).  LELA has no cost model, so this follows the model's
 * definition: h1 = relu(W1^T x + b1), h2 = relu(W2^T h1 + b2), y = W3^T h2 + b3,
 * L = (1/M) sum_{s,v} mask (y - t)^2; gradients by the chain rule, with the
 * ReLU derivative taken as 0 at 0.  Pinned in tests/test_oracle.py by central
 * finite differences of the loss and by closed-form cases. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { NI = 5, NH = 32, NO = 4 };
enum { OW1 = 0, OB1 = OW1 + NI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
       OB3 = OW3 + NH * NO, NP = OB3 + NO };

static void forward1(const double *w, const float *x, double *h1, double *h2, double *y)
{
	for (int j = 0; j < NH; ++j) {
		double a = w[OB1 + j];
		for (int f = 0; f < NI; ++f)
			a += (double) x[f] * w[OW1 + f * NH + j];
		h1[j] = a > 0.0 ? a : 0.0;
	}
	for (int j = 0; j < NH; ++j) {
		double a = w[OB2 + j];
		for (int i = 0; i < NH; ++i)
			a += h1[i] * w[OW2 + i * NH + j];
		h2[j] = a > 0.0 ? a : 0.0;
	}
	for (int v = 0; v < NO; ++v) {
		double a = w[OB3 + v];
		for (int i = 0; i < NH; ++i)
			a += h2[i] * w[OW3 + i * NO + v];
		y[v] = a;
	}
}

int orc_mlp_forward(const float *params, const float *x, size_t n, double *y)
{
	if (!params || (n && (!x || !y)))
		return ORC_E_ARG;
	double w[NP], h1[NH], h2[NH];
	for (int q = 0; q < NP; ++q)
		w[q] = params[q];
	for (size_t s = 0; s < n; ++s)
		forward1(w, x + s * NI, h1, h2, y + s * NO);
	return ORC_OK;
}

static void loss_grad_d(const double *w, const float *x, const float *t, const uint8_t *mask, size_t n,
			double *grad, double *loss)
{
	size_t cnt = 0;
	for (size_t i = 0; i < n * NO; ++i)
		cnt += mask[i] != 0;
	const double inv = cnt ? 1.0 / (double) cnt : 0.0;
	memset(grad, 0, NP * sizeof(double));
	double L = 0.0;
	for (size_t s = 0; s < n; ++s) {
		const float *xs = x + s * NI;
		double h1[NH], h2[NH], y[NO], dy[NO], dz2[NH], dz1[NH];
		forward1(w, xs, h1, h2, y);
		for (int v = 0; v < NO; ++v) {
			const double d = mask[s * NO + v] ? y[v] - (double) t[s * NO + v] : 0.0;
			L += d * d * inv;
			dy[v] = 2.0 * d * inv;
		}
		for (int i = 0; i < NH; ++i) {
			double g = 0.0;
			for (int v = 0; v < NO; ++v)
				g += w[OW3 + i * NO + v] * dy[v];
			dz2[i] = h2[i] > 0.0 ? g : 0.0;
		}
		for (int i = 0; i < NH; ++i) {
			double g = 0.0;
			for (int j = 0; j < NH; ++j)
				g += w[OW2 + i * NH + j] * dz2[j];
			dz1[i] = h1[i] > 0.0 ? g : 0.0;
		}
		for (int f = 0; f < NI; ++f)
			for (int j = 0; j < NH; ++j)
				grad[OW1 + f * NH + j] += (double) xs[f] * dz1[j];
		for (int j = 0; j < NH; ++j)
			grad[OB1 + j] += dz1[j];
		for (int i = 0; i < NH; ++i)
			for (int j = 0; j < NH; ++j)
				grad[OW2 + i * NH + j] += h1[i] * dz2[j];
		for (int j = 0; j < NH; ++j)
			grad[OB2 + j] += dz2[j];
		for (int i = 0; i < NH; ++i)
			for (int v = 0; v < NO; ++v)
				grad[OW3 + i * NO + v] += h2[i] * dy[v];
		for (int v = 0; v < NO; ++v)
			grad[OB3 + v] += dy[v];
	}
	*loss = L;
}

int orc_mlp_loss_grad(const float *params, const float *x, const float *t, const uint8_t *mask, size_t n,
		      double *grad, double *loss)
{
	if (!params || !grad || !loss || (n && (!x || !t || !mask)))
		return ORC_E_ARG;
	double w[NP];
	for (int q = 0; q < NP; ++q)
		w[q] = params[q];
	loss_grad_d(w, x, t, mask, n, grad, loss);
	return ORC_OK;
}

int orc_mlp_adam(float *params, const float *x, const float *t, const uint8_t *mask, size_t n, int steps,
		 double lr, double *loss_hist)
{
	if (!params || steps < 0 || (n && (!x || !t || !mask)))
		return ORC_E_ARG;
	double w[NP], g[NP], m1[NP], m2[NP];
	for (int q = 0; q < NP; ++q) {
		w[q] = params[q];
		m1[q] = m2[q] = 0.0;
	}
	double p1 = 1.0, p2 = 1.0;
	for (int s = 0; s < steps; ++s) {
		double L;
		loss_grad_d(w, x, t, mask, n, g, &L);
		if (loss_hist)
			loss_hist[s] = L;
		p1 *= 0.9;
		p2 *= 0.999;
		for (int q = 0; q < NP; ++q) {
			m1[q] = 0.9 * m1[q] + 0.1 * g[q];
			m2[q] = 0.999 * m2[q] + 0.001 * g[q] * g[q];
			w[q] -= lr * (m1[q] / (1.0 - p1)) / (sqrt(m2[q] / (1.0 - p2)) + 1e-8);
		}
	}
	for (int q = 0; q < NP; ++q)
		params[q] = (float) w[q];
	return ORC_OK;
}
